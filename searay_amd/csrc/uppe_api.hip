// uppe_api.hip -- extern "C" entry points of libsearay_uppe.so (declared in include/searay_uppe.h).
#include <cstring>
#include <exception>

#include "uppe_internal.h"

using namespace uppe;

namespace uppe {

void prof_take(uppe_handle* h, hipEvent_t* a, hipEvent_t* b) {
  if (h->prof_free.size() >= 2) {
    *a = h->prof_free.back();
    h->prof_free.pop_back();
    *b = h->prof_free.back();
    h->prof_free.pop_back();
  } else {
    UPPE_HIP(hipEventCreate(a));
    UPPE_HIP(hipEventCreate(b));
  }
}

void prof_push(uppe_handle* h, int kc, hipEvent_t a, hipEvent_t b) {
  h->prof_pending.push_back(ProfRecord{kc, a, b});
  if (h->prof_pending.size() > 4096) prof_collect(h);
}

ProfScope::ProfScope(uppe_handle* h_, int kc_) : h(h_), kc(kc_) {
  if (!h->prof) return;
  if (h->prof_pending.size() > 4096) prof_collect(h);  // the destructor queues pairs without collecting
  prof_take(h, &a, &b);
  UPPE_HIP(hipEventRecord(a, h->stream));
}

ProfScope::~ProfScope() noexcept {
  if (!a) return;
  // never throws: a launch that failed while this scope was open unwinds through here, and a failed record
  // (after a sticky device error) only loses this timing pair
  if (std::uncaught_exceptions() > 0 || hipEventRecord(b, h->stream) != hipSuccess) {
    h->prof_free.push_back(a);
    h->prof_free.push_back(b);
    return;
  }
  h->prof_pending.push_back(ProfRecord{kc, a, b});
}

void prof_collect(uppe_handle* h) {
  if (h->prof_pending.empty()) return;
  UPPE_HIP(hipStreamSynchronize(h->stream));
  for (auto& r : h->prof_pending) {
    float ms = 0.f;
    UPPE_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    h->prof_ms[r.kc] += ms;
    h->prof_n[r.kc] += 1;
    h->prof_free.push_back(r.a);
    h->prof_free.push_back(r.b);
  }
  h->prof_pending.clear();
}

}  // namespace uppe

static bool pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

template <typename F>
static int guarded(uppe_handle* h, F&& f) {
  try {
    f();
    return UPPE_OK;
  } catch (const Error& e) {
    if (h) h->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (h) h->err = e.what();
    return UPPE_ERR_DEVICE;
  }
}

static void destroy_impl(uppe_handle* h) {
  if (h->plain) plain_destroy(h);
  if (h->fused) fused_destroy(h);
  for (auto& r : h->prof_pending) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : h->prof_free) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
}

// a new handle for descriptor d; a batch member (slots non-null) keeps its per-instance arrays in the batch's slabs
static int create_handle(const uppe_desc* d, const MemberSlots* slots, uppe_handle** out) {
  if (!d || !out) return UPPE_ERR_ARG;
  *out = nullptr;
  if (d->model != UPPE_MODEL_UPPE && d->model != UPPE_MODEL_PARAXIAL) return UPPE_ERR_ARG;
  const bool par = d->model == UPPE_MODEL_PARAXIAL;
  // 2^n + 1 frequencies for real fields (uppe_kernel.py:354-360), 2^n for the envelope (paraxial_kernel.py:371-377)
  if (par ? (d->nw < 2 || !pow2(d->nw)) : (d->nw < 3 || !pow2(d->nw - 1))) return UPPE_ERR_ARG;
  if (!pow2(d->n1) || !pow2(d->n2)) return UPPE_ERR_ARG;
  if (d->geometry != UPPE_RADIAL && d->geometry != UPPE_CARTESIAN) return UPPE_ERR_ARG;
  if (d->geometry == UPPE_RADIAL && (d->nk < 1 || d->nk > d->n1)) return UPPE_ERR_ARG;
  if (d->precision < UPPE_PREC_FP64_PLAIN || d->precision > UPPE_PREC_BF16X3) return UPPE_ERR_ARG;
  auto* h = new uppe_handle();
  h->d = *d;
  h->nw = d->nw;
  h->n1 = d->n1;
  h->n2 = d->n2;
  h->nk = d->geometry == UPPE_RADIAL ? d->nk : d->n1;
  h->nm = d->n2;
  h->paraxial = par;
  h->nt = par ? d->nw : 2 * (d->nw - 1);
  h->P = (long)d->n1 * d->n2;
  h->npt = (long)h->nw * h->P;
  h->ntt = (long)h->nt * h->P;
  int rc = guarded(h, [&] {
    int ndev = 0;
    UPPE_HIP(hipGetDeviceCount(&ndev));
    if (d->device < 0 || d->device >= ndev) throw Error{UPPE_ERR_DEVICE, "no such HIP device"};
    UPPE_HIP(hipSetDevice(d->device));
    UPPE_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    if (d->precision == UPPE_PREC_FP64_PLAIN) {
      if (slots) throw Error{UPPE_ERR_UNSUPPORTED, "batches run on the fused paths"};
      plain_create(h);
    } else {
      fused_create(h, slots);
    }
  });
  if (rc != UPPE_OK) {
    destroy_impl(h);
    delete h;
    return rc;
  }
  *out = h;
  return UPPE_OK;
}

extern "C" {

int uppe_create(const uppe_desc* d, uppe_handle** out) { return create_handle(d, nullptr, out); }

int uppe_set_hankel(uppe_handle* h, const double* H, const double* lam) {
  if (!h) return UPPE_ERR_ARG;
  if (!H || !lam) {
    h->err = "null argument";
    return UPPE_ERR_ARG;
  }
  if (h->d.geometry != UPPE_RADIAL) {
    h->err = "Hankel tables given for a Cartesian grid";
    return UPPE_ERR_ARG;
  }
  return guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    h->H.assign(H, H + (size_t)h->nk * h->n1 * h->nm);
    h->lam.assign(lam, lam + h->n1);
    for (double v : h->lam)
      if (!(v > 0.0)) throw Error{UPPE_ERR_ARG, "root volumes must be positive"};
    if (h->plain)
      plain_set_hankel(h);
    else
      fused_set_hankel(h);
    h->have_hankel = true;
  });
}

int uppe_set_medium(uppe_handle* h, const uppe_medium* m) {
  if (!h) return UPPE_ERR_ARG;
  if (!m || !m->w || !m->kz || !m->kg || !m->ngas || !m->nl_filter || !m->cond_filter) {
    h->err = "null argument";
    return UPPE_ERR_ARG;
  }
  if (h->paraxial) {
    h->err = "uppe_set_medium on a paraxial handle: use uppe_set_paraxial_medium";
    return UPPE_ERR_STATE;
  }
  if (m->w[0] != 0.0) {
    h->err = "UPPE propagator requires frequency range [0,...]";
    return UPPE_ERR_ARG;
  }
  if (m->ion.kind < UPPE_ION_NONE || m->ion.kind > UPPE_ION_SERIES) {
    h->err = "unknown ionization kind";
    return UPPE_ERR_ARG;
  }
  return guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    h->w.assign(m->w, m->w + h->nw);
    if (h->plain)
      plain_set_medium(h, m);
    else
      fused_set_medium(h, m);
    h->have_medium = true;
  });
}

int uppe_set_paraxial_medium(uppe_handle* h, const uppe_paraxial_medium* m) {
  if (!h) return UPPE_ERR_ARG;
  if (!m || !m->w || !m->kz || !m->k0 || !m->ngas) {
    h->err = "null argument";
    return UPPE_ERR_ARG;
  }
  if (!h->paraxial) {
    h->err = "uppe_set_paraxial_medium on a UPPE handle: use uppe_set_medium";
    return UPPE_ERR_STATE;
  }
  if (m->ion.kind < UPPE_ION_NONE || m->ion.kind > UPPE_ION_SERIES) {
    h->err = "unknown ionization kind";
    return UPPE_ERR_ARG;
  }
  for (int f = 0; f < h->nw; f++)
    if (!(m->k0[f] != 0.0)) {
      h->err = "paraxial medium: k0 must be nonzero";
      return UPPE_ERR_ARG;
    }
  return guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    h->w.assign(m->w, m->w + h->nw);
    if (h->plain)
      plain_set_paraxial_medium(h, m);
    else
      fused_set_paraxial_medium(h, m);
    h->have_medium = true;
  });
}

int uppe_set_rotator(uppe_handle* h, const uppe_rotator* r) {
  if (!h) return UPPE_ERR_ARG;
  return guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    if (!r || r->states < 3) {
      h->rot_states = 0;
      h->rot_tjj.clear();
      h->rot_wdt.clear();
    } else {
      if (!r->tjj || !r->wdt) throw Error{UPPE_ERR_ARG, "null rotator table"};
      if (r->states > 4096) throw Error{UPPE_ERR_ARG, "rotator: at most 4096 states"};
      if (!(r->hbar != 0.0)) throw Error{UPPE_ERR_ARG, "rotator: hbar must be nonzero"};
      h->rot_states = r->states;
      h->rot_tjj.assign(r->tjj, r->tjj + r->states);
      h->rot_wdt.assign(r->wdt, r->wdt + 2 * r->states);
      h->rot_dalpha = r->dalpha;
      h->rot_hbar = r->hbar;
    }
    if (h->plain) plain_set_rotator(h);
    if (h->fused) fused_set_rotator(h);
  });
}

int uppe_set_field(uppe_handle* h, const double* a_r) {
  if (!h) return UPPE_ERR_ARG;
  if (!a_r) {
    h->err = "null argument";
    return UPPE_ERR_ARG;
  }
  if (h->d.geometry == UPPE_RADIAL && !h->have_hankel) {
    h->err = "Hankel tables not set";
    return UPPE_ERR_STATE;
  }
  return guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    if (h->plain)
      plain_set_field(h, a_r);
    else
      fused_set_field(h, a_r);
    h->have_field = true;
    h->steps = h->evals = 0;
    h->nan_steps = 0;
    h->z_last = 0.0;
  });
}

int uppe_advance(uppe_handle* h, const uppe_step_control* c, int* steps_out) {
  if (!h) return UPPE_ERR_ARG;
  if (!c) {
    h->err = "null argument";
    return UPPE_ERR_ARG;
  }
  if (!h->have_field || !h->have_medium) {
    h->err = "field or medium not set";
    return UPPE_ERR_STATE;
  }
  if (c->n_sub < 1 || !(c->length > 0.0) || (c->n_schedule > 0 && !c->dz_schedule) || c->n_schedule < 0) {
    h->err = "bad step control";
    return UPPE_ERR_ARG;
  }
  int taken = 0;
  int rc = guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    taken = h->plain ? plain_advance(h, c) : fused_advance(h, c);
  });
  if (steps_out) *steps_out = taken;
  return rc;
}

int uppe_advance_static(uppe_handle* h, const uppe_step_control* c, int* steps_out) {
  if (steps_out) *steps_out = 0;
  if (!h) return UPPE_ERR_ARG;
  if (!c) {
    h->err = "null argument";
    return UPPE_ERR_ARG;
  }
  if (!h->fused || h->batch) {
    h->err = "uppe_advance_static: standalone handles on the fused paths only (not fp64-plain, not batch members)";
    return UPPE_ERR_UNSUPPORTED;
  }
  if (c->n_sub < 1 || !(c->length > 0.0) || (c->n_schedule > 0 && !c->dz_schedule) || c->n_schedule < 0) {
    h->err = "bad step control";
    return UPPE_ERR_ARG;
  }
  const size_t need = fused_static_need(h, c), cap = fused_static_capacity();
  if (need > cap) {
    h->err = "uppe_advance_static: this grid and step control need " + std::to_string(need) +
             " bytes of workspace; the static workspace holds " + std::to_string(cap) + " bytes";
    return UPPE_ERR_ARG;
  }
  if (!h->have_field || !h->have_medium) {
    h->err = "field or medium not set";
    return UPPE_ERR_STATE;
  }
  int taken = 0;
  int rc = guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    taken = fused_advance_static(h, c);
  });
  if (steps_out) *steps_out = taken;
  return rc;
}

int uppe_advance_checkpointed(uppe_handle* h, const uppe_step_control* c, int checkpoint_every, int* steps_out) {
  if (!h) return UPPE_ERR_ARG;
  if (!c || checkpoint_every < 1) {
    h->err = "null step control or checkpoint interval below 1";
    return UPPE_ERR_ARG;
  }
  if (!h->have_field || !h->have_medium) {
    h->err = "field or medium not set";
    return UPPE_ERR_STATE;
  }
  if (!h->fused) {
    h->err = "adjoint: the fused paths only (not fp64-plain)";
    return UPPE_ERR_UNSUPPORTED;
  }
  if (c->n_sub < 1 || !(c->length > 0.0) || (c->n_schedule > 0 && !c->dz_schedule) || c->n_schedule < 0) {
    h->err = "bad step control";
    return UPPE_ERR_ARG;
  }
  int taken = 0;
  int rc = guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    taken = fused_advance_checkpointed(h, c, checkpoint_every);
  });
  if (steps_out) *steps_out = taken;
  return rc;
}

int uppe_adjoint(uppe_handle* h, double z_total, const double* g_a_r, double* g_a0_r) {
  if (!h) return UPPE_ERR_ARG;
  if (!g_a_r || !g_a0_r) {
    h->err = "null argument";
    return UPPE_ERR_ARG;
  }
  if (!h->fused) {
    h->err = "adjoint: the fused paths only (not fp64-plain)";
    return UPPE_ERR_UNSUPPORTED;
  }
  return guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    fused_adjoint(h, z_total, g_a_r, g_a0_r);
    h->have_field = false;  // the replays leave the stepper mid-run
  });
}

int uppe_finish(uppe_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t) {
  if (!h) return UPPE_ERR_ARG;
  if (!h->have_field || !h->have_medium) {
    h->err = "field or medium not set";
    return UPPE_ERR_STATE;
  }
  return guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    if (h->plain)
      plain_finish(h, z_total, a_r, j_r, chi_t, ne_t);
    else
      fused_finish(h, z_total, a_r, j_r, chi_t, ne_t);
    h->have_field = false;
  });
}

int uppe_get_state(uppe_handle* h, double* q) {
  if (!h) return UPPE_ERR_ARG;
  if (!q) {
    h->err = "null argument";
    return UPPE_ERR_ARG;
  }
  if (!h->have_field) {
    h->err = "no field";
    return UPPE_ERR_STATE;
  }
  return guarded(h, [&] {
    UPPE_HIP(hipSetDevice(h->d.device));
    if (h->plain)
      plain_get_state(h, q);
    else
      fused_get_state(h, q);
  });
}

int uppe_get_info(uppe_handle* h, uppe_info* info) {
  if (!h || !info) return UPPE_ERR_ARG;
  info->steps = h->steps;
  info->evals = h->evals;
  info->nan_steps = h->nan_steps;
  info->reserved = 0;
  info->z_last = h->z_last;
  return UPPE_OK;
}

int uppe_profile_enable(uppe_handle* h, int enable) {
  if (!h) return UPPE_ERR_ARG;
  return guarded(h, [&] {
    prof_collect(h);
    h->prof = enable != 0;
    for (int i = 0; i < UPPE_KC_COUNT; i++) {
      h->prof_ms[i] = 0.0;
      h->prof_n[i] = 0;
    }
  });
}

int uppe_profile_read(uppe_handle* h, int kc, double* total_ms, long long* launches) {
  if (!h) return UPPE_ERR_ARG;
  if (kc < 0 || kc >= UPPE_KC_COUNT) {
    h->err = "bad kernel class";
    return UPPE_ERR_ARG;
  }
  return guarded(h, [&] {
    prof_collect(h);
    if (total_ms) *total_ms = h->prof_ms[kc];
    if (launches) *launches = h->prof_n[kc];
  });
}

int uppe_synchronize(uppe_handle* h) {
  if (!h) return UPPE_ERR_ARG;
  return guarded(h, [&] { UPPE_HIP(hipStreamSynchronize(h->stream)); });
}

int uppe_destroy(uppe_handle* h) {
  if (!h) return UPPE_OK;
  if (h->batch) {
    h->err = "handle belongs to a batch: uppe_batch_destroy releases it";
    return UPPE_ERR_STATE;
  }
  (void)hipSetDevice(h->d.device);
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  destroy_impl(h);
  delete h;
  return UPPE_OK;
}

const char* uppe_last_error(const uppe_handle* h) { return h ? h->err.c_str() : "null handle"; }

}  // extern "C"

// ---- batches ------------------------------------------------------------------------------------------

template <typename F>
static int batch_guarded(uppe_batch* b, F&& f) {
  try {
    f();
    return UPPE_OK;
  } catch (const Error& e) {
    if (b) b->err = e.msg;
    return e.code;
  } catch (const std::exception& e) {
    if (b) b->err = e.what();
    return UPPE_ERR_DEVICE;
  }
}

extern "C" {

int uppe_batch_create(const uppe_desc* descs, int n, uppe_batch** out) {
  if (!descs || !out || n < 1) return UPPE_ERR_ARG;
  *out = nullptr;
  auto* b = new uppe_batch();
  b->device = descs[0].device;
  int rc = batch_guarded(b, [&] {
    UPPE_HIP(hipSetDevice(b->device));
    fused_batch_create(b, descs, n, create_handle);
  });
  if (rc != UPPE_OK) {
    fused_batch_destroy(b);
    delete b;
    return rc;
  }
  *out = b;
  return UPPE_OK;
}

int uppe_batch_member(uppe_batch* b, int i, uppe_handle** h) {
  if (!b || !h) return UPPE_ERR_ARG;
  if (i < 0 || i >= (int)b->members.size()) {
    b->err = "member index out of range";
    return UPPE_ERR_ARG;
  }
  *h = b->members[(size_t)i];
  return UPPE_OK;
}

int uppe_batch_advance(uppe_batch* b, const uppe_step_control* c, int* steps_out) {
  if (!b) return UPPE_ERR_ARG;
  if (!c) {
    b->err = "null step controls";
    return UPPE_ERR_ARG;
  }
  return batch_guarded(b, [&] {
    UPPE_HIP(hipSetDevice(b->device));
    fused_batch_advance(b, c, steps_out);
  });
}

int uppe_batch_set_autotune(uppe_batch* b, int enable) {
  if (!b) return UPPE_ERR_ARG;
  return batch_guarded(b, [&] { fused_batch_set_autotune(b, enable); });
}

int uppe_batch_set_graph(uppe_batch* b, int steps_per_graph) {
  if (!b) return UPPE_ERR_ARG;
  return batch_guarded(b, [&] { fused_batch_set_graph(b, steps_per_graph); });
}

int uppe_batch_plan(const uppe_batch* b, int member, int* plan) {
  if (!b || !plan || member < 0 || member >= (int)b->members.size()) return UPPE_ERR_ARG;
  return batch_guarded(const_cast<uppe_batch*>(b), [&] { fused_batch_plan(b, member, plan); });
}

int uppe_batch_destroy(uppe_batch* b) {
  if (!b) return UPPE_OK;
  (void)hipSetDevice(b->device);
  fused_batch_destroy(b);
  delete b;
  return UPPE_OK;
}

const char* uppe_batch_last_error(const uppe_batch* b) { return b ? b->err.c_str() : "null batch"; }

}  // extern "C"

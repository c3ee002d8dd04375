// uppe_fused.hip -- the fused MI355X UPPE stepper (FP64 / FP32 / BF16x3 Hankel).
//
// Layout: every spectral array is strip-major, [P][NWP] complex with P = N1*N2 transverse points and the
// frequency axis contiguous (NWP = Nw padded to a multiple of 32). In this layout
//   - the radial transforms are plain row-major GEMMs:  E_r = Hb . E_k,  J_k = Hf . J_r  (complex columns
//     interleaved), Hb[i][k] = H[k][i] L_k/L_i, Hf[k][j] = H[k][j] L_j/L_k (grid_tools.py:308-325);
//   - the time-domain source is one contiguous strip per transverse point.
// One RHS evaluation (uppe_kernel.py:188-216) is three kernels:
//   1. k2r GEMM; its B-operand load forms the RK stage input, propagates it linearly and converts it to the
//      electric field (uppe.cl:22-108, uppe_kernel.py:210-213);
//   2. source: one workgroup per strip does irfft -> rate -> plasma scan -> soft window -> Kerr ->
//      P, J_p -> one complex FFT for both rffts -> spectral current and filter (uppe_kernel.py:139-186);
//   3. r2k GEMM; its epilogue forms the ODE right-hand side (uppe.cl:73-91) and the RK bookkeeping, the
//      per-point step estimate (paraxial.cl:40-58) and the field conditioner (uppe_kernel.py:131-137).
// Step control (uppe_kernel.py:218-227, 299-325) runs on the device inside these three launches (step_pos,
// step_commit), so a whole subcycle runs without host synchronisation and without launches of its own.
#include <hip/hip_ext.h>

#include "uppe_internal.h"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace uppe {

// ---------------------------------------------------------------------------------------------------
// device step state
// ---------------------------------------------------------------------------------------------------
struct StepState {
  double zi, L, dzmin, dphi, rel_amin;
  double zoff;  // (z0 - zi) of the current subcycle
  double z;     // local z within the subcycle
  double dz;    // current step
  double z_last;
  unsigned long long dzraw;  // bits of the running minimum of the per-point estimate
  int sub, sub_end, active, fixed;
  int nanflag, nan_steps;
  int steps;              // steps taken in this advance (also the index of the current step's replayed dz)
  int sched_n, rec_cap, error, max_steps;
  int pend;               // the last step ran; its end (z advance, records, bookkeeping) is not committed yet
  const double* sched;    // replayed steps (device), or null
  // FP32 runs: the state of every column f < rb_f1 is kept as q e^{-Im kz qbase} (q the interaction-picture
  // state), so that the factors e^{+Im kz z} of the right-hand side stay within a span of the reference point
  double qbase;           // the reference point (0: the state is q itself)
  double rb_kmax;         // largest Im kz of the medium
  int rebase, rb_f1;      // the reference moves with the run (rebase), and the columns it applies to
};

// the exponent span Im kz (z - qbase) after which a step's update moves the reference point to the step's end
constexpr double kRebaseSpan = 8.0;

// per-instance scalars of the source kernel in a batched launch (chi3/n_ref, rate model and its prefactor)
struct SrcInst {
  double redchi3, log_pre;
  uppe_ionization ion;
  const void* rot_tab;  // the rotator rows (RotRow<T>), its state count and 0.5 dalpha / hbar, when it has one
  int rot_states;
  double rot_fc;
};

// arrays a batch member keeps inside its batch's slabs (one slab per kind, instances a fixed stride apart)
enum {
  SL_PHS, SL_KZC, SL_NG, SL_Q, SL_K1, SL_K2, SL_K3, SL_ER, SL_JR, SL_ST, SL_QMAX,
  SL_FRB_HI, SL_FRB_LO, SL_FRF_HI, SL_FRF_LO, SL_COUNT
};
struct MemberSlots {
  void* p[SL_COUNT] = {};
};

// 128-bit fingerprint of host arrays (two independent multiply-xorshift lanes over their 8-byte words, the tail
// zero-padded): a batch run lets its members share one copy of the operator or of the kz tables when their
// fingerprints agree
struct Fingerprint {
  uint64_t a = 0, b = 0;
  bool set = false;
  void mix(const void* p, size_t bytes) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    if (!set) {
      a = 0x9e3779b97f4a7c15ull;
      b = 0xc2b2ae3d27d4eb4full;
      set = true;
    }
    a ^= bytes;
    b += bytes;
    for (size_t i = 0; i < bytes; i += 8) {
      uint64_t w = 0;
      std::memcpy(&w, c + i, std::min<size_t>(8, bytes - i));
      a = (a ^ w) * 0xff51afd7ed558ccdull;
      a ^= a >> 29;
      b = (b + w) * 0xc4ceb9fe1a85ec53ull;
      b ^= b >> 32;
    }
  }
  bool operator==(const Fingerprint& o) const { return set && o.set && a == o.a && b == o.b; }
};

struct FusedState {
  // batch member: device addresses of the per-instance arrays inside the batch's slabs (not owned); null for
  // a standalone handle
  const MemberSlots* slots = nullptr;
  // host copies of the frequency tables and source time step (a batched launch shares them between instances)
  std::vector<double> w_h, nl_h, cond_h;
  // fingerprints of the Hankel operator (set_hankel) and of the kz and kg inputs of the propagation tables
  // (set_medium): equal fingerprints let a batch run read one member's copy
  Fingerprint op_fp, tab_fp;
  int prec = 0;
  int nwp = 0;
  size_t el = 0;  // sizeof(T)
  // tables
  double* phs = nullptr;   // [P][nwp] Re kz + kg (FP64)
  void* kzc = nullptr;     // [P][nwp] complex T
  void* w = nullptr;       // [nwp] T
  void* nlf = nullptr;     // [nwp] T
  void* cond = nullptr;    // [nwp] T
  double* ng = nullptr;    // [P]
  void* Hb = nullptr;      // [Nr][Nk] T
  void* Hf = nullptr;      // [Nk][Nr] T
  void* tw = nullptr;      // [Nt/2] complex T
  double redchi3 = 0, dt = 0;
  double max_ki = 0;  // largest Im kz of the loaded medium
  // FP32: the state's columns f < f1 are kept as q e^{-Im kz qbase} (the step state's reference point, brought
  // back after every advance; 0: the state is q itself)
  double qbase = 0;
  std::vector<double> ki_h;  // Im kz of the loaded medium, point order of the host arrays
  uppe_ionization ion{};
  // state and work buffers: complex T [P][nwp]
  void *q = nullptr, *k1 = nullptr, *k2 = nullptr, *k3 = nullptr, *er = nullptr, *jr = nullptr;
  void *chi_t = nullptr, *ne_t = nullptr;  // [P][Nt] T
  double* stage = nullptr;                 // host-layout FP64 staging (max of 2*npt, ntt)
  unsigned long long *qmax = nullptr, *qmax_next = nullptr;
  StepState* st = nullptr;
  StepState* st_host = nullptr;
  double *sched = nullptr, *rec = nullptr;
  double* rec_active = nullptr;  // rec while the current advance records steps, else null
  int sched_cap = 0, rec_cap = 0;
  // matrix-core contraction (cm < 0: SIMT GEMM)
  int cm = -1;
  void *frb_hi = nullptr, *frb_lo = nullptr;  // fragments of Hb (k2r)
  void *frf_hi = nullptr, *frf_lo = nullptr;  // fragments of Hf (r2k)
  int kp_b = 0, kp_f = 0;                     // padded contraction lengths
  // launch plan of each contraction direction (0 = k2r, 1 = r2k): column tiles of 8 complex columns (the
  // last one also takes the remainder), row blocks per tile, waves per workgroup, 16-row tiles per wave
  struct HkPlan {
    int tiles = 1, split = 1, nwv = 1, rt = 1, wide = 0;
  } plan[2];
  int n_cu = 256;
  int spt = 0;          // source samples per thread of the step (0: the default shape)
  int tuned_f1 = -1;    // nonlinear band the contraction plans were timed for (-1: not yet)
  int f1 = 0;  // the r2k direction covers complex columns [0, f1): where the nonlinear filter is nonzero
  // Cartesian: radix-8 pass tables of the x and y transforms (complex T)
  void *tw_x = nullptr, *tw_y = nullptr;
  int twlen_x = 0, twlen_y = 0;
  void* rot_tab = nullptr;  // rotator rows (device, RotRow<T> per state j >= 2), when the handle has a rotator
  // adjoint (radial MFMA path): fragments of Hf^T (the shape of Hb) and Hb^T (the shape of Hf)
  void *fab_hi = nullptr, *fab_lo = nullptr, *faf_hi = nullptr, *faf_lo = nullptr;
  // checkpointed forward record: q, the step state and qmax every ck_every steps; (zoff, z, dz) of every step;
  // the stage fields E_r of one segment (4 per step); cotangent work arrays
  int ck_every = 0, ck_n = 0, rec_steps = 0, er_cap = 0;
  int ckq_cap = 0, ckst_cap = 0, ckqm_cap = 0;  // segments
  int zrec_cap = 0;                             // doubles (3 per step)
  bool recorded = false, er_last_valid = false;
  void* ck_q = nullptr;
  StepState* ck_st = nullptr;
  unsigned long long* ck_qmax = nullptr;
  double* zrec = nullptr;
  double* zrec_active = nullptr;
  std::vector<double> zrec_host;
  void* er_store = nullptr;
  void *gs = nullptr, *gq = nullptr, *gk = nullptr, *g2 = nullptr, *g3 = nullptr;
};

// per-state constants of the rotator for the source kernel: the coefficient T_j S_j 2 dalpha (times n_g in the
// kernel), W_j dt, E = exp(i W_j dt / 2), E^2, and the binary powers E^(2 SPT 2^b) from which the thread with
// index t forms E^(2 SPT t) (rotations.cl:80-129, rotations.py:56-61)
constexpr int ROT_NB = 10;  // bits of the thread index (at most 1024 threads)
template <typename T>
struct RotRow {
  double wr64, wi64;
  T c, wr, wi, pad;
  cplx<T> e[2 + ROT_NB];
};

constexpr int BM = 64, BN = 64, BK = 16;

// k2r prologue: none, RK stage input with E = i kz A (UPPE), linear propagation only (finish), RK stage input kept
// as A (paraxial envelope), and the adjoint's k2r input conj((i/2) e^{-i kz z} / Re kz) g (the transposed RHS
// factor)
enum { PRO_NONE = 0, PRO_STAGE = 1, PRO_FINISH = 2, PRO_STAGE_ENV = 3, PRO_ADJ = 4 };
constexpr bool pro_stage(int pro) { return pro == PRO_STAGE || pro == PRO_STAGE_ENV; }
// r2k epilogue: none, the RK stage (forward), the adjoint RK stage (conj(i kz e^{i kz z}) times the contraction,
// then the reverse RK4 recurrence), the adjoint of finish's linear propagation
enum { EPI_NONE = 0, EPI_STAGE = 1, EPI_ADJ = 2, EPI_ADJ_FIN = 3 };

__device__ inline unsigned long long dbits(double x) { return (unsigned long long)__double_as_longlong(x); }
__device__ inline double bitsd(unsigned long long b) { return __longlong_as_double((long long)b); }

// The maxima of one run, kQW words: [0] the committed max|q|, [2] max|q| over the columns the r2k direction never
// updates, then kSlots words (kSlotStride apart) for the stage-4 maxima of the running step. Each workgroup of a
// stage-4 launch takes one slot, so that the workgroups do not serialise on one word; the commit takes the
// maximum over the slots. The values are bit patterns of non-negative doubles, whose integer order is their order.
// (The stage-1 step estimate keeps one word in the step state: every later launch of the step reads it.)
constexpr int kSlots = 16, kSlotStride = 8, kQmaxSlot = 8, kQW = kQmaxSlot + kSlots * kSlotStride;

__device__ inline int red_slot(int block) { return (block & (kSlots - 1)) * kSlotStride; }

// ---------------------------------------------------------------------------------------------------
// device step control (uppe_kernel.py:218-227, 299-325). No launch of its own: every launch of a step reads the
// step state. The step size is a function of the committed state (stage 1's r2k epilogue leaves the minimum
// per-point estimate in dzraw), so every launch from stage 2 on forms it itself. The end of a step (z advance,
// subcycle bookkeeping, records) is committed by the first block of the next step's stage-1 source launch, which
// is the first launch that reads nothing of the state; until then the step is marked pending, and the next
// step's stage-1 k2r launch advances a local copy. A chunk of steps ends with one commit launch (f_ctl_commit).
// ---------------------------------------------------------------------------------------------------

// dz of the current step: the replayed schedule, the subcycle's remainder for fixed steps, or the estimate
// clamped to [dzmin, L - z] (uppe_kernel.py:221-227, dzmin first, then the subcycle end: Q5)
__device__ inline double step_dz(const StepState& s) {
  if (s.sched_n > 0) return s.steps < s.sched_n ? s.sched[s.steps] : 0.0;
  if (s.fixed) return s.L - s.z;
  double dz = bitsd(s.dzraw);
  if (dz < s.dzmin) dz = s.dzmin;
  if (s.z + dz > s.L) dz = s.L - s.z;
  return dz;
}

// does the update of the step ending at zend move the reference point of the rebased columns
__device__ inline bool rebase_at(const StepState& s, double zend) {
  return s.rebase && (zend - s.qbase) * s.rb_kmax > kRebaseSpan;
}

// end of a step of size dz: the reference point of the rebased columns (moved to the step's end, zoff + (z + dz),
// the stage-4 phase origin, when its span is used up), z advance, NaN count, step count, the termination test
// (z - L)^2 > (L/1e6)^2 and the next subcycle (uppe_kernel.py:299-325)
__device__ inline void end_step(StepState& s, double dz) {
  const double zend = s.zoff + (s.z + dz);
  if (rebase_at(s, zend)) s.qbase = zend;
  s.z = s.z + dz;
  s.steps++;
  if (s.nanflag) {
    s.nan_steps++;
    s.nanflag = 0;
  }
  s.dzraw = dbits(INFINITY);
  const double z = s.z, L = s.L;
  if (s.max_steps > 0 && s.steps >= s.max_steps) {
    s.z_last = z;
    s.active = 0;
    return;
  }
  if (!(z < L && (z - L) * (z - L) > (L / 1e6) * (L / 1e6))) {
    s.z_last = z;
    s.sub++;
    if (s.sub >= s.sub_end) {
      s.active = 0;
    } else {
      s.z = 0.0;
      s.zoff = (s.zi + (double)s.sub * L) - s.zi;
    }
  }
}

// where a launch of `stage` stands: phase origin, z, dz (stages 2-4), whether the run is active, the reference
// point of the rebased columns f < f1, and (stage 4) whether this step's update moves it. Stage 1 advances a
// pending step in a local copy.
struct StepPos {
  double zoff, z, dz, qb;
  int active, f1, rb_end;
};
// The whole step state of a launch in one batch of scalar loads and one wait. Read field by field, the loads
// would sit along step_pos's branches, each waiting for the one before (about ten round trips). The empty asm
// takes every word in a scalar register, so that no load sinks into a branch.
__device__ inline StepState load_state(const StepState* st) {
  static_assert(sizeof(StepState) % 16 == 0, "the step state is read as 16-byte words");
  constexpr int NW = sizeof(StepState) / 16;
  const uint4* p = reinterpret_cast<const uint4*>(st);
  uint4 w[NW];
#pragma unroll
  for (int i = 0; i < NW; i++) w[i] = p[i];
#pragma unroll
  for (int i = 0; i < NW; i++) asm volatile("" ::"s"(w[i].x), "s"(w[i].y), "s"(w[i].z), "s"(w[i].w));
  StepState s;
  __builtin_memcpy(&s, w, sizeof(StepState));
  return s;
}

__device__ inline StepPos step_pos(const StepState* stp, int stage) {
  const StepState st = load_state(stp);
  if (stage == 1) {
    if (!st.pend) return StepPos{st.zoff, st.z, 0.0, st.qbase, st.active, st.rb_f1, 0};
    StepState loc = st;
    if (loc.active) {
      if (loc.sched_n > 0 && loc.steps >= loc.sched_n)
        loc.active = 0;
      else
        end_step(loc, step_dz(loc));
    }
    return StepPos{loc.zoff, loc.z, 0.0, loc.qbase, loc.active, loc.rb_f1, 0};
  }
  const double dz = step_dz(st);
  const int end = stage == 4 && rebase_at(st, st.zoff + (st.z + dz));
  return StepPos{st.zoff, st.z, dz, st.qbase, st.active, st.rb_f1, end};
}

// a value every lane of the wavefront holds, moved to scalar registers
__device__ inline double uniform_f64(double x) {
  const long long b = __double_as_longlong(x);
  const int lo = __builtin_amdgcn_readfirstlane((int)b), hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// the reference of one column in a launch at phase origin zph: the amplitude origin za (zph - qb for a rebased
// column, zph otherwise), the reference point qb (0 for a column that is q itself) and whether this step's update
// moves it to zph (stage 4 only)
struct ColRef {
  double za, qb;
  bool end;
};
__device__ inline ColRef col_ref(int f, int f1, double zph, double qb, int rb_end) {
  if (f < f1) return ColRef{zph - qb, qb, rb_end != 0};
  return ColRef{zph, 0.0, false};
}

// e^x for FP64 x as the hardware exp2 of the remainder of x log2(e) in [-1/2, 1/2] times an exact power of two:
// relative error ~2e-7 over the whole FP64 exponent range, in a few instructions
__device__ inline double exp_split(double x) {
  const double y = x * 1.4426950408889634;
  const double n = fmin(fmax(rint(y), -2100.0), 2100.0);
  return ldexp((double)__builtin_amdgcn_exp2f((float)(y - n)), (int)n);
}

// the stage-4 state of one point: moved to the new reference point zph (factor e^{-Im kz za}) when the step's
// update moves it; returns |q|^2 of the interaction-picture state, e^{2 Im kz qb} |q_kept|^2, in FP64
template <typename T>
__device__ inline double ref_update(cplx<T>& qn, T ki, const ColRef& cr, double zph) {
  double qb = cr.qb;
  if (cr.end) {
    const T g = (T)exp_split(-(double)ki * cr.za);
    qn = cmk(qn.re * g, qn.im * g);
    qb = zph;
  }
  double m2 = (double)qn.re * qn.re + (double)qn.im * qn.im;
  if (qb != 0.0) m2 *= exp_split(2.0 * (double)ki * qb);
  return m2;
}

// amin of the step estimate against a point kept relative to the reference point qb: amin e^{-Im kz qb}
__device__ inline double ref_amin(double amin, double ki, double qb) {
  return qb != 0.0 ? amin * exp_split(-ki * qb) : amin;
}

// commit a pending step (one thread): records of its dz and phase origin, the end of step, and the committed
// max|q| over the stage-4 slots and qmax[2] (max|q| over the columns the r2k direction never updates: the current
// is zero there), the slots cleared for the next step
__device__ inline void step_commit(StepState* st, unsigned long long* qmax, double* rec, double* zrec, int zrec_cap) {
  if (!st->pend) return;
  st->pend = 0;
  if (!st->active) return;
  if (st->sched_n > 0 && st->steps >= st->sched_n) {
    st->error = 1;
    st->active = 0;
    return;
  }
  StepState s = *st;
  const double dz = step_dz(s);
  if (rec && s.steps < s.rec_cap) rec[s.steps] = dz;
  if (zrec && s.steps < zrec_cap) {
    // the phase origin of the step, for the adjoint's replay of the stage phases
    zrec[3 * s.steps] = s.zoff;
    zrec[3 * s.steps + 1] = s.z;
    zrec[3 * s.steps + 2] = dz;
  }
  s.dz = dz;
  end_step(s, dz);
  *st = s;
  unsigned long long m = qmax[2];
#pragma unroll
  for (int i = 0; i < kSlots; i++) {
    const unsigned long long v = qmax[kQmaxSlot + i * kSlotStride];
    m = v > m ? v : m;
    qmax[kQmaxSlot + i * kSlotStride] = 0ull;
  }
  qmax[0] = m;
}

// exp(i phs z) exp(-ki z) with the phase formed in FP64 (long-z cancellation). The FP32 path reduces the
// phase to a fraction of a turn in FP64 and evaluates the hardware sine, cosine (their input is in turns)
// and base-2 exponential: error ~1e-6 in the factor, against an RHS that is itself a small correction to the
// state, well inside the FP32 path's 1e-5 field tolerance.
// The amplitude takes its own origin za: z itself, or z - qbase for a column kept relative to the reference point
// qbase.
template <typename T>
__device__ inline cplx<T> linear_factor(double phs, T ki, double z, double za, int sgn);

template <>
__device__ inline cplx<double> linear_factor<double>(double phs, double ki, double z, double za, int sgn) {
  // the reference forms ar = -Ki z, ai = (Kr + kg) z and exp(ar)(cos ai, sin ai) (uppe.cl:66-70, 86-90)
  double ar = sgn > 0 ? -ki * za : ki * za;
  double ai = sgn > 0 ? phs * z : -(phs * z);
  double ex = exp(ar);
  double s, c;
  sincos(ai, &s, &c);
  return cmk(c * ex, s * ex);
}

template <>
__device__ inline cplx<float> linear_factor<float>(double phs, float ki, double z, double za, int sgn) {
  const double turns = phs * z * 0.15915494309189535;
  float f = (float)(turns - rint(turns));  // in [-1/2, 1/2]
  float ar = (float)(-(double)ki * za);
  if (sgn < 0) {
    f = -f;
    ar = -ar;
  }
  const float ex = __builtin_amdgcn_exp2f(ar * 1.4426950408889634f);
  return cmk(__builtin_amdgcn_cosf(f) * ex, __builtin_amdgcn_sinf(f) * ex);
}

// ---------------------------------------------------------------------------------------------------
// GEMM with fused prologue / epilogue:  out[M][2 nwp] = A[M][K] . B[K][2 nwp]
// ---------------------------------------------------------------------------------------------------
template <typename T, int PRO, int EPI>
__global__ __launch_bounds__(256) void f_gemm(const T* __restrict__ A, const cplx<T>* __restrict__ B,
                                              cplx<T>* __restrict__ out, int M, int K, int nwp,
                                              const double* __restrict__ phs, const cplx<T>* __restrict__ kzc,
                                              const T* __restrict__ cond, cplx<T>* __restrict__ q,
                                              const cplx<T>* __restrict__ kprev, cplx<T>* __restrict__ k1,
                                              cplx<T>* __restrict__ k2, cplx<T>* __restrict__ k3,
                                              StepState* __restrict__ st, unsigned long long* __restrict__ qmax,
                                              unsigned long long* __restrict__ qmax_next, int stage, double zfix,
                                              double qbfix, int f1fix) {
  __shared__ T As[BK][BM + 1];
  __shared__ T Bs[BK][BN];
  __shared__ double red[8];
  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const int row0 = blockIdx.y * BM, col0 = blockIdx.x * BN;
  double zph = zfix, cu = 0.0, dz = 0.0, qb = sizeof(T) == 4 ? qbfix : 0.0;
  int f1r = f1fix, rb_end = 0;
  if (st) {
    const StepPos sp = step_pos(st, stage);
    if (!sp.active) return;
    dz = sp.dz;
    double cs = stage == 1 ? 0.0 : (stage == 4 ? 1.0 : 0.5);
    zph = sp.zoff + (sp.z + cs * dz);
    cu = stage == 1 ? 0.0 : (stage == 4 ? dz : dz / 2);
    if constexpr (sizeof(T) == 4) {  // only FP32 states are rebased
      qb = sp.qb;
      f1r = sp.f1;
      rb_end = sp.rb_end;
    }
  }
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 4; j++) acc[i][j] = T(0);

  for (int k0 = 0; k0 < K; k0 += BK) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      int e = tid + 256 * i, r = e >> 4, kk = e & 15, gr = row0 + r, gk = k0 + kk;
      As[kk][r] = (gr < M && gk < K) ? A[(long)gr * K + gk] : T(0);
    }
#pragma unroll
    for (int i = 0; i < 2; i++) {
      int e = tid + 256 * i, kk = e >> 5, cf = e & 31, gk = k0 + kk, f = (col0 >> 1) + cf;
      cplx<T> v = cmk(T(0), T(0));
      if (gk < K) {
        long idx = (long)gk * nwp + f;
        v = B[idx];
        if (pro_stage(PRO)) {
          if (stage > 1) {
            cplx<T> kp = kprev[idx];
            T c = (T)cu;
            v = cscale(cmk(v.re + c * kp.re, v.im + c * kp.im), cond[f]);
          }
          cplx<T> kz = kzc[idx];
          v = cmul(linear_factor<T>(phs[idx], kz.im, zph, col_ref(f, f1r, zph, qb, 0).za, +1), v);
          if (PRO == PRO_STAGE) v = cmul(cmk(-kz.im, kz.re), v);  // E = i kz A
        } else if (PRO == PRO_FINISH) {
          cplx<T> kz = kzc[idx];
          v = cmul(linear_factor<T>(phs[idx], kz.im, zph, col_ref(f, f1r, zph, qb, 0).za, +1), v);
        }
      }
      Bs[kk][2 * cf] = v.re;
      Bs[kk][2 * cf + 1] = v.im;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk++) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[i][j] += a[i] * b[j];
    }
    __syncthreads();
  }

  double loc = (EPI == EPI_STAGE && stage == 4) ? 0.0 : INFINITY;
  bool nan = false;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int r = row0 + ty * 4 + i;
    if (r >= M) continue;
#pragma unroll
    for (int jj = 0; jj < 2; jj++) {
      const int f = (col0 >> 1) + tx * 2 + jj;
      const long idx = (long)r * nwp + f;
      cplx<T> J = cmk(acc[i][2 * jj], acc[i][2 * jj + 1]);
      if (EPI == EPI_NONE) {
        out[idx] = J;
        continue;
      }
      // S = 0.5 i exp(-i(kz+kg) z) J / Re kz (uppe.cl:73-91)
      cplx<T> kz = kzc[idx];
      const ColRef cr = col_ref(f, f1r, zph, qb, rb_end);
      cplx<T> t = cmul(linear_factor<T>(phs[idx], kz.im, zph, cr.za, -1), J);
      cplx<T> S = cmk((T(-0.5) * t.im) / kz.re, (T(0.5) * t.re) / kz.re);
      if (stage == 1) {
        k1[idx] = S;
        cplx<T> Q = q[idx];
        double qr = (double)Q.re, qi = (double)Q.im;
        double mod = sqrt(qr * qr + qi * qi);
        if (mod != mod) nan = true;
        if (!st->fixed) {
          // LoadStepSize (paraxial.cl:40-58) with amin = rel_amin * max|q| (uppe_kernel.py:221)
          double amin = ref_amin(st->rel_amin * bitsd(*qmax), (double)kz.im, cr.qb);
          double sr = (double)S.re, si = (double)S.im;
          double den = qr * qr + qi * qi + amin * amin;
          double alpha = (qr * sr + qi * si) / den;
          double keff = (qr * si - qi * sr) / den;
          double Kr = 0.5 * st->dphi / st->L + fabs(keff) - alpha * (double)(alpha < 0.0);
          double d = st->dphi / Kr;
          loc = fmin(loc, d);
        }
      } else if (stage == 2) {
        k2[idx] = S;
      } else if (stage == 3) {
        k3[idx] = S;
      } else {
        // qf accumulation of RKStage x3 + RKFinish, then the conditioner (uppe.cl:22-52)
        T c6 = (T)(dz / 6), c3 = (T)(dz / 3);
        cplx<T> Q = q[idx], a1 = k1[idx], a2 = k2[idx], a3 = k3[idx];
        T re = Q.re + c6 * a1.re;
        T im = Q.im + c6 * a1.im;
        re = re + c3 * a2.re;
        im = im + c3 * a2.im;
        re = re + c3 * a3.re;
        im = im + c3 * a3.im;
        re = re + c6 * S.re;
        im = im + c6 * S.im;
        T cf = cond[f];
        cplx<T> qn = cmk(re * cf, im * cf);
        const double m2 = ref_update<T>(qn, kz.im, cr, zph);
        q[idx] = qn;
        loc = fmax(loc, sqrt(m2));
      }
    }
  }
  if (EPI == EPI_NONE) return;
  if (stage == 1 && nan) atomicOr(&st->nanflag, 1);
  if (stage == 2 || stage == 3) return;
  if (stage == 1 && st->fixed) return;
  // block reduction: min of the dz estimate (stage 1) or max |q_new| (stage 4)
  const bool is_min = stage == 1;
  for (int o = 32; o > 0; o >>= 1) {
    double v = __shfl_xor(loc, o);
    loc = is_min ? fmin(loc, v) : fmax(loc, v);
  }
  if ((tid & 63) == 0) red[tid >> 6] = loc;
  __syncthreads();
  if (tid == 0) {
    double v = red[0];
    for (int i = 1; i < 4; i++) v = is_min ? fmin(v, red[i]) : fmax(v, red[i]);
    if (is_min) {
      atomicMin(&st->dzraw, dbits(v));
    } else {
      atomicMax(&qmax_next[red_slot(blockIdx.x + gridDim.x * blockIdx.y)], dbits(v));
      // the step has run; the next step's stage-1 source launch commits its end
      if (blockIdx.x == 0 && blockIdx.y == 0) st->pend = 1;
    }
  }
}

// ---------------------------------------------------------------------------------------------------
// Hankel contraction on the matrix cores
//
// out[M][2 nwp] = A[M][K] . B[K][2 nwp] with the same prologue/epilogue as f_gemm. One workgroup (up to 16
// waves of rw = 16 or 32 rows) owns an output tile of all M rows x BN real columns and stages its whole B slab
// (K x BN) in LDS once, so the prologue runs exactly once per element. A is read as pre-swizzled MFMA fragments straight from L2/HBM:
//   CM_F64 / CM_F32: v_mfma_{f64,f32}_16x16x4, lane l holds A[16 rt + l%16][16 kq + 4 j + l/16], j < 4;
//   CM_BF16X3:       v_mfma_f32_16x16x32_bf16, lane l holds A[16 rt + l%16][32 kc + 8 (l/16) + j], j < 8,
//                    as hi = bf16(a) and lo = bf16(a - hi); the product is hi.hi + hi.lo + lo.hi in FP32.
// The accumulator of a 16x16 tile sits in lane l, element i (i < 4), at column l % 16 and row 4 (l/16) + i for
// the f32-output instructions but row 4 i + l/16 for v_mfma_f64_16x16x4 (both measured with one-hot operands).
// ---------------------------------------------------------------------------------------------------
enum { CM_F64 = 0, CM_F32 = 1, CM_BF16X3 = 2 };

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int CM>
struct MfmaTraits;
template <>
struct MfmaTraits<CM_F64> {
  using T = double;
  using acc_t = f64x4;
  using frag_t = f64x4;  // 4 consecutive K steps of one lane
};
template <>
struct MfmaTraits<CM_F32> {
  using T = float;
  using acc_t = f32x4;
  using frag_t = f32x4;
};
template <>
struct MfmaTraits<CM_BF16X3> {
  using T = float;
  using acc_t = f32x4;
  using frag_t = bf16x8;
};

// one point of the r2k epilogue, from values already loaded: S from the contracted current, then the stage
// bookkeeping (stage 1: k1, NaN check and the per-point step estimate; stages 2, 3: k2, k3; stage 4: the RK
// update of q with the conditioner and max|q|)
template <typename T>
__device__ inline void stage_point(cplx<T> J, cplx<T> kz, double ph, cplx<T> Q, cplx<T> a1, cplx<T> a2,
                                   cplx<T> a3, T cf, long idx, int stage, double zph, const ColRef& cr, double dz,
                                   cplx<T>* __restrict__ q, cplx<T>* __restrict__ k1, cplx<T>* __restrict__ k2,
                                   cplx<T>* __restrict__ k3, const StepState* st, double amin, double& loc,
                                   bool& nan) {
  cplx<T> t = cmul(linear_factor<T>(ph, kz.im, zph, cr.za, -1), J);
  cplx<T> S;
  if constexpr (sizeof(T) == 4) {
    const float rk = __builtin_amdgcn_rcpf(kz.re);  // 1 ulp
    S = cmk(-0.5f * t.im * rk, 0.5f * t.re * rk);
  } else {
    S = cmk((T(-0.5) * t.im) / kz.re, (T(0.5) * t.re) / kz.re);
  }
  if (stage == 1) {
    k1[idx] = S;
    double qr = (double)Q.re, qi = (double)Q.im;
    if (qr != qr || qi != qi) nan = true;  // |q| is NaN exactly when a part is
    if (!st->fixed) {
      // loc keeps the largest rate Kr: the smallest per-point step dphi / Kr is dphi over it (a correctly
      // rounded quotient falls monotonically with the divisor), formed once after the loop
      double sr = (double)S.re, si = (double)S.im;
      const double am = ref_amin(amin, (double)kz.im, cr.qb);
      const double rden = 1.0 / (qr * qr + qi * qi + am * am);
      double alpha = (qr * sr + qi * si) * rden;
      double keff = (qr * si - qi * sr) * rden;
      double Kr = 0.5 * st->dphi / st->L + fabs(keff) - alpha * (double)(alpha < 0.0);
      loc = fmax(loc, Kr);
    }
  } else if (stage == 2) {
    k2[idx] = S;
  } else if (stage == 3) {
    k3[idx] = S;
  } else {
    T c6 = (T)(dz / 6), c3 = (T)(dz / 3);
    T re = Q.re + c6 * a1.re;
    T im = Q.im + c6 * a1.im;
    re = re + c3 * a2.re;
    im = im + c3 * a2.im;
    re = re + c3 * a3.re;
    im = im + c3 * a3.im;
    re = re + c6 * S.re;
    im = im + c6 * S.im;
    cplx<T> qn = cmk(re * cf, im * cf);
    const double m2 = ref_update<T>(qn, kz.im, cr, zph);
    q[idx] = qn;
    loc = fmax(loc, m2);  // |q|^2: the root of the largest, after the loop
  }
}

// one point of the r2k epilogue (loads, then stage_point)
template <typename T>
__device__ inline void stage_epilogue(cplx<T> J, long idx, int f, int stage, double zph, const ColRef& cr, double dz,
                                      const double* __restrict__ phs, const cplx<T>* __restrict__ kzc,
                                      const T* __restrict__ cond, cplx<T>* __restrict__ q,
                                      cplx<T>* __restrict__ k1, cplx<T>* __restrict__ k2,
                                      cplx<T>* __restrict__ k3, const StepState* st, double amin, double& loc,
                                      bool& nan) {
  const cplx<T> z0 = cmk(T(0), T(0));
  cplx<T> Q = z0, a1 = z0, a2 = z0, a3 = z0;
  T cf = T(1);
  if (stage == 1 || stage == 4) Q = q[idx];
  if (stage == 4) {
    a1 = k1[idx];
    a2 = k2[idx];
    a3 = k3[idx];
    cf = cond[f];
  }
  stage_point<T>(J, kzc[idx], phs[idx], Q, a1, a2, a3, cf, idx, stage, zph, cr, dz, q, k1, k2, k3, st, amin, loc,
                 nan);
}

// one point of the adjoint r2k epilogue. x = conj(i kz e^{i kz z}) times the contraction is the cotangent of the
// stage input; y = C x for stages 2-4 (their inputs are conditioned), x for stage 1. With gs the conditioned
// cotangent of the step's output and gq the running cotangent of its input, the reverse of
// q' = C(q + dz/6 k1 + dz/3 k2 + dz/3 k3 + dz/6 k4), k_s = S(C(q + c_s dz k_{s-1})) is
//   stage 4: gq = gs + y, gk3 = dz/3 gs + dz y;   stage 3: gq += y, gk2 = dz/3 gs + dz/2 y;
//   stage 2: gq += y, gk1 = dz/6 gs + dz/2 y;     stage 1: gq += y, the cotangent of the step's input,
// which for an earlier step (dzprev > 0) becomes its gs = C gq and gk4 = dzprev/6 gs. EPI_ADJ_FIN takes the
// adjoint of finish's exp(i kz z): gs = C conj(e^{i kz z}) x, gk4 = dzprev/6 gs.
template <typename T, int EPI>
__device__ inline void adj_point(cplx<T> acc, cplx<T> kz, double ph, T cf, long idx, int stage, double zph, double dz,
                                 double dzprev, cplx<T>* __restrict__ gq, cplx<T>* __restrict__ gs,
                                 cplx<T>* __restrict__ gk) {
  const cplx<T> back = cmul(linear_factor<T>(ph, -kz.im, zph, zph, -1), acc);  // conj(e^{i (Re kz + kg) z} e^{-Im kz z})
  if (EPI == EPI_ADJ_FIN) {
    const cplx<T> g = cscale(back, cf);
    gs[idx] = g;
    gk[idx] = cscale(g, (T)(dzprev / 6));
    return;
  }
  const cplx<T> x = cmul(cmk(-kz.im, -kz.re), back);  // conj(i kz) = -Im kz - i Re kz
  cplx<T> y = x;
  if (stage > 1) y = cscale(x, cf);
  const T d = (T)dz;
  if (stage == 4) {
    const cplx<T> g = gs[idx];
    gq[idx] = cadd(g, y);
    gk[idx] = cadd(cscale(g, d / T(3)), cscale(y, d));
  } else if (stage == 3) {
    const cplx<T> g = gs[idx];
    gq[idx] = cadd(gq[idx], y);
    gk[idx] = cadd(cscale(g, d / T(3)), cscale(y, d / T(2)));
  } else if (stage == 2) {
    const cplx<T> g = gs[idx];
    gq[idx] = cadd(gq[idx], y);
    gk[idx] = cadd(cscale(g, d / T(6)), cscale(y, d / T(2)));
  } else {
    const cplx<T> g = cadd(gq[idx], y);
    if (dzprev > 0.0) {
      const cplx<T> gn = cscale(g, cf);
      gs[idx] = gn;
      gk[idx] = cscale(gn, (T)(dzprev / 6));
    } else {
      gq[idx] = g;
    }
  }
}

// B-operand element (k, f) with the prologue of the k2r direction applied
template <typename T, int PRO>
__device__ inline cplx<T> load_b(const cplx<T>* __restrict__ B, long idx, int f, int stage, double zph, double za,
                                 double cu, const double* __restrict__ phs, const cplx<T>* __restrict__ kzc,
                                 const T* __restrict__ cond, const cplx<T>* __restrict__ kprev) {
  cplx<T> v = B[idx];
  if (pro_stage(PRO)) {
    if (stage > 1) {
      cplx<T> kp = kprev[idx];
      T c = (T)cu;
      v = cscale(cmk(v.re + c * kp.re, v.im + c * kp.im), cond[f]);
    }
    cplx<T> kz = kzc[idx];
    v = cmul(linear_factor<T>(phs[idx], kz.im, zph, za, +1), v);
    if (PRO == PRO_STAGE) v = cmul(cmk(-kz.im, kz.re), v);  // E = i kz A
  } else if (PRO == PRO_FINISH) {
    cplx<T> kz = kzc[idx];
    v = cmul(linear_factor<T>(phs[idx], kz.im, zph, za, +1), v);
  }
  return v;
}

// Workgroups are dealt to the 8 XCDs round-robin; renumber them so that each XCD owns a contiguous run of
// column tiles, whose partial 128-byte lines then meet in that XCD's L2.
// pointer of instance z of a batched launch (null stays null)
template <typename P>
__device__ inline P* zshift(P* p, long o) {
  return p ? p + o : p;
}

__device__ inline int xcd_tile(int b, int G) {
  const int q = G >> 3, r = G & 7, x = b & 7, slot = b >> 3;
  return x < r ? x * (q + 1) + slot : r * (q + 1) + (x - r) * q + slot;
}

// B-slab elements each thread loads per batch (all loads of a batch are in flight together)
constexpr int hankel_batch(int cm, int rt) { return (cm == CM_F64 || rt == 2) ? 2 : 4; }

// where the r2k epilogue's reductions go: the running minimum of the step estimate, the NaN flag and the
// "step has run" flag (null: not kept)
struct StepOut {
  unsigned long long* dzraw;
  int* nanflag;
  int* pend;
};

// One output tile of the contraction (see f_hankel): column tile `tile` of `ntile`, row block `rowblk`,
// instance `zi` of a batched launch.
template <int CM, int PRO, int EPI, int BN, int RT>
__device__ inline void hankel_tile(
    const typename MfmaTraits<CM>::frag_t* __restrict__ Ahi, const typename MfmaTraits<CM>::frag_t* __restrict__ Alo,
    const cplx<typename MfmaTraits<CM>::T>* __restrict__ B, cplx<typename MfmaTraits<CM>::T>* __restrict__ out,
    int M, int K, int Kp, int rw, int nwp, int f0, int fend, const double* __restrict__ phs,
    const cplx<typename MfmaTraits<CM>::T>* __restrict__ kzc, const typename MfmaTraits<CM>::T* __restrict__ cond,
    cplx<typename MfmaTraits<CM>::T>* __restrict__ q, const cplx<typename MfmaTraits<CM>::T>* __restrict__ kprev,
    cplx<typename MfmaTraits<CM>::T>* __restrict__ k1, cplx<typename MfmaTraits<CM>::T>* __restrict__ k2,
    cplx<typename MfmaTraits<CM>::T>* __restrict__ k3, const StepState* st, StepOut so,
    const unsigned long long* __restrict__ qmax, unsigned long long* __restrict__ qmax_next, int stage, double zfix,
    double dzfix, double dzprev, int wide, long zs, int zst, int kch, int share, int tile, int ntile, int rowblk,
    int zi) {
  using T = typename MfmaTraits<CM>::T;
  using C = cplx<T>;
  using acc_t = typename MfmaTraits<CM>::acc_t;
  if (zs) {
    // batched launch: instance zi of equally shaped instances whose arrays lie zs elements apart
    // (state, tables), one fragment set apart (operator); batch members (zst = 1) also one step state and three
    // maxima apart, the azimuthal modes of one grid (zst = 0) share them. share bit 0: every instance reads
    // the first one's operator, bit 1: its kz and phase tables (their inputs are the same)
    const long z = zi, zo = z * zs;
    const long fs = (long)((M + 15) / 16) * (CM == CM_BF16X3 ? Kp / 32 : Kp / 16) * 64;
    Ahi = zshift(Ahi, (share & 1) ? 0 : z * fs);
    Alo = zshift(Alo, (share & 1) ? 0 : z * fs);
    B = zshift(B, zo);
    out = zshift(out, zo);
    phs = zshift(phs, (share & 2) ? 0 : zo);
    kzc = zshift(kzc, (share & 2) ? 0 : zo);
    q = zshift(q, zo);
    kprev = zshift(kprev, zo);
    k1 = zshift(k1, zo);
    k2 = zshift(k2, zo);
    k3 = zshift(k3, zo);
    if (zst) {
      st = zshift(st, z);
      qmax = zshift(qmax, (long)kQW * z);
      qmax_next = zshift(qmax_next, (long)kQW * z);
      if (so.dzraw) so.dzraw = &const_cast<StepState*>(st)->dzraw;
      if (so.nanflag) so.nanflag = &const_cast<StepState*>(st)->nanflag;
      if (so.pend) so.pend = &const_cast<StepState*>(st)->pend;
    }
  }
  // no early exit for a finished run: the contraction only writes scratch (E_r) or stops before its
  // epilogue, so its loads need not wait for the step state
  extern __shared__ __align__(16) char smem[];
  __shared__ double red[16];
  const int nth = blockDim.x, nwv = nth >> 6;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int bm = nwv * rw;
  const int row0 = rowblk * bm;
  // column tiles of HC complex columns (one 16-column MFMA tile); the last tile of the grid also takes the
  // remainder (fewer than HC columns), so that the grid never needs a workgroup for a sliver. Wide tiles take
  // HC << wide columns each (wide = 1: two MFMA column tiles; wide = 2 with BN = 32: four), the last one what is
  // left: every A fragment then feeds that many MFMA column tiles, which divides the operator reads.
  // nmt = MFMA column tiles holding columns below fend.
  constexpr int HC = 8;
  constexpr int BW = 2 * BN;       // LDS row width in real columns: BN / 8 MFMA column tiles
  const int tcw = HC << wide;      // complex columns per tile
  const int fc0 = f0 + tile * tcw; // first complex column of this tile
  const int hcw = wide ? tcw : ((tile == ntile - 1 && fend - fc0 > HC) ? 2 * HC : HC);  // staged columns
  const int nmt = min(BW / 16, (min(fend - fc0, hcw) + HC - 1) / HC);
  double zph = zfix, cu = 0.0, dz = dzfix, qb = 0.0;
  int f1r = 0, rb_end = 0;  // launches without a step state (set-up, adjoint) see the state itself

  // ---- A fragments: a ring of RING contraction chunks per row tile, the first RING issued before the B
  // slab is staged so their latency overlaps it
  using F = typename MfmaTraits<CM>::frag_t;
  constexpr int RING = (CM == CM_F64 || RT == 2) ? 2 : 4;
  constexpr int rt_n = RT;  // 16-row tiles per wave (rw = 16 RT)
  const int mt0 = (row0 + wave * rw) / 16;  // first 16-row tile of this wave
  const int M16 = (M + 15) / 16;
  const int NCH = CM == CM_BF16X3 ? Kp / 32 : Kp / 16;  // contraction chunks
  F ring_h[RING][RT], ring_l[RING][CM == CM_BF16X3 ? RT : 1];
  auto fetch_frag = [&](int ch, int slot) {
    if (ch >= NCH) return;
#pragma unroll
    for (int r = 0; r < RT; r++)
      if (mt0 + r < M16) {
        const long fi = ((long)(mt0 + r) * NCH + ch) * 64 + lane;
        ring_h[slot][r] = Ahi[fi];
        if constexpr (CM == CM_BF16X3) ring_l[slot][r] = Alo[fi];
      }
  };
#pragma unroll
  for (int d = 0; d < RING; d++) fetch_frag(d, d);
  // The step position, formed once the first operator fragments are in flight: its read of the step
  // state then overlaps their latency instead of preceding it. (Formed later still, after the first loads of the
  // B slab, it would keep those registers live across step_pos and spill.)
  if (st) {
    const StepPos sp = step_pos(st, stage);
    dz = sp.dz;
    double cs = stage == 1 ? 0.0 : (stage == 4 ? 1.0 : 0.5);
    zph = sp.zoff + (sp.z + cs * dz);
    cu = stage == 1 ? 0.0 : (stage == 4 ? dz : dz / 2);
    if constexpr (sizeof(T) == 4) {  // only FP32 states are rebased
      qb = sp.qb;
      f1r = sp.f1;
      rb_end = sp.rb_end;
    }
  }
  // the step position is uniform over the workgroup: keep it in scalar registers
  zph = uniform_f64(zph);
  cu = uniform_f64(cu);
  dz = uniform_f64(dz);
  qb = uniform_f64(qb);
  f1r = __builtin_amdgcn_readfirstlane(f1r);
  rb_end = __builtin_amdgcn_readfirstlane(rb_end);

  // accumulators of the MFMA loop: wave w owns rows [row0 + w rw, row0 + (w+1) rw), rw <= 32
  constexpr int CT = BW / 16;  // column tiles the accumulators can hold; nmt of them are in use
  acc_t acc[RT][CT];
#pragma unroll
  for (int r = 0; r < RT; r++)
#pragma unroll
    for (int c = 0; c < CT; c++) acc[r][c] = acc_t{0, 0, 0, 0};

  // the contraction runs over chunks of kch rows of the B slab (all Kp rows at once when the slab fits in LDS;
  // kch is a multiple of the fragment ring's span, so that chunk boundaries keep the ring's slots in step)
  const int lgh = __builtin_ctz(hcw);
  for (int kb = 0; kb < Kp; kb += kch) {
    const int kcn = min(kch, Kp - kb);  // rows of this chunk
    if (kb > 0) __syncthreads();        // the previous chunk's MFMA loop has read its slab
    // ---- stage rows [kb, kb + kcn) of the B slab (Kp x hcw) in LDS with the prologue applied; loads of a batch are
    // issued together
    constexpr int EB = hankel_batch(CM, RT);  // register room: elements per thread per batch
    const int ldb = kch + 8;    // bf16 row stride (k contiguous), padded
    const int ne = kcn * hcw;
    const int eb = min(EB, (ne + nth - 1) / nth);  // elements per thread per batch: as few as cover the slab
    for (int base = 0; base < ne; base += eb * nth) {
      C vb[EB], vk[EB], vz[EB];
      double vp[EB];
      T vc[EB];
#pragma unroll
      for (int j = 0; j < EB; j++) {
        const int e = base + tid + j * nth;
        const int k = kb + (e >> lgh), f = fc0 + (e & (hcw - 1));
        vb[j] = cmk(T(0), T(0));
        vk[j] = cmk(T(0), T(0));
        vz[j] = cmk(T(1), T(0));
        vp[j] = 0.0;
        vc[j] = T(1);
        if (j < eb && e < ne && k < K && f < fend) {
          const long idx = (long)k * nwp + f;
          vb[j] = B[idx];
          if (PRO != PRO_NONE) {
            vz[j] = kzc[idx];
            vp[j] = phs[idx];
          }
          if (pro_stage(PRO) && stage > 1) {
            vk[j] = kprev[idx];
            vc[j] = cond[f];
          }
        }
      }
#pragma unroll
      for (int j = 0; j < EB; j++) {
        const int e = base + tid + j * nth;
        if (j >= eb || e >= ne) break;
        const int k = e >> lgh, cf = e & (hcw - 1);  // k: row within the chunk
        C v = vb[j];
        if (pro_stage(PRO)) {
          if (stage > 1) {
            T c = (T)cu;
            v = cscale(cmk(v.re + c * vk[j].re, v.im + c * vk[j].im), vc[j]);
          }
          v = cmul(linear_factor<T>(vp[j], vz[j].im, zph, col_ref(fc0 + cf, f1r, zph, qb, 0).za, +1), v);
          if (PRO == PRO_STAGE) v = cmul(cmk(-vz[j].im, vz[j].re), v);  // E = i kz A
        } else if (PRO == PRO_FINISH) {
          v = cmul(linear_factor<T>(vp[j], vz[j].im, zph, col_ref(fc0 + cf, f1r, zph, qb, 0).za, +1), v);
        } else if (PRO == PRO_ADJ) {
          // conj of S = (i/2) e^{-i kz z} J / Re kz: (-i/2) e^{i (Re kz + kg) z} e^{Im kz z} g / Re kz
          const C a = cmul(linear_factor<T>(vp[j], -vz[j].im, zph, zph, +1), v);
          v = cmk((T(0.5) * a.im) / vz[j].re, (T(-0.5) * a.re) / vz[j].re);
        }
        if constexpr (CM == CM_BF16X3) {
          __bf16* bhi = reinterpret_cast<__bf16*>(smem);
          __bf16* blo = bhi + BW * ldb;
          __bf16 hr = (__bf16)v.re, hi_ = (__bf16)v.im;
          bhi[(2 * cf) * ldb + k] = hr;
          bhi[(2 * cf + 1) * ldb + k] = hi_;
          blo[(2 * cf) * ldb + k] = (__bf16)((float)v.re - (float)hr);
          blo[(2 * cf + 1) * ldb + k] = (__bf16)((float)v.im - (float)hi_);
        } else {
          T* bs = reinterpret_cast<T*>(smem);
          bs[k * BW + 2 * cf] = v.re;
          bs[k * BW + 2 * cf + 1] = v.im;
        }
      }
    }
    __syncthreads();

    // ---- MFMA main loop over the chunk's contraction chunks (global indices: the ring prefetches across chunks)
    if constexpr (CM == CM_BF16X3) {
      const __bf16* bhi = reinterpret_cast<const __bf16*>(smem);
      const __bf16* blo = bhi + BW * ldb;
      const int KC = (kb + kcn) / 32;
      for (int kc0 = kb / 32; kc0 < KC; kc0 += RING) {
#pragma unroll
        for (int d = 0; d < RING; d++) {
          const int kc = kc0 + d;
          if (kc >= KC) break;
          bf16x8 bh[CT], bl[CT];
#pragma unroll
          for (int c = 0; c < CT; c++) {
            if (c >= nmt) break;
            const int col = c * 16 + (lane & 15), ko = kc * 32 - kb + 8 * (lane >> 4);
            bh[c] = *reinterpret_cast<const bf16x8*>(bhi + col * ldb + ko);
            bl[c] = *reinterpret_cast<const bf16x8*>(blo + col * ldb + ko);
          }
#pragma unroll
          for (int r = 0; r < RT; r++) {
            if (mt0 + r >= M16) break;
#pragma unroll
            for (int c = 0; c < CT; c++) {
              if (c >= nmt) break;
              acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ring_h[d][r], bh[c], acc[r][c], 0, 0, 0);
              acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ring_h[d][r], bl[c], acc[r][c], 0, 0, 0);
              acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ring_l[d][r], bh[c], acc[r][c], 0, 0, 0);
            }
          }
          fetch_frag(kc + RING, d);
        }
      }
    } else {
      const T* bs = reinterpret_cast<const T*>(smem);
      const int KQ = (kb + kcn) / 16;
      for (int kq0 = kb / 16; kq0 < KQ; kq0 += RING) {
#pragma unroll
        for (int d = 0; d < RING; d++) {
          const int kq = kq0 + d;
          if (kq >= KQ) break;
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int krow = kq * 16 - kb + 4 * j + (lane >> 4);
            T b[CT];
#pragma unroll
            for (int c = 0; c < CT; c++) b[c] = c < nmt ? bs[krow * BW + c * 16 + (lane & 15)] : T(0);
#pragma unroll
            for (int r = 0; r < 2; r++) {
              if (r >= rt_n || mt0 + r >= M16) break;
#pragma unroll
              for (int c = 0; c < CT; c++) {
                if (c >= nmt) break;
                if constexpr (CM == CM_F64)
                  acc[r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(ring_h[d][r][j], b[c], acc[r][c], 0, 0, 0);
                else
                  acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(ring_h[d][r][j], b[c], acc[r][c], 0, 0, 0);
              }
            }
          }
          fetch_frag(kq + RING, d);
        }
      }
    }
  }  // contraction chunks
  __syncthreads();  // B slab no longer needed: reuse LDS for the output tile

  // ---- accumulators -> LDS tile [bm][BW]
  T* os = reinterpret_cast<T*>(smem);
#pragma unroll
  for (int r = 0; r < RT; r++) {
#pragma unroll
    for (int c = 0; c < CT; c++) {
      if (c >= nmt) break;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int rin = CM == CM_F64 ? 4 * i + (lane >> 4) : 4 * (lane >> 4) + i;
        const int row = wave * rw + r * 16 + rin, col = c * 16 + (lane & 15);
        os[row * BW + col] = (T)acc[r][c][i];
      }
    }
  }
  __syncthreads();

  // ---- epilogue over complex points of the tile (the r2k direction leaves the state alone once the run
  // has finished)
  if (EPI == EPI_ADJ || EPI == EPI_ADJ_FIN) {
    // adjoint: q = gq (cotangent of the step's input state, accumulated), k1 = gs (conditioned cotangent of
    // the step's output), k2 = gk (cotangent of the stage derivative the next adjoint stage takes)
    const int npts = bm * hcw;
    for (int e = tid; e < npts; e += nth) {
      const int row = e >> lgh, cf = e & (hcw - 1), r = row0 + row, f = fc0 + cf;
      if (r >= M || f >= fend) continue;
      const long idx = (long)r * nwp + f;
      adj_point<T, EPI>(cmk(os[row * BW + 2 * cf], os[row * BW + 2 * cf + 1]), kzc[idx], phs[idx], cond[f], idx,
                        stage, zph, dz, dzprev, q, k1, k2);
    }
    return;
  }
  if (EPI == EPI_STAGE && !st->active) return;
  const double amin = (EPI == EPI_STAGE && stage == 1 && !st->fixed) ? st->rel_amin * bitsd(*qmax) : 0.0;
  double loc = 0.0;  // stage 1: the largest step rate Kr; stage 4: the largest |q|^2
  bool nan = false;
  // every column of the epilogue lies below fend = f1, the rebased band (or the state is not rebased: qb = 0)
  const ColRef cr_band{zph - qb, qb, rb_end != 0};
  // points are taken EPB per thread per batch, all of a batch's loads issued before its arithmetic
  constexpr int EPB = sizeof(T) == 8 ? 1 : 2;
  const int npts = bm * hcw;
  for (int base = 0; base < npts; base += EPB * nth) {
    long idx[EPB];
    int fo[EPB];
    bool ok[EPB];
    C J[EPB], kz[EPB], Q[EPB], a1[EPB], a2[EPB], a3[EPB];
    double ph[EPB];
    T cfv[EPB];
#pragma unroll
    for (int j = 0; j < EPB; j++) {
      const int e = base + tid + j * nth;
      const int row = e >> lgh, cf = e & (hcw - 1), r = row0 + row, f = fc0 + cf;
      ok[j] = e < npts && r < M && f < fend;
      idx[j] = (long)r * nwp + f;
      fo[j] = f;
      if (!ok[j]) continue;
      J[j] = cmk(os[row * BW + 2 * cf], os[row * BW + 2 * cf + 1]);
      if (EPI == EPI_NONE) {
        out[idx[j]] = J[j];
        continue;
      }
      kz[j] = kzc[idx[j]];
      ph[j] = phs[idx[j]];
      if (stage == 1 || stage == 4) Q[j] = q[idx[j]];
      if (stage == 4) {
        a1[j] = k1[idx[j]];
        a2[j] = k2[idx[j]];
        a3[j] = k3[idx[j]];
        cfv[j] = cond[fo[j]];
      }
    }
    if (EPI == EPI_NONE) continue;
#pragma unroll
    for (int j = 0; j < EPB; j++)
      if (ok[j])
        stage_point<T>(J[j], kz[j], ph[j], Q[j], a1[j], a2[j], a3[j], cfv[j], idx[j], stage, zph, cr_band, dz, q, k1,
                       k2, k3, st, amin, loc, nan);
  }
  if (EPI == EPI_NONE) return;
  if (stage == 1 && nan) atomicOr(so.nanflag, 1);
  if (stage == 2 || stage == 3) return;
  if (stage == 1 && st->fixed) return;
  // max|q| feeds only the adaptive step estimate. A fixed step spans its whole subcycle, so a fixed run needs it
  // only from the step that ends the advance call (its last subcycle, or the step that reaches max_steps), for a
  // later adaptive call; the other fixed steps skip the reduction and its atomics.
  if (stage == 4 && st->fixed && st->sub + 1 < st->sub_end &&
      !(st->max_steps > 0 && st->steps + 1 >= st->max_steps)) {
    if (tid == 0 && so.pend && tile == 0 && rowblk == 0) *so.pend = 1;  // the step has run
    return;
  }
  const bool is_min = stage == 1;
  loc = is_min ? (loc > 0.0 ? st->dphi / loc : INFINITY) : sqrt(loc);  // smallest step, or max |q|
  for (int o = 32; o > 0; o >>= 1) {
    double v = __shfl_xor(loc, o);
    loc = is_min ? fmin(loc, v) : fmax(loc, v);
  }
  if (lane == 0) red[wave] = loc;
  __syncthreads();
  if (tid == 0) {
    double v = red[0];
    for (int i = 1; i < nwv; i++) v = is_min ? fmin(v, red[i]) : fmax(v, red[i]);
    if (is_min) {
      atomicMin(so.dzraw, dbits(v));
    } else {
      atomicMax(qmax_next + red_slot(tile + ntile * rowblk), dbits(v));
      // the step has run; the next step's stage-1 source launch commits its end
      if (so.pend && tile == 0 && rowblk == 0) *so.pend = 1;
    }
  }
}

template <int CM, int PRO, int EPI, int BN, int RT>
__global__ __launch_bounds__(BN > 16 ? 512 : 1024) void f_hankel(
    const typename MfmaTraits<CM>::frag_t* __restrict__ Ahi, const typename MfmaTraits<CM>::frag_t* __restrict__ Alo,
    const cplx<typename MfmaTraits<CM>::T>* __restrict__ B, cplx<typename MfmaTraits<CM>::T>* __restrict__ out,
    int M, int K, int Kp, int rw, int nwp, int f0, int fend, const double* __restrict__ phs,
    const cplx<typename MfmaTraits<CM>::T>* __restrict__ kzc, const typename MfmaTraits<CM>::T* __restrict__ cond,
    cplx<typename MfmaTraits<CM>::T>* __restrict__ q, const cplx<typename MfmaTraits<CM>::T>* __restrict__ kprev,
    cplx<typename MfmaTraits<CM>::T>* __restrict__ k1, cplx<typename MfmaTraits<CM>::T>* __restrict__ k2,
    cplx<typename MfmaTraits<CM>::T>* __restrict__ k3, StepState* __restrict__ st,
    unsigned long long* __restrict__ qmax, unsigned long long* __restrict__ qmax_next, int stage, double zfix,
    double dzfix, double dzprev, int wide, long zs, int zst, int kch, int share) {
  const StepOut so{st ? &st->dzraw : nullptr, st ? &st->nanflag : nullptr, st ? &st->pend : nullptr};
  hankel_tile<CM, PRO, EPI, BN, RT>(Ahi, Alo, B, out, M, K, Kp, rw, nwp, f0, fend, phs, kzc, cond, q, kprev, k1,
                                           k2, k3, st, so, qmax, qmax_next, stage, zfix, dzfix, dzprev, wide, zs, zst,
                                           kch, share, xcd_tile(blockIdx.x, gridDim.x), gridDim.x, blockIdx.y,
                                           blockIdx.z);
}

// ---------------------------------------------------------------------------------------------------
// Cartesian transverse transforms (grid_tools.py:218-250; FFT_axis1/2, IFFT_axis1/2: fft.cl:387-441)
//
// A 2-D transverse transform is two passes of 1-D FFTs, one per axis. A workgroup takes CW consecutive
// frequency columns of one line (all Na points along the axis at one index of the other axis), so every
// global access is a run of CW complex values; the tile sits in LDS as [Na][CW] and its CW columns are
// transformed together (consecutive threads touch consecutive columns, so the LDS accesses of every pass are
// contiguous). The first pass of k2r applies the stage prologue on load, the last pass of r2k the RK stage
// epilogue on store. Forward transforms are unnormalised, inverse ones scale by 1/Na per pass (numpy).
// ---------------------------------------------------------------------------------------------------
// twiddle-free radix-R first pass (span 1) over the CW columns: butterfly j reads elements j + r Na/R and
// writes element j R + r; nj threads per column take 8/R butterflies each
template <typename T, int CW, int R>
__device__ inline void cols_first_pass(cplx<T>* buf, int na, int nj, int sign) {
  const int t = threadIdx.x, c = t % CW, jt = t / CW, nb = na / R;
  constexpr int IT = 8 / R;
  cplx<T> v[IT][R];
#pragma unroll
  for (int i = 0; i < IT; i++) {
    const int j = jt + i * nj;
    if (j < nb)
#pragma unroll
      for (int r = 0; r < R; r++) v[i][r] = buf[(j + r * nb) * CW + c];
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < IT; i++) {
    const int j = jt + i * nj;
    if (j >= nb) continue;
    if constexpr (R == 4) {
      dft4(v[i][0], v[i][1], v[i][2], v[i][3], sign);
    } else {
      const cplx<T> a = cadd(v[i][0], v[i][1]), b = csub(v[i][0], v[i][1]);
      v[i][0] = a;
      v[i][1] = b;
    }
#pragma unroll
    for (int r = 0; r < R; r++) buf[(j * R + r) * CW + c] = v[i][r];
  }
  __syncthreads();
}

// Stockham FFT of the CW columns of an [na][CW] LDS tile (the pass sequence and table layout of lds_fft_r8)
template <typename T, int CW>
__device__ inline void lds_fft_cols(cplx<T>* buf, int na, int lg, const cplx<T>* tab, int sign) {
  const int t = threadIdx.x, c = t % CW, jt = t / CW, nj = blockDim.x / CW;
  int Ns = 1;
  if (lg % 3 == 2) {
    cols_first_pass<T, CW, 4>(buf, na, nj, sign);
    Ns = 4;
  } else if (lg % 3 == 1) {
    cols_first_pass<T, CW, 2>(buf, na, nj, sign);
    Ns = 2;
  }
  const cplx<T>* tp = tab;
  for (; Ns < na; Ns *= 8) {
    const int nb = na >> 3;
    const bool act = jt < nb;
    cplx<T> v[8], y[8];
    if (act)
#pragma unroll
      for (int r = 0; r < 8; r++) v[r] = buf[(jt + r * nb) * CW + c];
    __syncthreads();
    if (act) {
      const int k = jt & (Ns - 1);
#pragma unroll
      for (int r = 1; r < 8; r++) {
        cplx<T> w = tp[(r - 1) * Ns + k];
        if (sign > 0) w.im = -w.im;
        v[r] = cmul(v[r], w);
      }
      dft8(v, y, sign);
      const int o = (jt - k) * 8 + k;
#pragma unroll
      for (int r = 0; r < 8; r++) buf[(o + r * Ns) * CW + c] = y[r];
    }
    __syncthreads();
    tp += 7 * Ns;
  }
}

// one 1-D pass of a transverse transform: axis 1 (x, lines strided by Ny Nwp) or 2 (y, strided by Nwp); grid
// (column tiles, index along the other axis); block >= 64 threads, CW per column line
template <typename T, int CW, int PRO, int EPI>
__global__ __launch_bounds__(1024) void f_tfft(
    const cplx<T>* __restrict__ src, cplx<T>* __restrict__ dst, int axis, int na, int lg, int ny, int nwp, int fend,
    const cplx<T>* __restrict__ tab_g, int tab_len, int sign, T scale, const double* __restrict__ phs,
    const cplx<T>* __restrict__ kzc, const T* __restrict__ cond, cplx<T>* __restrict__ q,
    const cplx<T>* __restrict__ kprev, cplx<T>* __restrict__ k1, cplx<T>* __restrict__ k2,
    cplx<T>* __restrict__ k3, StepState* __restrict__ st, unsigned long long* __restrict__ qmax,
    unsigned long long* __restrict__ qmax_next, int stage, double zfix, double qbfix, int f1fix) {
  using C = cplx<T>;
  extern __shared__ __align__(16) char smem[];
  C* buf = reinterpret_cast<C*>(smem);
  C* tab = buf + na * CW;
  __shared__ double red[16];
  const int t = threadIdx.x, nth = blockDim.x, lane = t & 63, wave = t >> 6, nwv = nth >> 6;
  const int f0 = blockIdx.x * CW;
  const long sa = axis == 1 ? (long)ny * nwp : (long)nwp;
  const long base = (axis == 1 ? (long)blockIdx.y : (long)blockIdx.y * ny) * nwp + f0;
  double zph = zfix, cu = 0.0, dz = 0.0, qb = sizeof(T) == 4 ? qbfix : 0.0;
  int f1r = f1fix, rb_end = 0;
  if (st) {
    const StepPos sp = step_pos(st, stage);
    dz = sp.dz;
    const double cs = stage == 1 ? 0.0 : (stage == 4 ? 1.0 : 0.5);
    zph = sp.zoff + (sp.z + cs * dz);
    cu = stage == 1 ? 0.0 : (stage == 4 ? dz : dz / 2);
    if constexpr (sizeof(T) == 4) {  // only FP32 states are rebased
      qb = sp.qb;
      f1r = sp.f1;
      rb_end = sp.rb_end;
    }
  }
  for (int i = t; i < tab_len; i += nth) tab[i] = tab_g[i];
  // the tile, with the prologue applied; four elements per thread per batch, their loads issued together
  const int ne = na * CW;
  for (int b0 = 0; b0 < ne; b0 += 4 * nth) {
    C v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = b0 + t + j * nth, a = e / CW, c = e % CW, f = f0 + c;
      v[j] = cmk(T(0), T(0));
      if (e < ne && f < fend) {
        const long idx = base + a * sa + c;
        v[j] = PRO == PRO_NONE ? src[idx]
                               : load_b<T, PRO>(src, idx, f, stage, zph, col_ref(f, f1r, zph, qb, 0).za, cu, phs,
                                                kzc, cond, kprev);
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int e = b0 + t + j * nth;
      if (e < ne) buf[e] = v[j];
    }
  }
  __syncthreads();
  lds_fft_cols<T, CW>(buf, na, lg, tab, sign);
  // store, or the RK stage epilogue (left out once the run has finished)
  if (EPI == EPI_STAGE && !st->active) return;
  const double amin = (EPI == EPI_STAGE && stage == 1 && !st->fixed) ? st->rel_amin * bitsd(*qmax) : 0.0;
  double loc = 0.0;  // stage 1: the largest step rate Kr; stage 4: the largest |q|^2
  bool nan = false;
  // the epilogue covers f < fend = f1, the rebased band (or the state is not rebased: qb = 0)
  const ColRef cr_band{zph - qb, qb, rb_end != 0};
  for (int e = t; e < ne; e += nth) {
    const int a = e / CW, c = e % CW, f = f0 + c;
    if (f >= fend) continue;
    const long idx = base + a * sa + c;
    const C v = cscale(buf[e], scale);
    if (EPI == EPI_NONE)
      dst[idx] = v;
    else
      stage_epilogue<T>(v, idx, f, stage, zph, cr_band, dz, phs, kzc, cond, q, k1, k2, k3, st, amin, loc, nan);
  }
  if (EPI == EPI_NONE) return;
  if (stage == 1 && nan) atomicOr(&st->nanflag, 1);
  if (stage == 2 || stage == 3) return;
  if (stage == 1 && st->fixed) return;
  const bool is_min = stage == 1;
  loc = is_min ? (loc > 0.0 ? st->dphi / loc : INFINITY) : sqrt(loc);  // smallest step, or max |q|
  for (int o = 32; o > 0; o >>= 1) {
    double v = __shfl_xor(loc, o);
    loc = is_min ? fmin(loc, v) : fmax(loc, v);
  }
  if (lane == 0) red[wave] = loc;
  __syncthreads();
  if (t == 0) {
    double v = red[0];
    for (int i = 1; i < nwv; i++) v = is_min ? fmin(v, red[i]) : fmax(v, red[i]);
    if (is_min) {
      atomicMin(&st->dzraw, dbits(v));
    } else {
      atomicMax(&qmax_next[red_slot(blockIdx.x + gridDim.x * blockIdx.y)], dbits(v));
      // the step has run; the next step's stage-1 source launch commits its end
      if (blockIdx.x == 0 && blockIdx.y == 0) st->pend = 1;
    }
  }
}

// ---------------------------------------------------------------------------------------------------
// strip transforms on one wavefront (the wave-per-strip source kernel)
// ---------------------------------------------------------------------------------------------------
__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// twiddles of the radix-8 passes of an n = 2^LG point transform that lane `lane` applies: pass q (span
// Ns = 2^(LG mod 3) 8^q), butterfly b (j = lane + 64 b), factor r = 1..7, from the tables of lds_fft_r8
template <int LG>
struct WaveTw {
  static constexpr int NP = LG / 3;                               // radix-8 passes
  static constexpr int NB = (1 << LG) / 8;                        // butterflies per pass
  static constexpr int BPL = NB >= 64 ? NB / 64 : 1;              // per lane
  cplx<float> t[NP > 0 ? NP : 1][BPL][7];
};

template <int LG>
__device__ inline void wave_tw_load(WaveTw<LG>& W, const cplx<float>* __restrict__ tab, int lane) {
  using WT = WaveTw<LG>;
  int off = 0;
#pragma unroll
  for (int q = 0; q < WT::NP; q++) {
    const int Ns = (1 << (LG % 3)) << (3 * q);
#pragma unroll
    for (int b = 0; b < WT::BPL; b++) {
      const int j = lane + 64 * b, k = j & (Ns - 1);
#pragma unroll
      for (int r = 1; r < 8; r++) W.t[q][b][r - 1] = j < WT::NB ? tab[off + (r - 1) * Ns + k] : cmk(1.0f, 0.0f);
    }
    off += 7 * Ns;
  }
}

// The Stockham passes of lds_fft_r8 on one wavefront: the n = 2^LG point strip buf (lds_swz layout) is
// transformed in place by the 64 lanes, each pass's loads ordered before its stores by wavefront-scope fences
// (wave_sync) instead of workgroup barriers, the twiddles from registers (W)
template <int LG>
__device__ inline void wave_fft_r8(cplx<float>* buf, const WaveTw<LG>& W, int sign, int lane) {
  using WT = WaveTw<LG>;
  constexpr int n = 1 << LG, r0 = LG % 3;
  auto at = [&](int e) -> cplx<float>& { return buf[lds_swz<float>(e)]; };
  if constexpr (r0 > 0) {
    constexpr int R = 1 << r0, nb = n >> r0, BP = (nb + 63) / 64;
    cplx<float> v[BP][4];
#pragma unroll
    for (int b = 0; b < BP; b++) {
      const int j = lane + 64 * b;
#pragma unroll
      for (int r = 0; r < R; r++)
        if (j < nb) v[b][r] = at(j + r * nb);
    }
    wave_sync();
#pragma unroll
    for (int b = 0; b < BP; b++) {
      const int j = lane + 64 * b;
      if (j >= nb) continue;
      if constexpr (R == 4) {
        dft4(v[b][0], v[b][1], v[b][2], v[b][3], sign);
      } else {
        const cplx<float> s0 = cadd(v[b][0], v[b][1]), s1 = csub(v[b][0], v[b][1]);
        v[b][0] = s0;
        v[b][1] = s1;
      }
#pragma unroll
      for (int r = 0; r < R; r++) at(j * R + r) = v[b][r];
    }
    wave_sync();
  }
#pragma unroll
  for (int q = 0; q < WT::NP; q++) {
    const int Ns = (1 << r0) << (3 * q);
    cplx<float> v[WT::BPL][8];
#pragma unroll
    for (int b = 0; b < WT::BPL; b++) {
      const int j = lane + 64 * b;
      if (j < WT::NB)
#pragma unroll
        for (int r = 0; r < 8; r++) v[b][r] = at(j + r * WT::NB);
    }
    wave_sync();
#pragma unroll
    for (int b = 0; b < WT::BPL; b++) {
      const int j = lane + 64 * b;
      if (j >= WT::NB) continue;
      const int k = j & (Ns - 1);
      cplx<float> y[8];
#pragma unroll
      for (int r = 1; r < 8; r++) {
        cplx<float> w = W.t[q][b][r - 1];
        if (sign > 0) w.im = -w.im;
        v[b][r] = cmul(v[b][r], w);
      }
      dft8(v[b], y, sign);
      const int o = (j - k) * 8 + k;
#pragma unroll
      for (int r = 0; r < 8; r++) at(o + r * Ns) = y[r];
    }
    wave_sync();
  }
}

// ---------------------------------------------------------------------------------------------------
// source kernel: one workgroup per strip
// ---------------------------------------------------------------------------------------------------
// ionisation rate of one sample (ionization.cl:22-44, 71-97). The FP32 path forms the closed form as one
// base-2 exponential 2^(log2(rate_conv C_pre) + C_pow log2 E + C_exp log2(e) / E) with the hardware exp2,
// log2 and reciprocal (about 1 ulp each; the exponent's own rounding, |x| 2^-24, dominates), and keeps the
// series polynomial in FP64 (its terms cancel) before an FP32 exponential; the FP64 path follows the
// reference's expressions.
template <typename T>
__device__ inline double ion_rate(double ea, const uppe_ionization& ion, double log_pre);

template <>
__device__ inline double ion_rate<double>(double ea, const uppe_ionization& ion, double log_pre) {
  if (ion.kind == UPPE_ION_EXPLICIT) {
    double emag = ion.e_conv * ea + ion.e_cutoff;
    return ion.rate_conv * ion.c_pre * pow(emag, ion.c_pow) * exp(ion.c_exp / emag);
  }
  double lg = log(ion.e_conv * ea + ion.e_cutoff);
  return ion.rate_conv * exp(ion.c[0] + ion.c[1] * lg + ion.c[2] * lg * lg + ion.c[3] * (lg * lg * lg) +
                             ion.c[4] * (lg * lg * lg * lg) + ion.c[5] * (lg * lg * lg * lg * lg));
}

template <>
__device__ inline double ion_rate<float>(double ea, const uppe_ionization& ion, double log_pre) {
  const float emag = (float)(ion.e_conv * ea + ion.e_cutoff);
  if (ion.kind == UPPE_ION_EXPLICIT) {
    // log_pre here is log2(rate_conv C_pre)
    const float x = (float)log_pre + (float)ion.c_pow * __builtin_amdgcn_logf(emag) +
                    (float)(ion.c_exp * 1.4426950408889634) * __builtin_amdgcn_rcpf(emag);
    return (double)__builtin_amdgcn_exp2f(x);
  }
  const float lg = logf(emag);
  const double l = (double)lg;
  const double poly = ion.c[0] + l * (ion.c[1] + l * (ion.c[2] + l * (ion.c[3] + l * (ion.c[4] + l * ion.c[5]))));
  return ion.rate_conv * (double)expf((float)poly);
}

// FP32 rate at field modulus ea (sim units). The explicit form runs in FP32 throughout: forming the field in
// atomic units in FP32 adds one rounding (~1e-7 relative, times the rate's logarithmic slope), below the FP32
// evaluation error of the exponent itself. The series form is ion_rate<float>.
__device__ inline float ion_rate_f32(float ea, const uppe_ionization& ion, double log_pre) {
  if (ion.kind == UPPE_ION_EXPLICIT) {
    const float emag = fmaf((float)ion.e_conv, ea, (float)ion.e_cutoff);
    const float x = (float)log_pre + (float)ion.c_pow * __builtin_amdgcn_logf(emag) +
                    (float)(ion.c_exp * 1.4426950408889634) * __builtin_amdgcn_rcpf(emag);
    return __builtin_amdgcn_exp2f(x);
  }
  return (float)ion_rate<float>((double)ea, ion, log_pre);
}

// 1 - exp(-x) for x >= 0 in FP32 (the depletion factor of the electron density): below 1/16 the Taylor
// polynomial to x^5 (truncation below 3e-8 relative), above it the hardware base-2 exponential, whose ~1 ulp
// the cancellation enlarges at most 16-fold. The library expm1f costs about twice the instructions.
__device__ inline float one_minus_exp_neg(float x) {
  const float p = fmaf(x, fmaf(x, fmaf(x, fmaf(x, 1.0f / 120.0f, -1.0f / 24.0f), 1.0f / 6.0f), -0.5f), 1.0f);
  const float big = 1.0f - __builtin_amdgcn_exp2f(-x * 1.4426950408889634f);
  return x < 0.0625f ? x * p : big;
}

// exp(+2 pi i k/N), the split twiddle of the half-length inverse real transform: FP32 from the hardware
// sine and cosine of k/N turns (about 2 ulp), FP64 from sincospi
template <typename T>
__device__ inline cplx<T> twiddle_inv(int k, int N);
template <>
__device__ inline cplx<float> twiddle_inv<float>(int k, int N) {
  const float t = (float)k * (1.0f / (float)N);  // N is a power of two: the reciprocal and product are exact
  return cmk(__builtin_amdgcn_cosf(t), __builtin_amdgcn_sinf(t));
}
template <>
__device__ inline cplx<double> twiddle_inv<double>(int k, int N) {
  double sn, cs;
  sincospi(2.0 * (double)k / (double)N, &sn, &cs);
  return cmk(cs, sn);
}

// real sample n of a strip buffer in the lds_swz layout (the real and imaginary parts of complex element
// n / 2 hold samples n and n + 1 for even n)
template <typename T>
__device__ inline T& lds_sample(cplx<T>* Y, int n) {
  return reinterpret_cast<T*>(Y + lds_swz<T>(n >> 1))[n & 1];
}

// RUN consecutive real samples from n0 (a multiple of RUN), as 16-byte vectors when RUN allows: a vector holds
// whole elements that lds_swz keeps together
template <typename T, int RUN>
__device__ inline void lds_read_reals(const cplx<T>* Y, int n0, T* dst) {
  constexpr int V = 16 / sizeof(T);
  if constexpr (RUN % V == 0) {
    typedef T vec __attribute__((ext_vector_type(V)));
#pragma unroll
    for (int i = 0; i < RUN / V; i++) {
      vec v = *reinterpret_cast<const vec*>(Y + lds_swz<T>(n0 / 2 + i * (V / 2)));
#pragma unroll
      for (int k = 0; k < V; k++) dst[i * V + k] = v[k];
    }
  } else {
#pragma unroll
    for (int j = 0; j < RUN; j++) dst[j] = lds_sample(const_cast<cplx<T>*>(Y), n0 + j);
  }
}

template <typename T, int RUN>
__device__ inline void lds_write_reals(cplx<T>* Y, int n0, const T* src) {
  constexpr int V = 16 / sizeof(T);
  if constexpr (RUN % V == 0) {
    typedef T vec __attribute__((ext_vector_type(V)));
#pragma unroll
    for (int i = 0; i < RUN / V; i++) {
      vec v;
#pragma unroll
      for (int k = 0; k < V; k++) v[k] = src[i * V + k];
      *reinterpret_cast<vec*>(Y + lds_swz<T>(n0 / 2 + i * (V / 2))) = v;
    }
  } else {
#pragma unroll
    for (int j = 0; j < RUN; j++) lds_sample(Y, n0 + j) = src[j];
  }
}

// DPP lane move (gfx9 control codes: row_shr:n = 0x110 + n, wave_shr:1 = 0x138, row_bcast:15 = 0x142,
// row_bcast:31 = 0x143); lanes without a source, or in rows outside RM, read 0
template <int CTRL, int RM, typename T>
__device__ inline T dpp_move(T v) {
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, RM, 0xf, false));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, RM, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, RM, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// exclusive prefix sum over the 64 lanes of a wave: shift by one lane, then an inclusive scan (log steps within
// each 16-lane row, then the row totals carried by row_bcast)
template <typename T>
__device__ inline T wave_excl_prefix(T v) {
  v = dpp_move<0x138, 0xf>(v);
  v += dpp_move<0x111, 0xf>(v);
  v += dpp_move<0x112, 0xf>(v);
  v += dpp_move<0x114, 0xf>(v);
  v += dpp_move<0x118, 0xf>(v);
  v += dpp_move<0x142, 0xa>(v);
  v += dpp_move<0x143, 0xc>(v);
  return v;
}

// sum of v over lanes 0 .. k-1 (k wave-uniform, at most 16; v held by lanes 0 .. 15): an inclusive scan within the
// first row of 16 lanes, read at lane k-1
template <typename T>
__device__ inline T lanes_below(T v, int k) {
  v += dpp_move<0x111, 0xf>(v);
  v += dpp_move<0x112, 0xf>(v);
  v += dpp_move<0x114, 0xf>(v);
  v += dpp_move<0x118, 0xf>(v);
  if (k == 0) return T(0);
  if constexpr (sizeof(T) == 4) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k - 1));
  } else {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, k - 1), hi = __builtin_amdgcn_readlane((int)(b >> 32), k - 1);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
  }
}

// E^(2 m0) of a rotator state for the thread with index tid, m0 = SPT tid. FP32: the phase m0 W dt reduced to
// turns in FP64, then the hardware sine, cosine and exp2; FP64: the product of the staged powers E^(2 SPT 2^b) over
// the set bits of tid
template <typename R>
__device__ inline cplx<R> rot_start(const RotRow<R>& rw, int tid, int m0);
template <>
__device__ inline cplx<float> rot_start<float>(const RotRow<float>& rw, int tid, int m0) {
  (void)tid;
  const double turns = rw.wr64 * (double)m0 * 0.15915494309189535;
  const float f = (float)(turns - rint(turns));
  const float mag = __builtin_amdgcn_exp2f((float)(-rw.wi64 * (double)m0 * 1.4426950408889634));
  return cmk(__builtin_amdgcn_cosf(f) * mag, __builtin_amdgcn_sinf(f) * mag);
}
template <>
__device__ inline cplx<double> rot_start<double>(const RotRow<double>& rw, int tid, int m0) {
  (void)m0;
  cplx<double> pw = cmk(1.0, 0.0);
#pragma unroll
  for (int b = 0; b < ROT_NB; b++) {
    if ((tid >> b) & 1) pw = cmul(pw, rw.e[2 + b]);
  }
  return pw;
}

// rotator states whose constants are staged in LDS at a time
constexpr int ROT_CHUNK = 64;
// rotator states per barrier in the source kernel
constexpr int ROT_G32 = 8, ROT_G64 = 2;

// Rotational (Raman) susceptibility on the SPT samples a thread owns (rotations.cl:26-129, AddChiUPPE and
// AddChiParaxial). For each state j >= 2 the reference runs the recurrence rho <- rho - i dt F_m E^(2m+1) and
// cum <- cum + 2 Im(T_j S_j rho conj(E^(2m)) W_j dt) n_g dalpha, adding cum to chi, over m = N-1-n (time from the
// last sample), E = exp(i W_j dt / 2), F_m = dalpha/(2 hbar) drive_m with drive = E(t)^2 for real fields and
// |E|^2 / 2 for the envelope. Both running sums are prefix sums over m, and the thread order already follows m:
// rho_m is a DPP wave scan plus the totals of the earlier waves, one barrier per G states. The sum over states
// commutes with the outer prefix sum, so the per-state terms g_j(m) are summed first and scanned once.
template <typename R, int SPT>
__device__ inline void rot_chi(const R* drive, R ngp, R* dchi, const RotRow<R>* __restrict__ rot_tab, int rot_states,
                               double rot_fc, double dt, int N, int a0) {
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63, wave = tid >> 6, nwv = nth >> 6;
  __shared__ R rot_wsr[16];
  constexpr int G = sizeof(R) == 4 ? ROT_G32 : ROT_G64;  // states per barrier
  const R dtr = (R)dt, fc = (R)rot_fc;
  const int m0 = N - a0 - SPT;  // = SPT tid: m of this thread's last sample; m grows as the sample index falls
  const int wv_u = __builtin_amdgcn_readfirstlane(wave);
  R gs[SPT];
#pragma unroll
  for (int j = 0; j < SPT; j++) gs[j] = R(0);
  // per-state constants, staged in LDS for ROT_CHUNK states at a time; the wave totals of G states per
  // barrier, double-buffered
  __shared__ RotRow<R> rot_r[ROT_CHUNK];
  __shared__ cplx<R> rot_ws[2][G][16];
  int par = 0;
  for (int s0 = 2; s0 < rot_states; s0 += ROT_CHUNK) {
    const int nc = min(ROT_CHUNK, rot_states - s0);
    __syncthreads();  // the previous chunk's constants are no longer read
    {
      const unsigned* src = reinterpret_cast<const unsigned*>(rot_tab + (s0 - 2));
      unsigned* dst = reinterpret_cast<unsigned*>(rot_r);
      const int words = nc * (int)(sizeof(RotRow<R>) / sizeof(unsigned));
      for (int i = tid; i < words; i += nth) dst[i] = src[i];
    }
    __syncthreads();
    for (int q0 = 0; q0 < nc; q0 += G) {
      // pass 1: this thread's sum of F_m E^(2m+1), its exclusive scan over the wave, the wave totals
      cplx<R> pw0[G], run[G];
#pragma unroll
      for (int g = 0; g < G; g++) {
        const RotRow<R>& rw = rot_r[min(q0 + g, nc - 1)];  // past the chunk end: a repeat, not used
        const cplx<R> e1 = rw.e[0], e2 = rw.e[1];
        cplx<R> pw = rot_start(rw, tid, m0), tot = cmk(R(0), R(0));
        pw0[g] = pw;
#pragma unroll
        for (int jj = SPT - 1; jj >= 0; jj--) {
          tot = cadd(tot, cscale(cmul(pw, e1), fc * drive[jj]));
          pw = cmul(pw, e2);
        }
        run[g] = cmk(wave_excl_prefix(tot.re), wave_excl_prefix(tot.im));
        if (lane == 63) rot_ws[par][g][wave] = cadd(run[g], tot);
      }
      __syncthreads();
      // pass 2: rho_m = -i dt (sum of F E^(2m+1) over earlier m); g_j(m) = c_j Im(rho conj(E^(2m)) W dt)
#pragma unroll
      for (int g = 0; g < G; g++) {
        if (q0 + g < nc) {
          const RotRow<R>& rw = rot_r[q0 + g];
          const R cj = rw.c * ngp, wr = rw.wr, wi = rw.wi;
          const cplx<R> e1 = rw.e[0], e2 = rw.e[1];
          const cplx<R> wt = lane < nwv ? rot_ws[par][g][lane] : cmk(R(0), R(0));
          cplx<R> r = cadd(run[g], cmk(lanes_below(wt.re, wv_u), lanes_below(wt.im, wv_u)));
          cplx<R> pw = pw0[g];
#pragma unroll
          for (int jj = SPT - 1; jj >= 0; jj--) {
            const cplx<R> rho = cmk(dtr * r.im, -dtr * r.re);
            const cplx<R> qv = cmul(rho, cconj(pw));
            gs[jj] += cj * (qv.re * wi + qv.im * wr);
            r = cadd(r, cscale(cmul(pw, e1), fc * drive[jj]));
            pw = cmul(pw, e2);
          }
        }
      }
      par ^= 1;
    }
  }
  R gtot = R(0);
#pragma unroll
  for (int jj = 0; jj < SPT; jj++) gtot += gs[jj];
  R cum = wave_excl_prefix(gtot);
  if (lane == 63) rot_wsr[wave] = cum + gtot;
  __syncthreads();
  cum += lanes_below(lane < nwv ? rot_wsr[lane] : R(0), wv_u);  // the earlier waves' totals
#pragma unroll
  for (int jj = SPT - 1; jj >= 0; jj--) {
    cum += gs[jj];
    dchi[jj] = cum;
  }
}

// Adjoint of rot_chi, for a thread that owns the samples n = n0 .. n0 + SPT - 1 with n0 = SPT t (m = N - 1 - n
// then falls along the thread, the reverse of rot_chi's order). With gam the cotangent of chi_rot at these samples,
// it adds to gd the cotangent of the drive. Writing rot_chi's state j (w_j its row's (wr, wi)) as
//   P_j(m) = sum_{m' < m} F(m') E^(2m'+1),   g_j(m) = c_j n_g Im(-i dt P_j(m) conj(E^(2m)) w_j)
//          = -c_j n_g dt Re(P_j(m) conj(E^(2m)) w_j),   chi_rot(m) = sum_{m'' <= m} sum_j g_j(m''),   F = fc drive,
// the reverse sweep is, with Gamma(m) = sum_{m' >= m} gam(m') and U_j(m) = sum_{m' > m} Gamma(m') conj(E^(2m')) w_j,
//   d L / d drive(m) = -fc dt sum_j c_j n_g Re(E^(2m+1) U_j(m)).
// Sums over larger m are sums over smaller n: Gamma is an inclusive and U_j an exclusive prefix sum in this
// thread order (a DPP wave scan plus the earlier waves' totals, one barrier per G states). The per-state rows are
// read from L2 (the forward's staging buffer keeps its own LDS).
template <typename R, int SPT>
__device__ inline void rot_chi_adj(const R* gam, R ngp, R* gd, const RotRow<R>* __restrict__ rot_tab, int rot_states,
                                   double rot_fc, double dt, int N, int n0) {
  const int tid = threadIdx.x, nth = blockDim.x, lane = tid & 63, wave = tid >> 6, nwv = nth >> 6;
  constexpr int G = sizeof(R) == 4 ? ROT_G32 : ROT_G64;
  __shared__ R adj_wsr[16];
  __shared__ cplx<R> adj_ws[2][G][16];
  const R dtr = (R)dt, fc = (R)rot_fc;
  const int wv_u = __builtin_amdgcn_readfirstlane(wave);
  // Gamma at this thread's samples: the inclusive prefix sum of gam in n
  R gt = R(0);
#pragma unroll
  for (int jj = 0; jj < SPT; jj++) gt += gam[jj];
  R before = wave_excl_prefix(gt);
  if (lane == 63) adj_wsr[wave] = before + gt;
  __syncthreads();
  before += lanes_below(lane < nwv ? adj_wsr[lane] : R(0), wv_u);  // the earlier waves' totals
  R Gm[SPT];
#pragma unroll
  for (int jj = 0; jj < SPT; jj++) {
    before += gam[jj];
    Gm[jj] = before;
  }
  // this thread's smallest m (its last sample): SPT (nth - 1 - t), as for the thread nth - 1 - t of rot_chi
  const int tm = nth - 1 - tid, mlow = N - n0 - SPT;
  R out[SPT];
#pragma unroll
  for (int jj = 0; jj < SPT; jj++) out[jj] = R(0);
  int par = 0;
  for (int q0 = 2; q0 < rot_states; q0 += G) {
    // pass 1: this thread's sum of Gamma conj(E^(2m)) W, its exclusive scan over the wave, the wave totals
    cplx<R> run[G];
#pragma unroll
    for (int g = 0; g < G; g++) {
      const RotRow<R> rw = rot_tab[min(q0 + g, rot_states - 1) - 2];  // past the last state: a repeat, not used
      const cplx<R> e2 = rw.e[1], wz = cmk(rw.wr, rw.wi);
      cplx<R> pw = rot_start(rw, tm, mlow), tot = cmk(R(0), R(0));
#pragma unroll
      for (int jj = SPT - 1; jj >= 0; jj--) {  // m = mlow + (SPT - 1 - jj)
        tot = cadd(tot, cscale(cmul(cconj(pw), wz), Gm[jj]));
        pw = cmul(pw, e2);
      }
      run[g] = cmk(wave_excl_prefix(tot.re), wave_excl_prefix(tot.im));
      if (lane == 63) adj_ws[par][g][wave] = cadd(run[g], tot);
    }
    __syncthreads();
    // pass 2: U_j at each sample (the sum over smaller n), then its term of the drive's cotangent
#pragma unroll
    for (int g = 0; g < G; g++) {
      if (q0 + g < rot_states) {
        const RotRow<R> rw = rot_tab[q0 + g - 2];
        const R cj = rw.c * ngp;
        const cplx<R> e1 = rw.e[0], e2 = rw.e[1], wz = cmk(rw.wr, rw.wi);
        const cplx<R> wt = lane < nwv ? adj_ws[par][g][lane] : cmk(R(0), R(0));
        cplx<R> u = cadd(run[g], cmk(lanes_below(wt.re, wv_u), lanes_below(wt.im, wv_u)));
        cplx<R> pws[SPT];
        cplx<R> pw = rot_start(rw, tm, mlow);
#pragma unroll
        for (int jj = SPT - 1; jj >= 0; jj--) {
          pws[jj] = pw;
          pw = cmul(pw, e2);
        }
#pragma unroll
        for (int jj = 0; jj < SPT; jj++) {  // n ascending: u holds the sum over the smaller n
          out[jj] -= cj * cmul(cmul(pws[jj], e1), u).re;
          u = cadd(u, cscale(cmul(cconj(pws[jj]), wz), Gm[jj]));
        }
      }
    }
    par ^= 1;
  }
#pragma unroll
  for (int jj = 0; jj < SPT; jj++) gd[jj] += fc * dtr * out[jj];
}

// one workgroup per strip, Nt/4 threads (at least min(Nt, 64)); thread t owns SPT = Nt / threads samples
// LGC > 0: an instance for Nt = 2^LGC (block Nt / SPT), whose transforms unroll with constant spans
template <typename T, int SPT, bool DIAG, bool ROT, int LGC = 0>
__global__ __launch_bounds__(1024) void f_source(const cplx<T>* __restrict__ er, cplx<T>* __restrict__ jr,
                                                 T* __restrict__ chi_out, T* __restrict__ ne_out,
                                                 const cplx<T>* __restrict__ tw, const T* __restrict__ w,
                                                 const T* __restrict__ nlf, const double* __restrict__ ng, int nt,
                                                 int lgn, int nwp, double redchi3, double dt, uppe_ionization ion,
                                                 double log_pre, const RotRow<T>* __restrict__ rot_tab,
                                                 int rot_states, double rot_fc, const SrcInst* __restrict__ inst,
                                                 long zs, StepState* commit, unsigned long long* cq, double* rec,
                                                 double* zrec, int zrec_cap, int commit_only) {
  // no early exit for a finished run: the kernel only writes scratch (J_r, diagnostics). The stage-1 launch
  // commits the previous step (commit non-null; nothing else of this launch reads the step state): thread 0 of
  // the first workgroup, or, in a launch of one workgroup more than strips (commit_only = the strip count), that
  // workgroup alone, which commits and leaves, so that no strip's chain waits for the commit's round trips. (The
  // commit stays ahead of the strip's loads: behind them it would hold their registers and raise the kernel's
  // VGPR count, which costs occupancy where many workgroups share a CU.)
  if (commit && threadIdx.x == 0 && (int)blockIdx.x == commit_only) {
    const int z = inst ? blockIdx.y : 0;
    step_commit(commit + z, cq + (long)kQW * z, rec, zrec, zrec_cap);
  }
  if (commit_only > 0 && (int)blockIdx.x == commit_only) return;
  extern __shared__ __align__(16) char smem[];
  cplx<T>* Y = reinterpret_cast<cplx<T>*>(smem);
  __shared__ T wsum[16];
  if (LGC > 0) lgn = LGC;
  const int N = LGC > 0 ? (1 << LGC) : nt, M = N / 2, lgm = lgn - 1;
  const long p = blockIdx.x;
  if (inst) {
    // batched launch: instance blockIdx.y, strips zs elements apart, its own medium scalars
    const int z = blockIdx.y;
    er += z * zs;
    jr += z * zs;
    ng += (long)z * gridDim.x;
    redchi3 = inst[z].redchi3;
    log_pre = inst[z].log_pre;
    ion = inst[z].ion;
    if (ROT) {
      rot_tab = static_cast<const RotRow<T>*>(inst[z].rot_tab);
      rot_states = inst[z].rot_states;
      rot_fc = inst[z].rot_fc;
    }
  }
  constexpr int NTH = LGC > 0 ? (1 << LGC) / SPT : 0;
  const int tid = threadIdx.x, nth = NTH > 0 ? NTH : (int)blockDim.x, lane = tid & 63, wave = tid >> 6,
            nwv = nth >> 6;
  const cplx<T>* X = er + p * nwp;
  // every global load the kernel makes up front, issued together: (FP32) the twiddle tables bound for LDS first,
  // then this strip's spectrum and the frequency and filter of the bins this thread writes at the end. The tables
  // come from L2 and reach LDS while the strip is still on its way from HBM.
  // Per-pass twiddle tables of the N- and N/2-point transforms (fewer than N + N/2 entries): in LDS for FP32
  // strips up to 4096 samples, read from L2 for FP64 and the 8192-sample strips (LDS would not fit)
  constexpr bool kTabLds = sizeof(T) == 4 && SPT <= 4;
  constexpr int TPT = kTabLds ? (3 * SPT + 1) / 2 : 1;  // table entries per thread: tlen < N + N/2 = 1.5 SPT nth
  const int tlen = r8_table_len(lgn) + r8_table_len(lgm);
  cplx<T> tv[TPT];
  if constexpr (kTabLds) {
#pragma unroll
    for (int i = 0; i < TPT; i++) {
      const int k = tid + i * nth;
      if (k < tlen) tv[i] = tw[k];
    }
  }
  constexpr int KPT = SPT / 2 + 1;  // spectral bins per thread (M / nth rounded up, plus the Nyquist bin)
  cplx<T> xk[KPT], xm[KPT];
  T wv[KPT], fv[KPT];
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int k = tid + i * nth;
    if (k < M) {
      xk[i] = X[k];
      xm[i] = X[M - k];
    }
    if (k <= M) {
      wv[i] = w[k];
      fv[i] = nlf[k];
    }
  }
  const cplx<T>* tws = tw;
  if constexpr (kTabLds) {
    cplx<T>* tl = Y + N;
#pragma unroll
    for (int i = 0; i < TPT; i++) {
      const int k = tid + i * nth;
      if (k < tlen) tl[k] = tv[i];
    }
    tws = tl;
    __syncthreads();
  }

  // 1. irfft (numpy semantics; imaginary parts of the DC and Nyquist bins are ignored, fft.cl:271-278)
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int k = tid + i * nth;
    if (k >= M) break;
    cplx<T> xa = xk[i], xb = xm[i];
    if (k == 0) {
      xa.im = T(0);
      xb.im = T(0);
    }
    cplx<T> ze = cscale(cadd(xa, cconj(xb)), T(0.5));
    cplx<T> zo = cmul(cscale(csub(xa, cconj(xb)), T(0.5)), twiddle_inv<T>(k, N));
    Y[lds_swz<T>(k)] = cmk(ze.re - zo.im, ze.im + zo.re);
  }
  __syncthreads();
  lds_fft_r8<T, (LGC > 0 ? LGC - 1 : 0), NTH, (SPT >= 8 ? 4 : 2)>(Y, M, lgm, tws + r8_table_len(lgn), +1, tid);
  const T invm = T(1) / T(M);
  // thread t owns samples [a0, a0 + SPT) with a0 = (nth - 1 - t) SPT: the depletion integral runs from the
  // last sample backward, which makes it a prefix sum over the threads in lane order
  const int a0 = (nth - 1 - tid) * SPT;
  T e[SPT];
  lds_read_reals<T, SPT>(Y, a0, e);
#pragma unroll
  for (int j = 0; j < SPT; j++) e[j] *= invm;

  // 2. rate, backward trapezoid with depletion (ionization.cl:127-155), soft window (fft.cl:331-355). With the
  // last rate sample zeroed, the trapezoid from sample n to the end is I_n = dt (r_n / 2 + sum_{m > n} r_m).
  // The sum is formed as a tree of positive terms (per thread, a DPP scan over the wave, then across
  // waves) in the path's precision: FP32 rounding grows with the tree depth, ~12 ulp at Nt = 4096.
  using R = T;
  const R ngp = (R)ng[p];
  R ne[SPT];
  if (ion.kind != UPPE_ION_NONE) {
    const R dtr = (R)dt;
    R r[SPT];
    R tot = R(0);
#pragma unroll
    for (int j = 0; j < SPT; j++) {
      const int n = a0 + j;
      R v;
      if constexpr (sizeof(T) == 4)
        v = ion_rate_f32(fabsf(e[j]), ion, log_pre);
      else
        v = ion_rate<T>(fabs(e[j]), ion, log_pre);
      r[j] = (n >= N - 1) ? R(0) : v;  // last sample zeroed (ionization.cl:147)
      tot += r[j];
    }
    const R ex = wave_excl_prefix(tot);  // rates of the later samples held by lower lanes of this wave
    if (lane == 63) wsum[wave] = ex + tot;
    __syncthreads();
    R later = ex;
    later += lanes_below(lane < nwv ? wsum[lane] : R(0), __builtin_amdgcn_readfirstlane(wave));  // earlier waves
#pragma unroll
    for (int j = SPT - 1; j >= 0; j--) {
      const int n = a0 + j;
      const R integral = (later + R(0.5) * r[j]) * dtr;
      later += r[j];
      // n_e = n_g (1 - exp(-integral)); the FP32 path evaluates the depletion with one_minus_exp_neg (accurate
      // also where the integral is tiny), the FP64 path as the reference does
      R x;
      if constexpr (sizeof(T) == 4)
        x = ngp * one_minus_exp_neg(integral);
      else
        x = ngp * (1.0 - exp(-integral));
      // SoftTimeWindow(4, 64)
      if (n < 4 || n > N - 5) x = R(0);
      if (n >= 4 && n < 68) {
        R qq = (R)(n - 4) / R(64);
        x *= R(10) * qq * qq * qq - R(15) * qq * qq * qq * qq + R(6) * qq * qq * qq * qq * qq;
      }
      if (n <= N - 5 && n > N - 69) {
        R qq = (R)(N - 5 - n) / R(64);
        x *= R(10) * qq * qq * qq - R(15) * qq * qq * qq * qq + R(6) * qq * qq * qq * qq * qq;
      }
      ne[j] = x;
    }
  } else {
#pragma unroll
    for (int j = 0; j < SPT; j++) ne[j] = R(0);
    __syncthreads();  // every thread has read its samples before the strip is overwritten
  }

  // 2b. rotational (Raman) susceptibility, driven by E(t)^2
  R dchi[SPT];
#pragma unroll
  for (int j = 0; j < SPT; j++) dchi[j] = R(0);
  if (ROT) {
    R drive[SPT];
#pragma unroll
    for (int j = 0; j < SPT; j++) drive[j] = e[j] * e[j];
    rot_chi<R, SPT>(drive, ngp, dchi, rot_tab, rot_states, rot_fc, dt, N, a0);
  }

  // 3. Kerr chi = ng chi3/nref E^2 (uppe.cl:132-151), P = chi E, J_p = ne E, packed as P + i J_p
  const T kc = (T)(ngp * redchi3);
  T yrun[2 * SPT];
#pragma unroll
  for (int j = 0; j < SPT; j++) {
    int n = a0 + j;
    T chi = (T)dchi[j] + kc * e[j] * e[j];
    T nej = (T)ne[j];
    yrun[2 * j] = chi * e[j];
    yrun[2 * j + 1] = nej * e[j];
    if (DIAG) {
      chi_out[p * N + n] = chi;
      ne_out[p * N + n] = nej;
    }
  }
  lds_write_reals<T, 2 * SPT>(Y, 2 * a0, yrun);
  __syncthreads();
  lds_fft_r8<T, LGC, NTH, (SPT >= 8 ? 4 : 2)>(Y, N, lgn, tws, -1, tid);

  // 4. J = F_NL [ rfft(J_p)/(0.1 - i w) + (-i w) rfft(P) ] (fft.cl:282-329, uppe_kernel.py:170-186)
#pragma unroll
  for (int i = 0; i < KPT; i++) {
    const int k = tid + i * nth;
    if (k > M) break;
    cplx<T> yk = Y[lds_swz<T>(k)], ym = Y[lds_swz<T>((N - k) & (N - 1))];
    cplx<T> pk = cscale(cadd(yk, cconj(ym)), T(0.5));
    cplx<T> d = csub(yk, cconj(ym));
    cplx<T> jp = cmk(T(0.5) * d.im, T(-0.5) * d.re);
    T wk = wv[i];
    cplx<T> out;
    if (ion.kind != UPPE_ION_NONE) {
      // J_p / (0.1 - i w): FP32 multiplies by the hardware reciprocal of |0.1 - i w|^2 (1 ulp; the denominator
      // is at least 0.01), FP64 divides
      T cr = T(0.1), ci = -wk;
      T den = cr * cr + ci * ci;
      if constexpr (sizeof(T) == 4) {
        const float rd = __builtin_amdgcn_rcpf(den);
        out = cmk((jp.re * cr + jp.im * ci) * rd, (jp.im * cr - jp.re * ci) * rd);
      } else {
        out = cmk((jp.re * cr + jp.im * ci) / den, (jp.im * cr - jp.re * ci) / den);
      }
      out = cadd(out, cmk(wk * pk.im, -wk * pk.re));
    } else {
      out = cmk(wk * pk.im, -wk * pk.re);
    }
    jr[p * nwp + k] = cscale(out, fv[i]);
  }
}

// ---------------------------------------------------------------------------------------------------
// Wave-per-strip source (FP32 stepping, Nt = 2^LG, 256 <= Nt <= 1024, no rotator, no diagnostics)
//
// One wavefront owns one strip and a private LDS buffer; a workgroup holds WPB such waves and they share
// nothing, so no workgroup barrier is ever taken. The strip transforms run the Stockham passes of lds_fft_r8
// with each lane taking its share of the butterflies, ordered by wavefront-scope fences (the LDS processes
// one wave's instructions in order). The twiddles of both transforms are loaded into registers up front, with
// the strip, so that no pass waits on a global load. Same arithmetic as f_source up to the FFT summation
// order; the depletion scan is one DPP wave scan.
// ---------------------------------------------------------------------------------------------------
// one strip of the source on one wave: E_r spectrum X (Nw = Nt/2 + 1 bins) to the current spectrum J
// (uppe_kernel.py:139-186), as f_source
template <int LG>
__device__ inline void source_strip_wave(const cplx<float>* __restrict__ X, cplx<float>* __restrict__ J,
                                         cplx<float>* Y, const cplx<float>* __restrict__ tw,
                                         const float* __restrict__ w, const float* __restrict__ nlf, float ngp,
                                         double redchi3, double dt, const uppe_ionization& ion, double log_pre,
                                         int lane) {
  constexpr int N = 1 << LG, M = N / 2, SPT = N / 64, KPL = M / 64;
  // every global load up front: the strip, the frequencies and filter of the bins this lane writes, and the
  // twiddles of both transforms
  cplx<float> xk[KPL], xm[KPL];
  float wv[KPL + 1], fv[KPL + 1];
#pragma unroll
  for (int i = 0; i < KPL; i++) {
    const int k = lane + 64 * i;
    xk[i] = X[k];
    xm[i] = X[M - k];
  }
#pragma unroll
  for (int i = 0; i <= KPL; i++) {
    const int k = min(lane + 64 * i, M);
    wv[i] = w[k];
    fv[i] = nlf[k];
  }
  WaveTw<LG - 1> twh;
  WaveTw<LG> twf;
  wave_tw_load<LG - 1>(twh, tw + r8_table_len(LG), lane);
  wave_tw_load<LG>(twf, tw, lane);

  // 1. irfft: the half-length complex transform with the standard split (fft.cl:271-278 semantics)
#pragma unroll
  for (int i = 0; i < KPL; i++) {
    const int k = lane + 64 * i;
    cplx<float> xa = xk[i], xb = xm[i];
    if (k == 0) {
      xa.im = 0.0f;
      xb.im = 0.0f;
    }
    const cplx<float> ze = cscale(cadd(xa, cconj(xb)), 0.5f);
    const cplx<float> zo = cmul(cscale(csub(xa, cconj(xb)), 0.5f), twiddle_inv<float>(k, N));
    Y[lds_swz<float>(k)] = cmk(ze.re - zo.im, ze.im + zo.re);
  }
  wave_sync();
  wave_fft_r8<LG - 1>(Y, twh, +1, lane);
  // lane l owns samples [a0, a0 + SPT), a0 = (63 - l) SPT: the depletion integral runs from the last sample
  // backward, a prefix sum in lane order
  const int a0 = (63 - lane) * SPT;
  float e[SPT];
  lds_read_reals<float, SPT>(Y, a0, e);
  const float invm = 1.0f / (float)M;
#pragma unroll
  for (int j = 0; j < SPT; j++) e[j] *= invm;
  wave_sync();

  // 2. rate, backward trapezoid with depletion (ionization.cl:127-155), soft window (fft.cl:331-355)
  float ne[SPT];
  if (ion.kind != UPPE_ION_NONE) {
    const float dtr = (float)dt;
    float r[SPT];
    float tot = 0.0f;
#pragma unroll
    for (int j = 0; j < SPT; j++) {
      const int n = a0 + j;
      const float v = ion_rate_f32(fabsf(e[j]), ion, log_pre);
      r[j] = (n >= N - 1) ? 0.0f : v;  // last sample zeroed (ionization.cl:147)
      tot += r[j];
    }
    float later = wave_excl_prefix(tot);
#pragma unroll
    for (int j = SPT - 1; j >= 0; j--) {
      const int n = a0 + j;
      const float integral = (later + 0.5f * r[j]) * dtr;
      later += r[j];
      float x = ngp * one_minus_exp_neg(integral);
      if (n < 4 || n > N - 5) x = 0.0f;
      if (n >= 4 && n < 68) {
        const float qq = (float)(n - 4) / 64.0f;
        x *= 10.0f * qq * qq * qq - 15.0f * qq * qq * qq * qq + 6.0f * qq * qq * qq * qq * qq;
      }
      if (n <= N - 5 && n > N - 69) {
        const float qq = (float)(N - 5 - n) / 64.0f;
        x *= 10.0f * qq * qq * qq - 15.0f * qq * qq * qq * qq + 6.0f * qq * qq * qq * qq * qq;
      }
      ne[j] = x;
    }
  } else {
#pragma unroll
    for (int j = 0; j < SPT; j++) ne[j] = 0.0f;
  }

  // 3. Kerr chi = ng chi3/nref E^2 (uppe.cl:132-151), P = chi E, J_p = ne E, packed as P + i J_p
  const float kc = (float)((double)ngp * redchi3);
  float yrun[2 * SPT];
#pragma unroll
  for (int j = 0; j < SPT; j++) {
    const float chi = kc * e[j] * e[j];
    yrun[2 * j] = chi * e[j];
    yrun[2 * j + 1] = ne[j] * e[j];
  }
  lds_write_reals<float, 2 * SPT>(Y, 2 * a0, yrun);
  wave_sync();
  wave_fft_r8<LG>(Y, twf, -1, lane);

  // 4. J = F_NL [ rfft(J_p)/(0.1 - i w) + (-i w) rfft(P) ] (fft.cl:282-329, uppe_kernel.py:170-186)
#pragma unroll
  for (int i = 0; i <= KPL; i++) {
    const int k = lane + 64 * i;
    if (k > M) continue;
    const cplx<float> yk = Y[lds_swz<float>(k)], ym = Y[lds_swz<float>((N - k) & (N - 1))];
    const cplx<float> pk = cscale(cadd(yk, cconj(ym)), 0.5f);
    const cplx<float> d = csub(yk, cconj(ym));
    const cplx<float> jp = cmk(0.5f * d.im, -0.5f * d.re);
    const float wk = wv[i];
    cplx<float> out;
    if (ion.kind != UPPE_ION_NONE) {
      const float cr = 0.1f, ci = -wk;
      const float rd = __builtin_amdgcn_rcpf(cr * cr + ci * ci);
      out = cmk((jp.re * cr + jp.im * ci) * rd, (jp.im * cr - jp.re * ci) * rd);
      out = cadd(out, cmk(wk * pk.im, -wk * pk.re));
    } else {
      out = cmk(wk * pk.im, -wk * pk.re);
    }
    J[k] = cscale(out, fv[i]);
  }
}

template <int LG, int WPB>
__global__ __launch_bounds__(64 * WPB) void f_source_w(const cplx<float>* __restrict__ er, cplx<float>* __restrict__ jr,
                                                       const cplx<float>* __restrict__ tw, const float* __restrict__ w,
                                                       const float* __restrict__ nlf, const double* __restrict__ ng,
                                                       int nstrips, int nwp, double redchi3, double dt,
                                                       uppe_ionization ion, double log_pre,
                                                       const SrcInst* __restrict__ inst, long zs, StepState* commit,
                                                       unsigned long long* cq, double* rec, double* zrec,
                                                       int zrec_cap) {
  // the stage-1 launch commits the previous step (nothing else of this launch reads the step state); a batched
  // launch takes instance blockIdx.y, whose strips lie zs elements apart, with its own medium scalars
  if (commit && blockIdx.x == 0 && threadIdx.x == 0) {
    const int z = inst ? blockIdx.y : 0;
    step_commit(commit + z, cq + (long)kQW * z, inst ? nullptr : rec, inst ? nullptr : zrec, zrec_cap);
  }
  __shared__ cplx<float> ybuf[WPB][1 << LG];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long p = (long)blockIdx.x * WPB + wv;
  if (p >= nstrips) return;  // no workgroup barrier follows: an idle wave may leave
  if (inst) {
    const int z = blockIdx.y;
    er += z * zs;
    jr += z * zs;
    ng += (long)z * nstrips;
    redchi3 = inst[z].redchi3;
    log_pre = inst[z].log_pre;
    ion = inst[z].ion;
  }
  source_strip_wave<LG>(er + p * nwp, jr + p * nwp, ybuf[wv], tw, w, nlf, (float)ng[p], redchi3, dt, ion, log_pre,
                        lane);
}

// Paraxial envelope source (paraxial_kernel.py:140-184): one workgroup per strip of N = Nw complex samples,
// max(min(N, 64), N/4) threads, thread t owning the SPT frequencies t + i nth and the SPT time samples
// a0 .. a0 + SPT - 1, a0 = (nth - 1 - t) SPT (the density integral runs from the last sample backward). With
// E(w) = i w A(w) (DtSpectral, sgn -1) and A(w) inverse-transformed side by side in LDS, the rate of |E|, the
// depletion scan, the rotator driven by |E|^2 / 2 and the Kerr term 0.75 n_g chi3 |E|^2 give P = chi E and
// Q = n_e (|A|^2/4 - 1) A; the reference's J = FFT(IFFT(-i w FFT(P)) + Q) is formed as -i w FFT(P) + FFT(Q).
template <typename T, int SPT, bool DIAG, bool ROT, int LGC = 0>
__global__ __launch_bounds__(1024) void f_source_env(const cplx<T>* __restrict__ ar, cplx<T>* __restrict__ jr,
                                                     T* __restrict__ chi_out, T* __restrict__ ne_out,
                                                     const cplx<T>* __restrict__ tw, const T* __restrict__ w,
                                                     const double* __restrict__ ng, int nt, int lgn, int nwp,
                                                     double redchi3, double dt, uppe_ionization ion, double log_pre,
                                                     const RotRow<T>* __restrict__ rot_tab, int rot_states,
                                                     double rot_fc, StepState* commit, unsigned long long* cq,
                                                     double* rec, double* zrec, int zrec_cap) {
  // the stage-1 launch commits the previous step (commit non-null; nothing else of this launch reads it)
  if (commit && blockIdx.x == 0 && threadIdx.x == 0) step_commit(commit, cq, rec, zrec, zrec_cap);
  extern __shared__ __align__(16) char smem[];
  if (LGC > 0) lgn = LGC;
  const int N = LGC > 0 ? (1 << LGC) : nt;
  cplx<T>* Y1 = reinterpret_cast<cplx<T>*>(smem);  // E, then P
  cplx<T>* Y2 = Y1 + N;                             // A, then Q
  __shared__ T wsum[16];
  const long p = blockIdx.x;
  constexpr int NTH = LGC > 0 ? (1 << LGC) / SPT : 0;
  const int tid = threadIdx.x, nth = NTH > 0 ? NTH : (int)blockDim.x, lane = tid & 63, wave = tid >> 6,
            nwv = nth >> 6;
  const cplx<T>* X = ar + p * nwp;
  cplx<T> xa[SPT];
  T wv[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int k = tid + i * nth;
    xa[i] = X[k];
    wv[i] = w[k];
  }
  // the N-point pass tables: in LDS for FP32, read from L2 for FP64
  const cplx<T>* tws = tw;
  if constexpr (sizeof(T) == 4) {
    cplx<T>* tl = Y2 + N;
    const int tlen = r8_table_len(lgn);
    for (int k = tid; k < tlen; k += nth) tl[k] = tw[k];
    tws = tl;
  }
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int k = tid + i * nth;
    Y1[lds_swz<T>(k)] = cmk(-wv[i] * xa[i].im, wv[i] * xa[i].re);
    Y2[lds_swz<T>(k)] = xa[i];
  }
  __syncthreads();
  lds_fft_r8_pair<T, LGC, NTH>(Y1, Y2, N, lgn, tws, +1, tid);
  const T invn = T(1) / T(N);
  const int a0 = (nth - 1 - tid) * SPT;
  cplx<T> e[SPT];
#pragma unroll
  for (int j = 0; j < SPT; j++) e[j] = cscale(Y1[lds_swz<T>(a0 + j)], invn);

  // rate of the envelope modulus and the backward trapezoid with depletion, the last rate sample zeroed
  // (ionization.cl:46-69, 99-155): I_n = dt (r_n / 2 + sum_{m > n} r_m), no time window in this model
  using R = T;
  const R ngp = (R)ng[p];
  R ne[SPT];
  if (ion.kind != UPPE_ION_NONE) {
    const R dtr = (R)dt;
    R r[SPT];
    R tot = R(0);
#pragma unroll
    for (int j = 0; j < SPT; j++) {
      const int n = a0 + j;
      const double em = sqrt((double)e[j].re * e[j].re + (double)e[j].im * e[j].im);
      const R v = (R)ion_rate<T>(em, ion, log_pre);
      r[j] = (n >= N - 1) ? R(0) : v;
      tot += r[j];
    }
    const R ex = wave_excl_prefix(tot);
    if (lane == 63) wsum[wave] = ex + tot;
    __syncthreads();
    R later = ex;
    later += lanes_below(lane < nwv ? wsum[lane] : R(0), __builtin_amdgcn_readfirstlane(wave));  // earlier waves
#pragma unroll
    for (int j = SPT - 1; j >= 0; j--) {
      const R integral = (later + R(0.5) * r[j]) * dtr;
      later += r[j];
      if constexpr (sizeof(T) == 4)
        ne[j] = ngp * one_minus_exp_neg(integral);
      else
        ne[j] = ngp * (1.0 - exp(-integral));
    }
  } else {
#pragma unroll
    for (int j = 0; j < SPT; j++) ne[j] = R(0);
  }

  // rotational susceptibility driven by the cycle average |E|^2 / 2 (rotations.cl:26-78, AddChiParaxial)
  R dchi[SPT];
#pragma unroll
  for (int j = 0; j < SPT; j++) dchi[j] = R(0);
  if (ROT) {
    R drive[SPT];
#pragma unroll
    for (int j = 0; j < SPT; j++) drive[j] = R(0.5) * (e[j].re * e[j].re + e[j].im * e[j].im);
    rot_chi<R, SPT>(drive, ngp, dchi, rot_tab, rot_states, rot_fc, dt, N, a0);
  }

  // chi = rotator + 0.75 n_g chi3 |E|^2 (paraxial.cl:112-131), P = chi E, Q = n_e (|A|^2/4 - 1) A (:99-110);
  // A(t) is read from LDS only now, which keeps it out of registers through the scans
  const T kc = (T)(0.75 * (double)ngp * redchi3);
  cplx<T> a[SPT];
#pragma unroll
  for (int j = 0; j < SPT; j++) a[j] = cscale(Y2[lds_swz<T>(a0 + j)], invn);
  __syncthreads();  // every thread has read its samples of E and A
#pragma unroll
  for (int j = 0; j < SPT; j++) {
    const int n = a0 + j;
    const T chi = (T)dchi[j] + kc * (e[j].re * e[j].re + e[j].im * e[j].im);
    const T fq = (T)ne[j] * (T(0.25) * (a[j].re * a[j].re + a[j].im * a[j].im) - T(1));
    Y1[lds_swz<T>(n)] = cscale(e[j], chi);
    Y2[lds_swz<T>(n)] = cscale(a[j], fq);
    if (DIAG) {
      chi_out[p * N + n] = chi;
      ne_out[p * N + n] = (T)ne[j];
    }
  }
  __syncthreads();
  lds_fft_r8_pair<T, LGC, NTH>(Y1, Y2, N, lgn, tws, -1, tid);
  // J = -i w FFT(P) + FFT(Q) (DtSpectral, sgn +1)
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const int k = tid + i * nth;
    const cplx<T> pk = Y1[lds_swz<T>(k)], qk = Y2[lds_swz<T>(k)];
    jr[p * nwp + k] = cmk(wv[i] * pk.im + qk.re, -wv[i] * pk.re + qk.im);
  }
}

// Adjoint of the Kerr source on one strip: the vector-Jacobian product of J = F_NL (-i w) rfft(P), P = n_g chi3 e^3,
// e = irfft(E) (uppe.cl:132-165, fft.cl:229-299). Given the cotangent G of J (G = dL/dRe J + i dL/dIm J), with
// R = (i w F_NL) G:
//   g_P(t) = Re sum_{k=0}^{Nt/2} R_k e^{2 pi i k t / Nt}          (adjoint of rfft)
//   g_e(t) = 3 n_g chi3 e(t)^2 g_P(t)
//   G_E(k) = c_k rfft(g_e)_k / Nt, c = 2 inside, 1 at DC and Nyquist, whose imaginary parts carry no cotangent
//                                                                  (adjoint of irfft)
// e and g_P come from one unnormalised inverse transform of Z = E^ext + i Y^ext (Hermitian extensions, Y = R with
// its inner bins halved), g_e from one forward transform. One workgroup per strip, max(min(Nt, 64), Nt/4)
// threads.
template <typename T, int SPT, int LGC = 0>
__global__ __launch_bounds__(1024) void f_source_adj(const cplx<T>* __restrict__ er, const cplx<T>* __restrict__ gj,
                                                     cplx<T>* __restrict__ ge, const cplx<T>* __restrict__ tw,
                                                     const T* __restrict__ w, const T* __restrict__ nlf,
                                                     const double* __restrict__ ng, int nt, int lgn, int nwp,
                                                     double redchi3) {
  extern __shared__ __align__(16) char smem[];
  cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem);
  if (LGC > 0) lgn = LGC;
  const int N = LGC > 0 ? (1 << LGC) : nt, M = N / 2;
  const long p = blockIdx.x;
  constexpr int NTH = LGC > 0 ? (1 << LGC) / SPT : 0;
  const int tid = threadIdx.x, nth = NTH > 0 ? NTH : (int)blockDim.x;
  const cplx<T>* X = er + p * nwp;
  const cplx<T>* G = gj + p * nwp;
  const cplx<T>* tws = tw;
  if constexpr (sizeof(T) == 4) {
    cplx<T>* tl = Z + N;
    const int tlen = r8_table_len(lgn);
    for (int k = tid; k < tlen; k += nth) tl[k] = tw[k];
    tws = tl;
  }
  for (int k = tid; k < N; k += nth) {
    const int kk = k <= M ? k : N - k;
    cplx<T> x = X[kk];
    const cplx<T> g = G[kk];
    const T wf = w[kk] * nlf[kk];
    cplx<T> y = cmk(-wf * g.im, wf * g.re);  // i w F_NL G
    if (kk == 0 || kk == M) {
      x.im = T(0);
      y.im = T(0);
    } else {
      y = cscale(y, T(0.5));
    }
    if (k > M) {
      x = cconj(x);
      y = cconj(y);
    }
    Z[lds_swz<T>(k)] = cmk(x.re - y.im, x.im + y.re);
  }
  __syncthreads();
  lds_fft_r8<T, LGC, NTH>(Z, N, lgn, tws, +1, tid);
  const T kc3 = (T)(3.0 * ng[p] * redchi3), invn = T(1) / T(N);
  T gev[SPT];
#pragma unroll
  for (int i = 0; i < SPT; i++) {
    const cplx<T> z = Z[lds_swz<T>(tid + i * nth)];
    const T e = z.re * invn;
    gev[i] = kc3 * e * e * z.im;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < SPT; i++) Z[lds_swz<T>(tid + i * nth)] = cmk(gev[i], T(0));
  __syncthreads();
  lds_fft_r8<T, LGC, NTH>(Z, N, lgn, tws, -1, tid);
  for (int k = tid; k <= M; k += nth) {
    const cplx<T> f = Z[lds_swz<T>(k)];
    ge[p * nwp + k] = (k == 0 || k == M) ? cmk(f.re * invn, T(0)) : cscale(f, T(2) * invn);
  }
}

// d rate / d|e| of ion_rate at field modulus ea (sim units) with rate its value there: the explicit form
// rate_conv C_pre E^C_pow exp(C_exp / E) or the series rate_conv exp(sum_n c_n log(E)^n), E = e_conv ea + e_cutoff
// (ionization.cl:22-44, 71-97)
__device__ inline double ion_rate_slope(double ea, const uppe_ionization& ion, double rate) {
  const double emag = ion.e_conv * ea + ion.e_cutoff;
  if (ion.kind == UPPE_ION_EXPLICIT) return rate * (ion.c_pow / emag - ion.c_exp / (emag * emag)) * ion.e_conv;
  const double l = log(emag);
  const double dp = ion.c[1] + l * (2.0 * ion.c[2] + l * (3.0 * ion.c[3] + l * (4.0 * ion.c[4] + l * 5.0 * ion.c[5])));
  return rate * dp / emag * ion.e_conv;
}

// FP32 rate and d rate / d|e| at field modulus ea (sim units): the explicit form in FP32 with one hardware
// reciprocal, w = C E^p exp(c/E), dw/dE = w (p - c/E)/E times the field's unit factor; the series form through
// ion_rate<float> and ion_rate_slope
__device__ inline void ion_rate_and_slope_f32(float ea, const uppe_ionization& ion, double log_pre, float& rate,
                                              float& slope) {
  if (ion.kind == UPPE_ION_EXPLICIT) {
    const float emag = fmaf((float)ion.e_conv, ea, (float)ion.e_cutoff);
    const float re = __builtin_amdgcn_rcpf(emag);
    const float cp = (float)ion.c_pow, ce = (float)ion.c_exp;
    rate = __builtin_amdgcn_exp2f((float)log_pre + cp * __builtin_amdgcn_logf(emag) +
                                  (float)(ion.c_exp * 1.4426950408889634) * re);
    slope = rate * (cp - ce * re) * re * (float)ion.e_conv;
    return;
  }
  const double r = ion_rate<float>((double)ea, ion, log_pre);
  rate = (float)r;
  slope = (float)ion_rate_slope((double)ea, ion, r);
}

// the soft time window applied to the electron density, SoftTimeWindow(4, 64) (fft.cl:331-355)
template <typename T>
__device__ inline T soft_window(int n, int N) {
  T x = (n < 4 || n > N - 5) ? T(0) : T(1);
  if (n >= 4 && n < 68) {
    const T q = (T)(n - 4) / T(64);
    x *= T(10) * q * q * q - T(15) * q * q * q * q + T(6) * q * q * q * q * q;
  }
  if (n <= N - 5 && n > N - 69) {
    const T q = (T)(N - 5 - n) / T(64);
    x *= T(10) * q * q * q - T(15) * q * q * q * q + T(6) * q * q * q * q * q;
  }
  return x;
}

// exclusive prefix sum of v over the threads of the block in thread order, and the block total; ws holds 16
// values of LDS scratch
template <typename T>
__device__ inline T block_excl_prefix(T v, T* ws, T* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nwv = (blockDim.x + 63) >> 6;
  const T ex = wave_excl_prefix(v);
  if (lane == min(63, (int)blockDim.x - 1 - wave * 64)) ws[wave] = ex + v;
  __syncthreads();
  T before = T(0), tot = T(0);
  for (int w2 = 0; w2 < nwv; w2++) {
    const T t = ws[w2];
    if (w2 < wave) before += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return before + ex;
}

// Adjoint of the full source on one strip (Kerr, photoionisation rate, depletion integral, soft window, plasma
// current): the vector-Jacobian product of
//   J = F_NL [ (-i w) rfft(P) + rfft(J_p) / (0.1 - i w) ],   P = n_g chi3 e^3,   J_p = n_e e,
//   n_e(n) = W(n) n_g (1 - exp(-I_n)),   I_n = dt (r_n / 2 + sum_{m > n} r_m),   r_n = rate(|e_n|), r_{N-1} = 0,
// e = irfft(E) (uppe.cl:132-165, ionization.cl:22-44,71-97,127-155, fft.cl:282-355, uppe_kernel.py:139-186).
// Given the cotangent G of J, with R = (i w F_NL) G and R' = F_NL G / (0.1 + i w):
//   g_P = Re sum_k R_k e^{2 pi i k t/N},   g_Jp = Re sum_k R'_k e^{2 pi i k t/N}          (adjoints of rfft)
//   g_I(n) = e_n g_Jp(n) W(n) n_g exp(-I_n)
//   g_r(m) = dt (g_I(m) / 2 + sum_{n < m} g_I(n)) for m < N - 1, 0 at N - 1             (transposed scan)
//   g_e = 3 n_g chi3 e^2 g_P + n_e g_Jp + g_r rate'(|e|) sgn(e)
//   G_E(k) = c_k rfft(g_e)_k / N                                                         (adjoint of irfft)
// e and g_P come from one inverse transform of E^ext + i R^ext, g_Jp from a second of R'^ext (Hermitian extensions,
// inner bins halved). Thread t owns samples t SPT .. t SPT + SPT - 1, so the forward suffix sum and the transposed
// prefix sum are block scans in thread order. One workgroup per strip, max(min(Nt, 64), Nt/4) threads.
template <typename T, int SPT, int LGC = 0, bool ROT = false>
__global__ __launch_bounds__(1024) void f_source_adj_ion(const cplx<T>* __restrict__ er, const cplx<T>* __restrict__ gj,
                                                         cplx<T>* __restrict__ ge, const cplx<T>* __restrict__ tw,
                                                         const T* __restrict__ w, const T* __restrict__ nlf,
                                                         const double* __restrict__ ng, int nt, int lgn, int nwp,
                                                         double redchi3, double dt, uppe_ionization ion,
                                                         double log_pre, const RotRow<T>* __restrict__ rot_tab,
                                                         int rot_states, double rot_fc) {
  extern __shared__ __align__(16) char smem[];
  cplx<T>* Z = reinterpret_cast<cplx<T>*>(smem);
  __shared__ T ws[16];
  if (LGC > 0) lgn = LGC;
  const int N = LGC > 0 ? (1 << LGC) : nt, M = N / 2;
  cplx<T>* Z2 = Z + N;
  const long p = blockIdx.x;
  constexpr int NTH = LGC > 0 ? (1 << LGC) / SPT : 0;
  const int tid = threadIdx.x, nth = NTH > 0 ? NTH : (int)blockDim.x;
  const cplx<T>* X = er + p * nwp;
  const cplx<T>* G = gj + p * nwp;
  const cplx<T>* tws = tw;
  if constexpr (sizeof(T) == 4) {
    cplx<T>* tl = Z2 + N;
    const int tlen = r8_table_len(lgn);
    for (int k = tid; k < tlen; k += nth) tl[k] = tw[k];
    tws = tl;
  }
  for (int k = tid; k < N; k += nth) {
    const int kk = k <= M ? k : N - k;
    cplx<T> x = X[kk];
    const cplx<T> g = G[kk];
    const T f = nlf[kk], wk = w[kk];
    const T wf = wk * f;
    cplx<T> y = cmk(-wf * g.im, wf * g.re);  // i w F_NL G
    T d;                                      // F_NL G (0.1 - i w) / (0.01 + w^2)
    if constexpr (sizeof(T) == 4)
      d = f * __builtin_amdgcn_rcpf(0.01f + wk * wk);
    else
      d = f / (0.01 + wk * wk);
    cplx<T> y2 = cmk(d * (T(0.1) * g.re + wk * g.im), d * (T(0.1) * g.im - wk * g.re));
    if (kk == 0 || kk == M) {
      x.im = T(0);
      y.im = T(0);
      y2.im = T(0);
    } else {
      y = cscale(y, T(0.5));
      y2 = cscale(y2, T(0.5));
    }
    if (k > M) {
      x = cconj(x);
      y = cconj(y);
      y2 = cconj(y2);
    }
    Z[lds_swz<T>(k)] = cmk(x.re - y.im, x.im + y.re);
    Z2[lds_swz<T>(k)] = y2;
  }
  __syncthreads();
  lds_fft_r8_pair<T, LGC, NTH>(Z, Z2, N, lgn, tws, +1, tid);
  const double ngp = ng[p];
  const T kc3 = (T)(3.0 * ngp * redchi3), invn = T(1) / T(N), ngt = (T)ngp, dtt = (T)dt;
  const int a0 = tid * SPT;
  T e[SPT], gp[SPT], gjp[SPT], r[SPT], dr[SPT];
  T rtot = T(0);
#pragma unroll
  for (int j = 0; j < SPT; j++) {
    const int n = a0 + j;
    const cplx<T> z = Z[lds_swz<T>(n)];
    e[j] = z.re * invn;
    gp[j] = z.im;
    gjp[j] = Z2[lds_swz<T>(n)].re;
    T rate = T(0), slope = T(0);  // no rate model: a rotator's medium without photoionisation
    if (ion.kind != UPPE_ION_NONE) {
      if constexpr (sizeof(T) == 4) {
        ion_rate_and_slope_f32(fabsf(e[j]), ion, log_pre, rate, slope);
      } else {
        const double ea = fabs(e[j]);
        rate = ion_rate<T>(ea, ion, log_pre);
        slope = ion_rate_slope(ea, ion, rate);
      }
    }
    r[j] = n >= N - 1 ? T(0) : rate;
    dr[j] = n >= N - 1 ? T(0) : slope * (e[j] > T(0) ? T(1) : (e[j] < T(0) ? T(-1) : T(0)));
    rtot += r[j];
  }
  // forward depletion integral: samples after this thread's own hold the rest of the suffix sum
  T rall;
  const T rbefore = block_excl_prefix(rtot, ws, &rall);
  T later = rall - rbefore - rtot;
  T ne[SPT], gI[SPT];
  T gtot = T(0);
#pragma unroll
  for (int j = SPT - 1; j >= 0; j--) {
    const int n = a0 + j;
    const T I = (later + T(0.5) * r[j]) * dtt;
    later += r[j];
    const T win = soft_window<T>(n, N);
    T ex, dep;
    if constexpr (sizeof(T) == 4) {
      ex = __builtin_amdgcn_exp2f(-I * 1.4426950408889634f);
      dep = one_minus_exp_neg(I);
    } else {
      ex = exp(-I);
      dep = 1.0 - exp(-I);
    }
    ne[j] = win * ngt * dep;
    gI[j] = e[j] * gjp[j] * win * ngt * ex;
    gtot += gI[j];
  }
  // transposed scan: g_r(m) = dt (g_I(m)/2 + sum of g_I over earlier samples)
  T gall;
  T earlier = block_excl_prefix(gtot, ws, &gall);
  T gev[SPT];
#pragma unroll
  for (int j = 0; j < SPT; j++) {
    const T gr = (earlier + T(0.5) * gI[j]) * dtt;
    earlier += gI[j];
    gev[j] = kc3 * e[j] * e[j] * gp[j] + ne[j] * gjp[j] + gr * dr[j];
  }
  if constexpr (ROT) {
    // chi_rot of the forward, by rot_chi in its own thread order (thread t takes the samples of thread nth-1-t
    // here, read from the strip), handed over through the free second strip; then the cotangent of chi_rot,
    // gam = e g_P (P = (chi_rot + kc e^2) e), back to the drive e^2: g_e += chi_rot g_P + 2 e d L / d drive
    const int a0m = (nth - 1 - tid) * SPT;
    T drive[SPT], dchi[SPT];
#pragma unroll
    for (int j = 0; j < SPT; j++) {
      const T em = Z[lds_swz<T>(a0m + j)].re * invn;
      drive[j] = em * em;
      dchi[j] = T(0);
    }
    rot_chi<T, SPT>(drive, (T)ngp, dchi, rot_tab, rot_states, rot_fc, dt, N, a0m);
    __syncthreads();  // every thread has read its R' samples from the second strip
#pragma unroll
    for (int j = 0; j < SPT; j++) Z2[lds_swz<T>(a0m + j)] = cmk(dchi[j], T(0));
    __syncthreads();
    T gam[SPT], gd[SPT];
#pragma unroll
    for (int j = 0; j < SPT; j++) {
      const T chr = Z2[lds_swz<T>(a0 + j)].re;
      gev[j] += chr * gp[j];
      gam[j] = e[j] * gp[j];
      gd[j] = T(0);
    }
    rot_chi_adj<T, SPT>(gam, (T)ngp, gd, rot_tab, rot_states, rot_fc, dt, N, a0);
#pragma unroll
    for (int j = 0; j < SPT; j++) gev[j] += T(2) * e[j] * gd[j];
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < SPT; j++) Z[lds_swz<T>(a0 + j)] = cmk(gev[j], T(0));
  __syncthreads();
  lds_fft_r8<T, LGC, NTH>(Z, N, lgn, tws, -1, tid);
  for (int k = tid; k <= M; k += nth) {
    const cplx<T> f = Z[lds_swz<T>(k)];
    ge[p * nwp + k] = (k == 0 || k == M) ? cmk(f.re * invn, T(0)) : cscale(f, T(2) * invn);
  }
}

// ---------------------------------------------------------------------------------------------------
// control, layout and reduction kernels
// ---------------------------------------------------------------------------------------------------

// commit the pending step of a chunk's last step (one block per instance in a batched launch, which records
// nothing)
__global__ void f_ctl_commit(StepState* st, unsigned long long* qmax, double* rec, double* zrec, int zrec_cap) {
  step_commit(st + blockIdx.x, qmax + (long)kQW * blockIdx.x, rec, zrec, zrec_cap);
}

// zero n 8-byte words (a fill kernel, so that a stream the caller captures into a graph holds no memset node)
__global__ void f_fill_zero(unsigned long long* p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0ull;
}

// max |q| over the state (set_field)
template <typename T>
__global__ void f_qmax(const cplx<T>* q, long n, unsigned long long* out) {
  double m = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    cplx<T> v = q[i];
    m = fmax(m, sqrt((double)v.re * v.re + (double)v.im * v.im));
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, dbits(m));
}

// max |q| over complex columns [f_from, nwp) of every row
template <typename T>
__global__ void f_qmax_cols(const cplx<T>* q, long P, int nwp, int f_from, unsigned long long* out) {
  const int w = nwp - f_from;
  double m = 0.0;
  if (w > 0)
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < P * w; i += (long)gridDim.x * blockDim.x) {
      cplx<T> v = q[(i / w) * nwp + f_from + (i % w)];
      m = fmax(m, sqrt((double)v.re * v.re + (double)v.im * v.im));
    }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, dbits(m));
}

// host layout (nw, P) complex128 -> [P][nwp] complex T (zero padding)
template <typename T>
__global__ void f_from_host(const double* src, cplx<T>* dst, int nw, long P, int nwp) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= P * nwp) return;
  long p = i / nwp;
  int f = (int)(i % nwp);
  if (f < nw)
    dst[i] = cmk((T)src[2 * (f * P + p)], (T)src[2 * (f * P + p) + 1]);
  else
    dst[i] = cmk(T(0), T(0));
}

// (kzc non-null: a state kept relative to the reference point qb in its columns f < f1, returned as the
// interaction-picture state e^{Im kz qb} q in FP64)
template <typename T>
__global__ void f_to_host(const cplx<T>* src, double* dst, int nw, long P, int nwp, const cplx<T>* kzc,
                          int f1, double qb) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= P * nw) return;
  long f = i / P, p = i % P;
  cplx<T> v = src[p * nwp + f];
  const double g = (kzc && f < f1) ? exp((double)kzc[p * nwp + f].im * qb) : 1.0;
  dst[2 * i] = (double)v.re * g;
  dst[2 * i + 1] = (double)v.im * g;
}

// cos(2 pi e / 16), e < 16: the twiddles of the azimuthal transforms (N2 <= 16 points; e scaled by 16 / N2)
__constant__ double kCos16[16] = {1.0,
                                  0.92387953251128675613,
                                  0.70710678118654752440,
                                  0.38268343236508977173,
                                  0.0,
                                  -0.38268343236508977173,
                                  -0.70710678118654752440,
                                  -0.92387953251128675613,
                                  -1.0,
                                  -0.92387953251128675613,
                                  -0.70710678118654752440,
                                  -0.38268343236508977173,
                                  0.0,
                                  0.38268343236508977173,
                                  0.70710678118654752440,
                                  0.92387953251128675613};

// The azimuthal transform of a radial grid with N2 > 1 points in phi, in place on a [Nr N2][nwp] array: for every
// (r, f < ncols), the N2-point DFT over the rows r N2 + m with exponent sign `sign`, scaled by sc: numpy's fft
// (sign -1, sc 1) and ifft (sign +1, sc 1/N2), as the reference's FFT_axis2 / IFFT_axis2 (fft.cl:401-441) take
// before and after the per-mode radial transforms (grid_tools.py:308-325), and their adjoints for the reverse
// sweep. A direct N2-point DFT per thread.
template <typename T, int N2>
__global__ void f_azim(cplx<T>* a, int nr, int nwp, int ncols, int sign, T sc) {
  const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= (long)nr * ncols) return;
  const int r = (int)(i / ncols), f = (int)(i % ncols);
  cplx<T>* p = a + (long)r * N2 * nwp + f;
  cplx<T> v[N2];
#pragma unroll
  for (int m = 0; m < N2; m++) v[m] = p[(long)m * nwp];
  const T sg = (T)sign;
#pragma unroll
  for (int k = 0; k < N2; k++) {
    T re = T(0), im = T(0);
#pragma unroll
    for (int m = 0; m < N2; m++) {
      const int e = ((m * k) % N2) * (16 / N2);
      const T c = (T)kCos16[e], sn = sg * (T)kCos16[(e + 12) & 15];  // sin(x) = cos(x - pi/2)
      re += v[m].re * c - v[m].im * sn;
      im += v[m].re * sn + v[m].im * c;
    }
    p[(long)k * nwp] = cmk(re * sc, im * sc);
  }
}

// move a state from reference point qa of its columns f < f1a under the Im kz of kza to reference point qbn of
// its columns f < f1b under the Im kz of kzb: q e^{Im kza qa [f < f1a] - Im kzb qbn [f < f1b]}, the exponent in
// FP64 (after a change of medium, or back to the state itself)
template <typename T>
__global__ void f_reref(cplx<T>* q, const cplx<T>* kza, int f1a, double qa, const cplx<T>* kzb, int f1b, double qbn,
                        long n, int nwp) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int f = (int)(i % nwp);
  const double ex = (f < f1a ? (double)kza[i].im * qa : 0.0) - (f < f1b ? (double)kzb[i].im * qbn : 0.0);
  if (ex == 0.0) return;
  const T g = (T)exp(ex);
  cplx<T> v = q[i];
  q[i] = cmk(v.re * g, v.im * g);
}

template <typename T>
__global__ void f_time_to_host(const T* src, double* dst, int nt, long P) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= P * nt) return;
  long n = i / P, p = i % P;
  dst[i] = (double)src[p * nt + n];
}

// medium tables: phs = Re kz + kg (FP64, as uppe.cl:69), kz as T; padding gets kz = 1, phs = 0
template <typename T>
__global__ void f_medium(const double* kz_host, const double* kg, double* phs, cplx<T>* kzc, int nw, long P,
                         int nwp, const double* k0) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= P * nwp) return;
  long p = i / nwp;
  int f = (int)(i % nwp);
  if (f < nw) {
    double kr = kz_host[2 * (f * P + p)], ki = kz_host[2 * (f * P + p) + 1];
    if (k0) {
      // envelope: phase rate Re kz, and k0(w) in the slot the UPPE right-hand side divides by (Re kz)
      phs[i] = kr;
      kzc[i] = cmk((T)k0[f], (T)ki);
    } else {
      phs[i] = kr + kg[f];
      kzc[i] = cmk((T)kr, (T)ki);
    }
  } else {
    phs[i] = 0.0;
    kzc[i] = cmk(T(1), T(0));
  }
}

// E = i kz A in place with kz indexed by the r-space point (finish, uppe_kernel.py:252-253)
template <typename T>
__global__ void f_vp_to_e(cplx<T>* a, const cplx<T>* kzc, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  cplx<T> kz = kzc[i];
  a[i] = cmul(cmk(-kz.im, kz.re), a[i]);
}

// ---------------------------------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------------------------------

static int ilog2i(long v) {
  int l = 0;
  while ((1L << l) < v) l++;
  return l;
}

template <typename K, typename... A>
static void launchf(uppe_handle* h, int kc, K kernel, dim3 grid, dim3 block, size_t shm, A... args) {
  if (!h->prof) {
    hipLaunchKernelGGL(kernel, grid, block, shm, h->stream, args...);
  } else {
    // events stamped by the dispatch itself: they bracket the kernel's execution, not the gaps around it
    hipEvent_t a, b;
    prof_take(h, &a, &b);
    hipExtLaunchKernelGGL(kernel, grid, block, (std::uint32_t)shm, h->stream, a, b, 0u, args...);
    prof_push(h, kc, a, b);
  }
  UPPE_LAUNCH_CHECK();
}

// ---- host-side MFMA fragment packing -------------------------------------------------------------

static uint16_t bf16_rne(float x) {
  uint32_t b;
  std::memcpy(&b, &x, 4);
  if ((b & 0x7f800000u) == 0x7f800000u) return (uint16_t)(b >> 16);  // inf / nan
  b += 0x7fffu + ((b >> 16) & 1u);
  return (uint16_t)(b >> 16);
}

static double bf16_value(uint16_t h) {
  uint32_t b = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &b, 4);
  return (double)f;
}

// A is M x K row-major FP64. K4 layout (f32/f64 16x16x4): element j of lane l in block (rt, kq) is
// A[16 rt + l%16][16 kq + 4 j + l/16]. BF16 layout (16x16x32): element j of lane l in block (rt, kc) is
// A[16 rt + l%16][32 kc + 8 (l/16) + j], split into hi = bf16(a) and lo = bf16(a - hi).
template <typename T>
static void pack_k4(const std::vector<double>& A, int M, int K, std::vector<T>& out, int* kp) {
  const int M16 = (M + 15) / 16, Kp = ((K + 15) / 16) * 16, KQ = Kp / 16;
  out.assign((size_t)M16 * KQ * 64 * 4, T(0));
  for (int rt = 0; rt < M16; rt++)
    for (int kq = 0; kq < KQ; kq++)
      for (int l = 0; l < 64; l++)
        for (int j = 0; j < 4; j++) {
          const int r = 16 * rt + (l & 15), k = 16 * kq + 4 * j + (l >> 4);
          out[(((size_t)rt * KQ + kq) * 64 + l) * 4 + j] = (r < M && k < K) ? (T)A[(size_t)r * K + k] : T(0);
        }
  *kp = Kp;
}

static void pack_bf16x3(const std::vector<double>& A, int M, int K, std::vector<uint16_t>& hi,
                        std::vector<uint16_t>& lo, int* kp) {
  const int M16 = (M + 15) / 16, Kp = ((K + 31) / 32) * 32, KC = Kp / 32;
  hi.assign((size_t)M16 * KC * 64 * 8, 0);
  lo.assign(hi.size(), 0);
  for (int rt = 0; rt < M16; rt++)
    for (int kc = 0; kc < KC; kc++)
      for (int l = 0; l < 64; l++)
        for (int j = 0; j < 8; j++) {
          const int r = 16 * rt + (l & 15), k = 32 * kc + 8 * (l >> 4) + j;
          const double a = (r < M && k < K) ? A[(size_t)r * K + k] : 0.0;
          const uint16_t h = bf16_rne((float)a);
          const size_t o = (((size_t)rt * KC + kc) * 64 + l) * 8 + j;
          hi[o] = h;
          lo[o] = bf16_rne((float)(a - bf16_value(h)));
        }
  *kp = Kp;
}

template <typename V>
static void* upload(const std::vector<V>& v) {
  void* d = nullptr;
  UPPE_HIP(hipMalloc(&d, sizeof(V) * std::max<size_t>(1, v.size())));
  UPPE_HIP(hipMemcpy(d, v.data(), sizeof(V) * v.size(), hipMemcpyHostToDevice));
  return d;
}

// Launch plan of one contraction: ncols complex columns in tiles of 8, split into row blocks of at least 128
// rows while that keeps one workgroup per CU (and into blocks of at most 512 rows whatever the CU count). When a
// grid of ceil(ncols/8) tiles would not fit the CUs, the last tile takes the remainder instead (a 2049-column
// transform is 256 tiles, not 257), so that no workgroup waits for a second round.
static FusedState::HkPlan plan_hankel(int M, int ncols, int n_cu) {
  FusedState::HkPlan p;
  const int m16 = (M + 15) / 16;
  const int narrow = (ncols + 7) / 8, merged = std::max(1, ncols / 8);
  p.tiles = narrow;
  p.split = 1;
  for (int sp = 8; sp >= 1; sp /= 2) {
    if (sp > 1 && m16 / sp < 8) continue;
    if ((long)narrow * sp <= n_cu) {
      p.tiles = narrow;
      p.split = sp;
      break;
    }
    if ((long)merged * sp <= n_cu) {
      p.tiles = merged;
      p.split = sp;
      break;
    }
  }
  // a row block holds at most 16 waves of two 16-row tiles: more rows take more row blocks, even where the grid
  // then exceeds one workgroup per CU
  while ((m16 + p.split - 1) / p.split > 32) p.split *= 2;
  const int per = (m16 + p.split - 1) / p.split;  // 16-row tiles per row block
  p.nwv = std::min(16, per);
  p.rt = (per + p.nwv - 1) / p.nwv;
  if (p.rt > 2) throw Error{UPPE_ERR_UNSUPPORTED, "radial contraction: too many rows"};
  return p;
}

// whether p is one of a batch member's slab addresses (owned by the batch, never freed by the member)
static bool slotted(const FusedState* s, const void* p) {
  if (!s->slots) return false;
  for (void* q : s->slots->p)
    if (q == p) return true;
  return false;
}

// fragments of a packed m x k operator (pack_k4 / pack_bf16x3): 16-row tiles x contraction chunks x 64 lanes
static long frag_count(int cm, int m, int k) {
  const long m16 = (m + 15) / 16;
  return cm == CM_BF16X3 ? m16 * ((k + 31) / 32) * 64 : m16 * ((k + 15) / 16) * 64;
}

constexpr size_t kMaxLds = 159 * 1024;  // dynamic LDS of one CU (gfx950: 160 KB, less the kernels' static arrays)

// dynamic LDS of one contraction launch: a chunk of kch rows of the B slab (2 BN real columns; BF16x3 keeps hi
// and lo planes with padded rows) or the output tile (bm rows), whichever is larger
static size_t hk_lds_bytes(int cm, int kch, int bm, int bn) {
  const size_t bw = 2 * (size_t)bn, el = cm == CM_F64 ? 8 : 4;
  const size_t slab = cm == CM_BF16X3 ? 2 * bw * (kch + 8) * 2 : (size_t)kch * bw * el;
  return std::max(slab, (size_t)bm * bw * el);
}

// contraction rows staged per chunk: all kp when the slab fits in LDS, else the largest multiple of 128 that fits
// (128 is a whole number of fragment-ring spans in every mode: 32 x 4 for BF16x3, 16 x 4 or 16 x 2 otherwise)
static int hk_chunk(int cm, int kp, int bm, int bn) {
  if (hk_lds_bytes(cm, kp, bm, bn) <= kMaxLds) return kp;
  int kch = kp / 128 * 128;
  while (kch > 128 && hk_lds_bytes(cm, kch, bm, bn) > kMaxLds) kch -= 128;
  return kch;
}

template <typename T>
void tune_single(uppe_handle* h, FusedState* s, bool may_time = true);

// the wave-per-strip source kernel for many short FP32 strips (SEARAY_UPPE_WAVE_SOURCE=0 selects the workgroup
// kernel, =2 forces the wave kernel whatever the strip count)
static bool wave_source_enabled() {
  const char* env = getenv("SEARAY_UPPE_WAVE_SOURCE");
  return !(env && env[0] == '0');
}
static bool wave_source_forced() {
  const char* env = getenv("SEARAY_UPPE_WAVE_SOURCE");
  return env && env[0] == '2';
}

// the FP32 stepping source kernel (no diagnostics, no rotator) for Nt = 2^LG, where the block of Nt / S threads
// fits 64..1024; otherwise the kernel for any Nt
template <int S, int LG>
constexpr auto source_fixed() {
  if constexpr ((1 << LG) / S <= 1024 && (1 << LG) / S >= 64)
    return f_source<float, S, false, false, LG>;
  else
    return f_source<float, S, false, false, 0>;
}

// the same for the FP32 envelope source kernel (Nt = Nw = 256 .. 1024)
template <int S, int LG>
constexpr auto env_fixed() {
  if constexpr ((1 << LG) / S <= 1024 && (1 << LG) / S >= 64)
    return f_source_env<float, S, false, false, LG>;
  else
    return f_source_env<float, S, false, false, 0>;
}

// the same for the FP32 adjoint source kernels (Kerr, and with photoionisation)
template <int S, int LG>
constexpr auto adj_fixed() {
  if constexpr ((1 << LG) / S <= 1024 && (1 << LG) / S >= 64)
    return f_source_adj<float, S, LG>;
  else
    return f_source_adj<float, S, 0>;
}
template <int S, int LG>
constexpr auto adj_ion_fixed() {
  if constexpr ((1 << LG) / S <= 1024 && (1 << LG) / S >= 64)
    return f_source_adj_ion<float, S, LG>;
  else
    return f_source_adj_ion<float, S, 0>;
}

template <typename T>
struct Fused {
  using C = cplx<T>;

  static void create(uppe_handle* h, FusedState* s) {
    const long P = h->P;
    s->nwp = ((h->nw + 31) / 32) * 32;
    s->el = sizeof(T);
    s->f1 = h->nw;
    int dev = 0, ncu = 0;
    UPPE_HIP(hipGetDevice(&dev));
    UPPE_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    s->n_cu = std::max(1, ncu);
    const size_t cs = sizeof(C) * P * s->nwp;
    // a batch member's per-instance arrays live in its batch's slabs
    auto take = [&](void** dst, int kind, size_t bytes) {
      if (s->slots)
        *dst = s->slots->p[kind];
      else
        UPPE_HIP(hipMalloc(dst, bytes));
    };
    take((void**)&s->phs, SL_PHS, sizeof(double) * P * s->nwp);
    take(&s->kzc, SL_KZC, cs);
    UPPE_HIP(hipMalloc(&s->w, sizeof(T) * s->nwp));
    UPPE_HIP(hipMalloc(&s->nlf, sizeof(T) * s->nwp));
    UPPE_HIP(hipMalloc(&s->cond, sizeof(T) * s->nwp));
    take((void**)&s->ng, SL_NG, sizeof(double) * P);
    void** bufs[] = {&s->q, &s->k1, &s->k2, &s->k3, &s->er, &s->jr};
    const int kinds[] = {SL_Q, SL_K1, SL_K2, SL_K3, SL_ER, SL_JR};
    for (int i = 0; i < 6; i++) {
      take(bufs[i], kinds[i], cs);
      UPPE_HIP(hipMemset(*bufs[i], 0, cs));
    }
    UPPE_HIP(hipMalloc(&s->chi_t, sizeof(T) * h->ntt));
    UPPE_HIP(hipMalloc(&s->ne_t, sizeof(T) * h->ntt));
    UPPE_HIP(hipMalloc(&s->stage, sizeof(double) * std::max(2 * h->npt, h->ntt)));
    take((void**)&s->qmax, SL_QMAX, sizeof(unsigned long long) * kQW);
    s->qmax_next = s->qmax + kQmaxSlot;
    UPPE_HIP(hipMemset(s->qmax, 0, sizeof(unsigned long long) * kQW));
    take((void**)&s->st, SL_ST, sizeof(StepState));
    UPPE_HIP(hipMemset(s->st, 0, sizeof(StepState)));
    UPPE_HIP(hipHostMalloc(&s->st_host, sizeof(StepState), hipHostMallocDefault));
    std::memset(s->st_host, 0, sizeof(StepState));
    // per-pass Stockham twiddles for the Nt- and Nt/2-point transforms, computed in FP64 and rounded to T
    std::vector<C> t = r8_tables<T>(h->nt);
    UPPE_HIP(hipMalloc(&s->tw, sizeof(C) * t.size()));
    UPPE_HIP(hipMemcpy(s->tw, t.data(), sizeof(C) * t.size(), hipMemcpyHostToDevice));
    if (h->d.geometry == UPPE_RADIAL) {
      // one operator per azimuthal mode
      UPPE_HIP(hipMalloc(&s->Hb, sizeof(T) * (size_t)h->n1 * h->nk * h->n2));
      UPPE_HIP(hipMalloc(&s->Hf, sizeof(T) * (size_t)h->n1 * h->nk * h->n2));
    } else {
      std::vector<C> tx = r8_table_one<T>(h->n1), ty = r8_table_one<T>(h->n2);
      s->twlen_x = (int)tx.size();
      s->twlen_y = (int)ty.size();
      UPPE_HIP(hipMalloc(&s->tw_x, sizeof(C) * std::max<size_t>(1, tx.size())));
      UPPE_HIP(hipMalloc(&s->tw_y, sizeof(C) * std::max<size_t>(1, ty.size())));
      if (!tx.empty()) UPPE_HIP(hipMemcpy(s->tw_x, tx.data(), sizeof(C) * tx.size(), hipMemcpyHostToDevice));
      if (!ty.empty()) UPPE_HIP(hipMemcpy(s->tw_y, ty.data(), sizeof(C) * ty.size(), hipMemcpyHostToDevice));
    }
  }

  static void set_hankel(uppe_handle* h, FusedState* s) {
    const int nr = h->n1, nk = h->nk, nm = h->n2;
    // per azimuthal mode m (H is [Nk][Nr][N2]): Hb_m = (Lambda_k / Lambda_j) H_m^T and Hf_m = (Lambda_j / Lambda_k) H_m,
    // the root-volume weighting of RootVolumeMultiply / RootVolumeDivide folded in (grid_tools.py:308-325), the
    // modes' operators one after another
    const size_t mat = (size_t)nr * nk;
    std::vector<double> hb(mat * nm), hf(mat * nm);
    for (int m = 0; m < nm; m++)
      for (int k = 0; k < nk; k++)
        for (int j = 0; j < nr; j++) {
          double v = h->H[((size_t)k * nr + j) * nm + m];
          hb[m * mat + (size_t)j * nk + k] = v * h->lam[k] / h->lam[j];
          hf[m * mat + (size_t)k * nr + j] = v * h->lam[j] / h->lam[k];
        }
    s->op_fp = Fingerprint();
    s->op_fp.mix(hb.data(), sizeof(double) * hb.size());
    s->op_fp.mix(hf.data(), sizeof(double) * hf.size());
    std::vector<T> hbt(hb.begin(), hb.end()), hft(hf.begin(), hf.end());
    UPPE_HIP(hipMemcpy(s->Hb, hbt.data(), sizeof(T) * hbt.size(), hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->Hf, hft.data(), sizeof(T) * hft.size(), hipMemcpyHostToDevice));
    void* old[] = {s->frb_hi, s->frb_lo, s->frf_hi, s->frf_lo, s->fab_hi, s->fab_lo, s->faf_hi, s->faf_lo};
    for (void* p : old)
      if (p && !slotted(s, p)) UPPE_HIP(hipFree(p));
    s->frb_hi = s->frb_lo = s->frf_hi = s->frf_lo = nullptr;
    s->fab_hi = s->fab_lo = s->faf_hi = s->faf_lo = nullptr;
    if (s->cm < 0) return;
    // the adjoint's transposes, per mode: Hf^T has the shape of Hb, Hb^T that of Hf
    std::vector<double> hft_(mat * nm), hbt_(mat * nm);
    for (int m = 0; m < nm; m++)
      for (int k = 0; k < nk; k++)
        for (int j = 0; j < nr; j++) {
          hft_[m * mat + (size_t)j * nk + k] = hf[m * mat + (size_t)k * nr + j];
          hbt_[m * mat + (size_t)k * nr + j] = hb[m * mat + (size_t)j * nk + k];
        }
    int kpa = 0;
    // the forward operators of a batch member go to its slots (sizes fixed by the shape: frag_count)
    auto place = [&](const auto& v, int kind) -> void* {
      if (!s->slots) return upload(v);
      void* d = s->slots->p[kind];
      UPPE_HIP(hipMemcpy(d, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice));
      return d;
    };
    // the fragment sets of the modes, one after another (the contraction's instance stride)
    auto modes = [&](const std::vector<double>& all, auto pack) {
      std::vector<double> one(mat);
      decltype(pack(one)) out;
      for (int m = 0; m < nm; m++) {
        std::copy(all.begin() + m * mat, all.begin() + (m + 1) * mat, one.begin());
        auto part = pack(one);
        out.first.insert(out.first.end(), part.first.begin(), part.first.end());
        out.second.insert(out.second.end(), part.second.begin(), part.second.end());
      }
      return out;
    };
    if (s->cm == CM_BF16X3) {
      auto pb = modes(hb, [&](const std::vector<double>& v) {
        std::pair<std::vector<uint16_t>, std::vector<uint16_t>> r;
        pack_bf16x3(v, nr, nk, r.first, r.second, &s->kp_b);
        return r;
      });
      s->frb_hi = place(pb.first, SL_FRB_HI);
      s->frb_lo = place(pb.second, SL_FRB_LO);
      auto pf = modes(hf, [&](const std::vector<double>& v) {
        std::pair<std::vector<uint16_t>, std::vector<uint16_t>> r;
        pack_bf16x3(v, nk, nr, r.first, r.second, &s->kp_f);
        return r;
      });
      s->frf_hi = place(pf.first, SL_FRF_HI);
      s->frf_lo = place(pf.second, SL_FRF_LO);
      auto pab = modes(hft_, [&](const std::vector<double>& v) {
        std::pair<std::vector<uint16_t>, std::vector<uint16_t>> r;
        pack_bf16x3(v, nr, nk, r.first, r.second, &kpa);
        return r;
      });
      s->fab_hi = upload(pab.first);
      s->fab_lo = upload(pab.second);
      auto paf = modes(hbt_, [&](const std::vector<double>& v) {
        std::pair<std::vector<uint16_t>, std::vector<uint16_t>> r;
        pack_bf16x3(v, nk, nr, r.first, r.second, &kpa);
        return r;
      });
      s->faf_hi = upload(paf.first);
      s->faf_lo = upload(paf.second);
    } else {
      auto pb = modes(hb, [&](const std::vector<double>& v) {
        std::pair<std::vector<T>, std::vector<T>> r;
        pack_k4<T>(v, nr, nk, r.first, &s->kp_b);
        return r;
      });
      s->frb_hi = place(pb.first, SL_FRB_HI);
      auto pf = modes(hf, [&](const std::vector<double>& v) {
        std::pair<std::vector<T>, std::vector<T>> r;
        pack_k4<T>(v, nk, nr, r.first, &s->kp_f);
        return r;
      });
      s->frf_hi = place(pf.first, SL_FRF_HI);
      auto pab = modes(hft_, [&](const std::vector<double>& v) {
        std::pair<std::vector<T>, std::vector<T>> r;
        pack_k4<T>(v, nr, nk, r.first, &kpa);
        return r;
      });
      s->fab_hi = upload(pab.first);
      auto paf = modes(hbt_, [&](const std::vector<double>& v) {
        std::pair<std::vector<T>, std::vector<T>> r;
        pack_k4<T>(v, nk, nr, r.first, &kpa);
        return r;
      });
      s->faf_hi = upload(paf.first);
    }
    s->plan[0] = plan_hankel(nr, h->nw, s->n_cu);
    s->plan[1] = plan_hankel(nk, s->f1, s->n_cu);
    s->tuned_f1 = -1;
  }

  static void set_medium(uppe_handle* h, FusedState* s, const uppe_medium* m) {
    const long P = h->P;
    std::vector<T> w(s->nwp, T(0)), nl(s->nwp, T(0)), cd(s->nwp, T(0));
    int last_nl = -1;
    for (int f = 0; f < h->nw; f++) {
      w[f] = (T)m->w[f];
      nl[f] = (T)m->nl_filter[f];
      cd[f] = (T)m->cond_filter[f];
      if (m->nl_filter[f] != 0.0) last_nl = f;
    }
    UPPE_HIP(hipMemcpy(s->w, w.data(), sizeof(T) * s->nwp, hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->nlf, nl.data(), sizeof(T) * s->nwp, hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->cond, cd.data(), sizeof(T) * s->nwp, hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->ng, m->ngas, sizeof(double) * P, hipMemcpyHostToDevice));
    s->w_h.assign(m->w, m->w + h->nw);
    s->nl_h.assign(m->nl_filter, m->nl_filter + h->nw);
    s->cond_h.assign(m->cond_filter, m->cond_filter + h->nw);
    s->tab_fp = Fingerprint();
    s->tab_fp.mix(m->kz, sizeof(double) * 2 * h->npt);
    s->tab_fp.mix(m->kg, sizeof(double) * h->nw);
    double* kg = nullptr;
    UPPE_HIP(hipMalloc(&kg, sizeof(double) * h->nw));
    UPPE_HIP(hipMemcpy(kg, m->kg, sizeof(double) * h->nw, hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->stage, m->kz, sizeof(double) * 2 * h->npt, hipMemcpyHostToDevice));
    const bool moved = begin_medium_change(h, s, m->kz);
    launchf(h, UPPE_KC_POINT, f_medium<T>, dim3(grid_for(P * s->nwp)), dim3(kBlock), 0, (const double*)s->stage,
            (const double*)kg, s->phs, (C*)s->kzc, h->nw, P, s->nwp, (const double*)nullptr);
    UPPE_HIP(hipStreamSynchronize(h->stream));
    UPPE_HIP(hipFree(kg));
    s->redchi3 = m->redchi3;
    s->dt = m->dt;
    s->ion = m->ion;
    s->max_ki = 0.0;
    for (long i = 0; i < h->npt; i++) s->max_ki = std::max(s->max_ki, m->kz[2 * i + 1]);
    // columns at and above f1 carry zero current: the r2k contraction and the stage updates skip them.
    // f1 covers the conditioner's four bins.
    const int f1 = std::min(h->nw, std::max(last_nl + 1, 4));
    if (moved) end_medium_change(h, s, f1);
    if (f1 != s->f1) {
      s->f1 = f1;
      if (h->have_hankel) s->plan[1] = plan_hankel(h->nk, s->f1, s->n_cu);
      const size_t cs = sizeof(C) * P * s->nwp;
      UPPE_HIP(hipMemset(s->k1, 0, cs));
      UPPE_HIP(hipMemset(s->k2, 0, cs));
      UPPE_HIP(hipMemset(s->k3, 0, cs));
      if (h->have_field) static_qmax(h, s);
    }
  }

  // paraxial medium (paraxial_kernel.py:72-135): no nonlinear filter or conditioner (both one), Re kz as the
  // phase rate, k0(w) as the divisor of the right-hand side
  static void set_paraxial_medium(uppe_handle* h, FusedState* s, const uppe_paraxial_medium* m) {
    const long P = h->P;
    s->tab_fp = Fingerprint();  // not shared (batches step the UPPE model)
    std::vector<T> w(s->nwp, T(0)), one(s->nwp, T(0));
    for (int f = 0; f < h->nw; f++) {
      w[f] = (T)m->w[f];
      one[f] = T(1);
    }
    UPPE_HIP(hipMemcpy(s->w, w.data(), sizeof(T) * s->nwp, hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->nlf, one.data(), sizeof(T) * s->nwp, hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->cond, one.data(), sizeof(T) * s->nwp, hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->ng, m->ngas, sizeof(double) * P, hipMemcpyHostToDevice));
    double* k0 = nullptr;
    UPPE_HIP(hipMalloc(&k0, sizeof(double) * h->nw));
    UPPE_HIP(hipMemcpy(k0, m->k0, sizeof(double) * h->nw, hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(s->stage, m->kz, sizeof(double) * 2 * h->npt, hipMemcpyHostToDevice));
    const bool moved = begin_medium_change(h, s, m->kz);
    launchf(h, UPPE_KC_POINT, f_medium<T>, dim3(grid_for(P * s->nwp)), dim3(kBlock), 0, (const double*)s->stage,
            (const double*)nullptr, s->phs, (C*)s->kzc, h->nw, P, s->nwp, (const double*)k0);
    UPPE_HIP(hipStreamSynchronize(h->stream));
    UPPE_HIP(hipFree(k0));
    s->redchi3 = m->redchi3;
    s->dt = m->dt;
    s->ion = m->ion;
    s->max_ki = 0.0;
    for (long i = 0; i < h->npt; i++) s->max_ki = std::max(s->max_ki, m->kz[2 * i + 1]);
    if (moved) end_medium_change(h, s, h->nw);
    if (h->nw != s->f1) {
      s->f1 = h->nw;
      if (h->have_hankel) s->plan[1] = plan_hankel(h->nk, s->f1, s->n_cu);
      const size_t cs = sizeof(C) * P * s->nwp;
      UPPE_HIP(hipMemset(s->k1, 0, cs));
      UPPE_HIP(hipMemset(s->k2, 0, cs));
      UPPE_HIP(hipMemset(s->k3, 0, cs));
      if (h->have_field) static_qmax(h, s);
    }
  }

  // A change of medium under a state kept relative to its reference point (FP32, qbase != 0): the state is q
  // e^{-Im kz_old qbase} in the old band, and must become q e^{-Im kz_new qbase} in the new one, which keeps q
  // across the change as the reference's subcycles do (uppe_kernel.py:300-305). begin_medium_change saves the old
  // Im kz (complex kz into the zero stage slope k1) before the new medium is loaded and checks the factor's range
  // against the host copy; end_medium_change applies it and clears k1.
  static bool begin_medium_change(uppe_handle* h, FusedState* s, const double* kz_new) {
    const bool moved = sizeof(T) == 4 && s->qbase != 0.0 && h->have_field;
    if (moved) {
      double ex = 0.0;
      for (long i = 0; i < h->npt; i++) ex = std::max(ex, (s->ki_h[(size_t)i] - kz_new[2 * i + 1]) * s->qbase);
      if (ex > 80.0)
        throw Error{UPPE_ERR_UNSUPPORTED, "FP32 range: this change of medium scales the state by up to exp(" +
                                              std::to_string(ex) + "); use FP64"};
      UPPE_HIP(hipMemcpyAsync(s->k1, s->kzc, sizeof(C) * h->P * s->nwp, hipMemcpyDeviceToDevice, h->stream));
    }
    s->ki_h.resize((size_t)h->npt);
    for (long i = 0; i < h->npt; i++) s->ki_h[(size_t)i] = kz_new[2 * i + 1];
    return moved;
  }

  static void end_medium_change(uppe_handle* h, FusedState* s, int f1_new) {
    const long n = (long)h->P * s->nwp;
    launchf(h, UPPE_KC_POINT, f_reref<T>, dim3(grid_for(n)), dim3(kBlock), 0, (C*)s->q, (const C*)s->k1, s->f1,
            s->qbase, (const C*)s->kzc, f1_new, s->qbase, n, s->nwp);
    UPPE_HIP(hipMemsetAsync(s->k1, 0, sizeof(C) * n, h->stream));
    UPPE_HIP(hipStreamSynchronize(h->stream));
  }

  static void static_qmax(uppe_handle* h, FusedState* s) {
    UPPE_HIP(hipMemsetAsync(s->qmax + 2, 0, sizeof(unsigned long long), h->stream));
    launchf(h, UPPE_KC_REDUCE, f_qmax_cols<T>, dim3(256), dim3(kBlock), 0, (const C*)s->q, h->P, s->nwp, s->f1,
            s->qmax + 2);
    UPPE_HIP(hipStreamSynchronize(h->stream));
  }

  // SIMT contraction (the one-off transforms of set_field/finish): one launch per azimuthal mode m over the rows
  // r N2 + m, with that mode's operator
  static void gemm_r2k_plain(uppe_handle* h, FusedState* s, const C* in, C* out) {
    dim3 grid(2 * s->nwp / BN, (h->nk + BM - 1) / BM);
    const int row = h->n2 * s->nwp;
    for (int m = 0; m < h->n2; m++)
      launchf(h, UPPE_KC_HANKEL, f_gemm<T, PRO_NONE, EPI_NONE>, grid, dim3(256), 0,
              (const T*)s->Hf + (size_t)m * h->n1 * h->nk, in + (long)m * s->nwp, out + (long)m * s->nwp, h->nk,
              h->n1, row, (const double*)nullptr, (const C*)nullptr, (const T*)nullptr, (C*)nullptr,
              (const C*)nullptr, (C*)nullptr, (C*)nullptr, (C*)nullptr, (StepState*)nullptr,
              (unsigned long long*)nullptr, (unsigned long long*)nullptr, 0, 0.0, 0.0, 0);
  }

  // the DFT over the azimuthal points of every (r, f < ncols) with exponent sign `sign`, scaled by sc (default:
  // numpy's fft for sign -1, its ifft for sign +1); nothing for N2 = 1
  static void azim(uppe_handle* h, FusedState* s, C* a, int ncols, int sign, double sc = 0.0) {
    if (h->n2 == 1) return;
    if (sc == 0.0) sc = sign > 0 ? 1.0 / h->n2 : 1.0;
    const long n = (long)h->n1 * ncols;
    auto kern = h->n2 == 2 ? f_azim<T, 2> : (h->n2 == 4 ? f_azim<T, 4> : (h->n2 == 8 ? f_azim<T, 8> : f_azim<T, 16>));
    launchf(h, UPPE_KC_TRANSVERSE_FFT, kern, dim3(grid_for(n)), dim3(kBlock), 0, a, h->n1, s->nwp, ncols, sign,
            (T)sc);
  }

  // the source kernel over every strip; a batched launch (nz instances zs elements apart, per-instance scalars
  // inst) may ask for spt_req samples per thread
  static void source(uppe_handle* h, FusedState* s, const C* er, C* jr, bool diag, StepState* commit, int nz = 1,
                     const SrcInst* inst = nullptr, long zs = 0, int spt_req = 0) {
    // the stage-1 launch commits the previous step; a standalone handle's advance also records dz (and the
    // adjoint's phase origins) there
    double* rec = (commit && !inst) ? s->rec_active : nullptr;
    double* zrec = (commit && !inst) ? s->zrec_active : nullptr;
    const int zrec_cap = s->zrec_cap / 3;
    const int N = h->nt;
    // four samples per thread (Nt >= 256): short strips then get short workgroups, several per CU; eight for the
    // 8192-sample strips (a workgroup holds at most 1024 threads)
    const int nth = spt_req > 0 ? N / spt_req : std::max(std::min(N, 64), std::min(N / 4, 1024));
    const int spt = N / nth;
    if (spt_req > 0 && (h->rot_states > 2 || h->paraxial || nth > 1024 || nth < 64))
      throw Error{UPPE_ERR_ARG, "source kernel: unsupported samples per thread"};
    const size_t shm = sizeof(C) * (sizeof(T) == 4 && N <= kMaxNtTables ? N + N + N / 2 : N);
    const int lgn = ilog2i(N);
    // the explicit-rate prefactor, in the base of the exponential each precision uses
    const double log_pre = s->ion.kind != UPPE_ION_EXPLICIT ? 0.0
                           : sizeof(T) == 4                  ? std::log2(s->ion.rate_conv * s->ion.c_pre)
                                                             : std::log(s->ion.rate_conv * s->ion.c_pre);
    const bool rot = h->rot_states > 2;
    if constexpr (sizeof(T) == 4) {
      // many short strips (at least four per CU): one wavefront per strip, no workgroup barriers. With fewer
      // strips the workgroup kernel, whose strip transforms spread over more waves, is as fast or faster.
      if (!rot && !diag && !h->paraxial && lgn >= 8 && lgn <= 10 &&
          ((long)h->P * nz >= 4L * s->n_cu || wave_source_forced()) && wave_source_enabled()) {
        const dim3 g((unsigned)h->P, (unsigned)nz);
        auto kern = lgn == 8 ? f_source_w<8, 1> : (lgn == 9 ? f_source_w<9, 1> : f_source_w<10, 1>);
        launchf(h, UPPE_KC_SOURCE, kern, g, dim3(64), 0, (const cplx<float>*)er, (cplx<float>*)jr,
                (const cplx<float>*)s->tw, (const float*)s->w, (const float*)s->nlf, (const double*)s->ng, (int)h->P,
                s->nwp, s->redchi3, s->dt, s->ion, log_pre, inst, zs, commit, s->qmax, rec, zrec, zrec_cap);
        return;
      }
    }
    const dim3 grid((unsigned)h->P, (unsigned)nz), block(nth);
    if (h->paraxial) {
      // two N-point complex strips (E and A), plus the N-point tables in FP32
      const size_t shm_env = sizeof(C) * (sizeof(T) == 4 ? 3 * N : 2 * N);
#define UPPE_SRC_ENV(S)                                                                                          \
  if (spt == S) {                                                                                                \
    auto kern = rot ? (diag ? f_source_env<T, S, true, true> : f_source_env<T, S, false, true>)                 \
                    : (diag ? f_source_env<T, S, true, false> : f_source_env<T, S, false, false>);              \
    if constexpr (sizeof(T) == 4)                                                                                \
      if (!rot && !diag && lgn >= 8 && lgn <= 10)                                                                \
        kern = lgn == 8 ? env_fixed<S, 8>() : (lgn == 9 ? env_fixed<S, 9>() : env_fixed<S, 10>());             \
    launchf(h, UPPE_KC_SOURCE, kern, grid, block, shm_env, er, jr, diag ? (T*)s->chi_t : (T*)nullptr,             \
            diag ? (T*)s->ne_t : (T*)nullptr, (const C*)s->tw, (const T*)s->w, (const double*)s->ng, N, lgn,      \
            s->nwp, s->redchi3, s->dt, s->ion, log_pre, (const RotRow<T>*)s->rot_tab, h->rot_states,              \
            0.5 * h->rot_dalpha / h->rot_hbar, commit, s->qmax, rec, zrec, zrec_cap);                            \
    return;                                                                                                      \
  }
      UPPE_SRC_ENV(1) UPPE_SRC_ENV(2) UPPE_SRC_ENV(4)
#undef UPPE_SRC_ENV
      throw Error{UPPE_ERR_UNSUPPORTED, "fused envelope source kernel: unsupported Nt"};
    }
    // a standalone stage-1 launch over fewer strips than the device has CUs takes one more workgroup, which only
    // commits the previous step (on a CU the strips leave idle)
    const int commit_only = (commit && !inst && (long)h->P < s->n_cu) ? (int)h->P : 0;
    const dim3 grid_src((unsigned)(h->P + (commit_only > 0 ? 1 : 0)), (unsigned)nz);
#define UPPE_SRC(S)                                                                                              \
  if (spt == S) {                                                                                                \
    auto kern = rot ? (diag ? f_source<T, S, true, true> : f_source<T, S, false, true>)                         \
                    : (diag ? f_source<T, S, true, false> : f_source<T, S, false, false>);                      \
    if constexpr (sizeof(T) == 4 && S <= 4)                                                                      \
      if (!rot && !diag && lgn >= 10 && lgn <= 12)                                                               \
        kern = lgn == 10 ? source_fixed<S, 10>() : (lgn == 11 ? source_fixed<S, 11>() : source_fixed<S, 12>());  \
    launchf(h, UPPE_KC_SOURCE, kern, grid_src, block, shm, er, jr, diag ? (T*)s->chi_t : (T*)nullptr,             \
            diag ? (T*)s->ne_t : (T*)nullptr, (const C*)s->tw, (const T*)s->w, (const T*)s->nlf,                  \
            (const double*)s->ng, N, lgn, s->nwp, s->redchi3, s->dt, s->ion, log_pre,                             \
            (const RotRow<T>*)s->rot_tab, h->rot_states, 0.5 * h->rot_dalpha / h->rot_hbar, inst, zs, commit,   \
            s->qmax, rec, zrec, zrec_cap, commit_only);                                                          \
    return;                                                                                                      \
  }
    UPPE_SRC(1) UPPE_SRC(2) UPPE_SRC(4) UPPE_SRC(8)
#undef UPPE_SRC
    throw Error{UPPE_ERR_UNSUPPORTED, "fused source kernel: unsupported Nt"};
  }

  // adjoint of the Kerr source on every strip (f_source_adj): er = the stage's E_r, gj = cotangent of J_r
  static void source_adj(uppe_handle* h, FusedState* s, const C* er, const C* gj, C* ge) {
    const int N = h->nt;
    // the forward source kernel's shape (wide workgroups when the strips are fewer than the CUs)
    const int nth = s->spt > 0 ? N / s->spt : std::max(std::min(N, 64), N / 4);
    const int spt = N / nth;
    const size_t shm = sizeof(C) * (sizeof(T) == 4 ? 2 * N : N);
    const int lgn = ilog2i(N);
    const dim3 grid((unsigned)h->P), block(nth);
#define UPPE_SRC_ADJ(S)                                                                                          \
  if (spt == S) {                                                                                                \
    auto kern = f_source_adj<T, S>;                                                                              \
    if constexpr (sizeof(T) == 4)                                                                                \
      if (lgn >= 10 && lgn <= 12)                                                                                \
        kern = lgn == 10 ? adj_fixed<S, 10>() : (lgn == 11 ? adj_fixed<S, 11>() : adj_fixed<S, 12>());          \
    launchf(h, UPPE_KC_SOURCE, kern, grid, block, shm, er, gj, ge, (const C*)s->tw, (const T*)s->w,               \
            (const T*)s->nlf, (const double*)s->ng, N, lgn, s->nwp, s->redchi3);                                  \
    return;                                                                                                      \
  }
    const bool rot = h->rot_states > 2;
    const RotRow<T>* rtab = (const RotRow<T>*)s->rot_tab;
    const double rfc = rot ? 0.5 * h->rot_dalpha / h->rot_hbar : 0.0;
    if (s->ion.kind != UPPE_ION_NONE || rot) {
      // E, G and R' strips, and the FP32 tables; a rotator takes this kernel with or without photoionisation
      const size_t shm_ion = sizeof(C) * (sizeof(T) == 4 ? 3 * N : 2 * N);
      const double log_pre = s->ion.kind != UPPE_ION_EXPLICIT ? 0.0
                             : sizeof(T) == 4                  ? std::log2(s->ion.rate_conv * s->ion.c_pre)
                                                               : std::log(s->ion.rate_conv * s->ion.c_pre);
#define UPPE_SRC_ADJ_ION(S)                                                                                      \
  if (spt == S) {                                                                                                \
    auto kern = rot ? f_source_adj_ion<T, S, 0, true> : f_source_adj_ion<T, S>;                                  \
    if constexpr (sizeof(T) == 4)                                                                                \
      if (!rot && lgn >= 10 && lgn <= 12)                                                                        \
        kern = lgn == 10 ? adj_ion_fixed<S, 10>() : (lgn == 11 ? adj_ion_fixed<S, 11>() : adj_ion_fixed<S, 12>()); \
    launchf(h, UPPE_KC_SOURCE, kern, grid, block, shm_ion, er, gj, ge, (const C*)s->tw,                          \
            (const T*)s->w, (const T*)s->nlf, (const double*)s->ng, N, lgn, s->nwp, s->redchi3, s->dt, s->ion,    \
            log_pre, rtab, h->rot_states, rfc);                                                                  \
    return;                                                                                                      \
  }
      UPPE_SRC_ADJ_ION(1) UPPE_SRC_ADJ_ION(2) UPPE_SRC_ADJ_ION(4)
#undef UPPE_SRC_ADJ_ION
      throw Error{UPPE_ERR_UNSUPPORTED, "adjoint source kernel: unsupported Nt"};
    }
    UPPE_SRC_ADJ(1) UPPE_SRC_ADJ(2) UPPE_SRC_ADJ(4)
#undef UPPE_SRC_ADJ
    throw Error{UPPE_ERR_UNSUPPORTED, "adjoint source kernel: unsupported Nt"};
  }

  // one launch of the matrix-core contraction out[M][ncols] = A[M][K] B[K][ncols] with fragments (ahi, alo) of A,
  // its launch plan, prologue PRO and epilogue EPI; q, k1..k3 are the state arrays the epilogue updates
  template <int CM, int PRO, int EPI>
  static void hk_launch(uppe_handle* h, FusedState* s, const void* ahi, const void* alo, int kp, int M, int K,
                        int ncols, const FusedState::HkPlan& pl, const C* B, C* out, const C* kprev, C* q, C* k1,
                        C* k2, C* k3, StepState* st, int stage, double zfix, double dzfix, double dzprev,
                        int nz = 1, long zs = 0, int share = 0) {
    using F = typename MfmaTraits<CM>::frag_t;
    // a single grid with azimuthal modes: one contraction per mode m (grid z), each over the rows r N2 + m of the
    // [Nr N2][nwp] arrays with its own operator, all sharing the step state
    int zst = 1;
    long row = s->nwp;
    if (zs == 0 && h->n2 > 1) {
      nz = h->n2;
      zs = s->nwp;
      zst = 0;
      row = (long)h->n2 * s->nwp;
    }
    const int rw = 16 * pl.rt;
    const int bm = pl.nwv * rw;
    // narrow tiles: 8 columns each, the last one up to 15; wide tiles: 8 << wide columns each, the last one what
    // is left
    const long t8 = (long)pl.tiles * 8, tw = (long)pl.tiles * (8 << pl.wide);
    const bool fits = pl.wide ? (tw >= ncols && tw < ncols + (8 << pl.wide)) : (t8 + 8 > ncols && t8 < ncols + 8);
    if (!fits || pl.wide < 0 || pl.wide > 2 || (pl.wide == 2 && pl.nwv > 8))
      throw Error{UPPE_ERR_STATE, "radial contraction plan out of date"};
    dim3 grid((unsigned)pl.tiles, (unsigned)((M + bm - 1) / bm), (unsigned)nz);
    const int bn = pl.wide == 2 ? 32 : 16, kch = hk_chunk(CM, kp, bm, bn);
    const size_t shm = hk_lds_bytes(CM, kch, bm, bn);
    if (shm > kMaxLds) throw Error{UPPE_ERR_UNSUPPORTED, "radial contraction: tile does not fit in LDS"};
#define UPPE_HK(BNV, RTV)                                                                                        \
  launchf(h, EPI == EPI_NONE ? UPPE_KC_HANKEL : UPPE_KC_HANKEL_R2K, f_hankel<CM, PRO, EPI, BNV, RTV>, grid,      \
          dim3(64 * pl.nwv), shm, (const F*)ahi,                                                               \
          (const F*)alo, B, out, M, K, kp, rw, (int)row, 0, ncols, (const double*)s->phs, (const C*)s->kzc,       \
          (const T*)s->cond, q, kprev, k1, k2, k3, st, s->qmax, s->qmax_next, stage, zfix, dzfix, dzprev, pl.wide, \
          zs, zst, kch, zst ? share : 0)
    if (pl.wide == 2) {
      if (pl.rt == 1)
        UPPE_HK(32, 1);
      else
        UPPE_HK(32, 2);
    } else {
      if (pl.rt == 1)
        UPPE_HK(16, 1);
      else
        UPPE_HK(16, 2);
    }
#undef UPPE_HK
  }

  // matrix-core contraction of one RHS stage; dir 0 = k2r (prologue), 1 = r2k (epilogue)
  template <int CM, int PRO, int EPI>
  static void hankel_mfma(uppe_handle* h, FusedState* s, int dir, const C* B, C* out, const C* kprev, int stage,
                          double zfix, StepState* st) {
    const int M = dir == 0 ? h->n1 : h->nk, K = dir == 0 ? h->nk : h->n1;
    const int kp = dir == 0 ? s->kp_b : s->kp_f;
    const int ncols = dir == 0 ? h->nw : s->f1;  // complex columns covered
    hk_launch<CM, PRO, EPI>(h, s, dir == 0 ? s->frb_hi : s->frf_hi, dir == 0 ? s->frb_lo : s->frf_lo, kp, M, K,
                            ncols, s->plan[dir], B, out, kprev, (C*)s->q, (C*)s->k1, (C*)s->k2, (C*)s->k3, st,
                            stage, zfix, 0.0, 0.0);
  }

  // one RK4 step; with erslot, stage s leaves its E_r in erslot[s - 1] (slots of P nwp points) for the adjoint
  template <int CM>
  static void step_mfma(uppe_handle* h, FusedState* s, C* erslot = nullptr) {
    const C* kp[5] = {nullptr, nullptr, (const C*)s->k1, (const C*)s->k2, (const C*)s->k3};
    const long stride = h->P * (long)s->nwp;
    for (int stage = 1; stage <= 4; stage++) {
      C* er = erslot ? erslot + (stage - 1) * stride : (C*)s->er;
      if (h->paraxial)
        hankel_mfma<CM, PRO_STAGE_ENV, EPI_NONE>(h, s, 0, (const C*)s->q, er, kp[stage], stage, 0.0, s->st);
      else
        hankel_mfma<CM, PRO_STAGE, EPI_NONE>(h, s, 0, (const C*)s->q, er, kp[stage], stage, 0.0, s->st);
      azim(h, s, er, h->nw, +1);  // azimuthal modes to phi (IFFT_axis2 after the radial transform)
      source(h, s, (const C*)er, (C*)s->jr, false, stage == 1 ? s->st : nullptr, 1, nullptr, 0, s->spt);
      azim(h, s, (C*)s->jr, s->f1, -1);  // phi to modes over the nonlinear band (FFT_axis2 first)
      hankel_mfma<CM, PRO_NONE, EPI_STAGE>(h, s, 1, (const C*)s->jr, (C*)nullptr, nullptr, stage, 0.0, s->st);
    }
  }

  // one RK4 step: 4 x (k2r, source, r2k); the step control runs inside them
  // one 1-D pass of a Cartesian transverse transform over columns [0, fend); sign +1 inverse (scaled 1/Na).
  // CW columns per line: 128-byte runs (16 FP32 or 8 FP64 columns), 8 FP32 columns for lines of 1024 points
  // (the block holds CW Na/8 <= 1024 threads)
  template <int PRO, int EPI>
  static void tpass(uppe_handle* h, FusedState* s, int axis, const C* src, C* dst, int fend, int sign,
                    const C* kprev, int stage, double zfix, StepState* st) {
    if constexpr (sizeof(T) == 4) {
      if ((axis == 1 ? h->n1 : h->n2) <= 512) {
        tpass_cw<16, PRO, EPI>(h, s, axis, src, dst, fend, sign, kprev, stage, zfix, st);
        return;
      }
    }
    tpass_cw<8, PRO, EPI>(h, s, axis, src, dst, fend, sign, kprev, stage, zfix, st);
  }

  template <int CW, int PRO, int EPI>
  static void tpass_cw(uppe_handle* h, FusedState* s, int axis, const C* src, C* dst, int fend, int sign,
                       const C* kprev, int stage, double zfix, StepState* st) {
    const int na = axis == 1 ? h->n1 : h->n2, nother = axis == 1 ? h->n2 : h->n1;
    const int nth = std::max(64, CW * std::max(1, na / 8));
    const void* tab = axis == 1 ? s->tw_x : s->tw_y;
    const int tl = axis == 1 ? s->twlen_x : s->twlen_y;
    dim3 grid((unsigned)((fend + CW - 1) / CW), (unsigned)nother);
    const size_t shm = sizeof(C) * ((size_t)na * CW + tl);
    if (nth > 1024 || shm > kMaxLds) throw Error{UPPE_ERR_UNSUPPORTED, "transverse pass: line too long"};
    const T scale = sign > 0 ? T(1.0 / na) : T(1);
    launchf(h, UPPE_KC_TRANSVERSE_FFT, f_tfft<T, CW, PRO, EPI>, grid, dim3(nth), shm, src, dst, axis, na, ilog2i(na),
            h->n2, s->nwp, fend, (const C*)tab, tl, sign, scale, (const double*)s->phs, (const C*)s->kzc,
            (const T*)s->cond, (C*)s->q, kprev, (C*)s->k1, (C*)s->k2, (C*)s->k3, st, s->qmax, s->qmax_next, stage,
            zfix, s->qbase, s->f1);
  }

  // one RK4 step on a Cartesian grid: per RHS, the inverse transform (y with the stage prologue, then x), the
  // source, the forward transform over the nonlinear band (x, then y with the RK stage epilogue)
  static void step_cart(uppe_handle* h, FusedState* s) {
    const C* kp[5] = {nullptr, nullptr, (const C*)s->k1, (const C*)s->k2, (const C*)s->k3};
    C *q = (C*)s->q, *er = (C*)s->er, *jr = (C*)s->jr;
    for (int stage = 1; stage <= 4; stage++) {
      if (h->paraxial)
        tpass<PRO_STAGE_ENV, EPI_NONE>(h, s, 2, q, er, h->nw, +1, kp[stage], stage, 0.0, s->st);
      else
        tpass<PRO_STAGE, EPI_NONE>(h, s, 2, q, er, h->nw, +1, kp[stage], stage, 0.0, s->st);
      tpass<PRO_NONE, EPI_NONE>(h, s, 1, er, er, h->nw, +1, nullptr, stage, 0.0, s->st);
      source(h, s, er, jr, false, stage == 1 ? s->st : nullptr);
      tpass<PRO_NONE, EPI_NONE>(h, s, 1, jr, jr, s->f1, -1, nullptr, stage, 0.0, s->st);
      tpass<PRO_NONE, EPI_STAGE>(h, s, 2, jr, (C*)nullptr, s->f1, -1, nullptr, stage, 0.0, s->st);
    }
  }

  static void step(uppe_handle* h, FusedState* s) {
    if (h->d.geometry == UPPE_CARTESIAN) {
      step_cart(h, s);
      return;
    }
    if (s->cm >= 0) {
      if constexpr (sizeof(T) == 8) {
        step_mfma<CM_F64>(h, s);
      } else {
        if (s->cm == CM_BF16X3)
          step_mfma<CM_BF16X3>(h, s);
        else
          step_mfma<CM_F32>(h, s);
      }
      return;
    }
    dim3 g_k2r(2 * s->nwp / BN, (h->n1 + BM - 1) / BM), g_r2k((2 * s->f1 + BN - 1) / BN, (h->nk + BM - 1) / BM);
    C* kp[5] = {nullptr, nullptr, (C*)s->k1, (C*)s->k2, (C*)s->k3};
    for (int stage = 1; stage <= 4; stage++) {
      launchf(h, UPPE_KC_HANKEL, h->paraxial ? f_gemm<T, PRO_STAGE_ENV, EPI_NONE> : f_gemm<T, PRO_STAGE, EPI_NONE>,
              g_k2r, dim3(256), 0, (const T*)s->Hb,
              (const C*)s->q, (C*)s->er, h->n1, h->nk, s->nwp, (const double*)s->phs, (const C*)s->kzc,
              (const T*)s->cond, (C*)nullptr, (const C*)kp[stage], (C*)nullptr, (C*)nullptr, (C*)nullptr, s->st,
              s->qmax, s->qmax_next, stage, 0.0, 0.0, 0);
      source(h, s, (const C*)s->er, (C*)s->jr, false, stage == 1 ? s->st : nullptr);
      launchf(h, UPPE_KC_HANKEL_R2K, f_gemm<T, PRO_NONE, EPI_STAGE>, g_r2k, dim3(256), 0, (const T*)s->Hf,
              (const C*)s->jr, (C*)nullptr, h->nk, h->n1, s->nwp, (const double*)s->phs, (const C*)s->kzc,
              (const T*)s->cond, (C*)s->q, (const C*)nullptr, (C*)s->k1, (C*)s->k2, (C*)s->k3, s->st, s->qmax,
              s->qmax_next, stage, 0.0, 0.0, 0);
    }
  }

  static void set_field(uppe_handle* h, FusedState* s, const double* a_r) {
    const long P = h->P;
    s->qbase = 0.0;
    UPPE_HIP(hipMemcpy(s->stage, a_r, sizeof(double) * 2 * h->npt, hipMemcpyHostToDevice));
    launchf(h, UPPE_KC_POINT, f_from_host<T>, dim3(grid_for(P * s->nwp)), dim3(kBlock), 0, (const double*)s->stage,
            (C*)s->er, h->nw, P, s->nwp);
    UPPE_HIP(hipMemsetAsync(s->q, 0, sizeof(C) * P * s->nwp, h->stream));
    if (h->d.geometry == UPPE_CARTESIAN) {
      // numpy fft over x, then over y (grid_tools.py:220-229)
      tpass<PRO_NONE, EPI_NONE>(h, s, 1, (const C*)s->er, (C*)s->er, h->nw, -1, nullptr, 0, 0.0, nullptr);
      tpass<PRO_NONE, EPI_NONE>(h, s, 2, (const C*)s->er, (C*)s->q, h->nw, -1, nullptr, 0, 0.0, nullptr);
    } else {
      azim(h, s, (C*)s->er, h->nw, -1);  // FFT_axis2 first (grid_tools.py:308-316)
      gemm_r2k_plain(h, s, (const C*)s->er, (C*)s->q);
    }
    UPPE_HIP(hipMemsetAsync(s->qmax, 0, sizeof(unsigned long long) * kQW, h->stream));
    launchf(h, UPPE_KC_REDUCE, f_qmax<T>, dim3(256), dim3(kBlock), 0, (const C*)s->q, P * s->nwp, s->qmax);
    launchf(h, UPPE_KC_REDUCE, f_qmax_cols<T>, dim3(256), dim3(kBlock), 0, (const C*)s->q, P, s->nwp, s->f1,
            s->qmax + 2);
    UPPE_HIP(hipStreamSynchronize(h->stream));
  }

  // the host copy of the step state of a new advance call (uppe_kernel.py:299-325), with the replayed schedule
  // uploaded; the caller uploads the state itself
  // move the state to reference point qn under the loaded medium (a launch on the handle's stream, waited for)
  static void rereference(uppe_handle* h, FusedState* s, double qn) {
    if (s->qbase == qn) return;
    const long n = (long)h->P * s->nwp;
    launchf(h, UPPE_KC_POINT, f_reref<T>, dim3(grid_for(n)), dim3(kBlock), 0, (C*)s->q, (const C*)s->kzc, s->f1,
            s->qbase, (const C*)s->kzc, s->f1, qn, n, s->nwp);
    UPPE_HIP(hipStreamSynchronize(h->stream));
    s->qbase = qn;
  }

  // The right-hand side carries exp(+Im kz z) (uppe.cl:86-90). FP32 runs keep it in range by moving the state's
  // reference point with the run (rebase): the factors then span at most one subcycle plus kRebaseSpan. A run
  // that replays its own steps (the checkpointed forward run of the adjoint) keeps the state itself and takes
  // the whole interval's range.
  static StepState* init_state(uppe_handle* h, FusedState* s, const uppe_step_control* c, bool rebase_ok = true) {
    const bool rebase = sizeof(T) == 4 && rebase_ok;
    if (sizeof(T) == 4) {
      if (rebase) {
        if (s->max_ki * c->length > 40.0)
          throw Error{UPPE_ERR_UNSUPPORTED, "FP32 range: max Im kz times the subcycle length is " +
                                                std::to_string(s->max_ki * c->length) +
                                                " (at most 40); use FP64 or more subcycles"};
      } else {
        if (s->max_ki * std::fabs(s->qbase) > 80.0)
          throw Error{UPPE_ERR_UNSUPPORTED, "FP32 range: the state's reference point is too far for this run (" +
                                                std::to_string(s->max_ki * std::fabs(s->qbase)) +
                                                "); set the field again"};
        rereference(h, s, 0.0);
        const double zmax = ((c->zi + (double)(c->sub0 + c->n_sub) * c->length) - c->zi);
        if (s->max_ki * zmax > 80.0)
          throw Error{UPPE_ERR_UNSUPPORTED,
                      "FP32 range: exp(Im kz z) would exceed 3e38 in this interval (max Im kz * z = " +
                          std::to_string(s->max_ki * zmax) + "); use FP64 or more diagnostic planes"};
      }
    }
    StepState* sh = s->st_host;
    std::memset(sh, 0, sizeof(StepState));
    sh->zi = c->zi;
    sh->L = c->length;
    sh->dzmin = c->dzmin;
    sh->dphi = c->dphi;
    sh->rel_amin = c->rel_amin;
    sh->sub = c->sub0;
    sh->sub_end = c->sub0 + c->n_sub;
    sh->z = 0.0;
    sh->zoff = (c->zi + (double)c->sub0 * c->length) - c->zi;
    sh->active = 1;
    sh->sched_n = c->n_schedule;
    sh->fixed = (c->n_schedule == 0 && c->dzmin >= c->length) ? 1 : 0;
    sh->dzraw = 0x7ff0000000000000ull;  // +inf
    sh->rec_cap = 0;
    sh->max_steps = c->max_steps > 0 ? c->max_steps : 0;
    sh->qbase = rebase ? s->qbase : 0.0;
    sh->rb_kmax = s->max_ki;
    sh->rebase = rebase ? 1 : 0;
    sh->rb_f1 = rebase ? s->f1 : 0;
    if (c->n_schedule > s->sched_cap) {
      if (s->sched) UPPE_HIP(hipFree(s->sched));
      UPPE_HIP(hipMalloc(&s->sched, sizeof(double) * c->n_schedule));
      s->sched_cap = c->n_schedule;
    }
    if (c->n_schedule > 0)
      UPPE_HIP(hipMemcpy(s->sched, c->dz_schedule, sizeof(double) * c->n_schedule, hipMemcpyHostToDevice));
    sh->pend = 0;
    sh->sched = c->n_schedule > 0 ? s->sched : nullptr;
    return sh;
  }

  // commit the last step of a chunk (the steps before it were committed by the next step's source launch)
  static void commit(uppe_handle* h, FusedState* s) {
    launchf(h, UPPE_KC_REDUCE, f_ctl_commit, dim3(1), dim3(1), 0, s->st, s->qmax, s->rec_active, s->zrec_active,
            s->zrec_cap / 3);
  }

  template <typename V>
  static void grow(V** p, int* cap, int need, size_t per, hipStream_t stream) {
    if (need <= *cap) return;
    int n = std::max(need, 2 * *cap);
    V* q = nullptr;
    UPPE_HIP(hipMalloc(&q, per * n));
    if (*p) {
      UPPE_HIP(hipMemcpyAsync(q, *p, per * *cap, hipMemcpyDeviceToDevice, stream));
      UPPE_HIP(hipStreamSynchronize(stream));
      UPPE_HIP(hipFree(*p));
    }
    *p = q;
    *cap = n;
  }

  // forward run that keeps what the adjoint needs (radial MFMA path, no rotator): q, the step state and qmax at the
  // start of every segment of `every` steps, (zoff, z, dz) of every step, and the stage fields E_r of the last
  // segment
  template <int CM>
  static int advance_ck(uppe_handle* h, FusedState* s, const uppe_step_control* c, int every) {
    const long P = h->P, stride = P * (long)s->nwp;
    StepState* sh = init_state(h, s, c, false);  // the checkpoints hold the state itself
    s->rec_active = nullptr;
    UPPE_HIP(hipMemcpyAsync(s->st, sh, sizeof(StepState), hipMemcpyHostToDevice, h->stream));
    tune_single<T>(h, s);
    if (s->er_cap < 4 * every) {
      if (s->er_store) UPPE_HIP(hipFree(s->er_store));
      UPPE_HIP(hipMalloc(&s->er_store, sizeof(C) * stride * 4 * every));
      s->er_cap = 4 * every;
    }
    s->recorded = false;
    int seg = 0;
    while (true) {
      grow(&s->zrec, &s->zrec_cap, 3 * (seg + 1) * every, sizeof(double), h->stream);
      grow((char**)&s->ck_q, &s->ckq_cap, seg + 1, sizeof(C) * stride, h->stream);
      grow(&s->ck_st, &s->ckst_cap, seg + 1, sizeof(StepState), h->stream);
      grow(&s->ck_qmax, &s->ckqm_cap, (long)kQW * (seg + 1), sizeof(unsigned long long), h->stream);
      UPPE_HIP(hipMemcpyAsync((C*)s->ck_q + seg * stride, s->q, sizeof(C) * stride, hipMemcpyDeviceToDevice, h->stream));
      UPPE_HIP(hipMemcpyAsync(s->ck_st + seg, s->st, sizeof(StepState), hipMemcpyDeviceToDevice, h->stream));
      UPPE_HIP(hipMemcpyAsync(s->ck_qmax + (long)kQW * seg, s->qmax, kQW * sizeof(unsigned long long), hipMemcpyDeviceToDevice,
                              h->stream));
      s->zrec_active = s->zrec;
      for (int i = 0; i < every; i++) step_mfma<CM>(h, s, (C*)s->er_store + (long)i * 4 * stride);
      commit(h, s);
      s->zrec_active = nullptr;
      UPPE_HIP(hipMemcpyAsync(sh, s->st, sizeof(StepState), hipMemcpyDeviceToHost, h->stream));
      UPPE_HIP(hipStreamSynchronize(h->stream));
      seg++;
      if (!sh->active) break;
      if ((long long)seg * every > 100000000LL) throw Error{UPPE_ERR_STATE, "step loop did not terminate"};
    }
    if (sh->error) throw Error{UPPE_ERR_ARG, "dz schedule exhausted"};
    const int steps = sh->steps;
    s->ck_every = every;
    s->ck_n = (steps + every - 1) / every;
    s->rec_steps = steps;
    s->zrec_host.resize((size_t)3 * steps);
    if (steps > 0)
      UPPE_HIP(hipMemcpy(s->zrec_host.data(), s->zrec, sizeof(double) * 3 * steps, hipMemcpyDeviceToHost));
    if (c->dz_record)
      for (int n = 0; n < std::min(steps, c->record_cap); n++) c->dz_record[n] = s->zrec_host[3 * n + 2];
    // the stage fields in store belong to the last segment when it started at a segment boundary
    s->er_last_valid = steps > 0 && (seg == s->ck_n);
    s->recorded = true;
    h->steps += steps;
    h->evals += 4LL * steps;
    h->nan_steps += sh->nan_steps;
    h->z_last = sh->z_last;
    return steps;
  }

  // reverse sweep: g_a_r is the cotangent of finish's r-space field (dL/dRe a + i dL/dIm a), g_a0_r receives the
  // cotangent of the field given to set_field
  template <int CM>
  static void adjoint_cm(uppe_handle* h, FusedState* s, double z_total, const double* g_a_r, double* g_a0_r) {
    const long P = h->P, stride = P * (long)s->nwp;
    const int nr = h->n1, nk = h->nk, every = s->ck_every, steps = s->rec_steps;
    const std::vector<double>& zr = s->zrec_host;
    void** work[] = {&s->gs, &s->gq, &s->gk, &s->g2, &s->g3};
    for (void** w : work)
      if (!*w) UPPE_HIP(hipMalloc(w, sizeof(C) * stride));
    const FusedState::HkPlan pl_ak = plan_hankel(nr, s->f1, s->n_cu), pl_ar = plan_hankel(nk, h->nw, s->n_cu),
                             pl_a0 = plan_hankel(nr, h->nw, s->n_cu);
    C *gs = (C*)s->gs, *gq = (C*)s->gq, *gk = (C*)s->gk, *g2 = (C*)s->g2, *g3 = (C*)s->g3;
    UPPE_HIP(hipMemsetAsync(g2, 0, sizeof(C) * stride, h->stream));  // columns >= f1 stay zero
    UPPE_HIP(hipMemsetAsync(g3, 0, sizeof(C) * stride, h->stream));
    UPPE_HIP(hipMemcpyAsync(s->stage, g_a_r, sizeof(double) * 2 * h->npt, hipMemcpyHostToDevice, h->stream));
    launchf(h, UPPE_KC_POINT, f_from_host<T>, dim3(grid_for(P * s->nwp)), dim3(kBlock), 0, (const double*)s->stage,
            g3, h->nw, P, s->nwp);
    // azimuthal points: the adjoint of ifft over phi (fft with sign -1 and 1/N2) takes the cotangent to the modes
    azim(h, s, g3, h->nw, -1, 1.0 / h->n2);
    // finish: a = Hb (e^{i kz z} q)
    hk_launch<CM, PRO_NONE, EPI_ADJ_FIN>(h, s, s->faf_hi, s->faf_lo, s->kp_f, nk, nr, h->nw, pl_ar, g3, nullptr,
                                         nullptr, gq, gs, gk, nullptr, nullptr, 0, z_total, 0.0,
                                         zr[3 * (steps - 1) + 2]);
    for (int seg = s->ck_n - 1; seg >= 0; seg--) {
      const int n0 = seg * every, n1 = std::min(steps, n0 + every);
      if (seg != s->ck_n - 1 || !s->er_last_valid) {
        // replay the segment from its checkpoint, keeping the stage fields
        UPPE_HIP(hipMemcpyAsync(s->q, (C*)s->ck_q + seg * stride, sizeof(C) * stride, hipMemcpyDeviceToDevice,
                                h->stream));
        UPPE_HIP(hipMemcpyAsync(s->st, s->ck_st + seg, sizeof(StepState), hipMemcpyDeviceToDevice, h->stream));
        UPPE_HIP(hipMemcpyAsync(s->qmax, s->ck_qmax + (long)kQW * seg, kQW * sizeof(unsigned long long),
                                hipMemcpyDeviceToDevice, h->stream));
        s->rec_active = nullptr;
        s->zrec_active = nullptr;
        for (int i = 0; i < n1 - n0; i++) step_mfma<CM>(h, s, (C*)s->er_store + (long)i * 4 * stride);
      }
      for (int n = n1 - 1; n >= n0; n--) {
        const double zoff = zr[3 * n], z = zr[3 * n + 1], dz = zr[3 * n + 2];
        for (int stage = 4; stage >= 1; stage--) {
          const double cs = stage == 1 ? 0.0 : (stage == 4 ? 1.0 : 0.5);
          const double zph = zoff + (z + cs * dz);
          const C* ers = (const C*)s->er_store + ((long)(n - n0) * 4 + (stage - 1)) * stride;
          hk_launch<CM, PRO_ADJ, EPI_NONE>(h, s, s->fab_hi, s->fab_lo, s->kp_b, nr, nk, s->f1, pl_ak, gk, g2, nullptr,
                                           nullptr, nullptr, nullptr, nullptr, nullptr, stage, zph, 0.0, 0.0);
          azim(h, s, g2, s->f1, +1, 1.0);  // adjoint of the fft over phi of J_r: the unscaled sign +1 transform
          source_adj(h, s, ers, g2, g3);
          azim(h, s, g3, h->nw, -1, 1.0 / h->n2);  // adjoint of the ifft over phi of E_r
          const double dzprev = (stage == 1 && n > 0) ? zr[3 * (n - 1) + 2] : 0.0;
          hk_launch<CM, PRO_NONE, EPI_ADJ>(h, s, s->faf_hi, s->faf_lo, s->kp_f, nk, nr, h->nw, pl_ar, g3, nullptr,
                                           nullptr, gq, gs, gk, nullptr, nullptr, stage, zph, dz, dzprev);
        }
      }
    }
    // set_field: q0 = Hf fft_phi(a0)
    hk_launch<CM, PRO_NONE, EPI_NONE>(h, s, s->fab_hi, s->fab_lo, s->kp_b, nr, nk, h->nw, pl_a0, gq, g2, nullptr,
                                      nullptr, nullptr, nullptr, nullptr, nullptr, 0, 0.0, 0.0, 0.0);
    azim(h, s, g2, h->nw, +1, 1.0);
    launchf(h, UPPE_KC_POINT, f_to_host<T>, dim3(grid_for(P * h->nw)), dim3(kBlock), 0, (const C*)g2, s->stage,
            h->nw, P, s->nwp, (const C*)nullptr, 0, 0.0);
    UPPE_HIP(hipMemcpyAsync(g_a0_r, s->stage, sizeof(double) * 2 * h->npt, hipMemcpyDeviceToHost, h->stream));
    UPPE_HIP(hipStreamSynchronize(h->stream));
  }

  static void check_adjoint(uppe_handle* h, FusedState* s) {
    if (h->d.geometry != UPPE_RADIAL || s->cm < 0)
      throw Error{UPPE_ERR_UNSUPPORTED, "adjoint: radial geometry on the matrix-core path"};
    if (h->paraxial) throw Error{UPPE_ERR_UNSUPPORTED, "adjoint: the UPPE model only"};
    if (h->nt > kMaxNtTables) throw Error{UPPE_ERR_UNSUPPORTED, "adjoint: Nt <= 4096"};
  }

  static int advance_checkpointed(uppe_handle* h, FusedState* s, const uppe_step_control* c, int every) {
    check_adjoint(h, s);
    if constexpr (sizeof(T) == 8) {
      return advance_ck<CM_F64>(h, s, c, every);
    } else {
      if (s->cm == CM_BF16X3) return advance_ck<CM_BF16X3>(h, s, c, every);
      return advance_ck<CM_F32>(h, s, c, every);
    }
  }

  static void adjoint(uppe_handle* h, FusedState* s, double z_total, const double* g_a_r, double* g_a0_r) {
    check_adjoint(h, s);
    if (!s->recorded || s->rec_steps < 1) throw Error{UPPE_ERR_STATE, "adjoint: no checkpointed forward run"};
    if constexpr (sizeof(T) == 8) {
      adjoint_cm<CM_F64>(h, s, z_total, g_a_r, g_a0_r);
    } else {
      if (s->cm == CM_BF16X3)
        adjoint_cm<CM_BF16X3>(h, s, z_total, g_a_r, g_a0_r);
      else
        adjoint_cm<CM_F32>(h, s, z_total, g_a_r, g_a0_r);
    }
  }

  // bytes of the static workspace one uppe_advance_static call takes: the five work arrays (slopes k1..k3, E_r,
  // J_r), the replayed schedule and the dz record, each on a 256-byte boundary
  static size_t static_need(const uppe_handle* h, const FusedState* s, const uppe_step_control* c) {
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t cs = al(sizeof(C) * (size_t)h->P * s->nwp);
    const size_t ns = c->n_schedule > 0 ? (size_t)c->n_schedule : 0;
    const size_t nr = (c->dz_record && c->record_cap > 0) ? (size_t)c->record_cap : 0;
    return 5 * cs + al(sizeof(double) * ns) + al(sizeof(double) * nr);
  }

  // advance with the work arrays, schedule and record carved out of the static workspace ws (its device
  // address) instead of the handle's allocations; nothing is allocated or freed. The slopes are cleared first
  // (columns the r2k direction never writes must read zero), with a fill kernel.
  static int advance_static(uppe_handle* h, FusedState* s, const uppe_step_control* c, char* ws, size_t ws_bytes) {
    const size_t need = static_need(h, s, c);
    if (need > ws_bytes)
      throw Error{UPPE_ERR_ARG, "uppe_advance_static: this grid needs " + std::to_string(need) +
                                    " bytes of workspace; the static workspace holds " + std::to_string(ws_bytes)};
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t cs = al(sizeof(C) * (size_t)h->P * s->nwp);
    struct Swap {
      FusedState* s;
      uppe_handle* h;
      void *k1, *k2, *k3, *er, *jr;
      double *sched, *rec;
      int sched_cap, rec_cap;
      bool prof;
      ~Swap() {
        s->k1 = k1;
        s->k2 = k2;
        s->k3 = k3;
        s->er = er;
        s->jr = jr;
        s->sched = sched;
        s->rec = rec;
        s->sched_cap = sched_cap;
        s->rec_cap = rec_cap;
        s->rec_active = nullptr;
        h->prof = prof;
      }
    } keep{s, h, s->k1, s->k2, s->k3, s->er, s->jr, s->sched, s->rec, s->sched_cap, s->rec_cap, h->prof};
    h->prof = false;  // event timing would create events
    s->k1 = ws;
    s->k2 = ws + cs;
    s->k3 = ws + 2 * cs;
    s->er = ws + 3 * cs;
    s->jr = ws + 4 * cs;
    s->sched = reinterpret_cast<double*>(ws + 5 * cs);
    s->sched_cap = c->n_schedule > 0 ? c->n_schedule : 0;
    s->rec = reinterpret_cast<double*>(ws + 5 * cs + al(sizeof(double) * (size_t)s->sched_cap));
    s->rec_cap = (c->dz_record && c->record_cap > 0) ? c->record_cap : 0;
    const long words = (long)(3 * cs / 8);
    launchf(h, UPPE_KC_POINT, f_fill_zero, dim3((unsigned)std::min<long>(1024, (words + 255) / 256)), dim3(256), 0,
            reinterpret_cast<unsigned long long*>(ws), words);
    return advance(h, s, c, false);
  }

  static int advance(uppe_handle* h, FusedState* s, const uppe_step_control* c, bool may_time = true) {
    StepState* sh = init_state(h, s, c);
    sh->rec_cap = c->dz_record ? c->record_cap : 0;
    if (sh->rec_cap > s->rec_cap) {
      if (s->rec) UPPE_HIP(hipFree(s->rec));
      UPPE_HIP(hipMalloc(&s->rec, sizeof(double) * sh->rec_cap));
      s->rec_cap = sh->rec_cap;
    }
    double* rec_dev = sh->rec_cap > 0 ? s->rec : nullptr;
    s->rec_active = rec_dev;
    UPPE_HIP(hipMemcpyAsync(s->st, sh, sizeof(StepState), hipMemcpyHostToDevice, h->stream));
    tune_single<T>(h, s, may_time);
    long long guard = 0;
    int chunk = c->n_sub;
    if (sh->max_steps > 0) chunk = std::min(chunk, sh->max_steps);
    while (true) {
      for (int i = 0; i < chunk; i++) step(h, s);
      commit(h, s);
      UPPE_HIP(hipMemcpyAsync(sh, s->st, sizeof(StepState), hipMemcpyDeviceToHost, h->stream));
      UPPE_HIP(hipStreamSynchronize(h->stream));
      guard += chunk;
      if (!sh->active) break;
      if (guard > 100000000LL) throw Error{UPPE_ERR_STATE, "step loop did not terminate"};
      // remaining subcycles times the observed steps per subcycle
      int done_sub = sh->sub - c->sub0;
      double per = done_sub > 0 ? (double)sh->steps / done_sub : (double)sh->steps + 1.0;
      chunk = std::max(1, (int)std::min(1e6, per * (sh->sub_end - sh->sub) * 0.5));
      if (sh->max_steps > 0) chunk = std::max(1, std::min(chunk, sh->max_steps - sh->steps));
    }
    if (sh->error) throw Error{UPPE_ERR_ARG, "dz schedule exhausted"};
    if (rec_dev && c->dz_record) {
      int n = std::min(sh->steps, sh->rec_cap);
      UPPE_HIP(hipMemcpy(c->dz_record, rec_dev, sizeof(double) * n, hipMemcpyDeviceToHost));
    }
    h->steps += sh->steps;
    h->evals += 4LL * sh->steps;
    h->nan_steps += sh->nan_steps;
    h->z_last = sh->z_last;
    s->qbase = sh->qbase;
    return sh->steps;
  }

  static void finish(uppe_handle* h, FusedState* s, double z_total, double* a_r, double* j_r, double* chi_t,
                     double* ne_t) {
    const long P = h->P, n = P * s->nwp;
    if (h->d.geometry == UPPE_CARTESIAN) {
      // linear propagation to z_total on load, then numpy ifft over y and over x
      tpass<PRO_FINISH, EPI_NONE>(h, s, 2, (const C*)s->q, (C*)s->er, h->nw, +1, nullptr, 0, z_total, nullptr);
      tpass<PRO_NONE, EPI_NONE>(h, s, 1, (const C*)s->er, (C*)s->er, h->nw, +1, nullptr, 0, z_total, nullptr);
    } else {
      dim3 g_k2r(2 * s->nwp / BN, (h->n1 + BM - 1) / BM);
      const int row = h->n2 * s->nwp;
      for (int m = 0; m < h->n2; m++) {
        const long o = (long)m * s->nwp;
        launchf(h, UPPE_KC_HANKEL, f_gemm<T, PRO_FINISH, EPI_NONE>, g_k2r, dim3(256), 0,
                (const T*)s->Hb + (size_t)m * h->n1 * h->nk, (const C*)s->q + o, (C*)s->er + o, h->n1, h->nk, row,
                (const double*)s->phs + o, (const C*)s->kzc + o, (const T*)s->cond, (C*)nullptr, (const C*)nullptr,
                (C*)nullptr, (C*)nullptr, (C*)nullptr, (StepState*)nullptr, (unsigned long long*)nullptr,
                (unsigned long long*)nullptr, 0, z_total, s->qbase, s->f1);
      }
      azim(h, s, (C*)s->er, h->nw, +1);
    }
    if (h->paraxial) {
      // the envelope source takes A itself (paraxial_kernel.py:243-245)
      source(h, s, (const C*)s->er, (C*)s->jr, true, nullptr);
    } else {
      UPPE_HIP(hipMemcpyAsync(s->k1, s->er, sizeof(C) * n, hipMemcpyDeviceToDevice, h->stream));
      launchf(h, UPPE_KC_POINT, f_vp_to_e<T>, dim3(grid_for(n)), dim3(kBlock), 0, (C*)s->k1, (const C*)s->kzc, n);
      source(h, s, (const C*)s->k1, (C*)s->jr, true, nullptr);
    }
    auto out_c = [&](const void* src, double* dst) {
      launchf(h, UPPE_KC_POINT, f_to_host<T>, dim3(grid_for(P * h->nw)), dim3(kBlock), 0, (const C*)src, s->stage,
              h->nw, P, s->nwp, (const C*)nullptr, 0, 0.0);
      UPPE_HIP(hipMemcpyAsync(dst, s->stage, sizeof(double) * 2 * h->npt, hipMemcpyDeviceToHost, h->stream));
      UPPE_HIP(hipStreamSynchronize(h->stream));
    };
    auto out_t = [&](const void* src, double* dst) {
      launchf(h, UPPE_KC_POINT, f_time_to_host<T>, dim3(grid_for(P * h->nt)), dim3(kBlock), 0, (const T*)src,
              s->stage, h->nt, P);
      UPPE_HIP(hipMemcpyAsync(dst, s->stage, sizeof(double) * h->ntt, hipMemcpyDeviceToHost, h->stream));
      UPPE_HIP(hipStreamSynchronize(h->stream));
    };
    if (a_r) out_c(s->er, a_r);
    if (j_r) out_c(s->jr, j_r);
    if (chi_t) out_t(s->chi_t, chi_t);
    if (ne_t) out_t(s->ne_t, ne_t);
    UPPE_HIP(hipStreamSynchronize(h->stream));
    // finish leaves k1 holding the r-space field: clear it so the next run starts from zero stage buffers
    UPPE_HIP(hipMemset(s->k1, 0, sizeof(C) * n));
  }

  static void get_state(uppe_handle* h, FusedState* s, double* q) {
    const long P = h->P;
    launchf(h, UPPE_KC_POINT, f_to_host<T>, dim3(grid_for(P * h->nw)), dim3(kBlock), 0, (const C*)s->q, s->stage,
            h->nw, P, s->nwp, s->qbase != 0.0 ? (const C*)s->kzc : (const C*)nullptr, s->f1, s->qbase);
    UPPE_HIP(hipMemcpyAsync(q, s->stage, sizeof(double) * 2 * h->npt, hipMemcpyDeviceToHost, h->stream));
    UPPE_HIP(hipStreamSynchronize(h->stream));
  }
};

// ---------------------------------------------------------------------------------------------------
// Batches: many radial instances stepped together. Members with the same (Nw, Nr, Nk, precision) form a group
// whose per-instance arrays lie in one slab per kind, a fixed stride apart, so that one launch of each step
// kernel covers a run of consecutive members (grid z for the contractions, grid y for the source). A run
// also needs the same nonlinear band and frequency tables; members whose tables differ start a new run. The
// contraction plans and the source kernel's samples per thread of every run are chosen by timing the
// candidates on the run itself (autotune), or taken from the single-instance heuristic.
// ---------------------------------------------------------------------------------------------------
struct BatchGroup {
  int nw = 0, n1 = 0, nk = 0, prec = 0, cm = -1;
  std::vector<int> members;         // member index of each slot
  void* slab[SL_COUNT] = {};
  size_t stride[SL_COUNT] = {};     // bytes between consecutive slots
  SrcInst* src = nullptr;           // device, one per slot
  StepState* st_host = nullptr;     // pinned, one per slot
};

struct BatchRun {
  int group = 0, first = 0, count = 0;  // slots [first, first + count) of the group
  FusedState::HkPlan plan[2];           // k2r, r2k
  int spt = 0;                          // source samples per thread (0: the single-instance choice)
  int share = 0;  // bit 0: the members' operators agree, bit 1: their kz tables agree (one copy is read)
};

struct FusedBatch {
  std::vector<BatchGroup> groups;
  std::vector<MemberSlots> slots;  // per member (addresses handed to the member states)
  std::vector<int> group_of, slot_of;
  std::vector<BatchRun> runs;      // runs of the last advance
  // tuned launch shapes by run: (group, first, count, f1) -> plans and samples per thread
  std::vector<std::pair<std::vector<long>, BatchRun>> tuned;
  bool autotune = true;
  // HIP graph of graph_steps steps of every run (0: launch step by step), launched on `main`; the runs'
  // streams fork from it and join it again inside the graph
  int graph_steps = 8;
  hipStream_t main = nullptr;
  hipEvent_t fork = nullptr;
  std::vector<hipEvent_t> joins;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

static int cm_of(int prec) {
  return prec == UPPE_PREC_FP64 ? CM_F64 : (prec == UPPE_PREC_BF16X3 ? CM_BF16X3 : CM_F32);
}

void fused_batch_create(uppe_batch* b, const uppe_desc* descs, int n, MemberFactory make_member) {
  auto* fb = new FusedBatch();
  b->fb = fb;
  fb->slots.resize((size_t)n);
  fb->group_of.assign((size_t)n, -1);
  fb->slot_of.assign((size_t)n, -1);
  for (int i = 0; i < n; i++) {
    const uppe_desc& d = descs[i];
    if (d.geometry != UPPE_RADIAL || d.model != UPPE_MODEL_UPPE || d.n2 != 1)
      throw Error{UPPE_ERR_UNSUPPORTED, "batch member " + std::to_string(i) + ": radial UPPE grids only"};
    if (d.precision < UPPE_PREC_FP64 || d.precision > UPPE_PREC_BF16X3)
      throw Error{UPPE_ERR_UNSUPPORTED, "batch member " + std::to_string(i) + ": fused precisions only"};
    if (d.device != descs[0].device) throw Error{UPPE_ERR_ARG, "batch members must share one device"};
    if (d.nk < 1 || d.nk > d.n1 || d.n1 < 1 || d.nw < 3) throw Error{UPPE_ERR_ARG, "batch member grid"};
    int g = 0;
    for (; g < (int)fb->groups.size(); g++) {
      const BatchGroup& G = fb->groups[(size_t)g];
      if (G.nw == d.nw && G.n1 == d.n1 && G.nk == d.nk && G.prec == d.precision) break;
    }
    if (g == (int)fb->groups.size()) {
      BatchGroup G;
      G.nw = d.nw;
      G.n1 = d.n1;
      G.nk = d.nk;
      G.prec = d.precision;
      G.cm = cm_of(d.precision);
      fb->groups.push_back(G);
    }
    fb->group_of[(size_t)i] = g;
    fb->slot_of[(size_t)i] = (int)fb->groups[(size_t)g].members.size();
    fb->groups[(size_t)g].members.push_back(i);
  }
  for (BatchGroup& G : fb->groups) {
    const size_t el = G.prec == UPPE_PREC_FP64 ? 8 : 4, nwp = (size_t)((G.nw + 31) / 32) * 32, P = (size_t)G.n1;
    const size_t cs = 2 * el * P * nwp;
    const size_t fbytes = G.cm == CM_F64 ? 32 : 16;  // one fragment: 4 doubles, 4 floats or 8 bf16
    G.stride[SL_PHS] = 8 * P * nwp;
    G.stride[SL_NG] = 8 * P;
    for (int k : {SL_KZC, SL_Q, SL_K1, SL_K2, SL_K3, SL_ER, SL_JR}) G.stride[k] = cs;
    G.stride[SL_ST] = sizeof(StepState);
    G.stride[SL_QMAX] = kQW * sizeof(unsigned long long);
    G.stride[SL_FRB_HI] = fbytes * (size_t)frag_count(G.cm, G.n1, G.nk);
    G.stride[SL_FRF_HI] = fbytes * (size_t)frag_count(G.cm, G.nk, G.n1);
    G.stride[SL_FRB_LO] = G.cm == CM_BF16X3 ? G.stride[SL_FRB_HI] : 0;
    G.stride[SL_FRF_LO] = G.cm == CM_BF16X3 ? G.stride[SL_FRF_HI] : 0;
    const size_t cnt = G.members.size();
    for (int k = 0; k < SL_COUNT; k++) {
      if (!G.stride[k]) continue;
      UPPE_HIP(hipMalloc(&G.slab[k], G.stride[k] * cnt));
      UPPE_HIP(hipMemset(G.slab[k], 0, G.stride[k] * cnt));
    }
    UPPE_HIP(hipMalloc(&G.src, sizeof(SrcInst) * cnt));
    UPPE_HIP(hipHostMalloc(&G.st_host, sizeof(StepState) * cnt, hipHostMallocDefault));
    for (size_t j = 0; j < cnt; j++) {
      MemberSlots& ms = fb->slots[(size_t)G.members[j]];
      for (int k = 0; k < SL_COUNT; k++)
        ms.p[k] = G.stride[k] ? static_cast<char*>(G.slab[k]) + j * G.stride[k] : nullptr;
    }
  }
  b->members.assign((size_t)n, nullptr);
  for (int i = 0; i < n; i++) {
    uppe_handle* h = nullptr;
    const int rc = make_member(&descs[i], &fb->slots[(size_t)i], &h);
    if (rc != UPPE_OK) throw Error{rc, "batch member " + std::to_string(i) + ": create failed"};
    h->batch = b;
    b->members[(size_t)i] = h;
    if (h->fused->cm != fb->groups[(size_t)fb->group_of[(size_t)i]].cm)
      throw Error{UPPE_ERR_UNSUPPORTED, "batches run on the matrix-core contraction"};
  }
}

void fused_batch_destroy(uppe_batch* b) {
  for (uppe_handle*& h : b->members) {
    if (!h) continue;
    h->batch = nullptr;
    (void)uppe_destroy(h);
    h = nullptr;
  }
  FusedBatch* fb = b->fb;
  if (!fb) return;
  for (BatchGroup& G : fb->groups) {
    for (void* p : G.slab)
      if (p) (void)hipFree(p);
    if (G.src) (void)hipFree(G.src);
    if (G.st_host) (void)hipHostFree(G.st_host);
  }
  if (fb->exec) (void)hipGraphExecDestroy(fb->exec);
  if (fb->graph) (void)hipGraphDestroy(fb->graph);
  for (hipEvent_t e : fb->joins) (void)hipEventDestroy(e);
  if (fb->fork) (void)hipEventDestroy(fb->fork);
  if (fb->main) (void)hipStreamDestroy(fb->main);
  delete fb;
  b->fb = nullptr;
}

void fused_batch_set_autotune(uppe_batch* b, int enable) { b->fb->autotune = enable != 0; }

void fused_batch_set_graph(uppe_batch* b, int steps) {
  if (steps < 0 || steps > 4096) throw Error{UPPE_ERR_ARG, "graph steps must be in [0, 4096]"};
  b->fb->graph_steps = steps;
}

void fused_batch_plan(const uppe_batch* b, int member, int* out) {
  const FusedBatch* fb = b->fb;
  const int g = fb->group_of[(size_t)member], j = fb->slot_of[(size_t)member];
  for (const BatchRun& r : fb->runs) {
    if (r.group != g || j < r.first || j >= r.first + r.count) continue;
    const int v[10] = {r.plan[0].tiles, r.plan[0].nwv, r.plan[0].rt, r.plan[1].tiles, r.plan[1].nwv,
                       r.plan[1].rt,    r.spt,         r.count,      r.first,         g};
    std::memcpy(out, v, sizeof(v));
    return;
  }
  throw Error{UPPE_ERR_STATE, "member has not been advanced in this batch"};
}

template <typename T>
struct BatchOps {
  using F = Fused<T>;
  using C = cplx<T>;

  // one RK4 step of a run: step_mfma with every launch covering the run's instances
  template <int CM>
  static void step_run(uppe_batch* b, const BatchRun& r) {
    const BatchGroup& G = b->fb->groups[(size_t)r.group];
    uppe_handle* h = b->members[(size_t)G.members[(size_t)r.first]];
    FusedState* s = h->fused;
    const long zs = h->P * (long)s->nwp;
    const C* kp[5] = {nullptr, nullptr, (const C*)s->k1, (const C*)s->k2, (const C*)s->k3};
    C *q = (C*)s->q, *k1 = (C*)s->k1, *k2 = (C*)s->k2, *k3 = (C*)s->k3;
    for (int stage = 1; stage <= 4; stage++) {
      F::template hk_launch<CM, PRO_STAGE, EPI_NONE>(h, s, s->frb_hi, s->frb_lo, s->kp_b, h->n1, h->nk, h->nw,
                                                     r.plan[0], (const C*)q, (C*)s->er, kp[stage], q, k1, k2, k3,
                                                     s->st, stage, 0.0, 0.0, 0.0, r.count, zs, r.share);
      F::source(h, s, (const C*)s->er, (C*)s->jr, false, stage == 1 ? s->st : nullptr, r.count, G.src + r.first, zs,
                r.spt);
      F::template hk_launch<CM, PRO_NONE, EPI_STAGE>(h, s, s->frf_hi, s->frf_lo, s->kp_f, h->nk, h->n1, s->f1,
                                                     r.plan[1], (const C*)s->jr, (C*)nullptr, nullptr, q, k1, k2,
                                                     k3, s->st, stage, 0.0, 0.0, 0.0, r.count, zs, r.share);
    }
  }

  static void step_run(uppe_batch* b, const BatchRun& r) {
    const int cm = b->fb->groups[(size_t)r.group].cm;
    if constexpr (sizeof(T) == 8) {
      step_run<CM_F64>(b, r);
    } else {
      if (cm == CM_BF16X3)
        step_run<CM_BF16X3>(b, r);
      else
        step_run<CM_F32>(b, r);
    }
  }

  // candidate plans of an m-row contraction over ncols complex columns with contraction length kp: narrow,
  // merged, wide (16 columns) and widest (32 columns, at most 8 waves) column tilings, 1 or 2 row tiles per
  // wave, 1 to 16 waves per workgroup (no wave without rows), within the LDS of one CU.
  // SEARAY_UPPE_TILE_WIDTH (8, 16 or 32 complex columns) keeps only the tilings of that width.
  static std::vector<FusedState::HkPlan> candidates(int cm, int m, int ncols, int kp) {
    std::vector<FusedState::HkPlan> out;
    const int m16 = (m + 15) / 16;
    const char* env = getenv("SEARAY_UPPE_TILE_WIDTH");
    const int only = env ? std::atoi(env) : 0;
    for (int t = 0; t < 4; t++) {
      const int wide = t < 2 ? 0 : t - 1, width = 8 << wide;
      const int tiles = t == 0 ? (ncols + 7) / 8 : (t == 1 ? std::max(1, ncols / 8) : (ncols + width - 1) / width);
      if (t == 1 && tiles == (ncols + 7) / 8) continue;
      if (only && only != width) continue;
      for (int rt = 1; rt <= 2; rt++)
        for (int nwv = 1; nwv <= (wide == 2 ? 8 : 16); nwv *= 2) {
          if (rt * nwv > m16 && !(rt == 1 && nwv == 1)) continue;
          const int bmc = 16 * rt * nwv, bnc = wide == 2 ? 32 : 16;
          if (hk_lds_bytes(cm, hk_chunk(cm, kp, bmc, bnc), bmc, bnc) > kMaxLds) continue;
          if (wide < 2 && (size_t)16 * rt * nwv * 32 * sizeof(T) > 65536) continue;
          FusedState::HkPlan p;
          p.tiles = tiles;
          p.nwv = nwv;
          p.rt = rt;
          p.wide = wide;
          p.split = (m16 + rt * nwv - 1) / (rt * nwv);
          out.push_back(p);
        }
    }
    return out;
  }

  // average time of a launch sequence on stream st (one untimed call first)
  template <typename L>
  static float time_launch(hipStream_t st, hipEvent_t e0, hipEvent_t e1, L&& launch, int reps = 3) {
    launch();
    UPPE_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < reps; i++) launch();
    UPPE_HIP(hipEventRecord(e1, st));
    UPPE_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    UPPE_HIP(hipEventElapsedTime(&ms, e0, e1));
    return ms / reps;
  }

  // choose launch shapes by timing on nz instances zs elements apart starting at member (h, s), as the second
  // RK stage runs them: the k2r contraction with its stage prologue (writes E_r, scratch) and the r2k
  // contraction with its stage epilogue (writes the second slope, which every step writes before it reads it)
  // give the two plans; with src, the source kernel's samples per thread are timed too (it writes J_r,
  // scratch). The state is not touched, and the plans do not change any result (every output sums its
  // contraction in the same order). The incoming plans are the defaults: another candidate replaces one only
  // when it is 3 % faster, so that timing noise does not pick between equals.
  template <int CM>
  static void tune_shapes(uppe_handle* h, FusedState* s, int nz, long zs, const SrcInst* src,
                          FusedState::HkPlan plan[2], int* spt_out, int share = 0) {
    C *q = (C*)s->q, *k1 = (C*)s->k1, *k2 = (C*)s->k2, *k3 = (C*)s->k3;
    hipEvent_t e0, e1;
    UPPE_HIP(hipEventCreate(&e0));
    UPPE_HIP(hipEventCreate(&e1));
    constexpr float margin = 0.97f;
    constexpr int reps = 5;
    try {
      auto k2r = [&](const FusedState::HkPlan& p) {
        return time_launch(h->stream, e0, e1, [&] {
          if (h->paraxial)
            F::template hk_launch<CM, PRO_STAGE_ENV, EPI_NONE>(h, s, s->frb_hi, s->frb_lo, s->kp_b, h->n1, h->nk,
                                                               h->nw, p, (const C*)q, (C*)s->er, (const C*)k1, q, k1,
                                                               k2, k3, s->st, 2, 0.0, 0.0, 0.0, nz, zs, share);
          else
            F::template hk_launch<CM, PRO_STAGE, EPI_NONE>(h, s, s->frb_hi, s->frb_lo, s->kp_b, h->n1, h->nk, h->nw,
                                                           p, (const C*)q, (C*)s->er, (const C*)k1, q, k1, k2, k3,
                                                           s->st, 2, 0.0, 0.0, 0.0, nz, zs, share);
        }, reps);
      };
      auto r2k = [&](const FusedState::HkPlan& p) {
        return time_launch(h->stream, e0, e1, [&] {
          F::template hk_launch<CM, PRO_NONE, EPI_STAGE>(h, s, s->frf_hi, s->frf_lo, s->kp_f, h->nk, h->n1, s->f1, p,
                                                         (const C*)s->jr, (C*)nullptr, nullptr, q, k1, k2, k3, s->st,
                                                         2, 0.0, 0.0, 0.0, nz, zs, share);
        }, reps);
      };
      // with the tile width pinned, the default plan is not a candidate: the first candidate is taken
      const bool pinned = getenv("SEARAY_UPPE_TILE_WIDTH") != nullptr;
      float best = pinned ? 1e30f : k2r(plan[0]);
      for (const auto& p : candidates(s->cm, h->n1, h->nw, s->kp_b)) {
        const float t = k2r(p);
        if (t < margin * best) {
          best = t;
          plan[0] = p;
        }
      }
      best = pinned ? 1e30f : r2k(plan[1]);
      for (const auto& p : candidates(s->cm, h->nk, s->f1, s->kp_f)) {
        const float t = r2k(p);
        if (t < margin * best) {
          best = t;
          plan[1] = p;
        }
      }
      if (src && !getenv("SEARAY_UPPE_SOURCE_SPT") && h->rot_states <= 2) {  // a pinned shape is kept, and the
                                                                               // rotator's
        auto source = [&](int spt) {
          return time_launch(h->stream, e0, e1, [&] {
            F::source(h, s, (const C*)s->er, (C*)s->jr, false, nullptr, nz, src, zs, spt);
          }, reps);
        };
        // the incoming choice (0: the default shape, max(min(Nt, 64), min(Nt/4, 1024)) threads)
        int cur = *spt_out > 0 ? *spt_out : h->nt / std::max(std::min(h->nt, 64), std::min(h->nt / 4, 1024));
        best = source(cur);
        for (int spt = 1; spt <= 8; spt *= 2) {
          const int nth = h->nt / spt;
          if (spt == cur || nth > 1024 || nth < 64) continue;
          const float t = source(spt);
          if (t < margin * best) {
            best = t;
            cur = spt;
          }
        }
        *spt_out = cur;
      }
    } catch (...) {
      (void)hipEventDestroy(e0);
      (void)hipEventDestroy(e1);
      throw;
    }
    UPPE_HIP(hipEventDestroy(e0));
    UPPE_HIP(hipEventDestroy(e1));
  }

  static void tune_shapes(uppe_handle* h, FusedState* s, int nz, long zs, const SrcInst* src,
                          FusedState::HkPlan plan[2], int* spt_out, int share = 0) {
    if constexpr (sizeof(T) == 8) {
      tune_shapes<CM_F64>(h, s, nz, zs, src, plan, spt_out, share);
    } else {
      if (s->cm == CM_BF16X3)
        tune_shapes<CM_BF16X3>(h, s, nz, zs, src, plan, spt_out, share);
      else
        tune_shapes<CM_F32>(h, s, nz, zs, src, plan, spt_out, share);
    }
  }

  // the launch shapes of batch run r
  static void tune(uppe_batch* b, BatchRun& r) {
    const BatchGroup& G = b->fb->groups[(size_t)r.group];
    uppe_handle* h = b->members[(size_t)G.members[(size_t)r.first]];
    tune_shapes(h, h->fused, r.count, h->P * (long)h->fused->nwp, G.src + r.first, r.plan, &r.spt, r.share);
  }
};

// the source kernel's samples per thread without timing, for a launch over `strips` strips: the widest
// workgroup (Nt/1024 samples per thread, at least one) when there are fewer strips than the device has CUs,
// where a strip's latency rather than the device's throughput sets the time, else the default shape (0).
// SEARAY_UPPE_SOURCE_SPT overrides; rotator and envelope sources keep their shape.
static int default_spt(const uppe_handle* h, const FusedState* s, long strips) {
  if (h->rot_states > 2 || h->paraxial) return 0;
  const char* env = getenv("SEARAY_UPPE_SOURCE_SPT");
  const int want = env ? std::atoi(env) : (strips < s->n_cu ? std::max(1, h->nt / 1024) : 0);
  if (want > 0 && (want & (want - 1)) == 0 && want <= 8 && h->nt / want <= 1024 && h->nt / want >= 64) return want;
  return 0;
}

// a standalone radial handle on the matrix cores: the source shape by default_spt; with SEARAY_UPPE_TUNE=1 the
// contraction plans are also timed on the handle once per nonlinear band (the r2k plan depends on it). By
// default a standalone handle keeps plan_hankel's plans, which fit one workgroup per CU: timing them against
// the candidates on the single-instance workloads found no plan faster by more than the run-to-run spread, and
// fixed plans keep every run's launch shapes (and the measured traffic per launch) the same.
template <typename T>
void tune_single(uppe_handle* h, FusedState* s, bool may_time) {
  if (s->cm < 0 || h->d.geometry != UPPE_RADIAL || s->tuned_f1 == s->f1) return;
  if (!may_time) {
    // the source shape only; the plans stay untimed until an advance that may time them
    s->spt = default_spt(h, s, h->P);
    return;
  }
  const char* env = getenv("SEARAY_UPPE_TUNE");
  if (env && env[0] == '1') {
    int unused = 0;
    BatchOps<T>::tune_shapes(h, s, 1, 0, nullptr, s->plan, &unused);
  }
  s->spt = default_spt(h, s, h->P);
  s->tuned_f1 = s->f1;
}

static hipStream_t run_stream(const uppe_batch* b, const BatchRun& r) {
  return b->members[(size_t)b->fb->groups[(size_t)r.group].members[(size_t)r.first]]->stream;
}

static void step_all(uppe_batch* b) {
  for (const BatchRun& r : b->fb->runs) {
    if (b->fb->groups[(size_t)r.group].prec == UPPE_PREC_FP64)
      BatchOps<double>::step_run(b, r);
    else
      BatchOps<float>::step_run(b, r);
  }
}

// capture `steps` steps of every run into one executable graph: the runs' streams fork from the main stream and
// join it at the end, so runs of different groups still overlap inside the graph
static void build_graph(uppe_batch* b, int steps) {
  FusedBatch* fb = b->fb;
  if (fb->exec) UPPE_HIP(hipGraphExecDestroy(fb->exec));
  fb->exec = nullptr;
  if (fb->graph) UPPE_HIP(hipGraphDestroy(fb->graph));
  fb->graph = nullptr;
  if (!fb->main) UPPE_HIP(hipStreamCreateWithFlags(&fb->main, hipStreamNonBlocking));
  if (!fb->fork) UPPE_HIP(hipEventCreateWithFlags(&fb->fork, hipEventDisableTiming));
  while (fb->joins.size() < fb->runs.size()) {
    hipEvent_t e;
    UPPE_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    fb->joins.push_back(e);
  }
  UPPE_HIP(hipStreamBeginCapture(fb->main, hipStreamCaptureModeThreadLocal));
  try {
    UPPE_HIP(hipEventRecord(fb->fork, fb->main));
    for (const BatchRun& r : fb->runs) UPPE_HIP(hipStreamWaitEvent(run_stream(b, r), fb->fork, 0));
    for (int k = 0; k < steps; k++) step_all(b);
    for (size_t i = 0; i < fb->runs.size(); i++) {
      UPPE_HIP(hipEventRecord(fb->joins[i], run_stream(b, fb->runs[i])));
      UPPE_HIP(hipStreamWaitEvent(fb->main, fb->joins[i], 0));
    }
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(fb->main, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  UPPE_HIP(hipStreamEndCapture(fb->main, &fb->graph));
  UPPE_HIP(hipGraphInstantiate(&fb->exec, fb->graph, nullptr, nullptr, 0));
}

// whether members a and b can share one launch: the same nonlinear band, frequency tables and source step
static bool same_run(const uppe_handle* a, const uppe_handle* b) {
  const FusedState *x = a->fused, *y = b->fused;
  return x->f1 == y->f1 && x->w_h == y->w_h && x->nl_h == y->nl_h && x->cond_h == y->cond_h && x->dt == y->dt &&
         (a->rot_states > 2) == (b->rot_states > 2);
}

void fused_batch_advance(uppe_batch* b, const uppe_step_control* c, int* steps_out) {
  FusedBatch* fb = b->fb;
  const int n = (int)b->members.size();
  // member checks and the initial step state of every member (uploaded per group below)
  for (int i = 0; i < n; i++) {
    uppe_handle* h = b->members[(size_t)i];
    FusedState* s = h->fused;
    const std::string who = "batch member " + std::to_string(i) + ": ";
    if (!h->have_hankel || !h->have_medium || !h->have_field)
      throw Error{UPPE_ERR_STATE, who + "needs the Hankel basis, the medium and the field first"};
    if (c[i].n_schedule != 0 || c[i].dz_record)
      throw Error{UPPE_ERR_UNSUPPORTED, who + "batches run adaptive or fixed steps (no schedule, no record)"};
    if (c[i].n_sub < 1 || !(c[i].length > 0.0)) throw Error{UPPE_ERR_ARG, who + "bad step control"};
    const BatchGroup& G = fb->groups[(size_t)fb->group_of[(size_t)i]];
    StepState* sh = s->prec == UPPE_PREC_FP64 ? Fused<double>::init_state(h, s, &c[i])
                                              : Fused<float>::init_state(h, s, &c[i]);
    sh->rec_cap = 0;
    G.st_host[fb->slot_of[(size_t)i]] = *sh;
  }
  for (BatchGroup& G : fb->groups) {
    std::vector<SrcInst> src(G.members.size());
    for (size_t j = 0; j < G.members.size(); j++) {
      const uppe_handle* hm = b->members[(size_t)G.members[j]];
      const FusedState* s = hm->fused;
      src[j].redchi3 = s->redchi3;
      src[j].ion = s->ion;
      src[j].rot_tab = hm->rot_states > 2 ? s->rot_tab : nullptr;
      src[j].rot_states = hm->rot_states;
      src[j].rot_fc = hm->rot_states > 2 ? 0.5 * hm->rot_dalpha / hm->rot_hbar : 0.0;
      src[j].log_pre = s->ion.kind != UPPE_ION_EXPLICIT ? 0.0
                       : G.prec == UPPE_PREC_FP64       ? std::log(s->ion.rate_conv * s->ion.c_pre)
                                                        : std::log2(s->ion.rate_conv * s->ion.c_pre);
    }
    UPPE_HIP(hipMemcpy(G.src, src.data(), sizeof(SrcInst) * src.size(), hipMemcpyHostToDevice));
    UPPE_HIP(hipMemcpy(G.slab[SL_ST], G.st_host, sizeof(StepState) * G.members.size(), hipMemcpyHostToDevice));
  }
  // runs of consecutive slots that share their tables, with their launch shapes
  fb->runs.clear();
  for (int g = 0; g < (int)fb->groups.size(); g++) {
    const BatchGroup& G = fb->groups[(size_t)g];
    for (int j = 0; j < (int)G.members.size();) {
      BatchRun r;
      r.group = g;
      r.first = j;
      const uppe_handle* h0 = b->members[(size_t)G.members[(size_t)j]];
      while (j < (int)G.members.size() && same_run(h0, b->members[(size_t)G.members[(size_t)j]])) j++;
      r.count = j - r.first;
      r.plan[0] = h0->fused->plan[0];
      r.plan[1] = h0->fused->plan[1];
      r.spt = default_spt(h0, h0->fused, (long)r.count * h0->P);
      // members whose operators (or kz tables) have equal fingerprints read the first member's copy: one copy
      // stays in the caches where the run's instances would each stream their own
      r.share = r.count > 1 ? 3 : 0;
      for (int m = r.first + 1; m < r.first + r.count; m++) {
        const FusedState* sm = b->members[(size_t)G.members[(size_t)m]]->fused;
        if (!(sm->op_fp == h0->fused->op_fp)) r.share &= ~1;
        if (!(sm->tab_fp == h0->fused->tab_fp)) r.share &= ~2;
      }
      if (fb->autotune) {
        const std::vector<long> key = {g, r.first, r.count, h0->fused->f1, r.share};
        auto it = std::find_if(fb->tuned.begin(), fb->tuned.end(), [&](const auto& e) { return e.first == key; });
        if (it != fb->tuned.end()) {
          r = it->second;
        } else {
          if (G.prec == UPPE_PREC_FP64)
            BatchOps<double>::tune(b, r);
          else
            BatchOps<float>::tune(b, r);
          fb->tuned.emplace_back(key, r);
        }
      }
      fb->runs.push_back(r);
    }
  }
  // step every run on its first member's stream (runs of different groups overlap) until no member is
  // active; the states come back once per chunk of steps
  int chunk = 1 << 30;
  for (int i = 0; i < n; i++) {
    int m = c[i].n_sub;
    if (c[i].max_steps > 0) m = std::min(m, c[i].max_steps);
    chunk = std::min(chunk, m);
  }
  chunk = std::max(1, chunk);
  // graphs replay graph_steps steps per launch (steps past a member's end are no-ops on its state); event
  // timing of the launches (profiling) takes the step-by-step path
  bool prof = false;
  for (const uppe_handle* h : b->members) prof = prof || h->prof;
  const int gs = prof ? 0 : fb->graph_steps;
  if (gs > 0) build_graph(b, gs);
  long long guard = 0;
  while (true) {
    int ran = 0;
    if (gs > 0) {
      for (; ran < chunk; ran += gs) UPPE_HIP(hipGraphLaunch(fb->exec, fb->main));
    } else {
      for (; ran < chunk; ran++) step_all(b);
    }
    for (const BatchRun& r : fb->runs) {
      // commit the chunk's last step of every member of the run, then bring the states back
      const BatchGroup& G = fb->groups[(size_t)r.group];
      uppe_handle* h0 = b->members[(size_t)G.members[(size_t)r.first]];
      const hipStream_t sr = gs > 0 ? fb->main : run_stream(b, r);
      hipLaunchKernelGGL(f_ctl_commit, dim3((unsigned)r.count), dim3(1), 0, sr, h0->fused->st, h0->fused->qmax,
                         (double*)nullptr, (double*)nullptr, 0);
      UPPE_LAUNCH_CHECK();
      UPPE_HIP(hipMemcpyAsync(G.st_host + r.first, static_cast<StepState*>(G.slab[SL_ST]) + r.first,
                              sizeof(StepState) * r.count, hipMemcpyDeviceToHost, sr));
    }
    if (gs > 0)
      UPPE_HIP(hipStreamSynchronize(fb->main));
    else
      for (const BatchRun& r : fb->runs) UPPE_HIP(hipStreamSynchronize(run_stream(b, r)));
    guard += ran;
    // remaining steps: per member, its remaining subcycles times its observed steps per subcycle
    double left = 0.0;
    bool active = false;
    for (int i = 0; i < n; i++) {
      const StepState& sh = fb->groups[(size_t)fb->group_of[(size_t)i]].st_host[fb->slot_of[(size_t)i]];
      if (!sh.active) continue;
      active = true;
      const int done_sub = sh.sub - c[i].sub0;
      const double per = done_sub > 0 ? (double)sh.steps / done_sub : (double)sh.steps + 1.0;
      double rem = per * (sh.sub_end - sh.sub) * 0.5;
      if (sh.max_steps > 0) rem = std::min(rem, (double)(sh.max_steps - sh.steps));
      left = std::max(left, rem);
    }
    if (!active) break;
    if (guard > 100000000LL) throw Error{UPPE_ERR_STATE, "batch step loop did not terminate"};
    chunk = std::max(1, (int)std::min(1e6, left));
  }
  for (int i = 0; i < n; i++) {
    uppe_handle* h = b->members[(size_t)i];
    const StepState& sh = fb->groups[(size_t)fb->group_of[(size_t)i]].st_host[fb->slot_of[(size_t)i]];
    if (sh.error) throw Error{UPPE_ERR_STATE, "batch member " + std::to_string(i) + ": step control error"};
    h->fused->qbase = sh.qbase;
    h->steps += sh.steps;
    h->evals += 4LL * sh.steps;
    h->nan_steps += sh.nan_steps;
    h->z_last = sh.z_last;
    if (steps_out) steps_out[i] = sh.steps;
  }
}

static bool is_fp64(uppe_handle* h) { return h->fused->prec == UPPE_PREC_FP64; }

void fused_create(uppe_handle* h, const MemberSlots* slots) {
  if (h->d.geometry == UPPE_RADIAL && (h->n2 < 1 || h->n2 > 16 || (h->n2 & (h->n2 - 1))))
    throw Error{UPPE_ERR_UNSUPPORTED, "fused path: radial geometry with 1, 2, 4, 8 or 16 azimuthal points"};
  if (h->d.geometry == UPPE_CARTESIAN &&
      (h->n1 < 2 || h->n2 < 2 || h->n1 > kMaxNxFused || h->n2 > kMaxNxFused))
    throw Error{UPPE_ERR_UNSUPPORTED, "fused Cartesian path supports 2 <= Nx, Ny <= " + std::to_string(kMaxNxFused)};
  if (h->nt > kMaxNtFused || h->nt < 4) throw Error{UPPE_ERR_UNSUPPORTED, "fused path supports 4 <= Nt <= 8192"};
  if (h->paraxial && h->nt > kMaxNtTables)
    throw Error{UPPE_ERR_UNSUPPORTED, "fused envelope path supports Nw <= 4096"};
  if (h->d.geometry == UPPE_RADIAL && h->n1 > kMaxNrFused)
    throw Error{UPPE_ERR_UNSUPPORTED, "fused path supports up to " + std::to_string(kMaxNrFused) + " radial cells"};
  auto* s = new FusedState();
  h->fused = s;
  s->slots = slots;
  s->prec = h->d.precision;
  const char* env = getenv("SEARAY_UPPE_SIMT_HANKEL");
  const bool simt = env && env[0] == '1';
  if (s->prec == UPPE_PREC_FP64)
    s->cm = simt ? -1 : CM_F64;
  else if (s->prec == UPPE_PREC_BF16X3)
    s->cm = CM_BF16X3;
  else
    s->cm = simt ? -1 : CM_F32;
  if (s->cm < 0 && h->d.geometry == UPPE_RADIAL && h->n2 > 1)
    throw Error{UPPE_ERR_UNSUPPORTED, "fused path: azimuthal modes on the matrix-core contraction only"};
  if (is_fp64(h))
    Fused<double>::create(h, s);
  else
    Fused<float>::create(h, s);
}

void fused_destroy(uppe_handle* h) {
  FusedState* s = h->fused;
  if (!s) return;
  void* ptrs[] = {s->phs, s->kzc, s->w, s->nlf, s->cond, s->ng, s->Hb, s->Hf, s->tw, s->q, s->k1, s->k2, s->k3,
                  s->er, s->jr, s->chi_t, s->ne_t, s->stage, s->qmax, s->st, s->sched, s->rec, s->frb_hi,
                  s->frb_lo, s->frf_hi, s->frf_lo, s->tw_x, s->tw_y, s->rot_tab, s->fab_hi, s->fab_lo, s->faf_hi,
                  s->faf_lo, s->ck_q, s->ck_st, s->ck_qmax, s->zrec, s->er_store, s->gs, s->gq, s->gk, s->g2, s->g3};
  for (void* p : ptrs)
    if (p && !slotted(s, p)) (void)hipFree(p);
  if (s->st_host) (void)hipHostFree(s->st_host);
  delete s;
  h->fused = nullptr;
}

template <typename T>
static void rot_rows(uppe_handle* h, FusedState* s) {
  // the source kernel's launch shape: nth threads, SPT samples each
  const int N = h->nt, nth = std::max(std::min(N, 64), N / 4), spt = N / nth;
  const int n = h->rot_states - 2;
  std::vector<RotRow<T>> rows((size_t)n);
  for (int i = 0; i < n; i++) {
    const int sj = i + 2;
    const double wr = h->rot_wdt[2 * sj], wi = h->rot_wdt[2 * sj + 1];
    // E^k = exp(i k W dt / 2) for the complex W dt = wr + i wi (rotations.cl:107)
    auto epow = [&](double k) {
      const double mag = std::exp(-0.5 * wi * k);
      return cplx<T>{(T)(mag * std::cos(0.5 * wr * k)), (T)(mag * std::sin(0.5 * wr * k))};
    };
    RotRow<T>& r = rows[(size_t)i];
    r.wr64 = wr;
    r.wi64 = wi;
    r.c = (T)(h->rot_tjj[sj] * (2.0 * sj * (sj - 1) / (15.0 * (2 * sj - 1))) * 2.0 * h->rot_dalpha);
    r.wr = (T)wr;
    r.wi = (T)wi;
    r.pad = T(0);
    r.e[0] = epow(1.0);
    r.e[1] = epow(2.0);
    for (int b = 0; b < ROT_NB; b++) r.e[2 + b] = epow(2.0 * spt * (double)(1 << b));
  }
  UPPE_HIP(hipMalloc(&s->rot_tab, sizeof(RotRow<T>) * (size_t)n));
  UPPE_HIP(hipMemcpy(s->rot_tab, rows.data(), sizeof(RotRow<T>) * (size_t)n, hipMemcpyHostToDevice));
}

void fused_set_rotator(uppe_handle* h) {
  FusedState* s = h->fused;
  if (s->rot_tab) UPPE_HIP(hipFree(s->rot_tab));
  s->rot_tab = nullptr;
  if (h->rot_states < 3) return;
  if (is_fp64(h))
    rot_rows<double>(h, s);
  else
    rot_rows<float>(h, s);
}

void fused_set_hankel(uppe_handle* h) {
  if (is_fp64(h))
    Fused<double>::set_hankel(h, h->fused);
  else
    Fused<float>::set_hankel(h, h->fused);
}

void fused_set_medium(uppe_handle* h, const uppe_medium* m) {
  if (is_fp64(h))
    Fused<double>::set_medium(h, h->fused, m);
  else
    Fused<float>::set_medium(h, h->fused, m);
}

int fused_advance_checkpointed(uppe_handle* h, const uppe_step_control* c, int every) {
  return is_fp64(h) ? Fused<double>::advance_checkpointed(h, h->fused, c, every)
                    : Fused<float>::advance_checkpointed(h, h->fused, c, every);
}

void fused_adjoint(uppe_handle* h, double z_total, const double* g_a_r, double* g_a0_r) {
  if (is_fp64(h))
    Fused<double>::adjoint(h, h->fused, z_total, g_a_r, g_a0_r);
  else
    Fused<float>::adjoint(h, h->fused, z_total, g_a_r, g_a0_r);
}

void fused_set_paraxial_medium(uppe_handle* h, const uppe_paraxial_medium* m) {
  if (is_fp64(h))
    Fused<double>::set_paraxial_medium(h, h->fused, m);
  else
    Fused<float>::set_paraxial_medium(h, h->fused, m);
}

void fused_set_field(uppe_handle* h, const double* a_r) {
  if (is_fp64(h))
    Fused<double>::set_field(h, h->fused, a_r);
  else
    Fused<float>::set_field(h, h->fused, a_r);
}

int fused_advance(uppe_handle* h, const uppe_step_control* c) {
  return is_fp64(h) ? Fused<double>::advance(h, h->fused, c) : Fused<float>::advance(h, h->fused, c);
}

// ---------------------------------------------------------------------------------------------------
// Static workspace of uppe_advance_static: the work arrays of an advance that allocates nothing. Each GPU the
// library runs on holds one copy until the process ends; every call on that GPU shares it (calls that overlap
// on one GPU race). Host code never passes the array's name to a kernel: its device address comes from
// hipGetSymbolAddress, looked up once per device with that device current.
// ---------------------------------------------------------------------------------------------------
constexpr size_t kStaticWorkspaceBytes = (size_t)48 << 20;
static __device__ __align__(256) unsigned char scratch[kStaticWorkspaceBytes];

static std::mutex g_scratch_mu;
static std::vector<char*> g_scratch_ptr;  // per device ordinal; null until looked up

// the device address of this device's workspace (the current device is h->d.device)
static char* scratch_address(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (!g_scratch_ptr[(size_t)device]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw Error{UPPE_ERR_DEVICE, std::string("static workspace: hipGetSymbolAddress failed: ") +
                                       hipGetErrorString(e)};
    g_scratch_ptr[(size_t)device] = static_cast<char*>(ptr);
  }
  return g_scratch_ptr[(size_t)device];
}

size_t fused_static_need(uppe_handle* h, const uppe_step_control* c) {
  return is_fp64(h) ? Fused<double>::static_need(h, h->fused, c) : Fused<float>::static_need(h, h->fused, c);
}

size_t fused_static_capacity() { return kStaticWorkspaceBytes; }

int fused_advance_static(uppe_handle* h, const uppe_step_control* c) {
  char* ws = scratch_address(h->d.device);
  return is_fp64(h) ? Fused<double>::advance_static(h, h->fused, c, ws, kStaticWorkspaceBytes)
                    : Fused<float>::advance_static(h, h->fused, c, ws, kStaticWorkspaceBytes);
}

void fused_finish(uppe_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t) {
  if (is_fp64(h))
    Fused<double>::finish(h, h->fused, z_total, a_r, j_r, chi_t, ne_t);
  else
    Fused<float>::finish(h, h->fused, z_total, a_r, j_r, chi_t, ne_t);
}

void fused_get_state(uppe_handle* h, double* q) {
  if (is_fp64(h))
    Fused<double>::get_state(h, h->fused, q);
  else
    Fused<float>::get_state(h, h->fused, q);
}

}  // namespace uppe

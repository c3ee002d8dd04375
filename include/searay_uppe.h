/*
 * searay_uppe.h -- C ABI of the MI355X UPPE stepper (libsearay_uppe.so).
 *
 * This is the native boundary under the Python drop-ins for SeaRay's `modules/uppe_kernel.py` and
 * `modules/paraxial_kernel.py`. The Python `track`/`propagator` (searay_amd/uppe_kernel.py,
 * searay_amd/paraxial_kernel.py) keep the reference signatures; underneath, they call these entry points in
 * place of the pyopencl calls listed against each function. One handle steps either equation (uppe_desc.model).
 *
 * Conventions
 *  - Every function returns an int status: 0 = OK, otherwise an UPPE_ERR_* code; uppe_last_error(h) has
 *    the text (a static string for a NULL handle).
 *  - Host arrays use the reference layout: spectral arrays are complex128 interleaved (re, im) with shape
 *    (Nw, N1, N2) in C order, time-domain arrays are float64 (Nt, N1, N2), Nt = 2 (Nw - 1) for UPPE (real
 *    fields) and Nt = Nw for the paraxial envelope (frequencies in the reference's ifftshift order).
 *    Host pointers are owned by the caller and only read or written during the call.
 *  - The handle owns all device memory. A handle is not thread-safe; separate handles are independent.
 *  - No C++ types or exceptions cross this boundary.
 */
#ifndef SEARAY_UPPE_H
#define SEARAY_UPPE_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  UPPE_OK = 0,
  UPPE_ERR_ARG = 1,         /* bad argument: sizes, NULL pointers, grid not 2^n(+1) (uppe_kernel.py:282-285,354-360) */
  UPPE_ERR_DEVICE = 2,      /* HIP runtime failure or no device */
  UPPE_ERR_STATE = 3,       /* call out of order (e.g. advance before set_field) */
  UPPE_ERR_UNSUPPORTED = 4  /* valid request this build does not implement */
};

enum { UPPE_RADIAL = 0, UPPE_CARTESIAN = 1 };

enum {
  UPPE_PREC_FP64_PLAIN = 0, /* one kernel per reference operation, reference layout, FP64, host-driven dz */
  UPPE_PREC_FP64 = 1,       /* fused MI355X kernels in FP64 */
  UPPE_PREC_FP32 = 2,       /* fused kernels in FP32 (FP64 linear phase; FP32 plasma scan as a tree of sums) */
  UPPE_PREC_BF16X3 = 3      /* as FP32, Hankel contraction as three BF16 MFMA products, FP32 accumulation */
};

enum { UPPE_ION_NONE = 0, UPPE_ION_EXPLICIT = 1, UPPE_ION_SERIES = 2 };

/* The equation a handle steps: the UPPE of real fields (uppe_kernel.py), or the paraxial envelope equation
 * (paraxial_kernel.py:29-331) with the same transverse transforms, RK4 and step-size control. */
enum { UPPE_MODEL_UPPE = 0, UPPE_MODEL_PARAXIAL = 1 };

/* Grid description. Replaces the shapes pyopencl arrays carry (uppe_kernel.py:95-119, paraxial_kernel.py:92-117)
 * and the transverse tool choice (uppe_kernel.py:374-377). nw = 2^n + 1 (UPPE) or 2^n (paraxial); n1, n2
 * powers of 2; nk <= n1 radial modes (radial only). */
typedef struct uppe_desc {
  int geometry;   /* UPPE_RADIAL or UPPE_CARTESIAN */
  int nw;         /* frequency nodes */
  int n1;         /* radial cells or x cells */
  int n2;         /* azimuthal nodes or y cells */
  int nk;         /* radial modes kept; ignored for Cartesian */
  int precision;  /* UPPE_PREC_* */
  int device;     /* HIP device ordinal */
  int model;      /* UPPE_MODEL_* (0, the UPPE, for descriptors that leave it zero) */
} uppe_desc;

/* Rate-kernel parameters, exactly the scalars Ionizer.RateCL passes (ionization.py:131-150). */
typedef struct uppe_ionization {
  int kind;       /* UPPE_ION_* */
  int reserved;
  double e_conv, rate_conv, e_cutoff;
  double c_pre, c_pow, c_exp;   /* kind == EXPLICIT: rate = rate_conv c_pre E^c_pow exp(c_exp/E) */
  double c[6];                  /* kind == SERIES:   rate = rate_conv exp(sum_n c_n log(E)^n) */
} uppe_ionization;

/* Medium and filters. Replaces Material.UpdateMaterial's device uploads (uppe_kernel.py:67-90) and the
 * Source filter arrays (uppe_kernel.py:96,120-129). */
typedef struct uppe_medium {
  const double* w;            /* [nw] frequency nodes, w[0] must be 0 */
  const double* kz;           /* complex128 [nw][n1][n2] */
  const double* kg;           /* [nw] Galilean wavenumber -w/vg */
  const double* ngas;         /* [n1][n2] gas density in plasma units */
  const double* nl_filter;    /* [nw] nonlinear band filter */
  const double* cond_filter;  /* [nw] field conditioner (zero below bin 4) */
  double redchi3;             /* chi3 / n_ref */
  double dt;                  /* pi / max(w) */
  uppe_ionization ion;
} uppe_medium;

/* Paraxial medium. Replaces the paraxial Material.UpdateMaterial uploads (paraxial_kernel.py:72-102) and the
 * Source frequency array (:117-135): the envelope is advanced as q = exp(-i kz z) a with the complex
 * kz = (w^2 (1 + chi0 - ng^2) - w0^2 dn^2 - 2 w w0 ng dn - kperp^2) / (2 k0), k0 = w0 dn + w ng. */
typedef struct uppe_paraxial_medium {
  const double* w;            /* [nw] frequencies in the reference's ifftshift order */
  const double* kz;           /* complex128 [nw][n1][n2] */
  const double* k0;           /* [nw] */
  const double* ngas;         /* [n1][n2] gas density in plasma units */
  double redchi3;             /* chi3 / n_ref */
  double dt;                  /* 2 pi / (nw dw) */
  uppe_ionization ion;        /* rates of the envelope modulus (ionization.cl:46-69, 99-125) */
} uppe_paraxial_medium;

/* One call runs subcycles sub0 .. sub0+n_sub-1 of length `length`. Subcycle s starts at z0 = zi + s*length
 * and its steps use the phase coordinate (z0 - zi) + z, formed in that order as the reference does.
 * Replaces the subcycle/step loop of propagator (uppe_kernel.py:299-325). */
typedef struct uppe_step_control {
  double zi;            /* start of the diagnostic interval */
  double length;        /* subcycle length L = (zf - zi)/subcycles */
  int sub0;             /* index of the first subcycle run by this call */
  int max_steps;        /* stop after this many steps, leaving the subcycle unfinished (0 = no limit) */
  double dzmin;         /* 'minimum step' */
  double dphi;          /* 0.4 in the reference (uppe_kernel.py:105) */
  double rel_amin;      /* 1e-2 in the reference (uppe_kernel.py:106) */
  int n_sub;            /* subcycles to run (>= 1) */
  int n_schedule;       /* entries in dz_schedule; 0 = adaptive step control */
  const double* dz_schedule; /* replayed steps (nullable) */
  double* dz_record;    /* receives the steps taken (nullable) */
  int record_cap;       /* capacity of dz_record */
  int reserved;
} uppe_step_control;

typedef struct uppe_info {
  long long steps;      /* RK4 steps taken since set_field */
  long long evals;      /* right-hand-side evaluations (4 per step) */
  int nan_steps;        /* steps whose field modulus sum was NaN (printed as '!' by the reference) */
  int reserved;
  double z_last;        /* local z reached in the last subcycle */
} uppe_info;

typedef struct uppe_handle uppe_handle;

/* Allocate device buffers for one grid. Replaces Source.__init__ device arrays (uppe_kernel.py:108-119)
 * and HankelTransformTool.AllocateDeviceMemory (grid_tools.py:299-303). The fused precisions (FP64, FP32,
 * BF16X3) take 4 <= Nt <= 8192 time samples (UPPE: Nt = 2 (nw - 1); paraxial: Nt = nw <= 4096), up to 2048
 * radial cells with 1, 2, 4, 8 or 16 azimuthal points, and 2 <= n1, n2 <= 1024 Cartesian cells; other grids
 * return UPPE_ERR_UNSUPPORTED, and the plain FP64 stepper takes them. */
int uppe_create(const uppe_desc* desc, uppe_handle** out);

/* Radial basis: H[nk][nr][nm] eigenvectors and Lambda[nr] root volumes (grid_tools.py:294,263,300-301). */
int uppe_set_hankel(uppe_handle* h, const double* H, const double* lambda_);

/* Medium tables; may be called again between advance calls (uppe_kernel.py:303). UPPE model only. */
int uppe_set_medium(uppe_handle* h, const uppe_medium* m);

/* Paraxial medium tables; may be called again between advance calls (paraxial_kernel.py:317). Paraxial model
 * only. The paraxial stepper has no field conditioner and no nonlinear band filter, and its caller passes
 * dzmin = 0 (paraxial_kernel.py:209-220 has no minimum step). */
int uppe_set_paraxial_medium(uppe_handle* h, const uppe_paraxial_medium* m);

/* Rotational (Raman) susceptibility added to chi in every source evaluation (update_current's
 * rotator.AddChi, uppe_kernel.py:174-176; rotations.py:44-85; rotations.cl:80-129): `states` levels with the
 * transition weights tjj[states] and W_j dt as complex128 wdt[states] (Rotator.Tj and Rotator.Wj * dt), the
 * polarizability difference and the normalised hbar. A NULL rotator, or one with fewer than 3 states,
 * switches it off. */
typedef struct uppe_rotator {
  int states;
  int reserved;
  const double* tjj;
  const double* wdt;
  double dalpha;
  double hbar;
} uppe_rotator;
int uppe_set_rotator(uppe_handle* h, const uppe_rotator* r);

/* Upload the r-space field a[nw][n1][n2] (complex128) and transform it to k-space
 * (uppe_kernel.py:290-291: Source(...) then T.kspacex(src.qw_dev)). */
int uppe_set_field(uppe_handle* h, const double* a_r);

/* RK4 interaction-picture stepping with step control (uppe_kernel.py:299-325, load_source :188-232).
 * steps_out (nullable) receives the steps taken by this call. FP32 handles keep the state of the nonlinear band
 * relative to a reference point that moves with the run, so absorbing media (factors exp(+Im kz z)) stay in
 * range; UPPE_ERR_UNSUPPORTED when max(Im kz) times one subcycle's length exceeds 40. */
int uppe_advance(uppe_handle* h, const uppe_step_control* c, int* steps_out);

/* uppe_advance that allocates nothing. No counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture of the surrounding work, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, stepping and results as uppe_advance (the same kernels in the same order), but the
 * call's work arrays (the RK slopes k1..k3, the r-space field and current), the replayed schedule and the dz
 * record live in one static device workspace of the library, 48 MiB per GPU, held from the library's load until
 * the process ends. A call needs 5 * W + S + R bytes of it, each term rounded up to 256 bytes: W = P * NWP *
 * (8 bytes in FP32 / BF16x3, 16 in FP64) with P = n1 * n2 and NWP = nw rounded up to a multiple of 32, S = 8 *
 * n_schedule, R = 8 * record_cap (0 without dz_record); larger requests return UPPE_ERR_ARG with the sizes in
 * uppe_last_error. Standalone handles on the fused precisions (FP64, FP32, BF16X3) only; no event timing. Every
 * call on one GPU shares its workspace, which keeps the last call's contents (the call clears what must start
 * at zero): calls that overlap on one GPU, from several threads or streams, race. */
int uppe_advance_static(uppe_handle* h, const uppe_step_control* c, int* steps_out);

/* finish() semantics (uppe_kernel.py:234-258, paraxial_kernel.py:222-249): propagate to z_total = zf - zi,
 * transform to r-space and evaluate the diagnostics. Outputs (each nullable): a_r, j_r complex128
 * [nw][n1][n2]; chi_t, ne_t float64 [nt][n1][n2] (the caller forms rfft (UPPE) or fft (paraxial) along
 * axis 0 as the reference does on the host). */
int uppe_finish(uppe_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t);

/* Gradients (no reference counterpart; the adjoint the input spectral-phase objectives need). UPPE media with
 * Kerr, the rotational Raman susceptibility, photoionisation (explicit and series rates) and the plasma current,
 * on the radial matrix-core paths (FP64, FP32, BF16x3) with any azimuthal point count and Nt <= 4096.
 *
 * uppe_advance_checkpointed: uppe_advance that also keeps what the adjoint needs: the state every
 * checkpoint_every steps and the phase origin and dz of every step. The step sizes are treated as constants by
 * the adjoint (fixed or replayed schedules give exact gradients).
 *
 * uppe_adjoint: the reverse sweep over the recorded run. g_a_r is the cotangent dL/dRe a + i dL/dIm a of the
 * r-space field uppe_finish(z_total) returns (complex128 [nw][n1][n2]); g_a0_r receives the cotangent of the
 * r-space field given to uppe_set_field. Each segment is replayed from its checkpoint. The medium must be the
 * one of the recorded run; afterwards the stepper needs uppe_set_field again. */
int uppe_advance_checkpointed(uppe_handle* h, const uppe_step_control* c, int checkpoint_every, int* steps_out);
int uppe_adjoint(uppe_handle* h, double z_total, const double* g_a_r, double* g_a0_r);

/* Read the k-space state q[nw][n1][n2] (complex128). No reference counterpart: used for checks. */
int uppe_get_state(uppe_handle* h, double* q_k);

/* Counters since the last uppe_set_field. */
int uppe_get_info(uppe_handle* h, uppe_info* info);

/* Per-kernel-class timing with HIP events on the handle's stream (for measurement): in the fused paths each
 * dispatch stamps its own start and end events, the plain path records an event pair around each launch.
 * enable != 0 starts recording and clears totals; uppe_profile_read returns total milliseconds and launch
 * count of a class. The radial transforms are two classes: UPPE_KC_HANKEL for the k->r direction (and the
 * one-off transforms of set_field and finish), UPPE_KC_HANKEL_R2K for the r->k direction with the RK stage
 * epilogue. */
enum { UPPE_KC_HANKEL = 0, UPPE_KC_SOURCE = 1, UPPE_KC_POINT = 2, UPPE_KC_REDUCE = 3, UPPE_KC_TRANSVERSE_FFT = 4,
       UPPE_KC_HANKEL_R2K = 5, UPPE_KC_COUNT = 6 };
int uppe_profile_enable(uppe_handle* h, int enable);
int uppe_profile_read(uppe_handle* h, int kernel_class, double* total_ms, long long* launches);

/* Batches (no reference counterpart: the reference steps one grid per propagator call). Many radial UPPE
 * instances stepped together: uppe_batch_create makes one member handle per descriptor (radial, UPPE model,
 * FP64 / FP32 / BF16X3), whose per-instance arrays the batch keeps in one slab per (Nw, Nr, Nk, precision)
 * group. Members are loaded like any handle (uppe_set_hankel, uppe_set_medium, uppe_set_field) and read back
 * with uppe_finish / uppe_get_state / uppe_get_info; uppe_destroy refuses them (uppe_batch_destroy releases
 * them). uppe_batch_advance runs uppe_advance on every member at once, c[i] being member i's step control
 * (adaptive or fixed steps: no schedule, no record; members may differ in subcycles and step counts) and
 * steps_out[i] (nullable) its step count: consecutive members of a group with the same nonlinear band,
 * frequency tables and presence of a rotator (uppe_set_rotator, per member) share every launch. With autotune on
 * (the default) the first
 * advance of each run times the candidate contraction tilings and source launch shapes on the run and keeps
 * the fastest; uppe_batch_plan reports member i's run as {k2r tiles, waves, row tiles per wave, r2k tiles,
 * waves, row tiles per wave, source samples per thread (0 = default), run length, first slot, group}.
 * uppe_batch_set_graph sets how many steps of all runs one HIP graph replays (default 8; 0 launches kernel by
 * kernel); the graph is captured at every advance, whose step count is then rounded up to a multiple of it
 * (steps after a member's end leave it untouched). Runs with event timing on (uppe_profile_enable) launch
 * kernel by kernel. */
typedef struct uppe_batch uppe_batch;
int uppe_batch_create(const uppe_desc* descs, int n, uppe_batch** out);
int uppe_batch_member(uppe_batch* b, int i, uppe_handle** member);
int uppe_batch_advance(uppe_batch* b, const uppe_step_control* c, int* steps_out);
int uppe_batch_set_autotune(uppe_batch* b, int enable);
int uppe_batch_set_graph(uppe_batch* b, int steps_per_graph);
int uppe_batch_plan(const uppe_batch* b, int member, int* plan /* [10] */);
int uppe_batch_destroy(uppe_batch* b);
const char* uppe_batch_last_error(const uppe_batch* b);

/* Block until all work on the handle's stream is done. */
int uppe_synchronize(uppe_handle* h);

int uppe_destroy(uppe_handle* h);
const char* uppe_last_error(const uppe_handle* h);

#ifdef __cplusplus
}
#endif
#endif /* SEARAY_UPPE_H */

// uppe_internal.h -- shared declarations of the MI355X UPPE stepper library (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/searay_uppe.h"

namespace uppe {

constexpr int kBlock = 256;        // threads per workgroup for point and strip kernels (4 wavefronts)
constexpr int kMaxNtFused = 8192;  // fused source kernel keeps one Nt-point complex strip in LDS
constexpr int kMaxNtTables = 4096;  // FP32 strips up to this length also keep their twiddle tables in LDS
constexpr int kMaxNrFused = 2048;  // radial cells of the fused path (contraction rows in blocks of <= 512)
constexpr int kMaxNxFused = 1024;  // Cartesian points per transverse axis of the fused path (one line per block)
constexpr int kMaxNtPlain = 16384; // plain strip FFT keeps Nt/2 complex points in LDS (FP64)

struct Error {
  int code;
  std::string msg;
};

#define UPPE_HIP(call)                                                                          \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess)                                                                       \
      throw ::uppe::Error{UPPE_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(e_)}; \
  } while (0)

#define UPPE_LAUNCH_CHECK() UPPE_HIP(hipGetLastError())

// ---------------------------------------------------------------------------------------------------
// complex helpers (interleaved re, im)
// ---------------------------------------------------------------------------------------------------
// An FP32 complex value is aligned to its 8-byte size, so that the compiler moves it with one 8-byte access
// (ds_read_b64 / ds_write_b64 in LDS) instead of two 4-byte ones; FP64 keeps the alignment of double (8), which
// host buffers of interleaved doubles guarantee.
template <typename T>
struct alignas(2 * sizeof(T) <= 8 ? 2 * sizeof(T) : alignof(T)) cplx {
  T re, im;
};

template <typename T>
__host__ __device__ inline cplx<T> cmk(T r, T i) {
  return cplx<T>{r, i};
}
template <typename T>
__host__ __device__ inline cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return cplx<T>{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ inline cplx<T> cadd(cplx<T> a, cplx<T> b) {
  return cplx<T>{a.re + b.re, a.im + b.im};
}
template <typename T>
__host__ __device__ inline cplx<T> csub(cplx<T> a, cplx<T> b) {
  return cplx<T>{a.re - b.re, a.im - b.im};
}
template <typename T>
__host__ __device__ inline cplx<T> cconj(cplx<T> a) {
  return cplx<T>{a.re, -a.im};
}
template <typename T>
__host__ __device__ inline cplx<T> cscale(cplx<T> a, T s) {
  return cplx<T>{a.re * s, a.im * s};
}

// ---------------------------------------------------------------------------------------------------
// Block-cooperative radix-2 FFT in LDS. Data is n = 2^lg complex points (interleaved) at `buf`.
// Twiddle table tw[k] = exp(-2 pi i k / Nbase), k < Nbase/2, with Nbase >= 2n; sign = -1 forward,
// +1 inverse; no normalisation. All threads of the block must call it; it ends with a barrier.
// ---------------------------------------------------------------------------------------------------
template <typename T>
__device__ inline void lds_fft_radix2(cplx<T>* buf, int n, int lg, const cplx<T>* __restrict__ tw, int nbase,
                                      int sign) {
  if (n <= 1) return;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int j = (int)(__brev((unsigned)i) >> (32 - lg));
    if (j > i) {
      cplx<T> t = buf[i];
      buf[i] = buf[j];
      buf[j] = t;
    }
  }
  __syncthreads();
  for (int s = 0; s < lg; s++) {
    const int half = 1 << s;
    const int tstep = nbase >> (s + 1);  // exp(-2 pi i pos / (2 half)) = tw[pos * nbase/(2 half)]
    for (int b = threadIdx.x; b < (n >> 1); b += blockDim.x) {
      const int pos = b & (half - 1);
      const int a = ((b >> s) << (s + 1)) + pos;
      const int c = a + half;
      cplx<T> w = tw[pos * tstep];
      if (sign > 0) w.im = -w.im;
      cplx<T> u = buf[a];
      cplx<T> v = cmul(buf[c], w);
      buf[a] = cadd(u, v);
      buf[c] = csub(u, v);
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------------------------------
// LDS layout of the strip buffers of the FP32 source kernel: complex element i lives at slot
// i ^ (((i >> 3) & 15) << 1). The map permutes each aligned block of 128 elements, keeps element pairs
// (2m, 2m + 1) together (16-byte vector accesses stay aligned) and spreads the power-of-two-strided writes
// of the early Stockham passes (spans 1, 4, 16) and the 32-byte per-thread runs over the LDS banks. FP64
// elements are 16 bytes wide and keep the identity layout.
// ---------------------------------------------------------------------------------------------------
template <typename T>
__device__ inline int lds_swz(int i) {
  if constexpr (sizeof(T) == 4)
    return i ^ (((i >> 3) & 15) << 1);
  else
    return i;
}

// ---------------------------------------------------------------------------------------------------
// Block-cooperative in-place Stockham FFT in LDS (buffer in the lds_swz layout). An n = 2^lg point transform
// runs one twiddle-free radix-2^(lg mod 3) pass (span 1) when lg is not a multiple of 3, then radix-8 passes.
// Pass with span Ns and radix R: butterfly j < n/R reads element j + r n/R, multiplies input r by
// exp(sign 2 pi i r (j mod Ns)/(R Ns)), forms the R-point DFT and writes element
// (j - j mod Ns) R + j mod Ns + r Ns (autosorted). Each pass loads its inputs into registers, waits for the
// block, then stores, so no second buffer is needed; the block has at least n/8 threads for the radix-8 passes
// and n/(2 MAXB) for the first pass (up to MAXB butterflies per thread; MAXB = 2 by default, 4 for the
// 8192-sample strips of a 1024-thread block).
// Twiddles: per-pass tables, the radix-8 pass of span Ns holding exp(-2 pi i r k/(8 Ns)) at
// [(Ns - R0) + (r - 1) Ns + k], r = 1..7, k < Ns, with R0 = 2^(lg mod 3) the span of the first radix-8
// pass; the table of an n-point transform has R0 (8^(lg/3) - 1) entries. Threads of a pass read consecutive
// entries. Inverse transforms conjugate.
// ---------------------------------------------------------------------------------------------------
__host__ __device__ inline int r8_table_len(int lg) {
  const int r0 = 1 << (lg % 3);
  int p8 = 1;
  for (int i = 0; i < lg / 3; i++) p8 *= 8;
  return r0 * (p8 - 1);
}

// 4-point DFT in place (sign < 0 forward)
template <typename T>
__device__ __forceinline__ void dft4(cplx<T>& a0, cplx<T>& a1, cplx<T>& a2, cplx<T>& a3, int sign) {
  const cplx<T> A0 = cadd(a0, a2), A1 = csub(a0, a2), A2 = cadd(a1, a3), d = csub(a1, a3);
  const cplx<T> A3 = sign < 0 ? cmk(d.im, -d.re) : cmk(-d.im, d.re);  // -i d forward, +i d inverse
  a0 = cadd(A0, A2);
  a2 = csub(A0, A2);
  a1 = cadd(A1, A3);
  a3 = csub(A1, A3);
}

// 8-point DFT of v[0..7] into y[0..7]: radix-2 split into sums and twiddled differences, then two 4-point DFTs
template <typename T>
__device__ __forceinline__ void dft8(const cplx<T>* v, cplx<T>* y, int sign) {
  const T h = T(0.70710678118654752440);
  cplx<T> a0 = cadd(v[0], v[4]), a1 = cadd(v[1], v[5]), a2 = cadd(v[2], v[6]), a3 = cadd(v[3], v[7]);
  cplx<T> b0 = csub(v[0], v[4]), b1 = csub(v[1], v[5]), b2 = csub(v[2], v[6]), b3 = csub(v[3], v[7]);
  if (sign < 0) {  // times exp(-i pi m/4), m = 1, 2, 3
    b1 = cmk(h * (b1.re + b1.im), h * (b1.im - b1.re));
    b2 = cmk(b2.im, -b2.re);
    b3 = cmk(h * (b3.im - b3.re), -h * (b3.re + b3.im));
  } else {  // times exp(+i pi m/4)
    b1 = cmk(h * (b1.re - b1.im), h * (b1.im + b1.re));
    b2 = cmk(-b2.im, b2.re);
    b3 = cmk(-h * (b3.re + b3.im), h * (b3.re - b3.im));
  }
  dft4(a0, a1, a2, a3, sign);
  dft4(b0, b1, b2, b3, sign);
  y[0] = a0;
  y[2] = a1;
  y[4] = a2;
  y[6] = a3;
  y[1] = b0;
  y[3] = b1;
  y[5] = b2;
  y[7] = b3;
}

// LG > 0 fixes lg = LG (and n, and the block size NTH when nonzero) at compile time: the passes then unroll
// with their spans, strides and table offsets as constants. MAXB bounds the butterflies a thread takes in the
// first pass. The transform is always inlined: as a call it would keep sign at run time and save the caller's
// registers to scratch around it.
template <typename T, int LG = 0, int NTH = 0, int MAXB = 2>
__device__ __forceinline__ void lds_fft_r8(cplx<T>* buf, int n_rt, int lg_rt, const cplx<T>* tab, int sign, int tid,
                                           int nth_rt = 0) {
  const int lg = LG > 0 ? LG : lg_rt;
  const int n = LG > 0 ? (1 << LG) : n_rt;
  const int r0 = lg % 3;
  int Ns = 1;
  if (r0 > 0) {
    // up to MAXB butterflies per thread (a radix-2 pass has n/2 of them); the block's threads, or nth_rt of them
    const int R = 1 << r0, nb = n >> r0, nth = NTH > 0 ? NTH : (nth_rt > 0 ? nth_rt : (int)blockDim.x);
    cplx<T> v[MAXB][4];
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
      const int j = tid + b * nth;
      if (j < nb)
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (r < R) v[b][r] = buf[lds_swz<T>(j + r * nb)];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
      const int j = tid + b * nth;
      if (j >= nb) continue;
      if (R == 4) {
        dft4(v[b][0], v[b][1], v[b][2], v[b][3], sign);
      } else {
        const cplx<T> s0 = cadd(v[b][0], v[b][1]), s1 = csub(v[b][0], v[b][1]);
        v[b][0] = s0;
        v[b][1] = s1;
      }
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (r < R) buf[lds_swz<T>(j * R + r)] = v[b][r];
    }
    __syncthreads();
    Ns = R;
  }
  const cplx<T>* tp = tab;
  auto pass = [&](int Ns) {
    const int nb = n >> 3;
    cplx<T> v[8], y[8];
    if (tid < nb)
#pragma unroll
      for (int r = 0; r < 8; r++) v[r] = buf[lds_swz<T>(tid + r * nb)];
    __syncthreads();
    if (tid < nb) {
      const int k = tid & (Ns - 1);
#pragma unroll
      for (int r = 1; r < 8; r++) {
        cplx<T> w = tp[(r - 1) * Ns + k];
        if (sign > 0) w.im = -w.im;
        v[r] = cmul(v[r], w);
      }
      dft8(v, y, sign);
      const int o = (tid - k) * 8 + k;
#pragma unroll
      for (int r = 0; r < 8; r++) buf[lds_swz<T>(o + r * Ns)] = y[r];
    }
    __syncthreads();
    tp += 7 * Ns;
  };
  if constexpr (LG > 0) {
#pragma unroll
    for (int Nc = 1 << (LG % 3); Nc < (1 << LG); Nc *= 8) pass(Nc);
  } else {
    for (; Ns < n; Ns *= 8) pass(Ns);
  }
}

// Two transforms of the same length and direction (buffers a and b, one table), each on one half of the block
// when a half covers a pass (n/8 butterflies of the radix-8 passes, and the first pass's within 2 per thread): both
// then share every pass's barriers, where one after the other they would take twice as many. Otherwise in turn.
template <typename T, int LG = 0, int NTH = 0>
__device__ __forceinline__ void lds_fft_r8_pair(cplx<T>* a, cplx<T>* b, int n_rt, int lg_rt, const cplx<T>* tab,
                                                int sign, int tid) {
  const int lg = LG > 0 ? LG : lg_rt;
  const int n = LG > 0 ? (1 << LG) : n_rt;
  const int nth = NTH > 0 ? NTH : (int)blockDim.x, half = nth / 2;
  const int nb0 = (lg % 3) ? n >> (lg % 3) : 0;
  if (half >= n / 8 && 2 * half >= nb0 && half >= 64) {
    const bool second = tid >= half;
    lds_fft_r8<T, LG, NTH / 2>(second ? b : a, n, lg, tab, sign, second ? tid - half : tid, half);
  } else {
    lds_fft_r8<T, LG, NTH>(a, n, lg, tab, sign, tid);
    lds_fft_r8<T, LG, NTH>(b, n, lg, tab, sign, tid);
  }
}

// host: the radix-8 pass tables of an n-point transform, FP64 -> T (see lds_fft_r8 for the layout)
template <typename T>
inline std::vector<cplx<T>> r8_table_one(int n) {
  int lg = 0;
  while ((1 << lg) < n) lg++;
  std::vector<cplx<T>> tab;
  for (int Ns = 1 << (lg % 3); Ns < n; Ns *= 8)
    for (int r = 1; r < 8; r++)
      for (int k = 0; k < Ns; k++) {
        const double a = -2.0 * M_PI * r * k / (8.0 * Ns);
        tab.push_back(cplx<T>{(T)std::cos(a), (T)std::sin(a)});
      }
  return tab;
}

// host: the tables of an N-point transform followed by those of the N/2-point transform
template <typename T>
inline std::vector<cplx<T>> r8_tables(int N) {
  std::vector<cplx<T>> tab = r8_table_one<T>(N), half = r8_table_one<T>(N / 2);
  tab.insert(tab.end(), half.begin(), half.end());
  return tab;
}

// ---------------------------------------------------------------------------------------------------
// Profiling of kernel classes with HIP events on the handle's stream
// ---------------------------------------------------------------------------------------------------
struct ProfRecord {
  int kc;
  hipEvent_t a, b;
};

struct PlainState;
struct FusedState;
struct MemberSlots;
struct FusedBatch;

}  // namespace uppe

struct uppe_handle {
  uppe_desc d;
  int nw = 0, n1 = 0, n2 = 0, nk = 0, nm = 0, nt = 0;
  bool paraxial = false;  // UPPE_MODEL_PARAXIAL: complex envelope strips, Nt = Nw
  long P = 0;     // transverse points n1*n2
  long npt = 0;   // nw * P
  long ntt = 0;   // nt * P
  hipStream_t stream = nullptr;
  std::string err;
  bool have_hankel = false, have_medium = false, have_field = false;
  long long steps = 0, evals = 0;
  int nan_steps = 0;
  double z_last = 0.0;
  // host copies of the tables (the reference keeps these on the host too)
  std::vector<double> H, lam, w;
  // rotational susceptibility: states (0 = off), Tjj[states], W_j dt [states] complex, dalpha, hbar
  int rot_states = 0;
  std::vector<double> rot_tjj, rot_wdt;
  double rot_dalpha = 0.0, rot_hbar = 1.0;
  // profiling
  bool prof = false;
  std::vector<uppe::ProfRecord> prof_pending;
  std::vector<hipEvent_t> prof_free;
  double prof_ms[UPPE_KC_COUNT] = {0};
  long long prof_n[UPPE_KC_COUNT] = {0};
  // implementation state (exactly one is non-null)
  uppe::PlainState* plain = nullptr;
  uppe::FusedState* fused = nullptr;
  // the batch that owns this handle (its arrays live in the batch's slabs), or null
  struct uppe_batch* batch = nullptr;
};

// a set of radial UPPE instances stepped together (one launch per kernel for each run of equally shaped ones)
struct uppe_batch {
  std::vector<uppe_handle*> members;
  uppe::FusedBatch* fb = nullptr;
  int device = 0;
  std::string err;
};

namespace uppe {

// event pair around one launch when profiling is on
struct ProfScope {
  uppe_handle* h;
  int kc;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(uppe_handle* h_, int kc_);
  ~ProfScope() noexcept;
};

void prof_collect(uppe_handle* h);  // synchronise and fold finished events into totals
void prof_take(uppe_handle* h, hipEvent_t* a, hipEvent_t* b);      // an unused event pair
void prof_push(uppe_handle* h, int kc, hipEvent_t a, hipEvent_t b);  // queue a recorded pair for collection

inline unsigned grid_for(long n, int block = kBlock) { return (unsigned)((n + block - 1) / block); }

// plain FP64 path (reference layout and pass structure)
void plain_create(uppe_handle* h);
void plain_destroy(uppe_handle* h);
void plain_set_hankel(uppe_handle* h);
void plain_set_medium(uppe_handle* h, const uppe_medium* m);
void plain_set_paraxial_medium(uppe_handle* h, const uppe_paraxial_medium* m);
void plain_set_rotator(uppe_handle* h);
void plain_set_field(uppe_handle* h, const double* a_r);
int plain_advance(uppe_handle* h, const uppe_step_control* c);
void plain_finish(uppe_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t);
void plain_get_state(uppe_handle* h, double* q);

// fused path (strip-major layout, FP64 / FP32 / BF16x3 Hankel); a batch member takes its per-instance arrays
// from slots
void fused_create(uppe_handle* h, const MemberSlots* slots = nullptr);
void fused_destroy(uppe_handle* h);
void fused_set_hankel(uppe_handle* h);
void fused_set_rotator(uppe_handle* h);
void fused_set_medium(uppe_handle* h, const uppe_medium* m);
void fused_set_paraxial_medium(uppe_handle* h, const uppe_paraxial_medium* m);
int fused_advance_checkpointed(uppe_handle* h, const uppe_step_control* c, int every);
void fused_adjoint(uppe_handle* h, double z_total, const double* g_a_r, double* g_a0_r);
void fused_set_field(uppe_handle* h, const double* a_r);
int fused_advance(uppe_handle* h, const uppe_step_control* c);
// uppe_advance_static: advance with its work arrays in the library's static device workspace (no allocation);
// need = the workspace bytes the call takes, capacity = the workspace size
size_t fused_static_need(uppe_handle* h, const uppe_step_control* c);
size_t fused_static_capacity();
int fused_advance_static(uppe_handle* h, const uppe_step_control* c);
void fused_finish(uppe_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t);
void fused_get_state(uppe_handle* h, double* q);

// batches of radial instances (uppe_batch_*): create allocates the slabs and the member handles through
// make_member; advance steps every member; plan reports the launch shapes chosen for a member's run
using MemberFactory = int (*)(const uppe_desc* d, const MemberSlots* slots, uppe_handle** out);
void fused_batch_create(uppe_batch* b, const uppe_desc* descs, int n, MemberFactory make_member);
void fused_batch_destroy(uppe_batch* b);
void fused_batch_advance(uppe_batch* b, const uppe_step_control* c, int* steps_out);
void fused_batch_plan(const uppe_batch* b, int member, int* out);
void fused_batch_set_autotune(uppe_batch* b, int enable);
void fused_batch_set_graph(uppe_batch* b, int steps);

}  // namespace uppe

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
constexpr int kMaxNtFused = 4096;  // fused source kernel keeps one Nt-point complex strip in LDS
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
template <typename T>
struct cplx {
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
// Block-cooperative in-place Stockham FFT in LDS (buffer in the lds_swz layout): radix-4 passes and one
// radix-2 pass when log2(n) is odd. Each pass loads its butterfly inputs into registers, waits for the whole
// block, then stores the outputs in autosorted order, so no second buffer is needed. Pass with span Ns and
// radix R: butterfly j reads element j + r n/R, multiplies input r by exp(sign 2 pi i r (j mod Ns)/(Ns R)),
// forms the R-point DFT and writes element (j - j mod Ns) R + j mod Ns + r Ns. At most two butterflies per
// thread per pass (n/R <= 2 blockDim.x).
// Twiddles come from per-pass tables, so that the threads of a pass read consecutive table entries
// (the single exp(-2 pi i k/nbase) table is read at power-of-two strides, which puts a wave on one LDS bank).
// Layout of `tab` (built by stockham_tables): the radix-4 pass of span Ns uses tab[Ns - 1 + (r-1) Ns + k] =
// exp(-2 pi i r k/(4 Ns)), k < Ns, r = 1..3; the radix-2 pass of an n-point transform (log2 n odd) uses
// tab[r2off + k] = exp(-2 pi i k/n), k < n/2. Inverse transforms conjugate.
// ---------------------------------------------------------------------------------------------------
template <typename T>
__device__ inline void lds_fft_stockham_tab(cplx<T>* buf, int n, int lg, const cplx<T>* tab, int r2off, int sign) {
  const int nth = blockDim.x, tid = threadIdx.x;
  int Ns = 1;
  int rem = lg;
  while (rem > 0) {
    const int R = rem >= 2 ? 4 : 2;
    const int nb = n / R;
    cplx<T> v[2][4];
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int j = tid + b * nth;
      if (j < nb)
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (r < R) v[b][r] = buf[lds_swz<T>(j + r * nb)];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < 2; b++) {
      const int j = tid + b * nth;
      if (j >= nb) continue;
      const int k = j & (Ns - 1);
      cplx<T> y[4];
      if (R == 4) {
        cplx<T> w1 = tab[Ns - 1 + k], w2 = tab[2 * Ns - 1 + k], w3 = tab[3 * Ns - 1 + k];
        if (sign > 0) {
          w1.im = -w1.im;
          w2.im = -w2.im;
          w3.im = -w3.im;
        }
        cplx<T> v1 = cmul(v[b][1], w1), v2 = cmul(v[b][2], w2), v3 = cmul(v[b][3], w3);
        cplx<T> a0 = cadd(v[b][0], v2), a1 = csub(v[b][0], v2), a2 = cadd(v1, v3), d = csub(v1, v3);
        cplx<T> a3 = sign < 0 ? cmk(d.im, -d.re) : cmk(-d.im, d.re);
        y[0] = cadd(a0, a2);
        y[1] = cadd(a1, a3);
        y[2] = csub(a0, a2);
        y[3] = csub(a1, a3);
      } else {
        cplx<T> w = tab[r2off + k];
        if (sign > 0) w.im = -w.im;
        cplx<T> v1 = cmul(v[b][1], w);
        y[0] = cadd(v[b][0], v1);
        y[1] = csub(v[b][0], v1);
      }
      const int o = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < 4; r++)
        if (r < R) buf[lds_swz<T>(o + r * Ns)] = y[r];
    }
    __syncthreads();
    Ns *= R;
    rem -= R == 4 ? 2 : 1;
  }
}

// host: per-pass tables for an N-point and an N/2-point transform (N a power of two >= 4), FP64 -> T,
// N + N/2 entries with r2off = N
template <typename T>
inline std::vector<cplx<T>> stockham_tables(int N) {
  std::vector<cplx<T>> tab((size_t)N + N / 2, cplx<T>{T(0), T(0)});
  int lgN = 0;
  while ((1 << lgN) < N) lgN++;
  for (int Ns = 1; 4 * Ns <= N; Ns *= 4)
    for (int r = 1; r <= 3; r++)
      for (int k = 0; k < Ns; k++) {
        const double a = -2.0 * M_PI * r * k / (4.0 * Ns);
        tab[(size_t)Ns - 1 + (size_t)(r - 1) * Ns + k] = cplx<T>{(T)std::cos(a), (T)std::sin(a)};
      }
  const int nodd = (lgN & 1) ? N : N / 2;
  for (int k = 0; k < nodd / 2; k++) {
    const double a = -2.0 * M_PI * k / nodd;
    tab[(size_t)N + k] = cplx<T>{(T)std::cos(a), (T)std::sin(a)};
  }
  return tab;
}

// exp(-2 pi i k/N), k < N/2, read from the per-pass tables
template <typename T>
__device__ inline cplx<T> stockham_twn(const cplx<T>* tab, int N, bool lg_odd, int k) {
  if (lg_odd) return tab[N + k];
  const int q = N / 4;
  if (k < q) return tab[q - 1 + k];
  cplx<T> w = tab[q - 1 + k - q];
  return cmk(w.im, -w.re);  // -i w
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

}  // namespace uppe

struct uppe_handle {
  uppe_desc d;
  int nw = 0, n1 = 0, n2 = 0, nk = 0, nm = 0, nt = 0;
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
  // profiling
  bool prof = false;
  std::vector<uppe::ProfRecord> prof_pending;
  std::vector<hipEvent_t> prof_free;
  double prof_ms[UPPE_KC_COUNT] = {0};
  long long prof_n[UPPE_KC_COUNT] = {0};
  // implementation state (exactly one is non-null)
  uppe::PlainState* plain = nullptr;
  uppe::FusedState* fused = nullptr;
};

namespace uppe {

// event pair around one launch when profiling is on
struct ProfScope {
  uppe_handle* h;
  int kc;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(uppe_handle* h_, int kc_);
  ~ProfScope() noexcept(false);
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
void plain_set_field(uppe_handle* h, const double* a_r);
int plain_advance(uppe_handle* h, const uppe_step_control* c);
void plain_finish(uppe_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t);
void plain_get_state(uppe_handle* h, double* q);

// fused path (strip-major layout, FP64 / FP32 / BF16x3 Hankel)
void fused_create(uppe_handle* h);
void fused_destroy(uppe_handle* h);
void fused_set_hankel(uppe_handle* h);
void fused_set_medium(uppe_handle* h, const uppe_medium* m);
void fused_set_field(uppe_handle* h, const double* a_r);
int fused_advance(uppe_handle* h, const uppe_step_control* c);
void fused_finish(uppe_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t);
void fused_get_state(uppe_handle* h, double* q);

}  // namespace uppe

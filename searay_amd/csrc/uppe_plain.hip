// uppe_plain.hip -- the plain FP64 GPU stepper.
//
// One kernel per logical operation of SeaRay's OpenCL path, in the reference layout (spectral
// (Nw, N1, N2), time (Nt, N1, N2)), FP64 throughout, with the step size reduced on the host every step
// (modules/uppe_kernel.py:188-232). The strip FFTs are simple LDS radix-2 transforms (one workgroup per
// strip) in place of the reference's serial strip FFT; the radial transforms are the naive contraction
// (kernels/fft.cl:443-477). This path is the baseline the fused path is measured against. The paraxial
// envelope model (modules/paraxial_kernel.py, kernels/paraxial.cl) runs the same way on complex strips.
#include "uppe_internal.h"

#include <algorithm>
#include <cstring>

namespace uppe {

using c64 = cplx<double>;

struct PlainState {
  double *w = nullptr, *kz = nullptr, *kg = nullptr, *ng = nullptr, *nlf = nullptr, *cond = nullptr;
  double *H = nullptr, *lam = nullptr;
  c64 *tw_t = nullptr, *tw1 = nullptr, *tw2 = nullptr;  // exp(-2 pi i k/N) for N = Nt, N1, N2
  double redchi3 = 0, dt = 0;
  uppe_ionization ion{};
  double *qw = nullptr, *qi = nullptr, *qf = nullptr, *Ew = nullptr, *Jw = nullptr, *scratchw = nullptr,
         *hscratch = nullptr, *dzb = nullptr;
  double *Et = nullptr, *ne = nullptr, *chi = nullptr, *Jt = nullptr;
  double *rot_tjj = nullptr, *rot_wdt = nullptr;  // rotator tables (device), when h->rot_states > 0
  // paraxial model: k0(w) and complex time strips of E, A and the current
  double *k0 = nullptr, *Etc = nullptr, *Atc = nullptr, *Jtc = nullptr;
  double* part = nullptr;       // per-block reduction partials (device)
  double* part_host = nullptr;  // pinned host copy
  int nblk_red = 0;
};

// chi += rotational susceptibility (rotations.cl:80-129, AddChiUPPE): one thread per strip, serially over the
// states and over time from the last sample, as the reference's work-item does
__global__ void plain_add_chi_rotations(double* chi, const double* Et, const double* ng, const double* tjj,
                                        const double* wdt, double dt, double dalpha, double hbar, int steps,
                                        int states, long P) {
  const long xy = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (xy >= P) return;
  for (int j = 2; j < states; j++) {
    const double wr = wdt[2 * j], wi = wdt[2 * j + 1];
    const double Sjj = 2.0 * j * (j - 1) / (15 * (2 * j - 1));
    double rr = 0.0, ri = 0.0, er = 1.0, ei = 0.0;
    const double hr = cos(0.5 * wr) * exp(-0.5 * wi), hi = sin(0.5 * wr) * exp(-0.5 * wi);
    double cum = 0.0;
    for (int k = steps - 1; k >= 0; --k) {
      const long idx = (long)k * P + xy;
      const double ar = rr * er + ri * ei, ai = -rr * ei + ri * er;  // rho conj(exp_iwt)
      cum += 2 * (tjj[j] * Sjj * (ar * wi + ai * wr)) * ng[xy] * dalpha;
      chi[idx] += cum;
      const double e = Et[idx];
      const double F = 0.5 * (dalpha / hbar) * e * e;
      double nr = er * hr - ei * hi, ni = er * hi + ei * hr;
      er = nr;
      ei = ni;
      rr -= dt * F * (-ei);  // rho -= dt F i exp_iwt
      ri -= dt * F * er;
      nr = er * hr - ei * hi;
      ni = er * hi + ei * hr;
      er = nr;
      ei = ni;
    }
  }
}

// ---------------------------------------------------------------------------------------------------
// paraxial envelope kernels (kernels/paraxial.cl, ionization.cl:46-69,99-125, rotations.cl:26-78)
// ---------------------------------------------------------------------------------------------------

// A = exp(i kz z) q with complex kz (paraxial.cl:60-77)
__global__ void plain_propagate_linear_env(double* q, const double* kz, double z, long npt) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  double kr = kz[2 * idx], ki = kz[2 * idx + 1];
  double ar = -ki * z, ai = kr * z;
  double er = cos(ai) * exp(ar), ei = sin(ai) * exp(ar);
  double qr = q[2 * idx], qi = q[2 * idx + 1];
  q[2 * idx] = er * qr - ei * qi;
  q[2 * idx + 1] = er * qi + ei * qr;
}

// S = (i/2) exp(-i kz z) J / k0(w) (paraxial.cl:79-97)
__global__ void plain_current_to_rhs_env(double* J, const double* kz, const double* k0, double z, long npt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  double kr = kz[2 * idx], ki = kz[2 * idx + 1];
  double ar = ki * z, ai = -kr * z;
  double er = cos(ai) * exp(ar), ei = sin(ai) * exp(ar);
  double jr = J[2 * idx], ji = J[2 * idx + 1];
  double tr = er * jr - ei * ji, ti = er * ji + ei * jr;
  double k = k0[idx / P];
  J[2 * idx] = (0.0 * tr - 0.5 * ti) / k;
  J[2 * idx + 1] = (0.0 * ti + 0.5 * tr) / k;
}

// DtSpectral with its sign: A <- -sgn i w A (fft.cl:282-299)
__global__ void plain_dt_spectral_sgn(double* A, const double* w, double sgn, long npt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  double w0 = w[idx / P];
  double ar = A[2 * idx], ai = A[2 * idx + 1];
  A[2 * idx] = sgn * w0 * ai;
  A[2 * idx + 1] = -sgn * w0 * ar;
}

// rates of the envelope modulus (EnvelopeRate, EnvelopeRateSeries)
__global__ void plain_rate_env(double* rate, const double* E, uppe_ionization p, long ntt) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= ntt) return;
  double er = E[2 * idx], ei = E[2 * idx + 1];
  double e = sqrt(er * er + ei * ei);
  if (p.kind == UPPE_ION_EXPLICIT) {
    double eau = p.e_conv * e + p.e_cutoff;
    rate[idx] = p.rate_conv * p.c_pre * pow(eau, p.c_pow) * exp(p.c_exp / eau);
  } else {
    double lg = log(p.e_conv * e + p.e_cutoff);
    rate[idx] = p.rate_conv * exp(p.c[0] + p.c[1] * lg + p.c[2] * lg * lg + p.c[3] * pow(lg, 3.0) +
                                  p.c[4] * pow(lg, 4.0) + p.c[5] * pow(lg, 5.0));
  }
}

// chi += 0.75 ng chi3 |E|^2 (paraxial.cl:112-131)
__global__ void plain_kerr_chi_env(double* chi, const double* E, const double* ng, double chi3, long ntt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= ntt) return;
  double er = E[2 * idx], ei = E[2 * idx + 1];
  chi[idx] += 0.75 * ng[idx % P] * chi3 * (er * er + ei * ei);
}

// chi += rotational susceptibility driven by |E|^2 / 2 (rotations.cl:26-78, AddChiParaxial): one thread per
// strip, serially over the states and over time from the last sample
__global__ void plain_add_chi_rotations_env(double* chi, const double* E, const double* ng, const double* tjj,
                                            const double* wdt, double dt, double dalpha, double hbar, int steps,
                                            int states, long P) {
  const long xy = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (xy >= P) return;
  for (int j = 2; j < states; j++) {
    const double wr = wdt[2 * j], wi = wdt[2 * j + 1];
    const double Sjj = 2.0 * j * (j - 1) / (15 * (2 * j - 1));
    double rr = 0.0, ri = 0.0, er = 1.0, ei = 0.0, cum = 0.0;
    const double hr = cos(0.5 * wr) * exp(-0.5 * wi), hi = sin(0.5 * wr) * exp(-0.5 * wi);
    for (int k = steps - 1; k >= 0; --k) {
      const long idx = (long)k * P + xy;
      const double ar = rr * er + ri * ei, ai = -rr * ei + ri * er;
      cum += 2 * (tjj[j] * Sjj * (ar * wi + ai * wr)) * ng[xy] * dalpha;
      chi[idx] += cum;
      const double Er = E[2 * idx], Ei = E[2 * idx + 1];
      const double F = 0.5 * (dalpha / hbar) * 0.5 * (Er * Er + Ei * Ei);
      double nr = er * hr - ei * hi, ni = er * hi + ei * hr;
      er = nr;
      ei = ni;
      rr -= dt * F * (-ei);
      ri -= dt * F * er;
      nr = er * hr - ei * hi;
      ni = er * hi + ei * hr;
      er = nr;
      ei = ni;
    }
  }
}

// P = chi E, complex (paraxial.cl:133-147)
__global__ void plain_polarization_env(double* J, const double* chi, const double* E, long ntt) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= ntt) return;
  J[2 * idx] = chi[idx] * E[2 * idx];
  J[2 * idx + 1] = chi[idx] * E[2 * idx + 1];
}

// J += ne (|A|^2/4 - 1) A (paraxial.cl:99-110)
__global__ void plain_plasma_current_env(double* J, const double* A, const double* ne, long ntt) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= ntt) return;
  double ar = A[2 * idx], ai = A[2 * idx + 1];
  double f = ne[idx] * (0.25 * (ar * ar + ai * ai) - 1.0);
  J[2 * idx] += f * ar;
  J[2 * idx + 1] += f * ai;
}

// ---------------------------------------------------------------------------------------------------
// point kernels (kernels/uppe.cl, fft.cl, paraxial.cl)
// ---------------------------------------------------------------------------------------------------

// A = exp(i(kz+kg)z) q (uppe.cl:54-71)
__global__ void plain_propagate_linear(double* q, const double* kz, const double* kg, double z, long npt,
                                       long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  long f = idx / P;
  double kr = kz[2 * idx], ki = kz[2 * idx + 1];
  double ar = -ki * z, ai = (kr + kg[f]) * z;
  double ex = exp(ar);
  double er = cos(ai) * ex, ei = sin(ai) * ex;
  double qr = q[2 * idx], qi = q[2 * idx + 1];
  q[2 * idx] = er * qr - ei * qi;
  q[2 * idx + 1] = er * qi + ei * qr;
}

// S = 0.5 i exp(-i(kz+kg)z) J / Re kz (uppe.cl:73-91)
__global__ void plain_current_to_rhs(double* J, const double* kz, const double* kg, double z, long npt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  long f = idx / P;
  double kr = kz[2 * idx], ki = kz[2 * idx + 1];
  double ar = ki * z, ai = -((kr + kg[f]) * z);
  double ex = exp(ar);
  double er = cos(ai) * ex, ei = sin(ai) * ex;
  double jr = J[2 * idx], ji = J[2 * idx + 1];
  double tr = er * jr - ei * ji, ti = er * ji + ei * jr;
  J[2 * idx] = (-0.5 * ti) / kr;
  J[2 * idx + 1] = (0.5 * tr) / kr;
}

// E = i kz A (uppe.cl:94-108)
__global__ void plain_vp_to_e(double* A, const double* kz, long npt) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  double br = -kz[2 * idx + 1], bi = kz[2 * idx];
  double ar = A[2 * idx], ai = A[2 * idx + 1];
  A[2 * idx] = br * ar - bi * ai;
  A[2 * idx + 1] = br * ai + bi * ar;
}

// RKStage / RKFinish (uppe.cl:22-52)
__global__ void plain_rk_stage(double* qw, const double* qi, double* qf, const double* kn, double csol, double ceval,
                               long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  qw[i] = qi[i] + ceval * kn[i];
  qf[i] += csol * kn[i];
}

__global__ void plain_rk_finish(double* qw, const double* qf, const double* kn, double csol, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  qw[i] = qf[i] + csol * kn[i];
}

// Filter (fft.cl:318-329)
__global__ void plain_filter(double* A, const double* filt, long npt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  double f = filt[idx / P];
  A[2 * idx] *= f;
  A[2 * idx + 1] *= f;
}

// DtSpectral: A <- -i w A (fft.cl:282-299, sgn = 1)
__global__ void plain_dt_spectral(double* A, const double* w, long npt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  double w0 = w[idx / P];
  double ar = A[2 * idx], ai = A[2 * idx + 1];
  A[2 * idx] = w0 * ai;
  A[2 * idx + 1] = -w0 * ar;
}

// iDtSpectral: A <- A/(damping - i w) (fft.cl:301-316, sgn = 1)
__global__ void plain_idt_spectral(double* A, const double* w, double damping, long npt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  double cr = damping, ci = -w[idx / P];
  double ar = A[2 * idx], ai = A[2 * idx + 1];
  double den = cr * cr + ci * ci;
  A[2 * idx] = (ar * cr + ai * ci) / den;
  A[2 * idx + 1] = (ai * cr - ar * ci) / den;
}

__global__ void plain_add(double* a, const double* b, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  a[i] = a[i] + b[i];
}

__global__ void plain_mul(double* out, const double* a, const double* b, long n) {
  long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[i] = a[i] * b[i];
}

// SoftTimeWindow (fft.cl:331-355)
__global__ void plain_soft_window(double* A, int cutoff, int width, int ni, long ntt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= ntt) return;
  int i0 = (int)(idx / P);
  if (i0 < cutoff || i0 > ni - cutoff - 1) A[idx] = 0.0;
  if (i0 >= cutoff && i0 < cutoff + width) {
    double q = (double)(i0 - cutoff) / (double)width;
    A[idx] *= 10 * q * q * q - 15 * q * q * q * q + 6 * q * q * q * q * q;
  }
  if (i0 <= ni - cutoff - 1 && i0 > ni - cutoff - width - 1) {
    double q = (double)(ni - cutoff - 1 - i0) / (double)width;
    A[idx] *= 10 * q * q * q - 15 * q * q * q * q + 6 * q * q * q * q * q;
  }
}

// AddKerrChi on a zeroed chi: chi = ng chi3 E^2 (uppe.cl:132-151)
__global__ void plain_kerr_chi(double* chi, const double* E, const double* ng, double chi3, long ntt, long P) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= ntt) return;
  double e = E[idx];
  chi[idx] += ng[idx % P] * chi3 * e * e;
}

// ExplicitRate / ExplicitRateSeries (ionization.cl:22-44, 71-97)
__global__ void plain_rate(double* rate, const double* E, uppe_ionization p, long ntt) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= ntt) return;
  double e = fabs(E[idx]);
  if (p.kind == UPPE_ION_EXPLICIT) {
    double emag = p.e_conv * e + p.e_cutoff;
    rate[idx] = p.rate_conv * p.c_pre * pow(emag, p.c_pow) * exp(p.c_exp / emag);
  } else {
    double lg = log(p.e_conv * e + p.e_cutoff);
    rate[idx] = p.rate_conv * exp(p.c[0] + p.c[1] * lg + p.c[2] * lg * lg + p.c[3] * pow(lg, 3.0) +
                                  p.c[4] * pow(lg, 4.0) + p.c[5] * pow(lg, 5.0));
  }
}

// ComputePlasmaDensity: one work-item per strip, serial backward trapezoid (ionization.cl:127-155)
__global__ void plain_plasma_density(double* ne, const double* ng, double dt, int steps, long P) {
  long off = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (off >= P) return;
  double prev = 0.0, raw = 0.0;
  ne[off + (long)(steps - 1) * P] = 0.0;
  for (int k = steps - 1; k >= 0; --k) {
    double cur = ne[off + (long)k * P];
    raw += 0.5 * (cur + prev) * dt;
    ne[off + (long)k * P] = ng[off] * (1.0 - exp(-raw));
    prev = cur;
  }
}

// RootVolumeMultiply / Divide (fft.cl:363-385)
__global__ void plain_root_volume(double* a, const double* lam, int divide, long npt, int nr, int nm) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  int i = (int)((idx / nm) % nr);
  if (divide) {
    a[2 * idx] /= lam[i];
    a[2 * idx + 1] /= lam[i];
  } else {
    a[2 * idx] *= lam[i];
    a[2 * idx + 1] *= lam[i];
  }
}

// RadialTransform: vout[f,i,m] = sum_j H[i,j,m] vin[f,j,m], i < Nk (fft.cl:443-459)
__global__ void plain_radial_transform(const double* H, const double* vin, double* vout, int nw, int nk, int nr,
                                       int nm) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long total = (long)nw * nk * nm;
  if (idx >= total) return;
  int m = (int)(idx % nm);
  int i = (int)((idx / nm) % nk);
  long f = idx / ((long)nm * nk);
  double sr = 0.0, si = 0.0;
  for (int j = 0; j < nr; j++) {
    double t = H[(long)i * nr * nm + (long)j * nm + m];
    long v = f * nr * nm + (long)j * nm + m;
    sr += t * vin[2 * v];
    si += t * vin[2 * v + 1];
  }
  long o = f * nr * nm + (long)i * nm + m;
  vout[2 * o] = sr;
  vout[2 * o + 1] = si;
}

// InverseRadialTransform: vout[f,i,m] = sum_{k<Nk} H[k,i,m] vin[f,k,m] (fft.cl:461-477)
__global__ void plain_inverse_radial_transform(const double* H, const double* vin, double* vout, int nw, int nk,
                                               int nr, int nm) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  long total = (long)nw * nr * nm;
  if (idx >= total) return;
  int m = (int)(idx % nm);
  int i = (int)((idx / nm) % nr);
  long f = idx / ((long)nm * nr);
  double sr = 0.0, si = 0.0;
  for (int j = 0; j < nk; j++) {
    double t = H[(long)j * nr * nm + (long)i * nm + m];
    long v = f * nr * nm + (long)j * nm + m;
    sr += t * vin[2 * v];
    si += t * vin[2 * v + 1];
  }
  long o = f * nr * nm + (long)i * nm + m;
  vout[2 * o] = sr;
  vout[2 * o + 1] = si;
}

// LoadModulus (paraxial.cl:27-38)
__global__ void plain_modulus(const double* q, double* m, long npt) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  double a = q[2 * idx], b = q[2 * idx + 1];
  m[idx] = sqrt(a * a + b * b);
}

// LoadStepSize (paraxial.cl:40-58)
__global__ void plain_step_size(const double* q, const double* s, double* dz, double dz0, double dphi, double amin,
                                long npt) {
  long idx = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (idx >= npt) return;
  const double min_rate = 0.5 * dphi / dz0;
  double qr = q[2 * idx], qi = q[2 * idx + 1], sr = s[2 * idx], si = s[2 * idx + 1];
  double den = qr * qr + qi * qi + amin * amin;
  double alpha = (qr * sr + qi * si) / den;
  double keff = (qr * si - qi * sr) / den;
  double K = min_rate + fabs(keff) - alpha * (double)(alpha < 0.0);
  dz[idx] = dphi / K;
}

// per-block partial sum, max and min (the final stage runs on the host, as pyopencl's reductions do)
__global__ void plain_reduce(const double* v, long n, double* part) {
  __shared__ double ss[kBlock], smax[kBlock], smin[kBlock];
  double s = 0.0, mx = -INFINITY, mn = INFINITY;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double x = v[i];
    s += x;
    mx = fmax(mx, x);
    mn = fmin(mn, x);
    if (x != x) mx = x;
  }
  ss[threadIdx.x] = s;
  smax[threadIdx.x] = mx;
  smin[threadIdx.x] = mn;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      ss[threadIdx.x] += ss[threadIdx.x + o];
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + o]);
      smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = ss[0];
    part[3 * blockIdx.x + 1] = smax[0];
    part[3 * blockIdx.x + 2] = smin[0];
  }
}

// ---------------------------------------------------------------------------------------------------
// strip transforms: one workgroup per strip, LDS radix-2
// ---------------------------------------------------------------------------------------------------

// numpy irfft of each strip: in complex (Nw, P), out real (Nt, P) (fft.cl:257-280 semantics)
__global__ void plain_irfft(const double* in, double* out, int nw, long P, const c64* tw, int nt, int lgm) {
  extern __shared__ c64 z[];
  const long t = blockIdx.x;
  const int M = nw - 1;
  for (int k = threadIdx.x; k < M; k += blockDim.x) {
    c64 xk = cmk(in[2 * (k * P + t)], in[2 * (k * P + t) + 1]);
    long km = (long)(M - k);
    c64 xm = cmk(in[2 * (km * P + t)], in[2 * (km * P + t) + 1]);
    if (k == 0) {
      xk.im = 0.0;
      xm.im = 0.0;
    }
    c64 ze = cscale(cadd(xk, cconj(xm)), 0.5);
    c64 zo = cmul(cscale(csub(xk, cconj(xm)), 0.5), cconj(tw[k]));
    z[k] = cmk(ze.re - zo.im, ze.im + zo.re);
  }
  __syncthreads();
  lds_fft_radix2(z, M, lgm, tw, nt, +1);
  const double inv = 1.0 / (double)M;
  for (int n = threadIdx.x; n < M; n += blockDim.x) {
    out[(2L * n) * P + t] = z[n].re * inv;
    out[(2L * n + 1) * P + t] = z[n].im * inv;
  }
}

// numpy rfft of each strip: in real (Nt, P), out complex (Nw, P) (fft.cl:229-255 semantics)
__global__ void plain_rfft(const double* in, double* out, int nw, long P, const c64* tw, int nt, int lgm) {
  extern __shared__ c64 z[];
  const long t = blockIdx.x;
  const int M = nw - 1;
  for (int m = threadIdx.x; m < M; m += blockDim.x) z[m] = cmk(in[(2L * m) * P + t], in[(2L * m + 1) * P + t]);
  __syncthreads();
  lds_fft_radix2(z, M, lgm, tw, nt, -1);
  for (int k = threadIdx.x; k <= M; k += blockDim.x) {
    c64 zk = z[k % M], zm = z[(M - k) % M];
    c64 ze = cscale(cadd(zk, cconj(zm)), 0.5);
    c64 d = csub(zk, cconj(zm));
    c64 zo = cmk(0.5 * d.im, -0.5 * d.re);
    c64 wk = (k < M) ? tw[k] : cmk(-1.0, 0.0);
    c64 x = cadd(ze, cmul(wk, zo));
    if (k == 0 || k == M) x.im = 0.0;
    out[2 * (k * P + t)] = x.re;
    out[2 * (k * P + t) + 1] = x.im;
  }
}

// complex FFT along one transverse axis of (Nw, N1, N2): numpy sign, inverse scaled by 1/n
// (FFT_axis1/2, IFFT_axis1/2: fft.cl:387-441). Line l of (Nw, other) starts at base(l) with stride `stride`.
__global__ void plain_axis_fft(double* a, int n, int lg, int axis, int n1, int n2, const c64* tw, int inverse) {
  extern __shared__ c64 z[];
  const long l = blockIdx.x;
  long base, stride;
  if (axis == 1) {  // lines over x for each (f, y)
    long f = l / n2, y = l % n2;
    base = f * n1 * n2 + y;
    stride = n2;
  } else {  // lines over y for each (f, x)
    base = l * n2;
    stride = 1;
  }
  for (int i = threadIdx.x; i < n; i += blockDim.x) z[i] = cmk(a[2 * (base + i * stride)], a[2 * (base + i * stride) + 1]);
  __syncthreads();
  lds_fft_radix2(z, n, lg, tw, n, inverse ? +1 : -1);
  const double s = inverse ? 1.0 / (double)n : 1.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    a[2 * (base + i * stride)] = z[i].re * s;
    a[2 * (base + i * stride) + 1] = z[i].im * s;
  }
}

// complex FFT of every time strip of (Nt, P): numpy sign, inverse scaled by 1/n (FFT, IFFT: fft.cl:199-227)
__global__ void plain_strip_fft(double* a, int n, int lg, long P, const c64* tw, int inverse) {
  extern __shared__ c64 z[];
  const long t = blockIdx.x;
  for (int i = threadIdx.x; i < n; i += blockDim.x) z[i] = cmk(a[2 * (i * P + t)], a[2 * (i * P + t) + 1]);
  __syncthreads();
  lds_fft_radix2(z, n, lg, tw, n, inverse ? +1 : -1);
  const double s = inverse ? 1.0 / (double)n : 1.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    a[2 * (i * P + t)] = z[i].re * s;
    a[2 * (i * P + t) + 1] = z[i].im * s;
  }
}

// ---------------------------------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------------------------------

static int ilog2(long v) {
  int l = 0;
  while ((1L << l) < v) l++;
  return l;
}

static c64* make_twiddles(int n) {
  std::vector<c64> t(std::max(1, n / 2));
  for (int k = 0; k < n / 2; k++) t[k] = cmk(cos(2.0 * M_PI * k / n), -sin(2.0 * M_PI * k / n));
  c64* d = nullptr;
  UPPE_HIP(hipMalloc(&d, sizeof(c64) * t.size()));
  UPPE_HIP(hipMemcpy(d, t.data(), sizeof(c64) * t.size(), hipMemcpyHostToDevice));
  return d;
}

template <typename K, typename... A>
static void launch(uppe_handle* h, int kc, K kernel, unsigned grid, unsigned block, size_t shm, A... args) {
  ProfScope ps(h, kc);
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(block), shm, h->stream, args...);
  UPPE_LAUNCH_CHECK();
}

void plain_set_rotator(uppe_handle* h) {
  PlainState* s = h->plain;
  if (s->rot_tjj) UPPE_HIP(hipFree(s->rot_tjj));
  if (s->rot_wdt) UPPE_HIP(hipFree(s->rot_wdt));
  s->rot_tjj = s->rot_wdt = nullptr;
  if (h->rot_states < 3) return;
  UPPE_HIP(hipMalloc(&s->rot_tjj, sizeof(double) * h->rot_states));
  UPPE_HIP(hipMalloc(&s->rot_wdt, sizeof(double) * 2 * h->rot_states));
  UPPE_HIP(hipMemcpy(s->rot_tjj, h->rot_tjj.data(), sizeof(double) * h->rot_states, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->rot_wdt, h->rot_wdt.data(), sizeof(double) * 2 * h->rot_states, hipMemcpyHostToDevice));
}

void plain_create(uppe_handle* h) {
  if (h->nt > kMaxNtPlain) throw Error{UPPE_ERR_UNSUPPORTED, "plain path supports Nt <= 16384"};
  if (h->paraxial && h->nt > kMaxNtPlain / 2)
    throw Error{UPPE_ERR_UNSUPPORTED, "plain path supports Nw <= 8192 for the paraxial envelope"};
  auto* s = new PlainState();
  h->plain = s;
  const size_t cs = sizeof(double) * 2 * h->npt, ts = sizeof(double) * h->ntt;
  UPPE_HIP(hipMalloc(&s->w, sizeof(double) * h->nw));
  UPPE_HIP(hipMalloc(&s->kz, cs));
  UPPE_HIP(hipMalloc(&s->kg, sizeof(double) * h->nw));
  UPPE_HIP(hipMalloc(&s->ng, sizeof(double) * h->P));
  UPPE_HIP(hipMalloc(&s->nlf, sizeof(double) * h->nw));
  UPPE_HIP(hipMalloc(&s->cond, sizeof(double) * h->nw));
  double** cplxbufs[] = {&s->qw, &s->qi, &s->qf, &s->Ew, &s->Jw, &s->scratchw, &s->hscratch};
  for (auto p : cplxbufs) {
    UPPE_HIP(hipMalloc(p, cs));
    UPPE_HIP(hipMemset(*p, 0, cs));
  }
  UPPE_HIP(hipMalloc(&s->dzb, sizeof(double) * h->npt));
  double** tbufs[] = {&s->Et, &s->ne, &s->chi, &s->Jt};
  for (auto p : tbufs) {
    UPPE_HIP(hipMalloc(p, ts));
    UPPE_HIP(hipMemset(*p, 0, ts));
  }
  s->nblk_red = (int)std::min<long>(1024, grid_for(h->npt));
  UPPE_HIP(hipMalloc(&s->part, sizeof(double) * 3 * s->nblk_red));
  UPPE_HIP(hipHostMalloc(&s->part_host, sizeof(double) * 3 * s->nblk_red, hipHostMallocDefault));
  if (h->paraxial) {
    UPPE_HIP(hipMalloc(&s->k0, sizeof(double) * h->nw));
    double** ebufs[] = {&s->Etc, &s->Atc, &s->Jtc};
    for (auto p : ebufs) {
      UPPE_HIP(hipMalloc(p, 2 * ts));
      UPPE_HIP(hipMemset(*p, 0, 2 * ts));
    }
  }
  s->tw_t = make_twiddles(h->nt);
  s->tw1 = make_twiddles(h->n1);
  s->tw2 = make_twiddles(h->n2);
  if (h->d.geometry == UPPE_RADIAL) {
    UPPE_HIP(hipMalloc(&s->H, sizeof(double) * (size_t)h->nk * h->n1 * h->nm));
    UPPE_HIP(hipMalloc(&s->lam, sizeof(double) * h->n1));
  }
}

void plain_destroy(uppe_handle* h) {
  PlainState* s = h->plain;
  if (!s) return;
  void* ptrs[] = {s->w, s->kz, s->kg, s->ng, s->nlf, s->cond, s->H, s->lam, s->tw_t, s->tw1, s->tw2, s->qw,
                  s->qi, s->qf, s->Ew, s->Jw, s->scratchw, s->hscratch, s->dzb, s->Et, s->ne, s->chi, s->Jt,
                  s->part, s->rot_tjj, s->rot_wdt, s->k0, s->Etc, s->Atc, s->Jtc};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (s->part_host) (void)hipHostFree(s->part_host);
  delete s;
  h->plain = nullptr;
}

void plain_set_hankel(uppe_handle* h) {
  PlainState* s = h->plain;
  UPPE_HIP(hipMemcpy(s->H, h->H.data(), sizeof(double) * h->H.size(), hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->lam, h->lam.data(), sizeof(double) * h->n1, hipMemcpyHostToDevice));
}

void plain_set_medium(uppe_handle* h, const uppe_medium* m) {
  PlainState* s = h->plain;
  UPPE_HIP(hipMemcpy(s->w, m->w, sizeof(double) * h->nw, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->kz, m->kz, sizeof(double) * 2 * h->npt, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->kg, m->kg, sizeof(double) * h->nw, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->ng, m->ngas, sizeof(double) * h->P, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->nlf, m->nl_filter, sizeof(double) * h->nw, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->cond, m->cond_filter, sizeof(double) * h->nw, hipMemcpyHostToDevice));
  s->redchi3 = m->redchi3;
  s->dt = m->dt;
  s->ion = m->ion;
}

void plain_set_paraxial_medium(uppe_handle* h, const uppe_paraxial_medium* m) {
  PlainState* s = h->plain;
  UPPE_HIP(hipMemcpy(s->w, m->w, sizeof(double) * h->nw, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->kz, m->kz, sizeof(double) * 2 * h->npt, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->k0, m->k0, sizeof(double) * h->nw, hipMemcpyHostToDevice));
  UPPE_HIP(hipMemcpy(s->ng, m->ngas, sizeof(double) * h->P, hipMemcpyHostToDevice));
  s->redchi3 = m->redchi3;
  s->dt = m->dt;
  s->ion = m->ion;
}

// transverse transforms (grid_tools.py:220-229, 308-325)
static void kspacex(uppe_handle* h, double* a) {
  PlainState* s = h->plain;
  const long npt = h->npt;
  if (h->d.geometry == UPPE_CARTESIAN) {
    launch(h, UPPE_KC_TRANSVERSE_FFT, plain_axis_fft, (unsigned)(h->nw * h->n2), kBlock, sizeof(c64) * h->n1, a,
           h->n1, ilog2(h->n1), 1, h->n1, h->n2, (const c64*)s->tw1, 0);
    launch(h, UPPE_KC_TRANSVERSE_FFT, plain_axis_fft, (unsigned)(h->nw * h->n1), kBlock, sizeof(c64) * h->n2, a,
           h->n2, ilog2(h->n2), 2, h->n1, h->n2, (const c64*)s->tw2, 0);
    return;
  }
  if (h->n2 > 1)
    launch(h, UPPE_KC_TRANSVERSE_FFT, plain_axis_fft, (unsigned)(h->nw * h->n1), kBlock, sizeof(c64) * h->n2, a,
           h->n2, ilog2(h->n2), 2, h->n1, h->n2, (const c64*)s->tw2, 0);
  launch(h, UPPE_KC_POINT, plain_root_volume, grid_for(npt), kBlock, 0, a, (const double*)s->lam, 0, npt, h->n1,
         h->nm);
  UPPE_HIP(hipMemcpyAsync(s->hscratch, a, sizeof(double) * 2 * npt, hipMemcpyDeviceToDevice, h->stream));
  UPPE_HIP(hipMemsetAsync(a, 0, sizeof(double) * 2 * npt, h->stream));
  launch(h, UPPE_KC_HANKEL_R2K, plain_radial_transform, grid_for((long)h->nw * h->nk * h->nm), kBlock, 0,
         (const double*)s->H, (const double*)s->hscratch, a, h->nw, h->nk, h->n1, h->nm);
  launch(h, UPPE_KC_POINT, plain_root_volume, grid_for(npt), kBlock, 0, a, (const double*)s->lam, 1, npt, h->n1,
         h->nm);
}

static void rspacex(uppe_handle* h, double* a) {
  PlainState* s = h->plain;
  const long npt = h->npt;
  if (h->d.geometry == UPPE_CARTESIAN) {
    launch(h, UPPE_KC_TRANSVERSE_FFT, plain_axis_fft, (unsigned)(h->nw * h->n1), kBlock, sizeof(c64) * h->n2, a,
           h->n2, ilog2(h->n2), 2, h->n1, h->n2, (const c64*)s->tw2, 1);
    launch(h, UPPE_KC_TRANSVERSE_FFT, plain_axis_fft, (unsigned)(h->nw * h->n2), kBlock, sizeof(c64) * h->n1, a,
           h->n1, ilog2(h->n1), 1, h->n1, h->n2, (const c64*)s->tw1, 1);
    return;
  }
  launch(h, UPPE_KC_POINT, plain_root_volume, grid_for(npt), kBlock, 0, a, (const double*)s->lam, 0, npt, h->n1,
         h->nm);
  UPPE_HIP(hipMemcpyAsync(s->hscratch, a, sizeof(double) * 2 * npt, hipMemcpyDeviceToDevice, h->stream));
  launch(h, UPPE_KC_HANKEL, plain_inverse_radial_transform, grid_for((long)h->nw * h->n1 * h->nm), kBlock, 0,
         (const double*)s->H, (const double*)s->hscratch, a, h->nw, h->nk, h->n1, h->nm);
  launch(h, UPPE_KC_POINT, plain_root_volume, grid_for(npt), kBlock, 0, a, (const double*)s->lam, 1, npt, h->n1,
         h->nm);
  if (h->n2 > 1)
    launch(h, UPPE_KC_TRANSVERSE_FFT, plain_axis_fft, (unsigned)(h->nw * h->n1), kBlock, sizeof(c64) * h->n2, a,
           h->n2, ilog2(h->n2), 2, h->n1, h->n2, (const c64*)s->tw2, 1);
}

// update_current (uppe_kernel.py:139-186)
static void update_current(uppe_handle* h) {
  PlainState* s = h->plain;
  const long npt = h->npt, ntt = h->ntt, P = h->P;
  const int lgm = ilog2(h->nw - 1);
  const size_t shm = sizeof(c64) * (h->nw - 1);
  launch(h, UPPE_KC_SOURCE, plain_irfft, (unsigned)P, kBlock, shm, (const double*)s->Ew, s->Et, h->nw, P,
         (const c64*)s->tw_t, h->nt, lgm);
  if (s->ion.kind != UPPE_ION_NONE) {
    launch(h, UPPE_KC_POINT, plain_rate, grid_for(ntt), kBlock, 0, s->ne, (const double*)s->Et, s->ion, ntt);
    launch(h, UPPE_KC_SOURCE, plain_plasma_density, grid_for(P, 64), 64, 0, s->ne, (const double*)s->ng, s->dt,
           h->nt, P);
    launch(h, UPPE_KC_POINT, plain_soft_window, grid_for(ntt), kBlock, 0, s->ne, 4, 64, h->nt, ntt, P);
    launch(h, UPPE_KC_POINT, plain_mul, grid_for(ntt), kBlock, 0, s->Jt, (const double*)s->ne, (const double*)s->Et,
           ntt);
    launch(h, UPPE_KC_SOURCE, plain_rfft, (unsigned)P, kBlock, shm, (const double*)s->Jt, s->Jw, h->nw, P,
           (const c64*)s->tw_t, h->nt, lgm);
    launch(h, UPPE_KC_POINT, plain_idt_spectral, grid_for(npt), kBlock, 0, s->Jw, (const double*)s->w, 0.1, npt, P);
  } else {
    UPPE_HIP(hipMemsetAsync(s->ne, 0, sizeof(double) * ntt, h->stream));
    UPPE_HIP(hipMemsetAsync(s->Jw, 0, sizeof(double) * 2 * npt, h->stream));
  }
  UPPE_HIP(hipMemsetAsync(s->chi, 0, sizeof(double) * ntt, h->stream));
  if (h->rot_states > 2)
    launch(h, UPPE_KC_SOURCE, plain_add_chi_rotations, grid_for(P, 64), 64, 0, s->chi, (const double*)s->Et,
           (const double*)s->ng, (const double*)s->rot_tjj, (const double*)s->rot_wdt, s->dt, h->rot_dalpha,
           h->rot_hbar, h->nt, h->rot_states, P);
  launch(h, UPPE_KC_POINT, plain_kerr_chi, grid_for(ntt), kBlock, 0, s->chi, (const double*)s->Et,
         (const double*)s->ng, s->redchi3, ntt, P);
  launch(h, UPPE_KC_POINT, plain_mul, grid_for(ntt), kBlock, 0, s->Jt, (const double*)s->chi, (const double*)s->Et,
         ntt);
  launch(h, UPPE_KC_SOURCE, plain_rfft, (unsigned)P, kBlock, shm, (const double*)s->Jt, s->scratchw, h->nw, P,
         (const c64*)s->tw_t, h->nt, lgm);
  launch(h, UPPE_KC_POINT, plain_dt_spectral, grid_for(npt), kBlock, 0, s->scratchw, (const double*)s->w, npt, P);
  launch(h, UPPE_KC_POINT, plain_add, grid_for(2 * npt), kBlock, 0, s->Jw, (const double*)s->scratchw, 2 * npt);
  launch(h, UPPE_KC_POINT, plain_filter, grid_for(npt), kBlock, 0, s->Jw, (const double*)s->nlf, npt, P);
}

// paraxial update_current (paraxial_kernel.py:140-184): Ew holds A(w) in r-space; Jw receives J(w)
static void update_current_env(uppe_handle* h) {
  PlainState* s = h->plain;
  const long ntt = h->ntt, P = h->P;
  const int lg = ilog2(h->nt);
  const size_t shm = sizeof(c64) * h->nt;
  auto strips = [&](double* a, int inverse) {
    launch(h, UPPE_KC_SOURCE, plain_strip_fft, (unsigned)P, kBlock, shm, a, h->nt, lg, P, (const c64*)s->tw_t,
           inverse);
  };
  UPPE_HIP(hipMemcpyAsync(s->Etc, s->Ew, sizeof(double) * 2 * ntt, hipMemcpyDeviceToDevice, h->stream));
  UPPE_HIP(hipMemcpyAsync(s->Atc, s->Ew, sizeof(double) * 2 * ntt, hipMemcpyDeviceToDevice, h->stream));
  UPPE_HIP(hipMemsetAsync(s->ne, 0, sizeof(double) * ntt, h->stream));
  UPPE_HIP(hipMemsetAsync(s->chi, 0, sizeof(double) * ntt, h->stream));
  launch(h, UPPE_KC_POINT, plain_dt_spectral_sgn, grid_for(ntt), kBlock, 0, s->Etc, (const double*)s->w, -1.0, ntt, P);
  strips(s->Etc, 1);
  strips(s->Atc, 1);
  if (s->ion.kind != UPPE_ION_NONE) {
    launch(h, UPPE_KC_POINT, plain_rate_env, grid_for(ntt), kBlock, 0, s->ne, (const double*)s->Etc, s->ion, ntt);
    launch(h, UPPE_KC_SOURCE, plain_plasma_density, grid_for(P, 64), 64, 0, s->ne, (const double*)s->ng, s->dt,
           h->nt, P);
  }
  if (h->rot_states > 2)
    launch(h, UPPE_KC_SOURCE, plain_add_chi_rotations_env, grid_for(P, 64), 64, 0, s->chi, (const double*)s->Etc,
           (const double*)s->ng, (const double*)s->rot_tjj, (const double*)s->rot_wdt, s->dt, h->rot_dalpha,
           h->rot_hbar, h->nt, h->rot_states, P);
  launch(h, UPPE_KC_POINT, plain_kerr_chi_env, grid_for(ntt), kBlock, 0, s->chi, (const double*)s->Etc,
         (const double*)s->ng, s->redchi3, ntt, P);
  launch(h, UPPE_KC_POINT, plain_polarization_env, grid_for(ntt), kBlock, 0, s->Jtc, (const double*)s->chi,
         (const double*)s->Etc, ntt);
  strips(s->Jtc, 0);
  launch(h, UPPE_KC_POINT, plain_dt_spectral_sgn, grid_for(ntt), kBlock, 0, s->Jtc, (const double*)s->w, 1.0, ntt, P);
  strips(s->Jtc, 1);
  launch(h, UPPE_KC_POINT, plain_plasma_current_env, grid_for(ntt), kBlock, 0, s->Jtc, (const double*)s->Atc,
         (const double*)s->ne, ntt);
  strips(s->Jtc, 0);
  UPPE_HIP(hipMemcpyAsync(s->Jw, s->Jtc, sizeof(double) * 2 * ntt, hipMemcpyDeviceToDevice, h->stream));
}

// host-side final stage of the reductions: sum, max, min
static void reduce_host(uppe_handle* h, const double* v, double* sum, double* mx, double* mn) {
  PlainState* s = h->plain;
  launch(h, UPPE_KC_REDUCE, plain_reduce, (unsigned)s->nblk_red, kBlock, 0, v, h->npt, s->part);
  UPPE_HIP(hipMemcpyAsync(s->part_host, s->part, sizeof(double) * 3 * s->nblk_red, hipMemcpyDeviceToHost,
                          h->stream));
  UPPE_HIP(hipStreamSynchronize(h->stream));
  double a = 0.0, b = -INFINITY, c = INFINITY;
  bool nan = false;
  for (int i = 0; i < s->nblk_red; i++) {
    a += s->part_host[3 * i];
    if (s->part_host[3 * i + 1] != s->part_host[3 * i + 1]) nan = true;
    b = std::max(b, s->part_host[3 * i + 1]);
    c = std::min(c, s->part_host[3 * i + 2]);
  }
  if (nan) b = NAN;
  if (sum) *sum = a;
  if (mx) *mx = b;
  if (mn) *mn = c;
}

// load_source (uppe_kernel.py:188-232)
static double load_source(uppe_handle* h, double zphase, bool return_dz, double zlocal, double L, double dzmin,
                          double dphi, double rel_amin) {
  PlainState* s = h->plain;
  const long npt = h->npt, P = h->P;
  UPPE_HIP(hipMemcpyAsync(s->Ew, s->qw, sizeof(double) * 2 * npt, hipMemcpyDeviceToDevice, h->stream));
  if (h->paraxial) {
    // paraxial_kernel.py:186-208
    launch(h, UPPE_KC_POINT, plain_propagate_linear_env, grid_for(npt), kBlock, 0, s->Ew, (const double*)s->kz,
           zphase, npt);
    rspacex(h, s->Ew);
    update_current_env(h);
    kspacex(h, s->Jw);
    launch(h, UPPE_KC_POINT, plain_current_to_rhs_env, grid_for(npt), kBlock, 0, s->Jw, (const double*)s->kz,
           (const double*)s->k0, zphase, npt, P);
  } else {
    launch(h, UPPE_KC_POINT, plain_propagate_linear, grid_for(npt), kBlock, 0, s->Ew, (const double*)s->kz,
           (const double*)s->kg, zphase, npt, P);
    launch(h, UPPE_KC_POINT, plain_vp_to_e, grid_for(npt), kBlock, 0, s->Ew, (const double*)s->kz, npt);
    rspacex(h, s->Ew);
    update_current(h);
    kspacex(h, s->Jw);
    launch(h, UPPE_KC_POINT, plain_current_to_rhs, grid_for(npt), kBlock, 0, s->Jw, (const double*)s->kz,
           (const double*)s->kg, zphase, npt, P);
  }
  h->evals++;
  if (!return_dz) return 0.0;
  launch(h, UPPE_KC_POINT, plain_modulus, grid_for(npt), kBlock, 0, (const double*)s->qw, s->dzb, npt);
  double sum, mx;
  reduce_host(h, s->dzb, &sum, &mx, nullptr);
  if (std::isnan(sum)) h->nan_steps++;
  const double amin = rel_amin * mx;
  launch(h, UPPE_KC_POINT, plain_step_size, grid_for(npt), kBlock, 0, (const double*)s->qw, (const double*)s->Jw,
         s->dzb, L, dphi, amin, npt);
  double dz;
  reduce_host(h, s->dzb, nullptr, nullptr, &dz);
  if (dz < dzmin) dz = dzmin;
  if (zlocal + dz > L) dz = L - zlocal;
  return dz;
}

static void condition_field(uppe_handle* h) {
  if (h->paraxial) return;  // the paraxial stepper has no conditioner
  PlainState* s = h->plain;
  launch(h, UPPE_KC_POINT, plain_filter, grid_for(h->npt), kBlock, 0, s->qw, (const double*)s->cond, h->npt, h->P);
}

void plain_set_field(uppe_handle* h, const double* a_r) {
  PlainState* s = h->plain;
  UPPE_HIP(hipMemcpy(s->qw, a_r, sizeof(double) * 2 * h->npt, hipMemcpyHostToDevice));
  kspacex(h, s->qw);
  UPPE_HIP(hipStreamSynchronize(h->stream));
}

// propagator's subcycle/step loop (uppe_kernel.py:299-325)
int plain_advance(uppe_handle* h, const uppe_step_control* c) {
  PlainState* s = h->plain;
  const long n = 2 * h->npt;
  const double L = c->length;
  int taken = 0;
  for (int sub = 0; sub < c->n_sub && !(c->max_steps > 0 && taken >= c->max_steps); sub++) {
    double z0 = c->zi + (double)(c->sub0 + sub) * L;
    double zoff = z0 - c->zi;
    double z = 0.0;
    while (z < L && (z - L) * (z - L) > (L / 1e6) * (L / 1e6)) {
      if (c->max_steps > 0 && taken >= c->max_steps) break;
      UPPE_HIP(hipMemcpyAsync(s->qi, s->qw, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
      UPPE_HIP(hipMemcpyAsync(s->qf, s->qw, sizeof(double) * n, hipMemcpyDeviceToDevice, h->stream));
      double dz = load_source(h, zoff + z, true, z, L, c->dzmin, c->dphi, c->rel_amin);
      if (c->n_schedule > 0) {
        if (taken >= c->n_schedule) throw Error{UPPE_ERR_ARG, "dz schedule exhausted"};
        dz = c->dz_schedule[taken];
      }
      if (c->dz_record && taken < c->record_cap) c->dz_record[taken] = dz;
      launch(h, UPPE_KC_POINT, plain_rk_stage, grid_for(n), kBlock, 0, s->qw, (const double*)s->qi, s->qf,
             (const double*)s->Jw, dz / 6, dz / 2, n);
      condition_field(h);
      load_source(h, zoff + (z + 0.5 * dz), false, 0, L, 0, 0, 0);
      launch(h, UPPE_KC_POINT, plain_rk_stage, grid_for(n), kBlock, 0, s->qw, (const double*)s->qi, s->qf,
             (const double*)s->Jw, dz / 3, dz / 2, n);
      condition_field(h);
      load_source(h, zoff + (z + 0.5 * dz), false, 0, L, 0, 0, 0);
      launch(h, UPPE_KC_POINT, plain_rk_stage, grid_for(n), kBlock, 0, s->qw, (const double*)s->qi, s->qf,
             (const double*)s->Jw, dz / 3, dz, n);
      condition_field(h);
      load_source(h, zoff + (z + dz), false, 0, L, 0, 0, 0);
      launch(h, UPPE_KC_POINT, plain_rk_finish, grid_for(n), kBlock, 0, s->qw, (const double*)s->qf,
             (const double*)s->Jw, dz / 6, n);
      condition_field(h);
      z += dz;
      taken++;
      h->steps++;
    }
    h->z_last = z;
  }
  UPPE_HIP(hipStreamSynchronize(h->stream));
  return taken;
}

// finish (uppe_kernel.py:234-258)
void plain_finish(uppe_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t) {
  PlainState* s = h->plain;
  const long npt = h->npt;
  if (h->paraxial) {
    // paraxial_kernel.py:222-249
    UPPE_HIP(hipMemcpyAsync(s->Ew, s->qw, sizeof(double) * 2 * npt, hipMemcpyDeviceToDevice, h->stream));
    launch(h, UPPE_KC_POINT, plain_propagate_linear_env, grid_for(npt), kBlock, 0, s->Ew, (const double*)s->kz,
           z_total, npt);
    rspacex(h, s->Ew);
    update_current_env(h);
    UPPE_HIP(hipStreamSynchronize(h->stream));
    if (a_r) UPPE_HIP(hipMemcpy(a_r, s->Ew, sizeof(double) * 2 * npt, hipMemcpyDeviceToHost));
    if (j_r) UPPE_HIP(hipMemcpy(j_r, s->Jw, sizeof(double) * 2 * npt, hipMemcpyDeviceToHost));
    if (chi_t) UPPE_HIP(hipMemcpy(chi_t, s->chi, sizeof(double) * h->ntt, hipMemcpyDeviceToHost));
    if (ne_t) UPPE_HIP(hipMemcpy(ne_t, s->ne, sizeof(double) * h->ntt, hipMemcpyDeviceToHost));
    return;
  }
  launch(h, UPPE_KC_POINT, plain_propagate_linear, grid_for(npt), kBlock, 0, s->qw, (const double*)s->kz,
         (const double*)s->kg, z_total, npt, h->P);
  rspacex(h, s->qw);
  UPPE_HIP(hipMemcpyAsync(s->Ew, s->qw, sizeof(double) * 2 * npt, hipMemcpyDeviceToDevice, h->stream));
  launch(h, UPPE_KC_POINT, plain_vp_to_e, grid_for(npt), kBlock, 0, s->Ew, (const double*)s->kz, npt);
  update_current(h);
  UPPE_HIP(hipStreamSynchronize(h->stream));
  if (a_r) UPPE_HIP(hipMemcpy(a_r, s->qw, sizeof(double) * 2 * npt, hipMemcpyDeviceToHost));
  if (j_r) UPPE_HIP(hipMemcpy(j_r, s->Jw, sizeof(double) * 2 * npt, hipMemcpyDeviceToHost));
  if (chi_t) UPPE_HIP(hipMemcpy(chi_t, s->chi, sizeof(double) * h->ntt, hipMemcpyDeviceToHost));
  if (ne_t) UPPE_HIP(hipMemcpy(ne_t, s->ne, sizeof(double) * h->ntt, hipMemcpyDeviceToHost));
}

void plain_get_state(uppe_handle* h, double* q) {
  UPPE_HIP(hipStreamSynchronize(h->stream));
  UPPE_HIP(hipMemcpy(q, h->plain->qw, sizeof(double) * 2 * h->npt, hipMemcpyDeviceToHost));
}

}  // namespace uppe

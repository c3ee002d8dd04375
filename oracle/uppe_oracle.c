/*
 * uppe_oracle.c -- FP64 CPU restatement of SeaRay's UPPE stepper. TEST INFRASTRUCTURE ONLY.
 *
 * This is the checker for the HIP library, never the product: only tests/, the smoke check and the CPU
 * baseline leg of bench.py load it. It keeps the reference's algorithm and pass structure so that it is
 * also the CPU timing baseline:
 *   - reference layout, spectral (Nw, N1, N2) complex, time (Nt, N1, N2) real, strips strided by N1*N2;
 *   - one loop per reference kernel (one loop iteration per OpenCL work-item);
 *   - the strip radix-2 FFT with serial bit reversal and recurrence twiddles (kernels/fft.cl:66-197);
 *   - naive radial transforms (kernels/fft.cl:443-477);
 *   - the propagator/load_source/update_current/finish sequence of modules/uppe_kernel.py:131-331,
 *     including every per-step reduction (max, sum, min) of the step-size control.
 * Each function cites the reference lines it restates. Exported entry points mirror include/searay_uppe.h
 * with the prefix orc_.
 *
 * Pinning (tests/test_oracle.py): the strip FFTs against numpy; the linear propagator and transforms against
 * the exact interaction-picture solution and the reference's Gaussian-beam test; RK4 against the Taylor
 * polynomial of a linear right-hand side; the step-size rule against a brute-force restatement; the Kerr and
 * plasma currents, the density scan and the rates against closed forms and the host models.
 * Parity unpinned: the full nonlinear run as a whole (the reference stores no nonlinear output to compare
 * with); it is pinned only through the parts above.
 *
 * The paraxial envelope equation (modules/paraxial_kernel.py, kernels/paraxial.cl) is the handle's second
 * model: complex time strips of Nt = Nw samples, the paraxial linear propagator and RHS, envelope rates, the
 * 0.75 Kerr factor, the envelope rotator drive and the plasma current ne (|A|^2/4 - 1) A. Pinned by the
 * reference's paraxial Gaussian-beam test (test/gaussian_test.py:63-108), the self-phase modulation of a
 * monochromatic plane wave, and the same step-control and transform checks as the UPPE model.
 */
#include "../include/searay_uppe.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ---------------------------------------------------------------------------------------------------- */
/* Strip FFTs (kernels/fft.cl)                                                                           */
/* ---------------------------------------------------------------------------------------------------- */

/* bit-reversed index (fft.cl:48-64) */
static long reverse_bits(long v, int nbits) {
  long r = 0;
  for (int b = 0; b < nbits; b++) {
    r = (r << 1) | (v & 1);
    v >>= 1;
  }
  return r;
}

/* In-place radix-2 decimation-in-time transform of n points spaced `interval` apart (fft.cl:66-136).
 * Twiddles advance by the recurrence W <- W*R with R = (cos(pi*inv/half), expsign*sin(pi*inv/half));
 * inversion == -1 divides by n at the end. */
void orc_complex_fft(double* re, double* im, long n, long interval, double inversion, double expsign) {
  int nbits = 0;
  for (long t = n; t > 1; t >>= 1) nbits++;
  for (long i = 0; i < n; i++) {
    long j = reverse_bits(i, nbits);
    if (j > i) {
      double t = re[j * interval];
      re[j * interval] = re[i * interval];
      re[i * interval] = t;
      t = im[j * interval];
      im[j * interval] = im[i * interval];
      im[i * interval] = t;
    }
  }
  long half = 1;
  for (long groups = n; groups > 1; groups >>= 1) {
    double ang = M_PI * inversion / (double)half;
    double rr = cos(ang), ri = expsign * sin(ang);
    double wr = 1.0, wi = 0.0;
    for (long i = 0; i < half; i++) {
      for (long j = 0; j < groups; j += 2) {
        long a = interval * (j * half + i);
        long b = interval * ((j + 1) * half + i);
        double fr = re[b] * wr - im[b] * wi;
        double fi = re[b] * wi + im[b] * wr;
        double ar = re[a], ai = im[a];
        re[b] = ar - fr;
        im[b] = ai - fi;
        re[a] = ar + fr;
        im[a] = ai + fi;
      }
      double t = wr * rr - wi * ri;
      wi = wr * ri + wi * rr;
      wr = t;
    }
    half *= 2;
  }
  if (inversion == -1.0) {
    for (long i = 0; i < n; i++) {
      re[i * interval] /= (double)n;
      im[i * interval] /= (double)n;
    }
  }
}

/* Real transform of n samples in the packed two-for-one form (fft.cl:138-197). Forward leaves
 * a[0] = lowest and a[s] = highest frequency, then (a[2k s], a[(2k+1) s]) for 0 < k < n/2; invert undoes it
 * including the 1/n normalisation of the inner complex transform. */
void orc_real_fft(double* a, long n, long s, int invert) {
  double theta = M_PI / (double)(n / 2);
  const double c1 = 0.5;
  double c2;
  if (!invert) {
    c2 = -0.5;
    orc_complex_fft(a, a + s, n / 2, 2 * s, 1.0, 1.0);
  } else {
    c2 = 0.5;
    theta = -theta;
  }
  const double rr = cos(theta), ri = sin(theta);
  double wr = rr, wi = ri;
  for (long i = 1; i < n / 4; i++) {
    long i1 = s * (2 * i), i2 = s * (2 * i + 1), i3 = s * (n - 2 * i), i4 = s * (n + 1 - 2 * i);
    double h1r = c1 * (a[i1] + a[i3]), h1i = c1 * (a[i2] - a[i4]);
    double h2r = -c2 * (a[i2] + a[i4]), h2i = c2 * (a[i1] - a[i3]);
    double tr = wr * h2r - wi * h2i, ti = wr * h2i + wi * h2r;
    tr += h1r;
    ti += h1i;
    a[i1] = tr;
    a[i2] = ti;
    tr -= 2 * h1r;
    ti -= 2 * h1i;
    a[i3] = -tr;
    a[i4] = ti;
    double t = wr * rr - wi * ri;
    wi = wr * ri + wi * rr;
    wr = t;
  }
  double t0 = a[0];
  if (!invert) {
    a[0] = t0 + a[s];
    a[s] = t0 - a[s];
  } else {
    a[0] = c1 * (t0 + a[s]);
    a[s] = c1 * (t0 - a[s]);
    orc_complex_fft(a, a + s, n / 2, 2 * s, -1.0, 1.0);
  }
}

/* numpy-layout rfft of one strip (fft.cl:229-255): `in` real, `out` complex interleaved, both strided by s
 * elements. The reference then inverse-transforms `in` back in place; that work is kept here so the CPU
 * baseline does what the reference does. */
void orc_rfft_strip(double* in, double* out, long steps, long s) {
  orc_real_fft(in, steps, s, 0);
  out[0] = in[0];
  out[1] = 0.0;
  for (long i = 1; i < steps / 2; i++) {
    out[2 * s * i] = in[2 * i * s];
    out[2 * s * i + 1] = -in[(2 * i + 1) * s];
  }
  out[steps * s] = in[s];
  out[steps * s + 1] = 0.0;
  orc_real_fft(in, steps, s, 1);
}

/* numpy-layout irfft of one strip (fft.cl:257-280). */
void orc_irfft_strip(const double* in, double* out, long modes, long s) {
  long steps = 2 * modes - 2;
  out[0] = in[0];
  out[s] = in[2 * s * (modes - 1)];
  for (long i = 1; i < modes - 1; i++) {
    out[2 * i * s] = in[2 * s * i];
    out[(2 * i + 1) * s] = -in[2 * s * i + 1];
  }
  orc_real_fft(out, steps, s, 1);
}

/* ---------------------------------------------------------------------------------------------------- */
/* Stepper state                                                                                         */
/* ---------------------------------------------------------------------------------------------------- */

typedef struct orc_handle {
  uppe_desc d;
  long nw, n1, n2, nk, nm, nt, ntr, npt, ntt;
  int paraxial;  /* UPPE_MODEL_PARAXIAL */
  /* medium */
  double *w, *kz, *kg, *ng, *nlf, *cond;
  double* k0;            /* paraxial: k0(w) = w0 dn + w ng */
  double *Etc, *Atc, *Jtc; /* paraxial: complex time strips of E, A and the current */
  double redchi3, dt;
  uppe_ionization ion;
  /* rotational susceptibility (rotations.py:44-85) */
  int rot_states;
  double *rot_tjj, *rot_wdt, rot_dalpha, rot_hbar;
  /* transverse basis */
  double *H, *lam;
  /* device-array counterparts of Source (uppe_kernel.py:108-119) and the Hankel scratch (grid_tools.py:302) */
  double *qw, *qi, *qf, *Ew, *Jw, *scratchw, *hscratch, *dzb;
  double *Et, *ne, *chi, *Jt;
  int have_field, have_medium, have_hankel;
  int linear_rhs;          /* test hook: S = lambda q in place of the UPPE right-hand side */
  double lam_re, lam_im;
  long long steps, evals;
  int nan_steps;
  double z_last;
  char err[256];
} orc_handle;

static int fail(orc_handle* h, int code, const char* msg) {
  if (h) snprintf(h->err, sizeof(h->err), "%s", msg);
  return code;
}

static int is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

/* ---------------------------------------------------------------------------------------------------- */
/* Point and strip kernels (kernels/uppe.cl, fft.cl, ionization.cl, paraxial.cl)                         */
/* ---------------------------------------------------------------------------------------------------- */

/* A = exp(i(kz+kg)z) q with complex kz (uppe.cl:54-71) */
static void k_propagate_linear(orc_handle* h, double* q, double z) {
  const long per = h->ntr;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    long f = idx / per;
    double kr = h->kz[2 * idx], ki = h->kz[2 * idx + 1];
    double ar = -ki * z, ai = (kr + h->kg[f]) * z;
    double er = cos(ai) * exp(ar), ei = sin(ai) * exp(ar);
    double qr = q[2 * idx], qi = q[2 * idx + 1];
    q[2 * idx] = er * qr - ei * qi;
    q[2 * idx + 1] = er * qi + ei * qr;
  }
}

/* S = 0.5 i exp(-i(kz+kg)z) J / Re(kz) (uppe.cl:73-91) */
static void k_current_to_rhs(orc_handle* h, double* J, double z) {
  const long per = h->ntr;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    long f = idx / per;
    double kr = h->kz[2 * idx], ki = h->kz[2 * idx + 1];
    double ar = ki * z, ai = -(kr + h->kg[f]) * z; /* -iKz */
    double er = cos(ai) * exp(ar), ei = sin(ai) * exp(ar);
    double jr = J[2 * idx], ji = J[2 * idx + 1];
    double tr = er * jr - ei * ji, ti = er * ji + ei * jr;
    /* (0, 0.5) * t */
    J[2 * idx] = (-0.5 * ti) / kr;
    J[2 * idx + 1] = (0.5 * tr) / kr;
  }
}

/* E = i kz A (uppe.cl:94-108) */
static void k_vp_to_e(orc_handle* h, double* A) {
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    double kr = h->kz[2 * idx], ki = h->kz[2 * idx + 1];
    double br = -ki, bi = kr; /* i*kz */
    double ar = A[2 * idx], ai = A[2 * idx + 1];
    A[2 * idx] = br * ar - bi * ai;
    A[2 * idx + 1] = br * ai + bi * ar;
  }
}

/* qw = qi + Ceval k; qf += Csol k (uppe.cl:22-39) */
static void k_rk_stage(orc_handle* h, double csol, double ceval) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 2 * h->npt; i++) {
    h->qw[i] = h->qi[i] + ceval * h->Jw[i];
    h->qf[i] += csol * h->Jw[i];
  }
}

/* qw = qf + Csol k (uppe.cl:41-52) */
static void k_rk_finish(orc_handle* h, double csol) {
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 2 * h->npt; i++) h->qw[i] = h->qf[i] + csol * h->Jw[i];
}

/* A <- -sgn i w A (fft.cl:282-299) */
static void k_dt_spectral(orc_handle* h, double* A, double sgn) {
  const long per = h->ntr;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    double w0 = h->w[idx / per];
    double ar = A[2 * idx], ai = A[2 * idx + 1];
    A[2 * idx] = sgn * w0 * ai;
    A[2 * idx + 1] = -sgn * w0 * ar;
  }
}

/* A <- A / (damping - i sgn w) (fft.cl:301-316) */
static void k_idt_spectral(orc_handle* h, double* A, double damping, double sgn) {
  const long per = h->ntr;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    double cr = damping, ci = -sgn * h->w[idx / per];
    double ar = A[2 * idx], ai = A[2 * idx + 1];
    double den = cr * cr + ci * ci;
    /* a * conj(c) / |c|^2 */
    double tr = ar * cr + ai * ci, ti = ai * cr - ar * ci;
    A[2 * idx] = tr / den;
    A[2 * idx + 1] = ti / den;
  }
}

/* A[w,...] *= filter[w] (fft.cl:318-329) */
static void k_filter(orc_handle* h, double* A, const double* filt) {
  const long per = h->ntr;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    double f = filt[idx / per];
    A[2 * idx] *= f;
    A[2 * idx + 1] *= f;
  }
}

/* zero `cutoff` samples at each end, ramp over `width` with 10q^3-15q^4+6q^5 (fft.cl:331-355) */
static void k_soft_time_window(orc_handle* h, double* A, long cutoff, long width) {
  const long per = h->ntr, ni = h->nt;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->ntt; idx++) {
    long i0 = idx / per;
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
}

/* chi += ng chi3 E^2 (uppe.cl:132-151) */
static void k_add_kerr_chi(orc_handle* h) {
  const long per = h->ntr;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->ntt; idx++) {
    double e = h->Et[idx];
    h->chi[idx] += h->ng[idx % per] * h->redchi3 * e * e;
  }
}

/* chi += rotational susceptibility, one strip per iteration, serially over the states and over time from
 * the last sample (rotations.cl:80-129, AddChiUPPE) */
static void k_add_chi_rotations(orc_handle* h) {
  const long per = h->ntr, steps = h->nt;
  const double dt = h->dt, dalpha = h->rot_dalpha, hbar = h->rot_hbar;
#pragma omp parallel for schedule(static)
  for (long xy = 0; xy < per; xy++) {
    for (int j = 2; j < h->rot_states; j++) {
      const double wr = h->rot_wdt[2 * j], wi = h->rot_wdt[2 * j + 1];
      const double Sjj = 2.0 * j * (j - 1) / (15 * (2 * j - 1));
      double rr = 0.0, ri = 0.0;                 /* rhojj */
      double er = 1.0, ei = 0.0;                 /* exp_iwt */
      const double hr = cos(0.5 * wr) * exp(-0.5 * wi), hi = sin(0.5 * wr) * exp(-0.5 * wi); /* exp_iwdth */
      double cum = 0.0;
      for (long k = steps - 1; k >= 0; --k) {
        const long idx = k * per + xy;
        /* term = Tjj Sjj (rho conj(exp_iwt)) wdt */
        const double ar = rr * er + ri * ei, ai = -rr * ei + ri * er;
        const double ti = h->rot_tjj[j] * Sjj * (ar * wi + ai * wr);
        cum += 2 * ti * h->ng[xy] * dalpha;
        h->chi[idx] += cum;
        const double e = h->Et[idx];
        const double F = 0.5 * (dalpha / hbar) * e * e;
        double nr = er * hr - ei * hi, ni = er * hi + ei * hr;
        er = nr;
        ei = ni;
        /* rho -= dt F i exp_iwt */
        rr -= dt * F * (-ei);
        ri -= dt * F * er;
        nr = er * hr - ei * hi;
        ni = er * hi + ei * hr;
        er = nr;
        ei = ni;
      }
    }
  }
}

/* P = chi E (uppe.cl:153-165) */
static void k_source_polarization(orc_handle* h) {
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->ntt; idx++) h->Jt[idx] = h->chi[idx] * h->Et[idx];
}

/* closed-form and series rates (ionization.cl:22-44, 71-97) */
static void k_rate(orc_handle* h) {
  const uppe_ionization* p = &h->ion;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->ntt; idx++) {
    double e = fabs(h->Et[idx]);
    if (p->kind == UPPE_ION_EXPLICIT) {
      double emag = p->e_conv * e + p->e_cutoff;
      h->ne[idx] = p->rate_conv * p->c_pre * pow(emag, p->c_pow) * exp(p->c_exp / emag);
    } else {
      double eau = p->e_conv * e;
      double lg = log(eau + p->e_cutoff);
      h->ne[idx] = p->rate_conv * exp(p->c[0] + p->c[1] * lg + p->c[2] * lg * lg + p->c[3] * pow(lg, 3) +
                                      p->c[4] * pow(lg, 4) + p->c[5] * pow(lg, 5));
    }
  }
}

/* backward trapezoid with depletion; the rate at the last index is zeroed first (ionization.cl:127-155) */
static void k_plasma_density(orc_handle* h) {
  const long s = h->ntr, steps = h->nt;
#pragma omp parallel for schedule(static)
  for (long off = 0; off < s; off++) {
    double prev = 0.0, raw = 0.0;
    h->ne[off + (steps - 1) * s] = 0.0;
    for (long k = steps - 1; k >= 0; --k) {
      double cur = h->ne[off + k * s];
      raw += 0.5 * (cur + prev) * h->dt;
      h->ne[off + k * s] = h->ng[off] * (1.0 - exp(-raw));
      prev = cur;
    }
  }
}

/* |q| (paraxial.cl:27-38) */
static void k_load_modulus(orc_handle* h) {
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    double qr = h->qw[2 * idx], qi = h->qw[2 * idx + 1];
    h->dzb[idx] = sqrt(qr * qr + qi * qi);
  }
}

/* per-point step estimate (paraxial.cl:40-58) */
static void k_load_step_size(orc_handle* h, double dz0, double dphi, double amin) {
  const double min_rate = 0.5 * dphi / dz0;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    double qr = h->qw[2 * idx], qi = h->qw[2 * idx + 1];
    double sr = h->Jw[2 * idx], si = h->Jw[2 * idx + 1];
    double den = qr * qr + qi * qi + amin * amin;
    double alpha = (qr * sr + qi * si) / den;
    double keff = (qr * si - qi * sr) / den;
    double K = min_rate + fabs(keff) - alpha * (double)(alpha < 0.0);
    h->dzb[idx] = dphi / K;
  }
}

/* ---------------------------------------------------------------------------------------------------- */
/* Paraxial envelope kernels (kernels/paraxial.cl, ionization.cl:46-69,99-125, rotations.cl:26-78)       */
/* ---------------------------------------------------------------------------------------------------- */

/* A = exp(i kz z) q with complex kz: iKz = (-Ki z, Kr z), exp(iKz) = exp(-Ki z)(cos Kr z, sin Kr z)
 * (paraxial.cl:60-77) */
static void k_propagate_linear_env(orc_handle* h, double* q, double z) {
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    double kr = h->kz[2 * idx], ki = h->kz[2 * idx + 1];
    double ar = -ki * z, ai = kr * z;
    double er = cos(ai) * exp(ar), ei = sin(ai) * exp(ar);
    double qr = q[2 * idx], qi = q[2 * idx + 1];
    q[2 * idx] = er * qr - ei * qi;
    q[2 * idx + 1] = er * qi + ei * qr;
  }
}

/* S = (i/2) exp(-i kz z) J / k0(w) (paraxial.cl:79-97) */
static void k_current_to_rhs_env(orc_handle* h, double* J, double z) {
  const long per = h->ntr;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    double kr = h->kz[2 * idx], ki = h->kz[2 * idx + 1];
    double ar = ki * z, ai = -kr * z; /* -iKz */
    double er = cos(ai) * exp(ar), ei = sin(ai) * exp(ar);
    double jr = J[2 * idx], ji = J[2 * idx + 1];
    double tr = er * jr - ei * ji, ti = er * ji + ei * jr;
    double k0 = h->k0[idx / per];
    J[2 * idx] = (0.0 * tr - 0.5 * ti) / k0;
    J[2 * idx + 1] = (0.0 * ti + 0.5 * tr) / k0;
  }
}

/* complex strip transforms of every (x, y) strip: FFT (inversion 1) or IFFT with 1/N (fft.cl:199-227) */
static void strip_fft_all(orc_handle* h, double* a, double inversion) {
#pragma omp parallel for schedule(static)
  for (long t = 0; t < h->ntr; t++) orc_complex_fft(a + 2 * t, a + 2 * t + 1, h->nt, 2 * h->ntr, inversion, -1.0);
}

/* rates of the envelope modulus |E| (EnvelopeRate, EnvelopeRateSeries) */
static void k_rate_env(orc_handle* h) {
  const uppe_ionization* p = &h->ion;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->ntt; idx++) {
    double er = h->Etc[2 * idx], ei = h->Etc[2 * idx + 1];
    double e = sqrt(er * er + ei * ei);
    if (p->kind == UPPE_ION_EXPLICIT) {
      double eau = p->e_conv * e + p->e_cutoff;
      h->ne[idx] = p->rate_conv * p->c_pre * pow(eau, p->c_pow) * exp(p->c_exp / eau);
    } else {
      double lg = log(p->e_conv * e + p->e_cutoff);
      h->ne[idx] = p->rate_conv * exp(p->c[0] + p->c[1] * lg + p->c[2] * lg * lg + p->c[3] * pow(lg, 3) +
                                      p->c[4] * pow(lg, 4) + p->c[5] * pow(lg, 5));
    }
  }
}

/* chi += 0.75 ng chi3 |E|^2 (paraxial.cl:112-131) */
static void k_add_kerr_chi_env(orc_handle* h) {
  const long per = h->ntr;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->ntt; idx++) {
    double er = h->Etc[2 * idx], ei = h->Etc[2 * idx + 1];
    h->chi[idx] += 0.75 * h->ng[idx % per] * h->redchi3 * (er * er + ei * ei);
  }
}

/* rotational susceptibility driven by the cycle average |E|^2 / 2 (rotations.cl:26-78, AddChiParaxial) */
static void k_add_chi_rotations_env(orc_handle* h) {
  const long per = h->ntr, steps = h->nt;
  const double dt = h->dt, dalpha = h->rot_dalpha, hbar = h->rot_hbar;
#pragma omp parallel for schedule(static)
  for (long xy = 0; xy < per; xy++) {
    for (int j = 2; j < h->rot_states; j++) {
      const double wr = h->rot_wdt[2 * j], wi = h->rot_wdt[2 * j + 1];
      const double Sjj = 2.0 * j * (j - 1) / (15 * (2 * j - 1));
      double rr = 0.0, ri = 0.0, er = 1.0, ei = 0.0, cum = 0.0;
      const double hr = cos(0.5 * wr) * exp(-0.5 * wi), hi = sin(0.5 * wr) * exp(-0.5 * wi);
      for (long k = steps - 1; k >= 0; --k) {
        const long idx = k * per + xy;
        const double ar = rr * er + ri * ei, ai = -rr * ei + ri * er;
        cum += 2 * (h->rot_tjj[j] * Sjj * (ar * wi + ai * wr)) * h->ng[xy] * dalpha;
        h->chi[idx] += cum;
        const double Er = h->Etc[2 * idx], Ei = h->Etc[2 * idx + 1];
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
}

/* P = chi E, complex (paraxial.cl:133-147) */
static void k_source_polarization_env(orc_handle* h) {
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->ntt; idx++) {
    h->Jtc[2 * idx] = h->chi[idx] * h->Etc[2 * idx];
    h->Jtc[2 * idx + 1] = h->chi[idx] * h->Etc[2 * idx + 1];
  }
}

/* J += ne (|A|^2/4 - 1) A (paraxial.cl:99-110) */
static void k_plasma_current_env(orc_handle* h) {
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->ntt; idx++) {
    double ar = h->Atc[2 * idx], ai = h->Atc[2 * idx + 1];
    double f = h->ne[idx] * (0.25 * (ar * ar + ai * ai) - 1.0);
    h->Jtc[2 * idx] += f * ar;
    h->Jtc[2 * idx + 1] += f * ai;
  }
}

/* ---------------------------------------------------------------------------------------------------- */
/* Transverse transforms (grid_tools.py:220-229, 308-325; fft.cl:363-477)                               */
/* ---------------------------------------------------------------------------------------------------- */

static void axis1_fft(orc_handle* h, double* a, double inversion) {
  const long n1 = h->n1, n2 = h->n2;
#pragma omp parallel for collapse(2) schedule(static)
  for (long f = 0; f < h->nw; f++)
    for (long y = 0; y < n2; y++) {
      long off = 2 * (f * n1 * n2 + y);
      orc_complex_fft(a + off, a + off + 1, n1, 2 * n2, inversion, -1.0);
    }
}

static void axis2_fft(orc_handle* h, double* a, double inversion) {
  const long n1 = h->n1, n2 = h->n2;
#pragma omp parallel for collapse(2) schedule(static)
  for (long f = 0; f < h->nw; f++)
    for (long x = 0; x < n1; x++) {
      long off = 2 * (f * n1 * n2 + x * n2);
      orc_complex_fft(a + off, a + off + 1, n2, 2, inversion, -1.0);
    }
}

static void root_volume(orc_handle* h, double* a, int divide) {
  const long nr = h->n1, nm = h->nm;
#pragma omp parallel for schedule(static)
  for (long idx = 0; idx < h->npt; idx++) {
    long i = (idx / nm) % nr;
    if (divide) {
      a[2 * idx] /= h->lam[i];
      a[2 * idx + 1] /= h->lam[i];
    } else {
      a[2 * idx] *= h->lam[i];
      a[2 * idx + 1] *= h->lam[i];
    }
  }
}

/* vout[f,i,m] = sum_j H[i,j,m] vin[f,j,m] for i < Nk (fft.cl:443-459) */
static void radial_transform(orc_handle* h, const double* vin, double* vout) {
  const long nr = h->n1, nm = h->nm, nk = h->nk;
#pragma omp parallel for collapse(2) schedule(static)
  for (long f = 0; f < h->nw; f++)
    for (long i = 0; i < nk; i++)
      for (long m = 0; m < nm; m++) {
        double sr = 0.0, si = 0.0;
        for (long j = 0; j < nr; j++) {
          double t = h->H[i * nr * nm + j * nm + m];
          sr += t * vin[2 * (f * nr * nm + j * nm + m)];
          si += t * vin[2 * (f * nr * nm + j * nm + m) + 1];
        }
        vout[2 * (f * nr * nm + i * nm + m)] = sr;
        vout[2 * (f * nr * nm + i * nm + m) + 1] = si;
      }
}

/* vout[f,i,m] = sum_{k<Nk} H[k,i,m] vin[f,k,m] (fft.cl:461-477) */
static void inverse_radial_transform(orc_handle* h, const double* vin, double* vout) {
  const long nr = h->n1, nm = h->nm, nk = h->nk;
#pragma omp parallel for collapse(2) schedule(static)
  for (long f = 0; f < h->nw; f++)
    for (long i = 0; i < nr; i++)
      for (long m = 0; m < nm; m++) {
        double sr = 0.0, si = 0.0;
        for (long j = 0; j < nk; j++) {
          double t = h->H[j * nr * nm + i * nm + m];
          sr += t * vin[2 * (f * nr * nm + j * nm + m)];
          si += t * vin[2 * (f * nr * nm + j * nm + m) + 1];
        }
        vout[2 * (f * nr * nm + i * nm + m)] = sr;
        vout[2 * (f * nr * nm + i * nm + m) + 1] = si;
      }
}

/* T.kspacex (grid_tools.py:220-224, 308-317) */
static void kspacex(orc_handle* h, double* a) {
  if (h->d.geometry == UPPE_CARTESIAN) {
    axis1_fft(h, a, 1.0);
    axis2_fft(h, a, 1.0);
    return;
  }
  axis2_fft(h, a, 1.0);
  root_volume(h, a, 0);
  memcpy(h->hscratch, a, sizeof(double) * 2 * h->npt);
  memset(a, 0, sizeof(double) * 2 * h->npt);
  radial_transform(h, h->hscratch, a);
  root_volume(h, a, 1);
}

/* T.rspacex (grid_tools.py:225-229, 318-325) */
static void rspacex(orc_handle* h, double* a) {
  if (h->d.geometry == UPPE_CARTESIAN) {
    axis2_fft(h, a, -1.0);
    axis1_fft(h, a, -1.0);
    return;
  }
  root_volume(h, a, 0);
  memcpy(h->hscratch, a, sizeof(double) * 2 * h->npt);
  inverse_radial_transform(h, h->hscratch, a);
  root_volume(h, a, 1);
  axis2_fft(h, a, -1.0);
}

/* ---------------------------------------------------------------------------------------------------- */
/* update_current, load_source, propagator, finish (modules/uppe_kernel.py)                              */
/* ---------------------------------------------------------------------------------------------------- */

static void rfft_all(orc_handle* h, double* in, double* out) {
#pragma omp parallel for schedule(static)
  for (long t = 0; t < h->ntr; t++) orc_rfft_strip(in + t, out + 2 * t, h->nt, h->ntr);
}

/* uppe_kernel.py:139-186 */
static void update_current(orc_handle* h) {
#pragma omp parallel for schedule(static)
  for (long t = 0; t < h->ntr; t++) orc_irfft_strip(h->Ew + 2 * t, h->Et + t, h->nw, h->ntr);
  if (h->ion.kind != UPPE_ION_NONE) {
    k_rate(h);
    k_plasma_density(h);
    k_soft_time_window(h, h->ne, 4, 64);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < h->ntt; i++) h->Jt[i] = h->ne[i] * h->Et[i];
    rfft_all(h, h->Jt, h->Jw);
    k_idt_spectral(h, h->Jw, 0.1, 1.0);
  } else {
    memset(h->ne, 0, sizeof(double) * h->ntt);
    memset(h->Jw, 0, sizeof(double) * 2 * h->npt);
  }
  memset(h->chi, 0, sizeof(double) * h->ntt);
  if (h->rot_states > 2) k_add_chi_rotations(h);
  k_add_kerr_chi(h);
  k_source_polarization(h);
  rfft_all(h, h->Jt, h->scratchw);
  k_dt_spectral(h, h->scratchw, 1.0);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < 2 * h->npt; i++) h->Jw[i] = h->Jw[i] + h->scratchw[i];
  k_filter(h, h->Jw, h->nlf);
}

/* paraxial update_current (paraxial_kernel.py:140-184): Ew holds A(w) in r-space on entry; Jw receives J(w) */
static void update_current_env(orc_handle* h) {
  memcpy(h->Etc, h->Ew, sizeof(double) * 2 * h->ntt);
  memcpy(h->Atc, h->Ew, sizeof(double) * 2 * h->ntt);
  memset(h->ne, 0, sizeof(double) * h->ntt);
  memset(h->chi, 0, sizeof(double) * h->ntt);
  k_dt_spectral(h, h->Etc, -1.0);
  strip_fft_all(h, h->Etc, -1.0);
  strip_fft_all(h, h->Atc, -1.0);
  if (h->ion.kind != UPPE_ION_NONE) {
    k_rate_env(h);
    k_plasma_density(h);
  }
  if (h->rot_states > 2) k_add_chi_rotations_env(h);
  k_add_kerr_chi_env(h);
  k_source_polarization_env(h);
  strip_fft_all(h, h->Jtc, 1.0);
  k_dt_spectral(h, h->Jtc, 1.0);
  strip_fft_all(h, h->Jtc, -1.0);
  k_plasma_current_env(h);
  strip_fft_all(h, h->Jtc, 1.0);
  memcpy(h->Jw, h->Jtc, sizeof(double) * 2 * h->npt);
}

static double reduce_sum(const double* v, long n) {
  double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (long i = 0; i < n; i++) s += v[i];
  return s;
}

static double reduce_max(const double* v, long n) {
  double m = -INFINITY;
#pragma omp parallel for reduction(max : m) schedule(static)
  for (long i = 0; i < n; i++) m = v[i] > m ? v[i] : m;
  return m;
}

static double reduce_min(const double* v, long n) {
  double m = INFINITY;
#pragma omp parallel for reduction(min : m) schedule(static)
  for (long i = 0; i < n; i++) m = v[i] < m ? v[i] : m;
  return m;
}

/* RHS S(z, qw) into Jw; with return_dz also the step estimate (uppe_kernel.py:188-232) */
static double load_source(orc_handle* h, double zphase, int return_dz, double zlocal, double L, double dzmin,
                          double dphi, double rel_amin) {
  if (h->linear_rhs) {
    for (long i = 0; i < h->npt; i++) {
      double qr = h->qw[2 * i], qi = h->qw[2 * i + 1];
      h->Jw[2 * i] = h->lam_re * qr - h->lam_im * qi;
      h->Jw[2 * i + 1] = h->lam_re * qi + h->lam_im * qr;
    }
  } else if (h->paraxial) {
    /* paraxial_kernel.py:186-220 */
    memcpy(h->Ew, h->qw, sizeof(double) * 2 * h->npt);
    k_propagate_linear_env(h, h->Ew, zphase);
    rspacex(h, h->Ew);
    update_current_env(h);
    kspacex(h, h->Jw);
    k_current_to_rhs_env(h, h->Jw, zphase);
  } else {
    memcpy(h->Ew, h->qw, sizeof(double) * 2 * h->npt);
    k_propagate_linear(h, h->Ew, zphase);
    k_vp_to_e(h, h->Ew);
    rspacex(h, h->Ew);
    update_current(h);
    kspacex(h, h->Jw);
    k_current_to_rhs(h, h->Jw, zphase);
  }
  h->evals++;
  if (!return_dz) return 0.0;
  k_load_modulus(h);
  if (isnan(reduce_sum(h->dzb, h->npt))) h->nan_steps++;
  double amin = rel_amin * reduce_max(h->dzb, h->npt);
  k_load_step_size(h, L, dphi, amin);
  double dz = reduce_min(h->dzb, h->npt);
  if (dz < dzmin) dz = dzmin;
  if (zlocal + dz > L) dz = L - zlocal;
  return dz;
}

/* conditioner: zero the lowest bins after every stage (uppe_kernel.py:131-137); the paraxial stepper has none */
static void condition_field(orc_handle* h) {
  if (!h->paraxial) k_filter(h, h->qw, h->cond);
}

/* ---------------------------------------------------------------------------------------------------- */
/* Exported ABI                                                                                          */
/* ---------------------------------------------------------------------------------------------------- */

static void free_all(orc_handle* h) {
  double** ptrs[] = {&h->w, &h->kz, &h->kg, &h->ng, &h->nlf, &h->cond, &h->H, &h->lam, &h->qw, &h->qi,
                     &h->qf, &h->Ew, &h->Jw, &h->scratchw, &h->hscratch, &h->dzb, &h->Et, &h->ne, &h->chi,
                     &h->Jt, &h->rot_tjj, &h->rot_wdt, &h->k0, &h->Etc, &h->Atc, &h->Jtc};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) {
    free(*ptrs[i]);
    *ptrs[i] = NULL;
  }
}

int orc_uppe_create(const uppe_desc* d, orc_handle** out) {
  if (!d || !out) return UPPE_ERR_ARG;
  *out = NULL;
  if (d->model != UPPE_MODEL_UPPE && d->model != UPPE_MODEL_PARAXIAL) return UPPE_ERR_ARG;
  const int par = d->model == UPPE_MODEL_PARAXIAL;
  if (par ? (d->nw < 2 || !is_pow2(d->nw)) : (d->nw < 3 || !is_pow2(d->nw - 1))) return UPPE_ERR_ARG;
  if (!is_pow2(d->n1) || !is_pow2(d->n2)) return UPPE_ERR_ARG;
  if (d->geometry != UPPE_RADIAL && d->geometry != UPPE_CARTESIAN) return UPPE_ERR_ARG;
  if (d->geometry == UPPE_RADIAL && (d->nk < 1 || d->nk > d->n1)) return UPPE_ERR_ARG;
  orc_handle* h = (orc_handle*)calloc(1, sizeof(orc_handle));
  if (!h) return UPPE_ERR_DEVICE;
  h->d = *d;
  h->nw = d->nw;
  h->n1 = d->n1;
  h->n2 = d->n2;
  h->nk = d->geometry == UPPE_RADIAL ? d->nk : d->n1;
  h->nm = d->n2;
  h->paraxial = par;
  h->nt = par ? d->nw : 2 * (d->nw - 1);
  h->ntr = (long)d->n1 * d->n2;
  h->npt = h->nw * h->ntr;
  h->ntt = h->nt * h->ntr;
  size_t cs = sizeof(double) * 2 * h->npt, ts = sizeof(double) * h->ntt;
  h->w = malloc(sizeof(double) * h->nw);
  h->kz = malloc(cs);
  h->kg = malloc(sizeof(double) * h->nw);
  h->ng = malloc(sizeof(double) * h->ntr);
  h->nlf = malloc(sizeof(double) * h->nw);
  h->cond = malloc(sizeof(double) * h->nw);
  h->H = calloc((size_t)h->nk * h->n1 * h->nm, sizeof(double));
  h->lam = calloc((size_t)h->n1, sizeof(double));
  h->qw = calloc(1, cs);
  h->qi = calloc(1, cs);
  h->qf = calloc(1, cs);
  h->Ew = calloc(1, cs);
  h->Jw = calloc(1, cs);
  h->scratchw = calloc(1, cs);
  h->hscratch = calloc(1, cs);
  h->dzb = calloc(h->npt, sizeof(double));
  h->Et = calloc(1, ts);
  h->ne = calloc(1, ts);
  h->chi = calloc(1, ts);
  h->Jt = calloc(1, ts);
  if (par) {
    h->k0 = calloc((size_t)h->nw, sizeof(double));
    h->Etc = calloc(1, 2 * ts);
    h->Atc = calloc(1, 2 * ts);
    h->Jtc = calloc(1, 2 * ts);
    if (!h->k0 || !h->Etc || !h->Atc || !h->Jtc) {
      free_all(h);
      free(h);
      return UPPE_ERR_DEVICE;
    }
  }
  if (!h->w || !h->kz || !h->kg || !h->ng || !h->nlf || !h->cond || !h->H || !h->lam || !h->qw || !h->qi ||
      !h->qf || !h->Ew || !h->Jw || !h->scratchw || !h->hscratch || !h->dzb || !h->Et || !h->ne || !h->chi ||
      !h->Jt) {
    free_all(h);
    free(h);
    return UPPE_ERR_DEVICE;
  }
  *out = h;
  return UPPE_OK;
}

int orc_uppe_set_hankel(orc_handle* h, const double* H, const double* lam) {
  if (!h || !H || !lam) return fail(h, UPPE_ERR_ARG, "null argument");
  if (h->d.geometry != UPPE_RADIAL) return fail(h, UPPE_ERR_ARG, "Hankel tables given for a Cartesian grid");
  memcpy(h->H, H, sizeof(double) * h->nk * h->n1 * h->nm);
  memcpy(h->lam, lam, sizeof(double) * h->n1);
  h->have_hankel = 1;
  return UPPE_OK;
}

/* Rotator.AddChi's tables (rotations.py:56-61); NULL or fewer than 3 states switches the rotator off */
int orc_uppe_set_rotator(orc_handle* h, const uppe_rotator* r) {
  if (!h) return UPPE_ERR_ARG;
  free(h->rot_tjj);
  free(h->rot_wdt);
  h->rot_tjj = h->rot_wdt = NULL;
  h->rot_states = 0;
  if (!r || r->states < 3) return UPPE_OK;
  if (!r->tjj || !r->wdt) return fail(h, UPPE_ERR_ARG, "null rotator table");
  if (r->states > 4096) return fail(h, UPPE_ERR_ARG, "rotator: at most 4096 states");
  h->rot_tjj = (double*)malloc(sizeof(double) * r->states);
  h->rot_wdt = (double*)malloc(sizeof(double) * 2 * r->states);
  if (!h->rot_tjj || !h->rot_wdt) return fail(h, UPPE_ERR_ARG, "out of host memory");
  memcpy(h->rot_tjj, r->tjj, sizeof(double) * r->states);
  memcpy(h->rot_wdt, r->wdt, sizeof(double) * 2 * r->states);
  h->rot_states = r->states;
  h->rot_dalpha = r->dalpha;
  h->rot_hbar = r->hbar;
  return UPPE_OK;
}

int orc_uppe_set_medium(orc_handle* h, const uppe_medium* m) {
  if (!h || !m || !m->w || !m->kz || !m->kg || !m->ngas || !m->nl_filter || !m->cond_filter)
    return fail(h, UPPE_ERR_ARG, "null argument");
  if (h->paraxial) return fail(h, UPPE_ERR_STATE, "uppe_set_medium on a paraxial handle");
  if (m->w[0] != 0.0) return fail(h, UPPE_ERR_ARG, "UPPE propagator requires frequency range [0,...]");
  memcpy(h->w, m->w, sizeof(double) * h->nw);
  memcpy(h->kz, m->kz, sizeof(double) * 2 * h->npt);
  memcpy(h->kg, m->kg, sizeof(double) * h->nw);
  memcpy(h->ng, m->ngas, sizeof(double) * h->ntr);
  memcpy(h->nlf, m->nl_filter, sizeof(double) * h->nw);
  memcpy(h->cond, m->cond_filter, sizeof(double) * h->nw);
  h->redchi3 = m->redchi3;
  h->dt = m->dt;
  h->ion = m->ion;
  h->have_medium = 1;
  return UPPE_OK;
}

/* paraxial Material.UpdateMaterial uploads and the Source frequency table (paraxial_kernel.py:72-135) */
int orc_uppe_set_paraxial_medium(orc_handle* h, const uppe_paraxial_medium* m) {
  if (!h || !m || !m->w || !m->kz || !m->k0 || !m->ngas) return fail(h, UPPE_ERR_ARG, "null argument");
  if (!h->paraxial) return fail(h, UPPE_ERR_STATE, "uppe_set_paraxial_medium on a UPPE handle");
  memcpy(h->w, m->w, sizeof(double) * h->nw);
  memcpy(h->kz, m->kz, sizeof(double) * 2 * h->npt);
  memcpy(h->k0, m->k0, sizeof(double) * h->nw);
  memcpy(h->ng, m->ngas, sizeof(double) * h->ntr);
  h->redchi3 = m->redchi3;
  h->dt = m->dt;
  h->ion = m->ion;
  h->have_medium = 1;
  return UPPE_OK;
}

int orc_uppe_set_field(orc_handle* h, const double* a) {
  if (!h || !a) return fail(h, UPPE_ERR_ARG, "null argument");
  if (h->d.geometry == UPPE_RADIAL && !h->have_hankel) return fail(h, UPPE_ERR_STATE, "Hankel tables not set");
  memcpy(h->qw, a, sizeof(double) * 2 * h->npt);
  kspacex(h, h->qw);
  h->have_field = 1;
  h->steps = h->evals = 0;
  h->nan_steps = 0;
  h->z_last = 0.0;
  return UPPE_OK;
}

/* propagator's subcycle and step loop (uppe_kernel.py:299-325) */
int orc_uppe_advance(orc_handle* h, const uppe_step_control* c, int* steps_out) {
  if (!h || !c) return fail(h, UPPE_ERR_ARG, "null argument");
  if (!h->have_field || !h->have_medium) return fail(h, UPPE_ERR_STATE, "field or medium not set");
  if (c->n_sub < 1 || !(c->length > 0.0)) return fail(h, UPPE_ERR_ARG, "bad subcycle count or length");
  const double L = c->length;
  int taken = 0;
  for (int s = 0; s < c->n_sub && !(c->max_steps > 0 && taken >= c->max_steps); s++) {
    double z0 = c->zi + (double)(c->sub0 + s) * L;
    double zoff = z0 - c->zi;
    double z = 0.0;
    while (z < L && (z - L) * (z - L) > (L / 1e6) * (L / 1e6)) {
      if (c->max_steps > 0 && taken >= c->max_steps) break;
      memcpy(h->qi, h->qw, sizeof(double) * 2 * h->npt);
      memcpy(h->qf, h->qw, sizeof(double) * 2 * h->npt);
      double dz = load_source(h, zoff + z, 1, z, L, c->dzmin, c->dphi, c->rel_amin);
      if (c->n_schedule > 0) {
        if (taken >= c->n_schedule) return fail(h, UPPE_ERR_ARG, "dz schedule exhausted");
        dz = c->dz_schedule[taken];
      }
      if (c->dz_record && taken < c->record_cap) c->dz_record[taken] = dz;
      k_rk_stage(h, dz / 6, dz / 2);
      condition_field(h);
      load_source(h, zoff + z + 0.5 * dz, 0, 0, L, 0, 0, 0);
      k_rk_stage(h, dz / 3, dz / 2);
      condition_field(h);
      load_source(h, zoff + z + 0.5 * dz, 0, 0, L, 0, 0, 0);
      k_rk_stage(h, dz / 3, dz);
      condition_field(h);
      load_source(h, zoff + z + dz, 0, 0, L, 0, 0, 0);
      k_rk_finish(h, dz / 6);
      condition_field(h);
      z += dz;
      taken++;
      h->steps++;
    }
    h->z_last = z;
  }
  if (steps_out) *steps_out = taken;
  return UPPE_OK;
}

/* uppe_kernel.py:234-258 */
int orc_uppe_finish(orc_handle* h, double z_total, double* a_r, double* j_r, double* chi_t, double* ne_t) {
  if (!h) return UPPE_ERR_ARG;
  if (!h->have_field || !h->have_medium) return fail(h, UPPE_ERR_STATE, "field or medium not set");
  if (h->paraxial) {
    /* paraxial_kernel.py:222-249: A = rspace(exp(i kz z) q), then update_current on it */
    memcpy(h->Ew, h->qw, sizeof(double) * 2 * h->npt);
    k_propagate_linear_env(h, h->Ew, z_total);
    rspacex(h, h->Ew);
    update_current_env(h);
    if (a_r) memcpy(a_r, h->Ew, sizeof(double) * 2 * h->npt);
  } else {
    k_propagate_linear(h, h->qw, z_total);
    rspacex(h, h->qw);
    memcpy(h->Ew, h->qw, sizeof(double) * 2 * h->npt);
    k_vp_to_e(h, h->Ew);
    update_current(h);
    if (a_r) memcpy(a_r, h->qw, sizeof(double) * 2 * h->npt);
  }
  if (j_r) memcpy(j_r, h->Jw, sizeof(double) * 2 * h->npt);
  if (chi_t) memcpy(chi_t, h->chi, sizeof(double) * h->ntt);
  if (ne_t) memcpy(ne_t, h->ne, sizeof(double) * h->ntt);
  h->have_field = 0;
  return UPPE_OK;
}

int orc_uppe_get_state(orc_handle* h, double* q) {
  if (!h || !q) return UPPE_ERR_ARG;
  memcpy(q, h->qw, sizeof(double) * 2 * h->npt);
  return UPPE_OK;
}

int orc_uppe_get_info(orc_handle* h, uppe_info* info) {
  if (!h || !info) return UPPE_ERR_ARG;
  info->steps = h->steps;
  info->evals = h->evals;
  info->nan_steps = h->nan_steps;
  info->reserved = 0;
  info->z_last = h->z_last;
  return UPPE_OK;
}

int orc_uppe_destroy(orc_handle* h) {
  if (!h) return UPPE_OK;
  free_all(h);
  free(h);
  return UPPE_OK;
}

const char* orc_uppe_last_error(const orc_handle* h) { return h ? h->err : "null handle"; }

/* ---------------------------------------------------------------------------------------------------- */
/* Single-kernel entry points for the oracle's own checks (not part of the stepper ABI)                  */
/* ---------------------------------------------------------------------------------------------------- */

/* one RHS evaluation S(zphase, q) into s (both k-space complex, reference layout) */
int orc_uppe_rhs(orc_handle* h, const double* q, double zphase, double* s) {
  if (!h || !q || !s) return UPPE_ERR_ARG;
  memcpy(h->qw, q, sizeof(double) * 2 * h->npt);
  load_source(h, zphase, 0, 0, 1, 0, 0, 0);
  memcpy(s, h->Jw, sizeof(double) * 2 * h->npt);
  return UPPE_OK;
}

/* update_current on an r-space field (UPPE: E; paraxial: A): Jw (r-space spectral), chi(t), ne(t) */
int orc_uppe_source(orc_handle* h, const double* e_r, double* j_r, double* chi_t, double* ne_t) {
  if (!h || !e_r) return UPPE_ERR_ARG;
  memcpy(h->Ew, e_r, sizeof(double) * 2 * h->npt);
  if (h->paraxial)
    update_current_env(h);
  else
    update_current(h);
  if (j_r) memcpy(j_r, h->Jw, sizeof(double) * 2 * h->npt);
  if (chi_t) memcpy(chi_t, h->chi, sizeof(double) * h->ntt);
  if (ne_t) memcpy(ne_t, h->ne, sizeof(double) * h->ntt);
  return UPPE_OK;
}

/* transverse transforms of a spectral array in place: dir = +1 kspacex, -1 rspacex */
int orc_uppe_transverse(orc_handle* h, double* a, int dir) {
  if (!h || !a) return UPPE_ERR_ARG;
  if (dir > 0)
    kspacex(h, a);
  else
    rspacex(h, a);
  return UPPE_OK;
}

/* test hook: replace the UPPE right-hand side by S = lambda q, keeping the RK4 loop, conditioner and
 * step-size control; enable = 0 restores the UPPE right-hand side */
int orc_uppe_set_linear_rhs(orc_handle* h, int enable, double lam_re, double lam_im) {
  if (!h) return UPPE_ERR_ARG;
  h->linear_rhs = enable;
  h->lam_re = lam_re;
  h->lam_im = lam_im;
  return UPPE_OK;
}

/* plasma density scan of a rate array (Nt, N1, N2) with gas density ng (N1, N2) */
int orc_plasma_density(const double* rate, const double* ng, double dt, long nt, long ntr, double* ne) {
  for (long off = 0; off < ntr; off++) {
    double prev = 0.0, raw = 0.0;
    for (long k = nt - 1; k >= 0; --k) {
      double cur = (k == nt - 1) ? 0.0 : rate[off + k * ntr];
      raw += 0.5 * (cur + prev) * dt;
      ne[off + k * ntr] = ng[off] * (1.0 - exp(-raw));
      prev = cur;
    }
  }
  return UPPE_OK;
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

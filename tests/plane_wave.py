'''Monochromatic plane waves in a Kerr medium: closed-form solutions of the UPPE right-hand side.

The reference advances q = exp(-i(kz + kg) z) A (uppe_kernel.py:260-331) with
    E = i kz A                                   (VectorPotentialToElectricField, uppe.cl:93-108)
    chi = n_g chi3 / n_ref E(t)^2, P = chi E     (AddKerrChi, SourcePolarization, uppe.cl:132-165)
    J = F_NL (-i w) rfft(P)                      (update_current, uppe_kernel.py:139-186; DtSpectral, fft.cl:282-299)
    dq/dz = S = (i/2) exp(+Im kz z) exp(-i(Re kz + kg) z) J / Re kz     (CurrentToODERHS, uppe.cl:73-91)
For one frequency bin m carrying a transversely uniform field, r-space amplitude e_m = i kz A_m, the time
domain field is E0 cos(w t + phi) with E0 = 2 |e_m| / Nt, and cos^3 = (3 cos + cos 3)/4 gives:

* a nonlinear band holding only the fundamental: dq/dz = i c exp(-2 kappa z) |q|^2 (kz / Re kz) q with
  c = (3/8) w n_g chi3/n_ref (2 |kz| / (Nt nrm))^2 and kappa = Im kz (nrm: the transverse transform's scale of a
  uniform field, N1 N2 for the Cartesian FFT, 1 for an identity radial basis). For real kz the phase advances at
  gamma = (3/8) w (n_g chi3 / n_ref) E0^2 and |q| stays constant. With absorption (rho = Im kz / Re kz),
  u = |q|^2 and s = 1 - exp(-2 kappa z):  u(z) = u0 / (1 + u0 c rho s / kappa),
  arg q(z) - arg q(0) = ln(1 + u0 c rho s / kappa) / (2 rho);
* a full nonlinear band: the third harmonic grows from zero as
  q3(z) = (3 w (n_g chi3/n_ref) / (2 Nt^2 nrm^2 Re kz3)) (i kz1 q1)^3 (exp(i D z) - 1) / (i D), D = 3 kz1 - kz3 (kg cancels:
  3 kg(w) = kg(3w)), to first order in chi3 (relative corrections ~ gamma z and gamma / D).

A factor 1/2 dropped, 1/Re kz replaced by 1/w, E = i w A in place of i kz A, a wrong exponent sign or a missing
exp(+Im kz z) each changes these rates by a factor of order one, so the checks below fail on any of them.
'''
import numpy as np

from searay_amd import _abi, dispersion, ionization

MKS_LENGTH = 0.8e-6 / (2 * np.pi)
NW = 33                 # Nt = 64 samples, frequencies 0 .. 4 in steps of 1/8
DW = 0.125
M = 8                   # the fundamental: w = 1 (800 nm)


def grid():
    return np.arange(NW) * DW


def bk7_kz(absorb=0.0):
    '''kz(w) of BK7 (kr = 0); absorb adds i*absorb*w to chi (Im kz > 0). kz = 1 at w = 0 (uppe_kernel.py:84).'''
    w = grid()
    chi = dispersion.BK7(MKS_LENGTH).chi(w) + 1j * absorb * w
    kz = np.sqrt(0j + w ** 2 * (1 + chi))
    kz[0] = 1.0
    return kz


def make(factory, geometry, kz, nl_bins, redchi3=1e-3):
    '''A stepper on the plane-wave grid: Cartesian 16 x 16, or radial 16 cells with an identity basis.'''
    if geometry == 'cartesian':
        n1, n2, nk, geo, nrm = 16, 16, 16, _abi.UPPE_CARTESIAN, 256.0
    else:
        n1, n2, nk, geo, nrm = 16, 1, 16, _abi.UPPE_RADIAL, 1.0
    st = factory(geo, NW, n1, n2, nk)
    if geo == _abi.UPPE_RADIAL:
        st.set_hankel(np.eye(n1).reshape(n1, n1, 1), np.ones(n1))
    w = grid()
    kz3 = np.ascontiguousarray(np.broadcast_to(kz[:, None, None], (NW, n1, n2)))
    kg = -w / 0.6                        # any window speed: the plane-wave rates do not depend on it
    nl = np.zeros(NW)
    nl[nl_bins[0]:nl_bins[1]] = 1.0
    cond = np.ones(NW)
    cond[:4] = 0.0
    dt = np.pi / w[-1]
    st.set_medium(w, kz3, kg, np.ones((n1, n2)), nl, cond, redchi3, dt, ionization.no_ionization_params())
    return st, (n1, n2), nrm


def field(shape, amp):
    a = np.zeros((NW,) + shape, dtype=np.complex128)
    a[M] = amp
    return a


def spm_rate(kz1, nrm, redchi3=1e-3):
    '''c of the module docstring: the rate per |q|^2 of the k-space state.'''
    nt = 2 * (NW - 1)
    w1 = M * DW
    return 0.375 * w1 * redchi3 * (2 * abs(kz1) / (nt * nrm)) ** 2


def spm_solution(c, kz1, q0, z):
    '''q(z) of dq/dz = i c exp(-2 kappa z) |q|^2 (kz/Re kz) q.'''
    kappa = kz1.imag
    u0 = abs(q0) ** 2
    if kappa == 0.0:
        return q0 * np.exp(1j * c * u0 * z)
    rho = kz1.imag / kz1.real
    s = 1.0 - np.exp(-2 * kappa * z)
    g = 1.0 + u0 * c * rho * s / kappa
    return q0 / np.sqrt(g) * np.exp(1j * np.log(g) / (2 * rho))


def third_harmonic(kz, q1, redchi3, nrm, z):
    nt = 2 * (NW - 1)
    w1 = M * DW
    kz1, kz3 = kz[M], kz[3 * M]
    d = 3 * kz1 - kz3
    pref = 3 * w1 * redchi3 / (2 * nt ** 2 * nrm ** 2 * kz3.real)
    return pref * (1j * kz1 * q1) ** 3 * (np.exp(1j * d * z) - 1) / (1j * d)


def run(st, a_r, length, steps):
    '''Fixed RK4 steps over [0, length]; returns the k-space state.'''
    st.set_field(a_r)
    dz = length / steps
    n, _ = st.advance(0.0, dz, 0, steps, dz)
    assert n == steps
    return st.get_state()

'''Photoionization rate models in atomic units (host side, FP64).

These models only produce coefficients for the device rate kernel; they are evaluated once per run.
Restated from SeaRay's ``modules/photoionization_base.py``: the atomic-unit system (:173-211), the
standard form C_pre E^C_pow exp(C_exp/E) and its cycle averaging (:213-259), Landau-Lifshitz hydrogen
(:261-269), ADK (:271-287), PPT in the tunnelling limit (:314-335), the hybrid PPT/PMPB theories with the
Fourier series (:337-412) and the stitched PPT (:414-426). The ionisation-potential table is physical
data (:59-79).
'''
import numpy as np
from scipy import special as spsf
from scipy import constants as C

# first ionisation potentials in eV for the species the UPPE examples use (photoionization_base.py:59-79)
ip = {'H': [13.59844],
      'He': [24.58741, 54.41778],
      'N': [14.53414, 29.6013, 47.44924, 77.4735, 97.8902, 552.0718, 667.046],
      'O': [13.61806, 35.11730, 54.9355, 77.41353, 113.8990, 138.1197, 739.29, 871.4101],
      'Ne': [21.56454, 40.96328, 63.45, 97.12, 126.21, 157.93, 207.2759, 239.0989, 1195.8286, 1362.1995],
      'Ar': [15.75962, 27.62967, 40.74, 59.81, 75.02, 91.009, 124.323, 143.460, 422.45, 478.69,
             538.96, 618.26, 686.10, 755.74, 854.77, 918.03, 4120.8857, 4426.2296],
      'Kr': [13.99961, 24.35985, 36.950, 52.5, 64.7, 78.5, 111.0, 125.802],
      'Xe': [12.130, 20.975, 31.05, 40.9, 54.14, 66.703, 91.6, 105.976]}


def CycleAveragingFactor(Uion, E):
    '''Ratio of cycle-averaged to static tunnelling rate (photoionization_base.py:125-130).'''
    return np.sqrt(3.0 / np.pi) * np.sqrt(E) / (2 * Uion) ** 0.75


class AtomicUnits:
    '''Atomic units expressed in SI (photoionization_base.py:173-211).'''

    def __init__(self):
        self.xa = C.hbar / (C.m_e * C.c * C.alpha)
        self.ta = C.hbar / (C.m_e * C.c ** 2 * C.alpha ** 2)
        self.ua = C.m_e * C.c ** 2 * C.alpha ** 2
        self.ea = self.ua / self.xa / C.e

    def energy_to_eV(self, u):
        return u * self.ua / C.e

    def energy_to_au(self, u):
        return u * C.e / self.ua

    def rate_to_mks(self, rate):
        return rate / self.ta

    def rate_to_au(self, rate):
        return rate * self.ta

    def field_to_mks(self, field):
        return field * self.ea

    def field_to_au(self, field):
        return field / self.ea

    def time_to_au(self, t):
        return t / self.ta

    def time_to_mks(self, t):
        return t * self.ta

    def length_to_au(self, x):
        return x / self.xa

    def length_to_mks(self, x):
        return x * self.xa


class Ionization(AtomicUnits):
    '''Standard-form rate w(E) = C_pre (|E|+Ecut)^C_pow exp(C_exp/(|E|+Ecut)), atomic units.'''

    def __init__(self, averaging, Uion, Z, lstar=0, l=0, m=0, w0=0, terms=1):
        super().__init__()
        self.avg = averaging
        self.Uion = Uion
        self.Z = Z
        self.lstar = lstar
        self.l = l
        self.m = m
        self.w0 = w0
        self.terms = terms
        self.cutoff_field = 1e-3
        self.C_pre = 0.0
        self.C_pow = 0.0
        self.C_exp = 0.0

    def GetCoeff(self):
        return self.C_pre, self.C_pow, self.C_exp

    def ToAverage(self):
        self.C_pre *= np.sqrt(3.0 / np.pi) / (2 * self.Uion) ** 0.75
        self.C_pow += 0.5

    def ToStatic(self):
        self.C_pre *= (2 * self.Uion) ** 0.75 / np.sqrt(3.0 / np.pi)
        self.C_pow -= 0.5

    def Rate(self, Es):
        E = np.abs(Es) + self.cutoff_field
        return self.C_pre * E ** self.C_pow * np.exp(self.C_exp / E)


class Hydrogen(Ionization):
    '''Landau-Lifshitz static hydrogen rate 4/E exp(-2/(3E)) (photoionization_base.py:261-269).'''

    def __init__(self, averaging, Uion, Z, lstar=0, l=0, m=0, w0=0, terms=1):
        super().__init__(averaging, Uion, Z)
        self.C_pre, self.C_pow, self.C_exp = 4.0, -1.0, -2.0 / 3.0
        if averaging:
            self.ToAverage()


def _angular_factor(l, m):
    return (2 * l + 1) * spsf.factorial(l + m) / spsf.factorial(m) / spsf.factorial(l - m)


class ADK(Ionization):
    '''ADK (Ammosov, Delone, Krainov 1986) in standard form, cycle averaged by construction
    (photoionization_base.py:271-287).'''

    def __init__(self, averaging, Uion, Z, lstar=0, l=0, m=0, w0=0, terms=1):
        super().__init__(averaging, Uion, Z, lstar, l, m)
        ns = Z / np.sqrt(2 * Uion)
        m = abs(self.m)
        e = np.exp(1.0)
        root = np.sqrt(ns ** 2 - lstar ** 2)
        pre = np.sqrt(3 / np.pi ** 3) * _angular_factor(l, m)
        pre *= (e / root) ** (m + 1.5)
        pre *= ((ns + lstar) / (ns - lstar)) ** (lstar + 0.5)
        pre *= Z ** 2 / ns ** 3
        pre *= (4 * e * Z ** 3 / ns ** 3 / root) ** (2 * ns - m - 1.5)
        self.C_pre = pre
        self.C_pow = m + 1.5 - 2 * ns
        self.C_exp = -2 * Z ** 3 / (3 * ns ** 3)
        if not averaging:
            self.ToStatic()


class PPT_Tunneling(Ionization):
    '''PPT in the tunnelling limit (gamma -> 0) with the Coulomb factor and ADK bound-state coefficient
    (photoionization_base.py:314-335).'''

    def __init__(self, averaging, Uion, Z, lstar=0, l=0, m=0, w0=0, terms=1):
        super().__init__(averaging, Uion, Z, lstar, l, m, w0, terms)
        ns = Z / np.sqrt(2 * Uion)
        m = abs(self.m)
        F0 = Z ** 3 / ns ** 3
        pre = Uion
        pre *= 2 ** (2 * ns) / (ns * spsf.gamma(ns + lstar + 1) * spsf.gamma(ns - lstar))
        pre *= np.sqrt(6 / np.pi)
        pre *= (2 * l + 1) * spsf.factorial(l + m) / 2 ** m / spsf.factorial(m) / spsf.factorial(l - m)
        pre *= (0.5 / F0) ** (m + 1.5)
        self.C_pow = m + 1.5
        self.C_exp = -2 * F0 / 3
        pre *= (2 * F0) ** (2 * ns)
        self.C_pow -= 2 * ns
        self.C_pre = pre
        if not averaging:
            self.ToStatic()


class HybridTheories(Ionization):
    '''Multiphoton-to-tunnelling theories with m = 0 (photoionization_base.py:337-359).'''

    def __init__(self, averaging, Uion, Z, lstar=0, l=0, m=0, w0=0, terms=1):
        if m != 0.0:
            raise ValueError('Hybrid theories all require m=0')
        if w0 == 0.0:
            raise ValueError('Hybrid theories require w>0')
        super().__init__(averaging, Uion, Z, lstar, l, m, w0, terms)

    def Params(self, E, w, Uion, Z):
        ns = Z / np.sqrt(2 * Uion)
        F = E * ns ** 3 / Z ** 3
        gam = Z * w / (ns * E)
        g = (3 / (2 * gam)) * ((1 + 1 / (2 * gam ** 2)) * np.arcsinh(gam) - np.sqrt(1 + gam ** 2) / (2 * gam))
        nu = (self.Uion / w) * (1 + 1 / (2 * gam ** 2))
        return ns, F, gam, g, nu

    def FourierSum(self, gam, nu):
        total = 0.0 * gam
        alpha = 2 * (np.arcsinh(gam) - gam / np.sqrt(1 + gam ** 2))
        beta = 2 * gam / np.sqrt(1 + gam ** 2)
        dnu = np.ceil(nu) - nu
        for n in range(self.terms):
            total += np.exp(-alpha * (n + dnu)) * spsf.dawsn(np.sqrt(beta * (n + dnu)))
        return total


class PPT(HybridTheories):
    '''PPT Eq. 54 with the Coulomb factor, m = 0 (photoionization_base.py:361-385).'''

    def Rate(self, Es):
        E = np.abs(Es) + self.cutoff_field
        w = self.w0
        ns, F, gam, g, nu = self.Params(E, w, self.Uion, self.Z)
        ans = self.Uion
        ans *= 2 ** (2 * ns) / (ns * spsf.gamma(ns + self.lstar + 1) * spsf.gamma(ns - self.lstar))
        ans *= np.sqrt(6 / np.pi) * (2 * self.l + 1)
        ans *= (F * np.sqrt(1 + gam ** 2) / 2) ** 1.5
        ans *= (4 / np.sqrt(3 * np.pi)) * (gam ** 2 / (1 + gam ** 2)) * self.FourierSum(gam, nu)
        ans *= np.exp(-2 * g / (3 * F))
        ans *= (2 / F) ** (2 * ns)
        if self.avg:
            return ans
        return ans / CycleAveragingFactor(self.Uion, E)


class PMPB(HybridTheories):
    '''Popruzhenko-Mur-Popov-Bauer rate, m = 0 (photoionization_base.py:387-412).'''

    def Rate(self, Es):
        E = np.abs(Es) + self.cutoff_field
        w = self.w0
        ns, F, gam, g, nu = self.Params(E, w, self.Uion, self.Z)
        ans = (2 / np.pi) * self.Uion
        ans *= 2 ** (2 * ns - 2) / (ns * spsf.gamma(ns + self.lstar + 1) * spsf.gamma(ns - self.lstar))
        ans *= 2 * self.l + 1
        ans *= (self.Uion / w) ** -1.5
        ans *= np.sqrt(2 * gam / np.sqrt(1 + gam ** 2))
        ans *= self.FourierSum(gam, nu)
        ans *= np.exp(-2 * g / (3 * F))
        ans *= (2 / F) ** (2 * ns) * (1 + gam * 2 / np.exp(1)) ** (-2 * ns)
        if self.avg:
            return ans
        return ans / CycleAveragingFactor(self.Uion, E)


class StitchedPPT(PPT_Tunneling, PPT):
    '''Tunnelling-limit PPT for Keldysh gamma < 0.5, full PPT above (photoionization_base.py:414-426).

    The constructor is PPT_Tunneling's, so C_pre is nonzero. The Ionizer therefore hands the device the
    closed tunnelling-limit coefficients (ionization.py:104-109,134-141); ``Rate`` is used on the host only.'''

    def Rate(self, Es):
        E = np.abs(Es) + self.cutoff_field
        gam = np.sqrt(2.0 * self.Uion) * self.w0 / E
        tunnel = gam < 0.5
        rate = np.zeros(np.shape(E))
        Es = np.asarray(Es)
        rate[tunnel] = PPT_Tunneling.Rate(self, Es[tunnel])
        rate[~tunnel] = PPT.Rate(self, Es[~tunnel])
        return rate

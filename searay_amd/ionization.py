'''Ionisation coefficients for the device rate kernel (host side, once per run).

Restates SeaRay's ``modules/ionization.py``: the simulation/atomic unit translation (:36-59), the
sim-unit wrappers around the rate models (:65-86) and ``Ionizer`` (:88-150), which turns a rate model into
the parameters of either the closed-form kernel (C_pre E^C_pow exp(C_exp/E)) or the series kernel
(exp of a quintic in log E, fitted on 256 log-spaced fields in [2e-3, 1] a.u.). The device kernels are in
the HIP library; ``Ionizer.kernel_params`` produces exactly the scalars the reference passes to
``ExplicitRate``/``ExplicitRateSeries`` (ionization.py:131-150).
'''
import numpy as np
import scipy.optimize
from scipy import constants as C
from . import photoionization_base as izbase

ION_NONE = 0
ION_EXPLICIT = 1
ION_SERIES = 2


class UnitTranslation(izbase.AtomicUnits):
    '''Simulation units (length mks_length, time mks_length/c, energy m_e c^2) vs atomic units.'''

    def __init__(self, mks_length):
        super().__init__()
        self.xs = mks_length
        self.ts = mks_length / C.c
        self.us = C.m_e * C.c ** 2
        self.es = self.us / self.xs / C.e
        w0 = C.c / mks_length
        self.ncrit = w0 ** 2 * C.epsilon_0 * C.m_e / C.e ** 2

    def rate_sim_to_au(self, rate):
        return rate * self.ta / self.ts

    def rate_au_to_sim(self, rate):
        return rate * self.ts / self.ta

    def field_sim_to_au(self, field):
        return field * self.es / self.ea

    def field_au_to_sim(self, field):
        return field * self.ea / self.es

    def energy_sim_to_au(self, energy):
        return energy * self.us / self.ua

    def energy_au_to_sim(self, energy):
        # the reference's conversion (ionization.py:58-59) multiplies by the same factor in both directions
        return energy * self.us / self.ua


class rate_wrapper:
    '''Build a rate model from simulation-unit inputs (ionization.py:65-71).'''

    def __init__(self, mks_length, averaging, Uion, Z, lstar=0, l=0, m=0, w0=0, terms=1):
        self.mks_length = mks_length
        uc = UnitTranslation(mks_length)
        super().__init__(averaging, uc.energy_sim_to_au(Uion), Z, lstar, l, m, uc.rate_sim_to_au(w0), terms)

    def get_mks_length(self):
        return self.mks_length


class ADK(rate_wrapper, izbase.ADK):
    pass


class PPT_Tunneling(rate_wrapper, izbase.PPT_Tunneling):
    pass


class PPT(rate_wrapper, izbase.PPT):
    pass


class PMPB(rate_wrapper, izbase.PMPB):
    pass


class StitchedPPT(rate_wrapper, izbase.StitchedPPT):
    pass


def _quintic(x, c0, c1, c2, c3, c4, c5):
    return c0 + c1 * x + c2 * x ** 2 + c3 * x ** 3 + c4 * x ** 4 + c5 * x ** 5


class Ionizer(UnitTranslation):
    '''Rate-model adapter (ionization.py:88-150).

    If the model's C_pre is nonzero the closed form is used; otherwise log(rate) is fitted by a quintic in
    log(E) with scipy.optimize.curve_fit over 256 log-spaced fields from 2*cutoff to 1 a.u.'''

    def __init__(self, rate_generator, pts=256, mks_length=0.0):
        if mks_length == 0.0:
            super().__init__(rate_generator.get_mks_length())
        else:
            super().__init__(mks_length)
        self.cutoff_field = 1e-3
        self.rate_gen = rate_generator
        self.C1, self.C2, self.C3 = self.rate_gen.GetCoeff()
        self.Clog = None
        if self.C1 == 0.0:
            E_table = np.logspace(np.log10(self.cutoff_field * 2), np.log10(1), pts)
            rate_table = self.rate_gen.Rate(E_table)
            self.Clog, _ = scipy.optimize.curve_fit(_quintic, np.log(E_table), np.log(rate_table))

    def Rate(self, Es):
        '''Host FP64 rate in simulation units (ionization.py:116-130).'''
        E = self.field_sim_to_au(np.abs(Es)) + self.cutoff_field
        if self.C1 > 0.0:
            return self.rate_au_to_sim(self.C1 * E ** self.C2 * np.exp(self.C3 / E))
        logE = np.log(E)
        acc = np.zeros(np.shape(E))
        for i in range(6):
            acc += self.Clog[i] * logE ** i
        return self.rate_au_to_sim(np.exp(acc))

    def kernel_params(self):
        '''Scalars for the device rate kernel: dict with kind, e_conv, rate_conv, e_cutoff and either
        (c_pre, c_pow, c_exp) or c[0..5] (ionization.py:131-150).'''
        p = dict(e_conv=self.field_sim_to_au(1.0), rate_conv=self.rate_au_to_sim(1.0),
                 e_cutoff=self.cutoff_field, c_pre=0.0, c_pow=0.0, c_exp=0.0, c=[0.0] * 6)
        if self.C1 > 0.0:
            p.update(kind=ION_EXPLICIT, c_pre=float(self.C1), c_pow=float(self.C2), c_exp=float(self.C3))
        else:
            p.update(kind=ION_SERIES, c=[float(v) for v in self.Clog])
        return p


def no_ionization_params():
    return dict(kind=ION_NONE, e_conv=0.0, rate_conv=0.0, e_cutoff=0.0, c_pre=0.0, c_pow=0.0, c_exp=0.0,
                c=[0.0] * 6)

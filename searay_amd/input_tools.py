'''Unit helper for building UPPE input dictionaries (host side).

Restates the parts of SeaRay's ``modules/input_tools.py`` needed to construct UPPE runs: unit parsing
``dnum`` (input_tools.py:17-61), ``a0`` (:97-108), ``chi3`` (:110-111) and ``TransformLimitedBandwidth``
(:89-95). Simulation units: length mks_length, time mks_length/c.
'''
import numpy as np
from scipy import constants as C


class InputHelper:
    def __init__(self, mks_length):
        self.x1 = mks_length
        self.t1 = mks_length / C.c
        self.w1 = 1 / self.t1
        self.E1 = C.m_e * C.c * self.w1 / C.e
        self.n1 = C.epsilon_0 * C.m_e * self.w1 ** 2 / C.e ** 2
        self.P1 = self.n1 * C.e * mks_length
        self.u1 = C.m_e * C.c ** 2
        self.N1 = self.n1 * self.x1 ** 3
        self._mult = {
            'fs': 1e-15 / self.t1, 'ps': 1e-12 / self.t1, 'ns': 1e-9 / self.t1, 'us': 1e-6 / self.t1,
            'ms': 1e-3 / self.t1, 's': 1 / self.t1,
            'um': 1e-6 / self.x1, 'mm': 1e-3 / self.x1, 'cm': 1e-2 / self.x1, 'm': 1 / self.x1,
            'W/m2': self.t1 * self.x1 ** 2 / self.N1 / self.u1,
            'W/cm2': 1e4 * self.t1 * self.x1 ** 2 / self.N1 / self.u1,
            'm2/W': self.N1 * self.u1 / self.x1 ** 2 / self.t1,
            'm2/V2': C.epsilon_0 * self.E1 ** 3 / self.P1,
            'mJ': 1e-3 / self.u1, 'eV': C.e / self.u1, 'cm-3': 1e6 / self.n1,
        }

    def dnum(self, arg):
        '''Parse "value units" into simulation units; numbers pass through (input_tools.py:17-61).'''
        if not isinstance(arg, str):
            return arg
        parts = arg.split(' ')
        val = np.double(parts[0])
        units = parts[1] if len(parts) > 1 else ''
        if units not in self._mult:
            raise ValueError('Unrecognized units in input.')
        return val * self._mult[units]

    def TransformLimitedBandwidth(self, w00, t00, sigmas):
        t00 = self.dnum(t00)
        sigma_w = 2 / t00
        band = (w00 - sigmas * sigma_w, w00 + sigmas * sigma_w)
        if band[0] < 0.0:
            raise ValueError('Bandwidth calculation led to negative frequency.')
        return t00, band

    def a0(self, energy, t0, r0, w0):
        '''Peak normalised vector potential of a Gaussian pulse; t0, r0 are 1/e amplitude half widths.'''
        energy = self.dnum(energy)
        t0 = self.dnum(t0)
        r0 = self.dnum(r0)
        w0 = self.dnum(w0)
        P0 = np.sqrt(2 / np.pi) * energy / t0
        I0 = 2 * P0 / (np.pi * r0 ** 2) / (self.n1 * self.x1 ** 3)
        return np.sqrt(2 * I0) / w0

    def chi3(self, n0, n2):
        '''chi3 from the Kerr index n2: (4/3) n0^2 n2 (input_tools.py:110-111).'''
        return (4 / 3) * n0 ** 2 * self.dnum(n2)

'''Linear susceptibility chi(w) of the propagation medium (host side, FP64, once per run).

The UPPE stepper needs chi(w) at the reference density on the frequency nodes (uppe_kernel.py:389) and the
group velocity that sets the moving window (examples/uppe/air_fil.py:108). This module restates those
parts of SeaRay's ``modules/dispersion.py``: the Sellmeier families (dispersion.py:178-369), the
phenomenological opacity (dispersion.py:117-164), the named media (dispersion.py:371-470) and
``GroupVelocityMagnitude`` (dispersion.py:80-91). The coefficient tables are physical data and are the
same numbers as the reference's. The OpenCL ``Dxk`` generators are for ray tracing and are not provided.

Frequencies are in simulation units: w = 1 corresponds to the wavelength 2 pi mks_length.
'''
import numpy as np
from scipy import constants as C


def _as_array(w):
    arr = np.asarray(w, dtype=float)
    return arr, arr.ndim == 0


class Vacuum:
    '''chi = 0; group velocity 1 (dispersion.py:58-91).'''

    def chi(self, w):
        arr, scalar = _as_array(w)
        return 0.0 if scalar else np.zeros(arr.shape)

    def dchidw(self, w):
        arr, scalar = _as_array(w)
        return 0.0 if scalar else np.zeros(arr.shape)

    def vg(self, xp):
        vg = np.array(xp[..., :4], copy=True)
        vg[..., 0] = 1.0
        vg[..., 1:4] = xp[..., 5:8] / xp[..., 4:5]
        return vg

    def GroupVelocityMagnitude(self, w):
        '''|v_g| for a plane wave along z with k = w sqrt(1+Re chi) (dispersion.py:80-91).'''
        arr, scalar = _as_array(w)
        arr = np.atleast_1d(arr)
        xp = np.zeros(arr.shape + (8,))
        xp[..., 4] = arr
        xp[..., 5] = arr * np.real(np.sqrt(1 + np.asarray(self.chi(arr), dtype=complex)))
        v = self.vg(xp)[..., 1:]
        mag = np.sqrt(np.sum(v * v, axis=-1))
        return mag[0] if scalar else mag


class ColdPlasma(Vacuum):
    def chi(self, w):
        return -1 / np.asarray(w, dtype=float) ** 2

    def dchidw(self, w):
        return 2 / np.asarray(w, dtype=float) ** 3


class attenuator(Vacuum):
    '''Phenomenological opacity bands: Im chi = 2 alpha/(w + weps) inside (w1, w2) (dispersion.py:117-164).'''

    def __init__(self, mks_length):
        self.mks_length = mks_length
        self.alpha = []
        self.wstop1 = []
        self.wstop2 = []

    def reset_opacity(self):
        self.alpha, self.wstop1, self.wstop2 = [], [], []

    def add_opacity_region(self, attenuationLength, wavelength1, wavelength2):
        self.alpha.append(1 / attenuationLength)
        self.wstop1.append(2 * np.pi / wavelength2)
        self.wstop2.append(2 * np.pi / wavelength1)

    def chi(self, w):
        arr, scalar = _as_array(w)
        arr = np.atleast_1d(arr)
        if len(self.alpha) == 0:
            ans = np.zeros(arr.shape)
        else:
            a = np.array(self.alpha)
            w1 = np.array(self.wstop1)
            w2 = np.array(self.wstop2)
            weps = 1e-4 * np.min(w1 + w2)
            wx = arr[..., None]
            inside = np.logical_and(wx > w1, wx < w2)
            terms = np.where(inside, 1j * (2 * a) / (weps + wx), 0j)
            ans = np.sum(terms, axis=-1)
        return ans[0] if scalar else ans


class isotropic_medium(attenuator):
    def vg(self, xp):
        vg = np.array(xp[..., :4], copy=True)
        w = xp[..., 4]
        denom = w * (1.0 + self.chi(w) + 0.5 * w * self.dchidw(w))
        vg[..., 0] = 1.0
        for i in (1, 2, 3):
            vg[..., i] = np.real(xp[..., 4 + i] / denom)
        return vg


class sellmeier_medium(isotropic_medium):
    '''chi = sum_i A_i lambda^2/(lambda^2 - L_i), L in um^2 (dispersion.py:178-234).'''

    def __init__(self, mks_length, A, L):
        super().__init__(mks_length)
        self.A = np.asarray(A, dtype=float)
        self.D = np.asarray(L, dtype=float) * 1e-12 / self.mks_length ** 2 / (2 * np.pi) ** 2

    def chi(self, w):
        arr, scalar = _as_array(w)
        opacity = super().chi(w)
        res = np.sum(self.A / (1 - np.multiply.outer(arr ** 2, self.D)), axis=-1)
        return opacity + res

    def dchidw(self, w):
        arr, _ = _as_array(w)
        num = np.multiply.outer(2 * arr, self.A * self.D)
        den = (1.0 - np.multiply.outer(arr ** 2, self.D)) ** 2
        return np.sum(num / den, axis=-1)


class sellmeier_medium_alt1(isotropic_medium):
    '''chi = sum_i B_i/(C_i - lambda^-2), B and C in um^-2 (dispersion.py:236-294).'''

    def __init__(self, mks_length, B, C):
        super().__init__(mks_length)
        scale = (2.0 * np.pi) ** 2 * self.mks_length ** 2 * 1e12
        self.Bn = scale * np.asarray(B, dtype=float)
        self.Cn = scale * np.asarray(C, dtype=float)

    def chi(self, w):
        arr, scalar = _as_array(w)
        opacity = super().chi(w)
        res = np.sum(self.Bn / (self.Cn - np.multiply.outer(arr ** 2, np.ones_like(self.Cn))), axis=-1)
        return opacity + res

    def dchidw(self, w):
        arr, _ = _as_array(w)
        num = np.multiply.outer(2 * arr, self.Bn)
        den = (self.Cn - np.multiply.outer(arr ** 2, np.ones_like(self.Cn))) ** 2
        return np.sum(num / den, axis=-1)


class sellmeier_medium_alt2(isotropic_medium):
    '''chi = 2 sum_r (N_r/N_c) A_r lambda_r^2/(lambda^2-lambda_r^2) (dispersion.py:296-369).

    N_r in cm^-3, lambda_r in nm. The linewidth term is inactive in the reference (its attenuation factor is 0).'''

    def __init__(self, mks_length, Nr, Ar, lr, rel_linewidth):
        super().__init__(mks_length)
        ncrit = C.m_e * C.epsilon_0 * (C.c / self.mks_length) ** 2 / C.e ** 2
        Nn = np.asarray(Nr, dtype=float) * 1e6 / ncrit
        self.ln2 = (1e-9 * np.asarray(lr, dtype=float) / (2.0 * np.pi * self.mks_length)) ** 2
        self.Neff = 2 * Nn * np.asarray(Ar, dtype=float) * self.ln2
        self.wres = 1 / np.sqrt(self.ln2)
        if np.ndim(rel_linewidth) == 0:
            self.linewidth = np.ones(self.Neff.shape[0]) * rel_linewidth
        else:
            self.linewidth = np.asarray(rel_linewidth, dtype=float)

    def chi(self, w):
        arr, scalar = _as_array(w)
        opacity = super().chi(w)
        trans = 1 - np.multiply.outer(arr ** 2, self.ln2)
        res = np.sum(self.Neff / trans, axis=-1)
        return opacity + res

    def dchidw(self, w):
        arr, _ = _as_array(w)
        num = np.multiply.outer(2 * arr, self.Neff * self.ln2)
        den = (1 - np.multiply.outer(arr ** 2, self.ln2)) ** 2
        return np.sum(num / den, axis=-1)


class SimpleAir(sellmeier_medium_alt1):
    '''Two-term air Sellmeier (Voronin & Zheltikov, Sci. Rep. 7, 46111, Eq. 2) (dispersion.py:371-377).'''

    def __init__(self, mks_length):
        super().__init__(mks_length, B=2 * np.array([0.05792105, 0.00167917]), C=np.array([238.0185, 57.362]))


_DRY_N = 1e15 * np.array([9.4136, 9.4136, 9.4136, 9.4136, 19870, 5329.1, 237.63])
_DRY_A1 = np.array([4.051e-6, 2.897e-5, 8.573e-7, 1.55e-8, 1.2029482, 0.26507582, 0.93132145])
_DRY_L1 = np.array([15131, 4290.9, 2684.9, 2011.3, 85, 127, 87])
_DRY_A2 = np.array([1.01e-6, 2.728e-5, 6.62e-7, 5.532e-9, 5.796725, 7.734925, 7.217322])
_DRY_L2 = np.array([14218, 4223.1, 2769.1, 1964.6, 24.546, 29.469, 22.645])


class DryAir(sellmeier_medium_alt2):
    '''Generalised Sellmeier for dry air, table 1 rows 1-4, 12-14 of Voronin & Zheltikov (dispersion.py:379-391).'''

    def __init__(self, mks_length, rel_linewidth=0.0):
        super().__init__(mks_length, Nr=np.concatenate((_DRY_N, _DRY_N)), Ar=np.concatenate((_DRY_A1, _DRY_A2)),
                         lr=np.concatenate((_DRY_L1, _DRY_L2)), rel_linewidth=rel_linewidth)


class HumidAir(sellmeier_medium_alt2):
    '''Dry air plus water vapour at 296 K and relative humidity rel_hum (dispersion.py:393-429).'''

    def __init__(self, mks_length, rel_hum, rel_linewidth=0.0):
        T, Tc = 296.0, 647.096
        a = (-7.85951783, 1.84408259, -11.7866497, 22.6807411, -15.9618719, 1.80122502)
        t = 1 - T / Tc
        ps = 22.064e6 * np.exp((Tc / T) * (a[0] * t + a[1] * t ** 1.5 + a[2] * t ** 3 + a[3] * t ** 3.5
                                         + a[4] * t ** 4 + a[5] * t ** 7.5))
        self.NH2O_mks = rel_hum * ps / (C.k * T)
        Nw = np.ones(8) * self.NH2O_mks * 1e-6
        A1 = np.array([2.945e-5, 3.273e-6, 1.862e-6, 2.544e-7, 1.126e-7, 6.856e-9, 1.985e-9, 0.25787285])
        l1 = np.array([47862, 6719, 2775.6, 1835.6, 1417.6, 1145.3, 947.73, 128])
        A2 = np.array([6.583e-8, 3.094e-6, 2.788e-6, 2.181e-7, 2.336e-7, 9.479e-9, 2.882e-9, 4.742131])
        l2 = np.array([16603, 5729.9, 2598.5, 1904.8, 1364.7, 1123.2, 935.09, 34.924])
        super().__init__(mks_length,
                         Nr=np.concatenate((_DRY_N, _DRY_N, Nw, Nw)),
                         Ar=np.concatenate((_DRY_A1, _DRY_A2, A1, A2)),
                         lr=np.concatenate((_DRY_L1, _DRY_L2, l1, l2)), rel_linewidth=rel_linewidth)

    def WaterVaporDensity(self):
        return self.NH2O_mks


class LiquidWater(sellmeier_medium):
    def __init__(self, mks_length):
        super().__init__(mks_length, A=np.array([5.689e-1, 1.720e-1, 2.063e-2, 1.124e-1]),
                         L=np.array([5.110e-3, 1.825e-2, 2.624e-2, 1.068e1]))


class BK7(sellmeier_medium):
    def __init__(self, mks_length):
        super().__init__(mks_length, A=np.array([1.03961212, 0.231792344, 1.01046945]),
                         L=np.array([0.00600069867, 0.0200179144, 103.560653]))


class NaCl(sellmeier_medium):
    def __init__(self, mks_length):
        super().__init__(mks_length, A=np.array([0.00055, 0.198, 0.48398, 0.38696, 0.25998, 0.08796, 3.17064, 0.30038]),
                         L=np.array([0.0, 0.05, 0.1, 0.128, 0.158, 40.5, 60.98, 120.34]) ** 2)


class KCl(sellmeier_medium):
    def __init__(self, mks_length):
        super().__init__(mks_length, A=np.array([0.26486, 0.30523, 0.41620, 0.18870, 2.6200]),
                         L=np.array([0.0, 0.1, 0.131, 0.162, 70.42]) ** 2)


class ZnSe(sellmeier_medium):
    def __init__(self, mks_length):
        super().__init__(mks_length, A=np.array([4.45813734, 0.467216334, 2.89566290]),
                         L=np.array([0.200859853, 0.391371166, 47.1362108]) ** 2)


class Ge(sellmeier_medium):
    def __init__(self, mks_length):
        super().__init__(mks_length, A=np.array([0.4886331, 14.5142535, 0.0091224]),
                         L=np.array([1.393959, 0.1626427, 752.190]))

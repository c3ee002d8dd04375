'''Python access to the FP64 CPU oracle (liboracle_uppe.so). TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py use this module. It checks the HIP
library; it is never the product path.
'''
import ctypes as ct
import os

import numpy as np

from searay_amd import _abi
from searay_amd import uppe_kernel

from . import build as _build

_lib = None


def library():
    global _lib
    if _lib is None:
        if not os.path.exists(_build.LIB):
            _build.build_oracle()
        lib = ct.CDLL(_build.LIB)
        dp = ct.POINTER(ct.c_double)
        lib.orc_complex_fft.argtypes = [dp, dp, ct.c_long, ct.c_long, ct.c_double, ct.c_double]
        lib.orc_real_fft.argtypes = [dp, ct.c_long, ct.c_long, ct.c_int]
        lib.orc_rfft_strip.argtypes = [dp, dp, ct.c_long, ct.c_long]
        lib.orc_irfft_strip.argtypes = [dp, dp, ct.c_long, ct.c_long]
        lib.orc_plasma_density.argtypes = [dp, dp, ct.c_double, ct.c_long, ct.c_long, dp]
        lib.orc_uppe_rhs.argtypes = [ct.c_void_p, dp, ct.c_double, dp]
        lib.orc_uppe_source.argtypes = [ct.c_void_p, dp, dp, dp, dp]
        lib.orc_uppe_transverse.argtypes = [ct.c_void_p, dp, ct.c_int]
        lib.orc_num_threads.restype = ct.c_int
        lib.orc_uppe_set_linear_rhs.argtypes = [ct.c_void_p, ct.c_int, ct.c_double, ct.c_double]
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class OracleStepper(_abi.Stepper):
    def __init__(self, geometry, nw, n1, n2, nk, precision=_abi.PREC_FP64, device=0, model=_abi.MODEL_UPPE):
        super().__init__(library(), 'orc_uppe_', geometry, nw, n1, n2, nk, precision, device, model)

    def rhs(self, q, zphase):
        '''One right-hand-side evaluation S(zphase, q) (k-space in and out).'''
        q = np.ascontiguousarray(q, dtype=np.complex128)
        s = np.zeros_like(q)
        self._check(self.lib.orc_uppe_rhs(self.h, _p(q.view(np.float64)), zphase, _p(s.view(np.float64))), 'rhs')
        return s

    def source(self, e_r):
        '''update_current on an r-space field (UPPE: E; paraxial: A): (J_r, chi(t), ne(t)).'''
        e_r = np.ascontiguousarray(e_r, dtype=np.complex128)
        j = np.zeros_like(e_r)
        tshape = (self.nt,) + self.shape[1:]
        chi = np.zeros(tshape)
        ne = np.zeros(tshape)
        self._check(self.lib.orc_uppe_source(self.h, _p(e_r.view(np.float64)), _p(j.view(np.float64)), _p(chi),
                                             _p(ne)), 'source')
        return j, chi, ne

    def set_linear_rhs(self, lam, enable=True):
        '''Test hook: the stepper integrates dq/dz = lam q with the same RK4 and step control.'''
        self._check(self.lib.orc_uppe_set_linear_rhs(self.h, int(enable), float(np.real(lam)), float(np.imag(lam))),
                    'set_linear_rhs')

    def transverse(self, a, direction):
        a = np.array(a, dtype=np.complex128, copy=True, order='C')
        self._check(self.lib.orc_uppe_transverse(self.h, _p(a.view(np.float64)), int(direction)), 'transverse')
        return a


def oracle_factory(geometry, nw, n1, n2, nk, model=_abi.MODEL_UPPE):
    return OracleStepper(geometry, nw, n1, n2, nk, model=model)


def oracle_track(vol_dict, record=None, xp=None, eikonal=None, vg=None):
    '''The reference's track() on the oracle, UPPE or paraxial by vol_dict['propagator']:
    (A, J, chiNL, ne, dom4d, evals); ray data as track() takes it.'''
    if vol_dict.get('propagator', 'uppe') == 'paraxial':
        from searay_amd import paraxial_kernel
        return paraxial_kernel.track_with(vol_dict, oracle_factory, record, xp, eikonal, vg)
    return uppe_kernel.track_with(vol_dict, oracle_factory, record, xp, eikonal, vg)


# -- strip FFT restatements (kernels/fft.cl) exposed for their own checks --------------------------------

def complex_fft(z, inversion=1.0, expsign=-1.0):
    buf = np.array(np.asarray(z, dtype=np.complex128), copy=True)
    v = buf.view(np.float64)
    library().orc_complex_fft(_p(v), _p(v[1:]), len(buf), 2, inversion, expsign)
    return buf


def rfft(x):
    x = np.array(x, dtype=np.float64, copy=True)
    out = np.zeros(len(x) // 2 + 1, dtype=np.complex128)
    library().orc_rfft_strip(_p(x), _p(out.view(np.float64)), len(x), 1)
    return out


def irfft(X):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    out = np.zeros(2 * (len(X) - 1))
    library().orc_irfft_strip(_p(X.view(np.float64)), _p(out), len(X), 1)
    return out


def plasma_density(rate, ng, dt):
    rate = np.ascontiguousarray(rate, dtype=np.float64)
    ng = np.ascontiguousarray(ng, dtype=np.float64)
    nt = rate.shape[0]
    ntr = int(np.prod(rate.shape[1:]))
    out = np.zeros_like(rate)
    library().orc_plasma_density(_p(rate), _p(ng), dt, nt, ntr, _p(out))
    return out


def num_threads():
    return library().orc_num_threads()

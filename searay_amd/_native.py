'''Loader for the HIP library libsearay_uppe.so (built in-tree by searay_amd.build).

There is no fallback: if the library is missing or does not load, an ImportError says how to build it.
'''
import ctypes as ct
import os

import numpy as np

from . import _abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), '_lib', 'libsearay_uppe.so')
_lib = None


def library():
    '''Return the loaded HIP library, loading it on first use.'''
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError('libsearay_uppe.so not found at %s; build it with `python -m searay_amd.build`'
                              % LIB_PATH)
        lib = ct.CDLL(LIB_PATH)
        for name in _abi.HIP_ONLY_ENTRY_POINTS + _abi.BATCH_ENTRY_POINTS:
            getattr(lib, 'uppe_' + name).restype = ct.c_int
        lib.uppe_batch_last_error.restype = ct.c_char_p
        _lib = lib
    return _lib


class HipStepper(_abi.Stepper):
    '''Stepper handle on the HIP library, with the timing entry points.'''

    def __init__(self, geometry, nw, n1, n2, nk, precision, device=0, model=_abi.MODEL_UPPE):
        super().__init__(library(), 'uppe_', geometry, nw, n1, n2, nk, precision, device, model)

    def profile_enable(self, enable=True):
        self._check(self.lib.uppe_profile_enable(self.h, int(bool(enable))), 'profile_enable')

    def profile_read(self, kernel_class):
        ms = ct.c_double(0.0)
        n = ct.c_longlong(0)
        kc = _abi.KERNEL_CLASSES[kernel_class] if isinstance(kernel_class, str) else int(kernel_class)
        self._check(self.lib.uppe_profile_read(self.h, kc, ct.byref(ms), ct.byref(n)), 'profile_read')
        return ms.value, n.value

    def synchronize(self):
        self._check(self.lib.uppe_synchronize(self.h), 'synchronize')

    def advance_static(self, zi, length, sub0, n_sub, dzmin, dphi=0.4, rel_amin=1e-2, dz_schedule=None,
                       record_cap=0, max_steps=0):
        '''advance() through uppe_advance_static: the same stepping with the work arrays in the library's static
        device workspace (no device allocation). Returns (steps, recorded dz or None).'''
        return self._advance_with(self.lib.uppe_advance_static, 'advance_static', zi, length, sub0, n_sub, dzmin,
                                  dphi, rel_amin, dz_schedule, record_cap, max_steps)

    def advance_checkpointed(self, zi, length, sub0, n_sub, dzmin, every, dphi=0.4, rel_amin=1e-2, dz_schedule=None,
                             record_cap=0):
        '''advance() keeping what adjoint() needs: the state every `every` steps and each step's phase origin and
        dz. Returns (steps, recorded dz or None).'''
        sched = None if dz_schedule is None else _abi._f64(dz_schedule)
        rec = np.zeros(max(record_cap, 0)) if record_cap > 0 else None
        c = _abi.uppe_step_control(float(zi), float(length), int(sub0), 0, float(dzmin), float(dphi),
                                   float(rel_amin), int(n_sub), 0 if sched is None else len(sched), _abi._ptr(sched),
                                   _abi._ptr(rec), record_cap if rec is not None else 0, 0)
        n = ct.c_int(0)
        self._check(self.lib.uppe_advance_checkpointed(self.h, ct.byref(c), int(every), ct.byref(n)),
                    'advance_checkpointed')
        return n.value, (rec[:min(n.value, record_cap)].copy() if rec is not None else None)

    def adjoint(self, z_total, g_a):
        '''Cotangent of the r-space field given to set_field, from the cotangent g_a = dL/dRe a + i dL/dIm a of the
        field finish(z_total) returns, over the last checkpointed run.'''
        g = _abi._c128(g_a).reshape(self.shape)
        out = np.zeros(self.shape, dtype=np.complex128)
        self._check(self.lib.uppe_adjoint(self.h, ct.c_double(z_total), _abi._ptr(g.view(np.float64)),
                                          _abi._ptr(out.view(np.float64))), 'adjoint')
        return out


class BatchMember(HipStepper):
    '''A member handle of a batch (searay_amd.batch): a stepper like any other, owned by its batch.'''

    def __init__(self, handle, shape, model=_abi.MODEL_UPPE):
        self.lib = library()
        self.prefix = 'uppe_'
        self.shape = tuple(shape)
        self.model = model
        self.nt = 2 * (self.shape[0] - 1)
        self._bind()
        self._keep = []
        self.h = handle

    def close(self):
        # the batch releases its members
        self.h = None

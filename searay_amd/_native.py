'''Loader for the HIP library libsearay_uppe.so (built in-tree by searay_amd.build).

There is no fallback: if the library is missing or does not load, an ImportError says how to build it.
'''
import ctypes as ct
import os

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
        for name in _abi.HIP_ONLY_ENTRY_POINTS:
            getattr(lib, 'uppe_' + name).restype = ct.c_int
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

'''ctypes mirror of include/searay_uppe.h and a thin stepper wrapper over one handle.

``Stepper`` binds the entry points of a library that implements the header (by default the HIP library
``libsearay_uppe.so``) and converts between numpy arrays in the reference layout and the C calls.
'''
import ctypes as ct
import numpy as np

UPPE_OK = 0
UPPE_ERR_ARG = 1
UPPE_ERR_DEVICE = 2
UPPE_ERR_STATE = 3
UPPE_ERR_UNSUPPORTED = 4
UPPE_RADIAL = 0
UPPE_CARTESIAN = 1
PREC_FP64_PLAIN = 0
PREC_FP64 = 1
PREC_FP32 = 2
PREC_BF16X3 = 3
PRECISIONS = {'fp64-plain': PREC_FP64_PLAIN, 'fp64': PREC_FP64, 'fp32': PREC_FP32, 'bf16x3': PREC_BF16X3}
MODEL_UPPE = 0
MODEL_PARAXIAL = 1

KC_HANKEL, KC_SOURCE, KC_POINT, KC_REDUCE, KC_TRANSVERSE_FFT, KC_HANKEL_R2K = range(6)
KERNEL_CLASSES = {'hankel': KC_HANKEL, 'source': KC_SOURCE, 'point': KC_POINT, 'reduce': KC_REDUCE,
                  'transverse_fft': KC_TRANSVERSE_FFT, 'hankel_r2k': KC_HANKEL_R2K}

_dp = ct.POINTER(ct.c_double)


class uppe_desc(ct.Structure):
    _fields_ = [('geometry', ct.c_int), ('nw', ct.c_int), ('n1', ct.c_int), ('n2', ct.c_int),
                ('nk', ct.c_int), ('precision', ct.c_int), ('device', ct.c_int), ('model', ct.c_int)]


class uppe_ionization(ct.Structure):
    _fields_ = [('kind', ct.c_int), ('reserved', ct.c_int), ('e_conv', ct.c_double), ('rate_conv', ct.c_double),
                ('e_cutoff', ct.c_double), ('c_pre', ct.c_double), ('c_pow', ct.c_double),
                ('c_exp', ct.c_double), ('c', ct.c_double * 6)]


class uppe_medium(ct.Structure):
    _fields_ = [('w', _dp), ('kz', _dp), ('kg', _dp), ('ngas', _dp), ('nl_filter', _dp), ('cond_filter', _dp),
                ('redchi3', ct.c_double), ('dt', ct.c_double), ('ion', uppe_ionization)]


class uppe_paraxial_medium(ct.Structure):
    _fields_ = [('w', _dp), ('kz', _dp), ('k0', _dp), ('ngas', _dp), ('redchi3', ct.c_double), ('dt', ct.c_double),
                ('ion', uppe_ionization)]


class uppe_rotator(ct.Structure):
    _fields_ = [('states', ct.c_int), ('reserved', ct.c_int), ('tjj', _dp), ('wdt', _dp), ('dalpha', ct.c_double),
                ('hbar', ct.c_double)]


class uppe_step_control(ct.Structure):
    _fields_ = [('zi', ct.c_double), ('length', ct.c_double), ('sub0', ct.c_int), ('max_steps', ct.c_int),
                ('dzmin', ct.c_double), ('dphi', ct.c_double), ('rel_amin', ct.c_double),
                ('n_sub', ct.c_int), ('n_schedule', ct.c_int), ('dz_schedule', _dp), ('dz_record', _dp),
                ('record_cap', ct.c_int), ('reserved', ct.c_int)]


class uppe_info(ct.Structure):
    _fields_ = [('steps', ct.c_longlong), ('evals', ct.c_longlong), ('nan_steps', ct.c_int),
                ('reserved', ct.c_int), ('z_last', ct.c_double)]


ENTRY_POINTS = ('create', 'set_hankel', 'set_medium', 'set_paraxial_medium', 'set_rotator', 'set_field', 'advance',
                'finish', 'get_state', 'get_info', 'destroy', 'last_error')
HIP_ONLY_ENTRY_POINTS = ('profile_enable', 'profile_read', 'synchronize', 'advance_checkpointed', 'adjoint',
                         'advance_static')
BATCH_ENTRY_POINTS = ('batch_create', 'batch_member', 'batch_advance', 'batch_set_autotune', 'batch_set_graph',
                      'batch_plan', 'batch_destroy')


def _ptr(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class UppeError(RuntimeError):
    pass


def _ion_struct(ion):
    io = uppe_ionization()
    io.kind = int(ion['kind'])
    for key in ('e_conv', 'rate_conv', 'e_cutoff', 'c_pre', 'c_pow', 'c_exp'):
        setattr(io, key, ion[key])
    for i in range(6):
        io.c[i] = ion['c'][i]
    return io


class Stepper:
    '''One stepper handle. The caller keeps numpy arrays in the reference layout:
    spectral (Nw, N1, N2) complex128, time (Nt, N1, N2) float64 with Nt = 2 (Nw - 1) for the UPPE and Nt = Nw
    for the paraxial envelope (model MODEL_PARAXIAL).'''

    def __init__(self, lib, prefix, geometry, nw, n1, n2, nk, precision, device=0, model=MODEL_UPPE):
        self.lib = lib
        self.prefix = prefix
        self.shape = (nw, n1, n2)
        self.model = model
        self.nt = nw if model == MODEL_PARAXIAL else 2 * (nw - 1)
        self._bind()
        self._keep = []
        d = uppe_desc(geometry, nw, n1, n2, nk if nk is not None else n1, precision, device, model)
        h = ct.c_void_p()
        rc = self._f['create'](ct.byref(d), ct.byref(h))
        if rc != UPPE_OK:
            raise UppeError('%screate failed with status %d (grid %s, precision %d)' % (prefix, rc, self.shape,
                                                                                      precision))
        self.h = h

    def _bind(self):
        f = {}
        for name in ENTRY_POINTS:
            fn = getattr(self.lib, self.prefix + name)
            if name == 'last_error':
                fn.restype = ct.c_char_p
            else:
                fn.restype = ct.c_int
            f[name] = fn
        self._f = f

    def _check(self, rc, what):
        if rc != UPPE_OK:
            msg = self._f['last_error'](self.h)
            raise UppeError('%s%s failed (status %d): %s' % (self.prefix, what, rc,
                                                            msg.decode() if msg else ''))

    def set_hankel(self, H, Lambda):
        H = _f64(H)
        Lambda = _f64(Lambda)
        self._check(self._f['set_hankel'](self.h, _ptr(H), _ptr(Lambda)), 'set_hankel')

    def set_medium(self, w, kz, kg, ngas, nl_filter, cond_filter, redchi3, dt, ion):
        arrs = dict(w=_f64(w), kz=_c128(kz).view(np.float64), kg=_f64(kg).reshape(-1), ngas=_f64(ngas),
                    nl=_f64(nl_filter), cond=_f64(cond_filter))
        io = _ion_struct(ion)
        m = uppe_medium(_ptr(arrs['w']), _ptr(arrs['kz']), _ptr(arrs['kg']), _ptr(arrs['ngas']),
                        _ptr(arrs['nl']), _ptr(arrs['cond']), float(redchi3), float(dt), io)
        self._check(self._f['set_medium'](self.h, ct.byref(m)), 'set_medium')

    def set_paraxial_medium(self, w, kz, k0, ngas, redchi3, dt, ion):
        '''Paraxial medium (paraxial_kernel.py:72-102): frequencies in ifftshift order, complex kz (k-space),
        k0(w), gas density, chi3 / n_ref, time step, envelope rate parameters.'''
        arrs = dict(w=_f64(w), kz=_c128(kz).view(np.float64), k0=_f64(k0).reshape(-1), ngas=_f64(ngas))
        m = uppe_paraxial_medium(_ptr(arrs['w']), _ptr(arrs['kz']), _ptr(arrs['k0']), _ptr(arrs['ngas']),
                                 float(redchi3), float(dt), _ion_struct(ion))
        self._check(self._f['set_paraxial_medium'](self.h, ct.byref(m)), 'set_paraxial_medium')

    def set_rotator(self, rotator, dt):
        '''Load a rotator's tables for source time step dt, or switch the rotator off (None). Any object with the
        attributes of the reference's Rotator works (rotations.py:27-35): Wj, Tj, dalpha, hbar; the device gets
        Tj and W_j dt as AddChi uploads them (rotations.py:56-61).'''
        if rotator is None:
            self._check(self._f['set_rotator'](self.h, None), 'set_rotator')
            return
        wj = np.asarray(rotator.Wj)
        tjj = _f64(rotator.Tj)
        wdt = _c128(wj * dt).view(np.float64)
        if tjj.shape[0] != wj.shape[0]:
            raise ValueError('rotator: Tj and Wj differ in length')
        r = uppe_rotator(int(wj.shape[0]), 0, _ptr(tjj), _ptr(wdt), float(rotator.dalpha), float(rotator.hbar))
        self._check(self._f['set_rotator'](self.h, ct.byref(r)), 'set_rotator')

    def set_field(self, a_r):
        a = _c128(a_r).reshape(self.shape)
        self._check(self._f['set_field'](self.h, _ptr(a.view(np.float64))), 'set_field')

    def advance(self, zi, length, sub0, n_sub, dzmin, dphi=0.4, rel_amin=1e-2, dz_schedule=None, record_cap=0,
                max_steps=0):
        return self._advance_with(self._f['advance'], 'advance', zi, length, sub0, n_sub, dzmin, dphi, rel_amin,
                                  dz_schedule, record_cap, max_steps)

    def _advance_with(self, fn, what, zi, length, sub0, n_sub, dzmin, dphi, rel_amin, dz_schedule, record_cap,
                      max_steps):
        sched = None if dz_schedule is None else _f64(dz_schedule)
        rec = np.zeros(max(record_cap, 0)) if record_cap > 0 else None
        c = uppe_step_control(float(zi), float(length), int(sub0), int(max_steps), float(dzmin), float(dphi), float(rel_amin),
                              int(n_sub), 0 if sched is None else len(sched), _ptr(sched), _ptr(rec),
                              record_cap if rec is not None else 0, 0)
        n = ct.c_int(0)
        self._check(fn(self.h, ct.byref(c), ct.byref(n)), what)
        if rec is not None:
            return n.value, rec[:min(n.value, record_cap)].copy()
        return n.value, None

    def finish(self, z_total, want_diagnostics=True):
        a = np.zeros(self.shape, dtype=np.complex128)
        j = np.zeros(self.shape, dtype=np.complex128) if want_diagnostics else None
        tshape = (self.nt,) + self.shape[1:]
        chi = np.zeros(tshape) if want_diagnostics else None
        ne = np.zeros(tshape) if want_diagnostics else None
        rc = self._f['finish'](self.h, ct.c_double(z_total), _ptr(a.view(np.float64)),
                               _ptr(j.view(np.float64)) if j is not None else None, _ptr(chi), _ptr(ne))
        self._check(rc, 'finish')
        return a, j, chi, ne

    def get_state(self):
        q = np.zeros(self.shape, dtype=np.complex128)
        self._check(self._f['get_state'](self.h, _ptr(q.view(np.float64))), 'get_state')
        return q

    def info(self):
        i = uppe_info()
        self._check(self._f['get_info'](self.h, ct.byref(i)), 'get_info')
        return dict(steps=i.steps, evals=i.evals, nan_steps=i.nan_steps, z_last=i.z_last)

    def close(self):
        if getattr(self, 'h', None):
            self._f['destroy'](self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

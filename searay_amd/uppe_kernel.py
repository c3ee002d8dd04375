'''Drop-in replacement for SeaRay's ``modules/uppe_kernel.py`` running on MI355X.

The module keeps the reference's call stack and signatures:

* ``track(cl, xp, eikonal, vg, vol_dict) -> (A, J, chiNL, ne, dom4d)`` (uppe_kernel.py:333-424)
* ``propagator(cl, ctool, vg, a, mat, ionizer, rotator, NL_band, subcycles, zi, zf, dzmin)
  -> (a, J, rfft(chi), rfft(ne), evaluations)`` (uppe_kernel.py:260-331)
* ``Material`` (uppe_kernel.py:32-90) builds kz, kg and the gas density on the host.

Everything that the reference runs as OpenCL kernels runs in the HIP library (``libsearay_uppe.so``) behind
the C ABI of include/searay_uppe.h. ``cl`` is accepted for compatibility and ignored. Two optional input
dictionary keys select the device path: 'uppe precision' in {'fp64', 'fp32', 'bf16x3', 'fp64-plain'}
(default 'fp64') and 'uppe device' (HIP ordinal, default 0).

The ray boundary follows the reference: without 'incoming wave' the initial field comes from the primary rays
(``field_tool.GetBoundaryFields(xp[:, 0, :], eikonal, 1)``), and after the last plane the rays are relaunched
from the exit field (``RelaunchRays``, or ``RelaunchRays1`` with 'full relaunch'), updating xp, eikonal and vg in
place. The rotational Raman ``rotator`` is loaded into the stepper and added to chi in every source evaluation.
'''
import logging

import numpy as np

from . import _abi
from . import caustic_tools
from . import grid_tools
from . import ionization

PRECISION_KEY = 'uppe precision'
DEVICE_KEY = 'uppe device'


class Material:
    '''Host-side medium: density plane, chi0(w), kz = sqrt(w^2(1+chi0) - kr^2), kg = -w/vg and the gas density
    in plasma units (uppe_kernel.py:32-90).'''

    def __init__(self, queue, ctool, N, coords, vol, chi, chi3, nref, vg):
        w, x1, x2, ext = ctool.GetGridInfo()
        self.T = ctool.GetTransverseTool()
        self.N = N
        self.vol = vol
        self.chi = chi
        self.w = w
        self.redchi3 = chi3 / nref
        self.n_ref = nref
        self.vg = vg
        self.xp_eff = np.zeros((N[1], N[2], 8))
        if coords == 'cartesian':
            self.xp_eff[..., 1] = np.outer(x1, np.ones(N[2]))
            self.xp_eff[..., 2] = np.outer(np.ones(N[1]), x2)
        else:
            self.xp_eff[..., 1] = np.outer(x1, np.cos(x2))
            self.xp_eff[..., 2] = np.outer(x1, np.sin(x2))
        self.kz = None
        self.kg = None
        self.ngas = None
        self.dens = None

    def density(self, z):
        self.xp_eff[..., 3] = z
        return np.asarray(self.vol.GetRelDensity(self.xp_eff), dtype=float)

    def UpdateMaterial(self, z, dens=None):
        '''Tables at longitudinal position z (uppe_kernel.py:67-90).'''
        if dens is None:
            dens = self.density(z)
        chi0 = np.copy(self.chi)
        # w with Re chi > 0 take the smallest density, w with Re chi < 0 the largest (uppe_kernel.py:72-77)
        chi0[np.where(np.real(self.chi) > 0.0)] *= np.min(dens)
        chi0[np.where(np.real(self.chi) < 0.0)] *= np.max(dens)
        w2n2 = self.w ** 2 * (1 + chi0)
        kz = np.sqrt(0j + w2n2[..., np.newaxis, np.newaxis] - self.T.kr2()[np.newaxis, ...])
        kz[np.where(np.imag(kz) < 0)] *= -1
        kz[np.where(np.real(kz) == 0.0)] = 1.0
        self.kz = kz
        self.kg = -self.w / self.vg
        self.ngas = self.n_ref * dens
        self.dens = dens


def source_tables(w, NL_band):
    '''dt, the nonlinear band filter and the conditioner of the reference's Source (uppe_kernel.py:95-129).'''
    dt = np.pi / np.max(w)
    dw = w[1] - w[0]
    freqs = w.shape[0]
    nl = np.ones(freqs)
    nl[:int(NL_band[0] / dw)] = 0.0
    nl[int(NL_band[1] / dw):] = 0.0
    cond = np.ones(freqs)
    cond[:4] = 0.0
    return dt, nl, cond


def hip_stepper_factory(precision='fp64', device=0):
    from . import _native
    prec = _abi.PRECISIONS[precision] if isinstance(precision, str) else int(precision)

    def make(geometry, nw, n1, n2, nk, model=_abi.MODEL_UPPE):
        return _native.HipStepper(geometry, nw, n1, n2, nk, prec, device, model)
    return make


class _RunContext:
    '''Per-track state shared by the diagnostic planes: one stepper handle for the whole run.'''

    def __init__(self, stepper_factory, geometry, ctool, N):
        T = ctool.GetTransverseTool()
        nk = getattr(T, 'Nk', N[1])
        self.stepper = stepper_factory(geometry, N[0], N[1], N[2], nk)
        self.geometry = geometry
        if geometry == _abi.UPPE_RADIAL:
            self.stepper.set_hankel(T.H, T.Lambda)
        self.medium_key = None
        self.rotator_key = (id(None), None)


def _propagate(ctx, ctool, vg, a, mat, ionizer, NL_band, subcycles, zi, zf, dzmin, record=None, rotator=None):
    w, x, y, ext = ctool.GetGridInfo()
    if w[0] != 0.0:
        raise ValueError('UPPE propagator requires frequency range [0,...]')
    if w.shape[0] == 1:
        raise ValueError('UPPE propagator cannot be used for monochromatic radiation.')
    st = ctx.stepper
    L = (zf - zi) / subcycles
    dt, nl, cond = source_tables(w, NL_band)
    ion = ionizer.kernel_params() if ionizer is not None else ionization.no_ionization_params()
    if ctx.rotator_key != (id(rotator), dt):
        st.set_rotator(rotator, dt)
        ctx.rotator_key = (id(rotator), dt)
    st.set_field(a)
    # density at every subcycle midpoint; consecutive subcycles with the same density share one device call
    dens = [mat.density(zi + sub * L + 0.5 * L) for sub in range(subcycles)]
    iterations = 0
    sub = 0
    while sub < subcycles:
        n = 1
        while sub + n < subcycles and np.array_equal(dens[sub + n], dens[sub]):
            n += 1
        # everything set_medium uploads: the density, the filters, the source time step, chi3 and the rate model
        key = (dens[sub].tobytes(), nl.tobytes(), cond.tobytes(), float(dt), float(mat.redchi3), repr(sorted(ion.items())))
        if key != ctx.medium_key:
            mat.UpdateMaterial(zi + sub * L + 0.5 * L, dens[sub])
            st.set_medium(w, mat.kz, mat.kg, mat.ngas, nl, cond, mat.redchi3, dt, ion)
            ctx.medium_key = key
        steps, rec = st.advance(zi, L, sub, n, dzmin, record_cap=(0 if record is None else 1 << 20))
        if record is not None:
            record.extend(list(rec))
        iterations += steps
        sub += n
    a_r, J, chi_t, ne_t = st.finish(zf - zi)
    return a_r, J, np.fft.rfft(chi_t, axis=0), np.fft.rfft(ne_t, axis=0), 4 * iterations


def propagator(cl, ctool, vg, a, mat, ionizer, rotator, NL_band, subcycles, zi, zf, dzmin):
    '''Advance the field a(w, x1, x2) from zi to the next diagnostic plane zf (uppe_kernel.py:260-331).

    Returns (a_r, J_r, rfft(chi), rfft(ne), RHS evaluations). The stepper handle is cached on ``mat``.'''
    ctx = getattr(mat, '_uppe_ctx', None)
    if ctx is None:
        geometry = _abi.UPPE_CARTESIAN if isinstance(ctool, caustic_tools.FourierTool) else _abi.UPPE_RADIAL
        ctx = _RunContext(getattr(mat, 'stepper_factory', None) or hip_stepper_factory(), geometry, ctool,
                          mat.N)
        mat._uppe_ctx = ctx
    return _propagate(ctx, ctool, vg, a, mat, ionizer, NL_band, subcycles, zi, zf, dzmin, rotator=rotator)


class RunSetup:
    '''Host-side set-up of one track() call (uppe_kernel.py:341-404): grids, initial field, medium, walls.'''

    def __init__(self, vol_dict, xp=None, eikonal=None):
        band = vol_dict['frequency band']
        self.NL_band = vol_dict['nonlinear band']
        size = (band[1] - band[0],) + tuple(vol_dict['size'])
        N = tuple(vol_dict['wave grid'])
        self.N = N
        self.diagnostic_steps = N[3] - 1
        self.subcycles = vol_dict['subcycles']
        steps = self.diagnostic_steps * self.subcycles
        powersof2 = [2 ** i for i in range(32)]
        if N[0] - 1 not in powersof2:
            raise ValueError('UPPE propagator requires 2**n+1 w-nodes')
        if N[1] not in powersof2:
            raise ValueError('UPPE propagator requires 2**n x-nodes')
        if N[2] not in powersof2:
            raise ValueError('UPPE propagator requires 2**n y-nodes')
        self.size = size
        self.window_speed = vol_dict.get('window speed', 1.0)
        chi3 = vol_dict.get('chi3', 0.0)
        self.cartesian = vol_dict['wave coordinates'] == 'cartesian'
        if self.cartesian:
            self.field_tool = caustic_tools.FourierTool(N, band, (0, 0, 0), size[1:])
        else:
            self.field_tool = caustic_tools.BesselBeamTool(N, band, (0, 0, 0), size[1:], None,
                                                           vol_dict['radial modes'])
        self.geometry = _abi.UPPE_CARTESIAN if self.cartesian else _abi.UPPE_RADIAL
        w_nodes, x1_nodes, x2_nodes, dom3d = self.field_tool.GetGridInfo()
        self.w = w_nodes
        if 'incoming wave' not in vol_dict:
            if xp is None or eikonal is None:
                raise ValueError("no 'incoming wave' and no ray data for the UPPE boundary")
            logging.info('Start UPPE propagation using ray data.')
            logging.warning('Polarization information is lost upon entering UPPE region.')
            self.A0 = np.asarray(self.field_tool.GetBoundaryFields(xp[:, 0, :], eikonal, 1), dtype=np.complex128)
        else:
            logging.info('Start UPPE propagation using wave data')
            self.A0 = np.asarray(vol_dict['incoming wave'](w_nodes, x1_nodes, x2_nodes), dtype=np.complex128)
        chi = np.asarray(vol_dict['dispersion inside'].chi(w_nodes)).astype(np.cdouble)
        self.ionizer = None
        if vol_dict.get('ionizer', None) is not None:
            self.ionizer = ionization.Ionizer(vol_dict['ionizer'])
        self.rotator = vol_dict.get('rotator', None)
        self.mat = Material(None, self.field_tool, N, vol_dict['wave coordinates'], vol_dict['object'], chi,
                            chi3, vol_dict['density reference'], self.window_speed)
        dens_nodes = grid_tools.cell_centers(-size[3] / 2, size[3] / 2, steps)
        field_walls = grid_tools.cell_walls(dens_nodes[0], dens_nodes[-1], steps)
        self.diagnostic_walls = np.linspace(-size[3] / 2, size[3] / 2, N[3])
        self.dzmin = vol_dict['minimum step']
        self.dom4d = np.concatenate((dom3d, [field_walls[0], field_walls[-1]]))
        self.dt, self.nl_filter, self.cond_filter = source_tables(w_nodes, self.NL_band)
        self.ion = self.ionizer.kernel_params() if self.ionizer is not None else ionization.no_ionization_params()

    def stepper_field(self):
        '''The initial field in the stepper's frequency order.'''
        return self.A0

    def plane(self, k):
        '''(zi, zf, subcycle length) of diagnostic interval k.'''
        zi = self.diagnostic_walls[k]
        zf = self.diagnostic_walls[k + 1]
        return zi, zf, (zf - zi) / self.subcycles

    def medium_is_uniform(self, plane):
        '''Whether every subcycle of diagnostic interval `plane` sees the density make_stepper loads (the midpoint of
        plane 0's first subcycle). The batched and adjoint paths load the medium once, where the reference updates
        it for every subcycle (uppe_kernel.py:300).'''
        zi0, _, L0 = self.plane(0)
        d0 = self.mat.density(zi0 + 0.5 * L0)
        zi, _, L = self.plane(plane)
        return all(np.array_equal(self.mat.density(zi + sub * L + 0.5 * L), d0) for sub in range(self.subcycles))

    def make_stepper(self, stepper_factory):
        '''A stepper with the transverse basis and the medium at the first subcycle loaded.'''
        T = self.field_tool.GetTransverseTool()
        st = stepper_factory(self.geometry, self.N[0], self.N[1], self.N[2], getattr(T, 'Nk', self.N[1]))
        if self.geometry == _abi.UPPE_RADIAL:
            st.set_hankel(T.H, T.Lambda)
        zi, zf, L = self.plane(0)
        self.mat.UpdateMaterial(zi + 0.5 * L)
        st.set_medium(self.w, self.mat.kz, self.mat.kg, self.mat.ngas, self.nl_filter, self.cond_filter,
                      self.mat.redchi3, self.dt, self.ion)
        st.set_rotator(self.rotator, self.dt)
        return st


def track_with(vol_dict, stepper_factory, record=None, xp=None, eikonal=None, vg=None):
    '''The body of ``track`` with an explicit stepper factory. Returns (A, J, chiNL, ne, dom4d, evals). With ray
    data, the rays are relaunched from the exit field as ``track`` does (uppe_kernel.py:419-423).'''
    rs = RunSetup(vol_dict, xp, eikonal)
    N = rs.N
    A = np.zeros(N, dtype=np.complex128)
    J = np.zeros(N, dtype=np.complex128)
    chiNL = np.zeros(N, dtype=np.complex128)
    ne = np.zeros(N, dtype=np.complex128)
    A[..., 0] = rs.A0
    ctx = _RunContext(stepper_factory, rs.geometry, rs.field_tool, N)
    rs.mat._uppe_ctx = ctx
    A0 = np.copy(A[..., 0])
    evals = []
    for k in range(rs.diagnostic_steps):
        zi, zf, _ = rs.plane(k)
        logging.info('Advancing to diagnostic plane %d', k + 1)
        A0, J0, chi0, ne0, ev = _propagate(ctx, rs.field_tool, rs.window_speed, A0, rs.mat, rs.ionizer, rs.NL_band,
                                           rs.subcycles, zi, zf, rs.dzmin, record, rotator=rs.rotator)
        logging.info('%d evaluations of j(w,kx,ky)', ev)
        evals.append(ev)
        A[..., k + 1] = A0
        J[..., k + 1] = J0
        chiNL[..., k + 1] = chi0
        ne[..., k + 1] = ne0
    ctx.stepper.close()
    if xp is not None:
        relaunch = rs.field_tool.RelaunchRays1 if vol_dict.get('full relaunch', False) else rs.field_tool.RelaunchRays
        relaunch(xp, eikonal, vg, A[..., -1], rs.size[3], vol_dict['dispersion inside'])
    return A, J, chiNL, ne, rs.dom4d, evals


def track(cl, xp, eikonal, vg, vol_dict):
    '''Propagate with the UPPE stepper on the GPU (uppe_kernel.py:333-424).

    ``cl`` is accepted for signature compatibility. ``xp`` (bundles, rays, 8), ``eikonal`` (bundles, 4) and ``vg``
    (bundles, rays, 4) are the rays on the entrance plane: they give the initial field when vol_dict has no
    'incoming wave', and are relaunched from the exit field in place. They may be None with 'incoming wave'.'''
    factory = hip_stepper_factory(vol_dict.get(PRECISION_KEY, 'fp64'), vol_dict.get(DEVICE_KEY, 0))
    A, J, chiNL, ne, dom4d, _ = track_with(vol_dict, factory, None, xp, eikonal, vg)
    return A, J, chiNL, ne, dom4d

'''Drop-in replacement for SeaRay's ``modules/paraxial_kernel.py`` running on MI355X.

The module keeps the reference's call stack and signatures:

* ``track(cl, xp, eikonal, vg, vol_dict) -> (A, J, chiNL, ne, dom4d)`` (paraxial_kernel.py:346-443)
* ``propagator(cl, ctool, a, mat, ionizer, rotator, subcycles, zi, zf) -> (a, J, chiNL, ne, evaluations)``
  (paraxial_kernel.py:251-344)
* ``Material`` (paraxial_kernel.py:27-102) builds the envelope's kz and k0 and the gas density on the host.

The paraxial envelope equation is stepped by the same HIP library as the UPPE (one handle with
``model = MODEL_PARAXIAL``, include/searay_uppe.h): the same transverse transforms, RK4 and step-size control,
with the envelope's linear propagator exp(i kz z), RHS (i/2) exp(-i kz z) J / k0 and source (envelope rates,
0.75 Kerr factor, the plasma current ne (|A|^2/4 - 1) A). ``cl`` is accepted for compatibility and ignored.
The device path is chosen with the same optional keys as the UPPE module: 'uppe precision' and 'uppe device'.
Frequencies are kept in the reference's ifftshift order on the device and sorted back on return.
'''
import logging

import numpy as np

from . import _abi
from . import caustic_tools
from . import grid_tools
from . import ionization
from .uppe_kernel import DEVICE_KEY, PRECISION_KEY, hip_stepper_factory


class Material:
    '''Host-side medium of the paraxial equation (paraxial_kernel.py:27-102): the density plane, chi0(w), and

        kz = (w^2 (1 + chi0 - ng^2) - w0^2 dn^2 - 2 w w0 ng dn - kperp^2) / (2 k0),  k0 = w0 dn + w ng,

    dn = n0 - ng, with frequencies in ifftshift order.'''

    def __init__(self, queue, ctool, N, coords, vol, chi, chi3, nref, w0, ng, n0):
        logging.info('creating material for paraxial solver')
        w, x1, x2, ext = ctool.GetGridInfo()
        self.T = ctool.GetTransverseTool()
        self.N = N
        self.vol = vol
        self.chi = np.fft.ifftshift(chi, axes=0)
        self.w = np.fft.ifftshift(w, axes=0)
        self.redchi3 = chi3 / nref
        self.n_ref = nref
        self.w0 = w0
        self.ng = ng
        self.n0 = n0
        self.xp_eff = np.zeros((N[1], N[2], 8))
        if coords == 'cartesian':
            self.xp_eff[..., 1] = np.outer(x1, np.ones(N[2]))
            self.xp_eff[..., 2] = np.outer(np.ones(N[1]), x2)
        else:
            self.xp_eff[..., 1] = np.outer(x1, np.cos(x2))
            self.xp_eff[..., 2] = np.outer(x1, np.sin(x2))
        self.kz = None
        self.k0 = None
        self.ngas = None

    def density(self, z):
        self.xp_eff[..., 3] = z
        return np.asarray(self.vol.GetRelDensity(self.xp_eff), dtype=float)

    def UpdateMaterial(self, z, dens=None):
        '''Tables at longitudinal position z (paraxial_kernel.py:72-102).'''
        if dens is None:
            dens = self.density(z)
        chi0 = np.copy(self.chi)
        chi0[np.where(np.real(self.chi) > 0.0)] *= np.min(dens)
        chi0[np.where(np.real(self.chi) < 0.0)] *= np.max(dens)
        w, w0, ng = self.w, self.w0, self.ng
        dn = self.n0 - self.ng
        fw = w ** 2 * (1 + chi0 - ng ** 2) - w0 ** 2 * dn ** 2 - 2 * w * w0 * ng * dn
        kappa2 = fw[..., np.newaxis, np.newaxis] - self.T.kr2()[np.newaxis, ...]
        k0 = (w0 * dn + w * ng)[..., np.newaxis, np.newaxis]
        self.kz = 0.5 * kappa2 / k0
        self.k0 = k0
        self.ngas = self.n_ref * dens


def source_dt(w):
    '''Time step of the envelope strips: 2 pi over the full band of N nodes (paraxial_kernel.py:110-114).'''
    if w.shape[0] == 1:
        bandwidth = 1.0
    else:
        bandwidth = (np.max(w) - np.min(w)) * (1 + 1 / (w.shape[0] - 1))
    return 2 * np.pi / bandwidth


class _RunContext:
    '''Per-track state shared by the diagnostic planes: one stepper handle for the whole run.'''

    def __init__(self, stepper_factory, geometry, ctool, N):
        T = ctool.GetTransverseTool()
        nk = getattr(T, 'Nk', N[1])
        self.stepper = stepper_factory(geometry, N[0], N[1], N[2], nk, model=_abi.MODEL_PARAXIAL)
        self.geometry = geometry
        if geometry == _abi.UPPE_RADIAL:
            self.stepper.set_hankel(T.H, T.Lambda)
        self.medium_key = None
        self.rotator_key = (id(None), None)


def _propagate(ctx, ctool, a, mat, ionizer, rotator, subcycles, zi, zf, record=None):
    w, x, y, ext = ctool.GetGridInfo()
    st = ctx.stepper
    a = np.fft.ifftshift(a, axes=0)
    w = np.fft.ifftshift(w, axes=0)
    L = (zf - zi) / subcycles
    dt = source_dt(w)
    ion = ionizer.kernel_params() if ionizer is not None else ionization.no_ionization_params()
    if ctx.rotator_key != (id(rotator), dt):
        st.set_rotator(rotator, dt)
        ctx.rotator_key = (id(rotator), dt)
    st.set_field(a)
    # UpdateMaterial at every subcycle midpoint (paraxial_kernel.py:317); subcycles with the same density
    # share one device call
    dens = [mat.density(zi + sub * L + 0.5 * L) for sub in range(subcycles)]
    iterations = 0
    sub = 0
    while sub < subcycles:
        n = 1
        while sub + n < subcycles and np.array_equal(dens[sub + n], dens[sub]):
            n += 1
        # everything set_paraxial_medium uploads: the density, the source time step, chi3 and the rate model
        key = (dens[sub].tobytes(), float(dt), float(mat.redchi3), repr(sorted(ion.items())))
        if key != ctx.medium_key:
            mat.UpdateMaterial(zi + sub * L + 0.5 * L, dens[sub])
            st.set_paraxial_medium(w, mat.kz, mat.k0, mat.ngas, mat.redchi3, dt, ion)
            ctx.medium_key = key
        # no minimum step in the paraxial step control (paraxial_kernel.py:209-220)
        steps, rec = st.advance(zi, L, sub, n, 0.0, record_cap=(0 if record is None else 1 << 20))
        if record is not None:
            record.extend(list(rec))
        iterations += steps
        sub += n
    a_r, J, chi_t, ne_t = st.finish(zf - zi)
    # sort frequencies (paraxial_kernel.py:338-342)
    return (np.fft.fftshift(a_r, axes=0), np.fft.fftshift(J, axes=0),
            np.fft.fftshift(np.fft.fft(chi_t, axis=0), axes=0), np.fft.fftshift(np.fft.fft(ne_t, axis=0), axes=0),
            4 * iterations)


def propagator(cl, ctool, a, mat, ionizer, rotator, subcycles, zi, zf):
    '''Advance the envelope a(w, x1, x2) from zi to the next diagnostic plane zf (paraxial_kernel.py:251-344).

    Returns (a, J, fft(chi), fft(ne), RHS evaluations), frequencies sorted. The stepper handle is cached on
    ``mat``.'''
    ctx = getattr(mat, '_paraxial_ctx', None)
    if ctx is None:
        geometry = _abi.UPPE_CARTESIAN if isinstance(ctool, caustic_tools.FourierTool) else _abi.UPPE_RADIAL
        ctx = _RunContext(getattr(mat, 'stepper_factory', None) or hip_stepper_factory(), geometry, ctool,
                          mat.N)
        mat._paraxial_ctx = ctx
    return _propagate(ctx, ctool, a, mat, ionizer, rotator, subcycles, zi, zf)


class RunSetup:
    '''Host-side set-up of one track() call (paraxial_kernel.py:354-420): grids, initial field, medium, walls.'''

    def __init__(self, vol_dict, xp=None, eikonal=None):
        band = vol_dict['frequency band']
        size = (band[1] - band[0],) + tuple(vol_dict['size'])
        self.size = size
        N = tuple(vol_dict['wave grid'])
        self.N = N
        self.diagnostic_steps = N[3] - 1
        self.subcycles = vol_dict['subcycles']
        steps = self.diagnostic_steps * self.subcycles
        powersof2 = [2 ** i for i in range(32)]
        if N[0] not in powersof2:
            raise ValueError('Paraxial propagator requires 2**n w-nodes')
        if N[1] not in powersof2:
            raise ValueError('Paraxial propagator requires 2**n x-nodes')
        if N[2] not in powersof2:
            raise ValueError('Paraxial propagator requires 2**n y-nodes')
        window_speed = vol_dict.get('window speed', 1.0)
        chi3 = vol_dict.get('chi3', 0.0)
        self.cartesian = vol_dict['wave coordinates'] == 'cartesian'
        if self.cartesian:
            self.field_tool = caustic_tools.FourierTool(N, band, (0, 0, 0), size[1:])
        else:
            self.field_tool = caustic_tools.BesselBeamTool(N, band, (0, 0, 0), size[1:], None,
                                                           vol_dict['radial modes'])
        self.geometry = _abi.UPPE_CARTESIAN if self.cartesian else _abi.UPPE_RADIAL
        w_nodes, x1_nodes, x2_nodes, dom3d = self.field_tool.GetGridInfo()
        if 'incoming wave' not in vol_dict:
            if xp is None or eikonal is None:
                raise ValueError("no 'incoming wave' and no ray data for the paraxial boundary")
            logging.info('Start paraxial propagation using ray data.')
            logging.warning('Polarization information is lost upon entering paraxial region.')
            self.A0 = np.asarray(self.field_tool.GetBoundaryFields(xp[:, 0, :], eikonal, 1), dtype=np.complex128)
        else:
            logging.info('Start paraxial propagation using wave data')
            self.A0 = np.asarray(vol_dict['incoming wave'](w_nodes, x1_nodes, x2_nodes), dtype=np.complex128)
        # envelope reference: the central frequency node and the window's index (paraxial_kernel.py:399-403)
        w0 = w_nodes[int(N[0] / 2)]
        n0 = 1 / window_speed
        ng = 1 / window_speed
        chi = np.asarray(vol_dict['dispersion inside'].chi(w_nodes)).astype(np.cdouble)
        self.ionizer = None
        if vol_dict.get('ionizer', None) is not None:
            self.ionizer = ionization.Ionizer(vol_dict['ionizer'])
        self.rotator = vol_dict.get('rotator', None)
        self.mat = Material(None, self.field_tool, N, vol_dict['wave coordinates'], vol_dict['object'], chi,
                            chi3, vol_dict['density reference'], w0, ng, n0)
        dens_nodes = grid_tools.cell_centers(-size[3] / 2, size[3] / 2, steps)
        field_walls = grid_tools.cell_walls(dens_nodes[0], dens_nodes[-1], steps)
        self.diagnostic_walls = np.linspace(-size[3] / 2, size[3] / 2, N[3])
        self.dom4d = np.concatenate((dom3d, [field_walls[0], field_walls[-1]]))
        self.dzmin = 0.0  # no minimum step in the paraxial step control (paraxial_kernel.py:209-220)

    def plane(self, k):
        '''(zi, zf, subcycle length) of diagnostic interval k.'''
        zi = self.diagnostic_walls[k]
        zf = self.diagnostic_walls[k + 1]
        return zi, zf, (zf - zi) / self.subcycles

    def make_stepper(self, stepper_factory):
        '''A paraxial stepper with the transverse basis, the medium at the first subcycle and the rotator loaded.
        Its fields are in ifftshift frequency order: see ``shifted`` and ``sorted``.'''
        T = self.field_tool.GetTransverseTool()
        st = stepper_factory(self.geometry, self.N[0], self.N[1], self.N[2], getattr(T, 'Nk', self.N[1]),
                             model=_abi.MODEL_PARAXIAL)
        if self.geometry == _abi.UPPE_RADIAL:
            st.set_hankel(T.H, T.Lambda)
        zi, zf, L = self.plane(0)
        self.mat.UpdateMaterial(zi + 0.5 * L)
        w = self.mat.w
        dt = source_dt(w)
        ion = self.ionizer.kernel_params() if self.ionizer is not None else ionization.no_ionization_params()
        st.set_paraxial_medium(w, self.mat.kz, self.mat.k0, self.mat.ngas, self.mat.redchi3, dt, ion)
        st.set_rotator(self.rotator, dt)
        return st

    def stepper_field(self):
        '''The initial field in the stepper's frequency order.'''
        return self.shifted(self.A0)

    @staticmethod
    def shifted(a):
        '''Sorted frequencies to the stepper's ifftshift order.'''
        return np.fft.ifftshift(a, axes=0)

    @staticmethod
    def sorted(a):
        '''The stepper's ifftshift order back to sorted frequencies.'''
        return np.fft.fftshift(a, axes=0)


def track_with(vol_dict, stepper_factory, record=None, xp=None, eikonal=None, vg=None):
    '''The body of ``track`` with an explicit stepper factory. Returns (A, J, chiNL, ne, dom4d, evals). With ray
    data, the rays are relaunched from the exit field as ``track`` does (paraxial_kernel.py:438-442).'''
    rs = RunSetup(vol_dict, xp, eikonal)
    N = rs.N
    A = np.zeros(N, dtype=np.complex128)
    J = np.zeros(N, dtype=np.complex128)
    chiNL = np.zeros(N, dtype=np.complex128)
    ne = np.zeros(N, dtype=np.complex128)
    A[..., 0] = rs.A0
    ctx = _RunContext(stepper_factory, rs.geometry, rs.field_tool, N)
    rs.mat._paraxial_ctx = ctx
    A0 = np.copy(A[..., 0])
    evals = []
    for k in range(rs.diagnostic_steps):
        zi, zf, _ = rs.plane(k)
        logging.info('Advancing to diagnostic plane %d', k + 1)
        A0, J0, chi0, ne0, ev = _propagate(ctx, rs.field_tool, A0, rs.mat, rs.ionizer, rs.rotator, rs.subcycles,
                                           zi, zf, record)
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
    '''Propagate the paraxial envelope with the stepper on the GPU (paraxial_kernel.py:346-443).

    ``cl`` is accepted for signature compatibility; the rays (or 'incoming wave') give the initial field and the
    rays are relaunched from the exit field in place, as ``uppe_kernel.track`` does.'''
    factory = hip_stepper_factory(vol_dict.get(PRECISION_KEY, 'fp64'), vol_dict.get(DEVICE_KEY, 0))
    A, J, chiNL, ne, dom4d, _ = track_with(vol_dict, factory, None, xp, eikonal, vg)
    return A, J, chiNL, ne, dom4d

'''Synthetic UPPE workloads: input dictionaries in the reference's format.

Each builder returns a ``vol_dict`` for ``uppe_kernel.track``. The seeded inputs are:

* ``radial_kerr_air``: radial Kerr-only propagation in air, 800 nm, 50 fs FWHM Gaussian, w0 = 1 mm, 0.5 P_cr,
  1 % random spectral phase, Nw = 513 over a 1 ps window, Nr = 128 Hankel modes, 100 fixed steps over 1 m.
* ``radial_o2_ionization``: radial Kerr + O2 photoionisation (Ip = 12.07 eV, Z = 0.53, 5.4e18 cm^-3) with plasma
  current, 800 nm, 40 fs, w0 = 100 um, 3 P_cr, Nw = 2049 over band (0, 8), Nr = 256, 400 adaptive
  subcycles over 20 cm; rate model 'explicit' (PPT tunnelling limit) or 'series' (PPT quintic fit).
* ``cartesian_filamentation``: 3-D Cartesian, super-Gaussian beam w0 = 1.2 mm with 10 % amplitude noise at
  20 P_cr, Kerr + O2, Nw = 513, 256 x 256 over 4 mm x 4 mm, 200 steps over 10 cm.
* ``gaussian_vacuum``: the linear vacuum Gaussian beam of the reference's UPPE test (test/gaussian_test.py:
  110-131), Cartesian or radial.

``scale`` shrinks a workload for tests: (Nw-1) and the transverse grid are divided by it (powers of 2), keeping
the physical window, and ``steps`` overrides the number of subcycles.
'''
import numpy as np
from scipy import constants as C

from . import dispersion
from . import input_tools
from . import ionization
from . import rotations
from . import volume

MKS_LENGTH = 0.8e-6 / (2 * np.pi)
N2_AIR = 1e-23        # m^2/W (examples/uppe/air_fil.py:28)
GAS_DENSITY = '5.4e18 cm-3'


def critical_power(n2=N2_AIR, wavelength=0.8e-6):
    '''P_cr = 3.77 lambda^2 / (8 pi n2) in W.'''
    return 3.77 * wavelength ** 2 / (8 * np.pi * n2)


def _box(vol):
    obj = volume.AnalyticBox('air')
    obj.Initialize(vol)
    return obj


def _gaussian_spectrum(w, t0):
    sigma = 2.0 / t0
    return np.exp(-(w - 1.0) ** 2 / sigma ** 2)


def _normalise(spec_w, a0):
    '''Scale so that the real time-domain field has peak a0.'''
    peak = np.max(np.abs(np.fft.irfft(spec_w)))
    return a0 / peak


def _base(helper, band, size, grid, nl_band, subcycles, dzmin, coords, modes=None):
    vol = {
        'frequency band': band,
        'nonlinear band': nl_band,
        'size': size,
        'wave grid': grid,
        'wave coordinates': coords,
        'subcycles': subcycles,
        'minimum step': dzmin,
        'density reference': helper.dnum(GAS_DENSITY),
        'density function': '1.0',
        'density lambda': lambda x, y, z, r2: np.ones(r2.shape),
        'propagator': 'uppe',
    }
    if modes is not None:
        vol['radial modes'] = modes
    vol['object'] = _box(vol)
    return vol


def kerr_pulse_sigma():
    '''Spectral 1/e half-width sigma = 2/t0 of the radial Kerr workload's pulse (50 fs FWHM), in units of w00.'''
    helper = input_tools.InputHelper(MKS_LENGTH)
    return 2.0 / (helper.dnum('50 fs') / np.sqrt(2 * np.log(2)))


def radial_kerr_air(scale=1, steps=None, seed=1000, precision=None):
    helper = input_tools.InputHelper(MKS_LENGTH)
    nw = (512 // scale) + 1
    nr = max(8, 128 // scale)
    T = helper.dnum('1 ps')
    dw = 2 * np.pi / T
    band = (0.0, 513 * dw)   # smaller grids keep the band and shorten the time window
    L = helper.dnum('1 m')
    subcycles = 100 if steps is None else steps
    air = dispersion.SimpleAir(MKS_LENGTH)
    vol = _base(helper, band, (helper.dnum('8 mm'), helper.dnum('8 mm'), L), (nw, nr, 1, 2), band, subcycles,
                L / subcycles, 'cylindrical', nr)
    vol['dispersion inside'] = air
    vol['dispersion outside'] = dispersion.Vacuum()
    vol['window speed'] = air.GroupVelocityMagnitude(1.0)
    vol['chi3'] = helper.chi3(1.0, '%g m2/W' % N2_AIR)
    t0 = helper.dnum('50 fs') / np.sqrt(2 * np.log(2))
    w0 = helper.dnum('1 mm')
    P0 = 0.5 * critical_power()
    energy = P0 * (t0 * helper.t1) / np.sqrt(2 / np.pi) / helper.u1
    a0 = helper.a0(energy, t0, w0, 1.0)
    def incoming(w, r, phi):
        rng = np.random.default_rng(seed)
        phase = 2 * np.pi * 0.01 * (rng.random(w.shape) - 0.5)
        spec = _gaussian_spectrum(w, t0)
        amp = _normalise(spec, a0)
        return amp * np.einsum('i,j,k', spec * np.exp(1j * phase), np.exp(-r ** 2 / w0 ** 2), np.ones(phi.shape))
    vol['incoming wave'] = incoming
    if precision:
        vol['uppe precision'] = precision
    return vol


def n2_rotator(states=32):
    '''The N2 rotator of the reference's examples (examples/uppe/rot_benchmark.py:50-55): B = 2 cm^-1,
    polarizability difference 4 pi 6.7e-25 cm^3, kT = 0.025 eV, nuclear degeneracy 6 - 3 (j mod 2), level
    damping 1/(100 ps), hbar normalised to the 800 nm carrier.'''
    helper = input_tools.InputHelper(MKS_LENGTH)
    j = np.arange(0, states)
    damping = np.ones(states) / helper.dnum('100 ps')
    dnuc = 6 - 3 * (j % 2)
    w00 = 1.0  # 2 pi / (800 nm) in units of 1/MKS_LENGTH
    hbar_norm = C.hbar * (w00 * C.c / MKS_LENGTH) / (C.m_e * C.c ** 2)
    return rotations.Rotator(2 / helper.dnum('1 cm'), 4 * np.pi * 6.7e-25 / helper.dnum('1 cm-3'),
                             helper.dnum('.025 eV'), dnuc, damping, hbar_norm)


def radial_raman_air(scale=1, steps=None, seed=1003, precision=None, states=32):
    '''The radial Kerr workload in N2 with the rotational (Raman) susceptibility added.'''
    vol = radial_kerr_air(scale=scale, steps=steps, seed=seed, precision=precision)
    vol['rotator'] = n2_rotator(states)
    return vol


def ionizing_pulse_sigma():
    '''Spectral 1/e half-width sigma = 2/t0 of the photoionisation workloads' pulse (40 fs FWHM), in units of w00.'''
    helper = input_tools.InputHelper(MKS_LENGTH)
    return 2.0 / (helper.dnum('40 fs') / np.sqrt(2 * np.log(2)))


def o2_ionizer(kind='explicit'):
    helper = input_tools.InputHelper(MKS_LENGTH)
    Uion = helper.dnum('12.07 eV')
    if kind == 'explicit':
        return ionization.PPT_Tunneling(MKS_LENGTH, False, Uion, 0.53, lstar=0, l=0, m=0, w0=1.0, terms=80)
    return ionization.PPT(MKS_LENGTH, False, Uion, 0.53, lstar=0, l=0, m=0, w0=1.0, terms=80)


def radial_o2_ionization(scale=1, steps=None, seed=1001, rate='explicit', precision=None, power=3.0,
                         length='20 cm', uv_opacity=False, nr=None):
    helper = input_tools.InputHelper(MKS_LENGTH)
    nw = (2048 // scale) + 1
    nr = max(8, 256 // scale) if nr is None else nr
    band = (0.0, 8.0 / scale)
    L = helper.dnum(length)
    subcycles = 400 if steps is None else steps
    air = dispersion.SimpleAir(MKS_LENGTH)
    if uv_opacity:
        # the UV absorption band of examples/uppe/air_fil.py:36 (Im kz up to 1.56e-4: exp(Im kz z) in the RHS
        # would leave the FP32 range after ~6.5 cm; the FP32 stepper keeps its state relative to a reference
        # point that moves with the run)
        air.add_opacity_region(helper.dnum('0.1 cm'), helper.dnum('0.01 um'), helper.dnum('0.3 um'))
    vol = _base(helper, band, (helper.dnum('2 mm'), helper.dnum('2 mm'), L), (nw, nr, 1, 2),
                (0.0, min(4.0, band[1])), subcycles, 1.0, 'cylindrical', nr)
    vol['dispersion inside'] = air
    vol['dispersion outside'] = dispersion.Vacuum()
    vol['window speed'] = air.GroupVelocityMagnitude(1.0)
    vol['chi3'] = helper.chi3(1.0, '%g m2/W' % N2_AIR)
    vol['ionizer'] = o2_ionizer(rate)
    t0 = helper.dnum('40 fs') / np.sqrt(2 * np.log(2))
    w0 = helper.dnum('100 um')
    P0 = power * critical_power()
    energy = P0 * (t0 * helper.t1) / np.sqrt(2 / np.pi) / helper.u1
    a0 = helper.a0(energy, t0, w0, 1.0)
    def incoming(w, r, phi):
        rng = np.random.default_rng(seed)
        phase = 2 * np.pi * 0.01 * (rng.random(w.shape) - 0.5)
        spec = _gaussian_spectrum(w, t0)
        amp = _normalise(spec, a0)
        return amp * np.einsum('i,j,k', spec * np.exp(1j * phase), np.exp(-r ** 2 / w0 ** 2), np.ones(phi.shape))
    vol['incoming wave'] = incoming
    if precision:
        vol['uppe precision'] = precision
    return vol


def cartesian_filamentation(scale=1, steps=None, seed=1002, precision=None, length='10 cm'):
    helper = input_tools.InputHelper(MKS_LENGTH)
    nw = (512 // scale) + 1
    nx = max(8, 256 // scale)
    T = helper.dnum('1 ps')
    dw = 2 * np.pi / T
    band = (0.0, 513 * dw)
    L = helper.dnum(length)
    subcycles = 200 if steps is None else steps
    air = dispersion.SimpleAir(MKS_LENGTH)
    vol = _base(helper, band, (helper.dnum('4 mm'), helper.dnum('4 mm'), L), (nw, nx, nx, 2), band, subcycles,
                L / subcycles, 'cartesian')
    vol['dispersion inside'] = air
    vol['dispersion outside'] = dispersion.Vacuum()
    vol['window speed'] = air.GroupVelocityMagnitude(1.0)
    vol['chi3'] = helper.chi3(1.0, '%g m2/W' % N2_AIR)
    vol['ionizer'] = o2_ionizer('explicit')
    t0 = helper.dnum('50 fs') / np.sqrt(2 * np.log(2))
    w0 = helper.dnum('1.2 mm')
    P0 = 20 * critical_power()
    energy = P0 * (t0 * helper.t1) / np.sqrt(2 / np.pi) / helper.u1
    a0 = helper.a0(energy, t0, w0, 1.0)
    def incoming(w, x, y):
        rng = np.random.default_rng(seed)
        spec = _gaussian_spectrum(w, t0)
        amp = _normalise(spec, a0)
        r2 = x[:, None] ** 2 + y[None, :] ** 2
        noise = 1.0 + 0.1 * (2 * rng.random(r2.shape) - 1.0)
        return amp * np.einsum('i,jk', spec, np.exp(-(r2 / w0 ** 2) ** 4) * noise)
    vol['incoming wave'] = incoming
    if precision:
        vol['uppe precision'] = precision
    return vol


def gaussian_vacuum(coords='cartesian', n=128, nw=65, precision=None):
    '''The reference's linear UPPE test setup: vacuum, a00 = 1, w00 = 1, r00 = 0.01 cm, band (0, 2),
    size 36 r00 x 36 r00 x 4 zR, three planes (test/gaussian_test.py:18-21, 60-63, 110-131).'''
    cm = 100 * MKS_LENGTH
    a00, w00, r00 = 1.0, 1.0, 0.01 / cm
    zR = 0.5 * w00 * r00 ** 2
    vol = {
        'density reference': 1.0,
        'density function': '1.0',
        'density lambda': lambda x, y, z, r2: np.ones(x.shape),
        'frequency band': (0.0, 2.0),
        'nonlinear band': (0.0, 2.0),
        'wave grid': (nw, n, n if coords == 'cartesian' else 1, 3),
        'wave coordinates': coords,
        'dispersion inside': dispersion.Vacuum(),
        'dispersion outside': dispersion.Vacuum(),
        'size': (36 * r00, 36 * r00, 4 * zR),
        'propagator': 'uppe',
        'subcycles': 1,
        'minimum step': 1,
    }
    if coords != 'cartesian':
        vol['radial modes'] = n

    def f(x, x0, dx):
        return np.exp(-(x - x0) ** 2 / dx ** 2)

    def incoming(w, x, y):
        if coords == 'cartesian':
            return a00 * np.einsum('i,j,k', f(w, w00, 0.1), f(x, 0, r00), f(y, 0, r00))
        return a00 * np.einsum('i,j,k', f(w, w00, 0.1), f(x, 0, r00), np.ones(y.shape))
    vol['incoming wave'] = incoming
    vol['object'] = _box(vol)
    if precision:
        vol['uppe precision'] = precision
    return vol, dict(a00=a00, w00=w00, r00=r00, zR=zR)


def gaussian_paraxial(coords='cartesian', n=128, nw=64, precision=None):
    '''The reference's paraxial Gaussian-beam test (test/gaussian_test.py:63-108): vacuum, a00 = 1, w00 = 1,
    r00 = 0.01 cm, band (0.95, 1.05) with a 0.01-wide spectrum, size 36 r00 x 36 r00 x 4 zR, three planes; the
    on-axis amplitude falls to 1/sqrt(5) and 1/sqrt(17) at the second and third planes.'''
    cm = 100 * MKS_LENGTH
    a00, w00, r00 = 1.0, 1.0, 0.01 / cm
    zR = 0.5 * w00 * r00 ** 2
    vol = {
        'density reference': 1.0,
        'density function': '1.0',
        'density lambda': lambda x, y, z, r2: np.ones(x.shape),
        'frequency band': (0.95, 1.05),
        'wave grid': (nw, n, n if coords == 'cartesian' else 1, 3),
        'wave coordinates': coords,
        'dispersion inside': dispersion.Vacuum(),
        'dispersion outside': dispersion.Vacuum(),
        'size': (36 * r00, 36 * r00, 4 * zR),
        'propagator': 'paraxial',
        'subcycles': 1,
    }
    if coords != 'cartesian':
        vol['radial modes'] = n

    def f(x, x0, dx):
        return np.exp(-(x - x0) ** 2 / dx ** 2)

    def incoming(w, x, y):
        if coords == 'cartesian':
            return a00 * np.einsum('i,j,k', f(w, w00, 0.01), f(x, 0, r00), f(y, 0, r00))
        return a00 * np.einsum('i,j,k', f(w, w00, 0.01), f(x, 0, r00), np.ones(y.shape))
    vol['incoming wave'] = incoming
    vol['object'] = _box(vol)
    if precision:
        vol['uppe precision'] = precision
    return vol, dict(a00=a00, w00=w00, r00=r00, zR=zR)


def paraxial_air_filament(scale=1, steps=None, seed=1004, precision=None, rate='explicit', power=1.0,
                          length='40 cm', planes=20, coords='cylindrical', rotator_states=0):
    '''The paraxial air filament of examples/paraxial/air-fil.py:25-109: 800 nm, 1 mJ in 50 fs, focused by a
    lens of 40 cm onto the centre of a 40 cm air box (entrance spot r00/2 with 20 cm radius of curvature),
    envelope band of 25 transform-limited widths either side of w00, Nw = 256, Nr = 256 with 128 Hankel modes
    (Cartesian: 256 x 256), Kerr (1e-23 m^2/W) and O2 photoionisation (12.1 eV, Z = 0.53, cycle-averaged rate,
    5.4e18 cm^-3), 20 diagnostic planes, one subcycle per plane. ``power`` scales the pulse energy; 1 %
    random spectral phase breaks the symmetry of the test inputs.'''
    helper = input_tools.InputHelper(MKS_LENGTH)
    um = 1e6 * MKS_LENGTH
    cm = 100 * MKS_LENGTH
    mm = 1000 * MKS_LENGTH
    wavelength = 0.8 / um
    waist = 150 / um
    w00 = 2 * np.pi / wavelength
    t00 = helper.dnum('50 fs')
    L = helper.dnum(length)
    time_to_focus = 40 / cm
    r00 = time_to_focus * wavelength / (np.pi * waist)
    sigma_w = 2 / t00
    band = (w00 - 25 * sigma_w, w00 + 25 * sigma_w)
    rbox = 3 / mm
    nw = 256 // scale
    nr = max(8, 256 // scale)
    air = dispersion.SimpleAir(MKS_LENGTH)
    vol = {
        'frequency band': band,
        'size': (2 * rbox, 2 * rbox, L),
        'wave grid': (nw, nr, nr if coords == 'cartesian' else 1, planes),
        'wave coordinates': coords,
        'subcycles': 1 if steps is None else steps,
        'density reference': helper.dnum(GAS_DENSITY),
        'density function': '1.0',
        'density lambda': lambda x, y, z, r2: np.ones(r2.shape),
        'propagator': 'paraxial',
        'dispersion inside': air,
        'dispersion outside': dispersion.Vacuum(),
        'chi3': helper.chi3(1.0, '%g m2/W' % N2_AIR),
        'window speed': air.GroupVelocityMagnitude(1.0),
    }
    if coords != 'cartesian':
        vol['radial modes'] = max(4, 128 // scale)
    if rate is not None:
        Uion = helper.dnum('12.1 eV')
        cls = ionization.PPT_Tunneling if rate == 'explicit' else ionization.PPT
        vol['ionizer'] = cls(MKS_LENGTH, True, Uion, 0.53, lstar=0, l=0, m=0, w0=w00, terms=80)
    if rotator_states:
        vol['rotator'] = n2_rotator(rotator_states)
    vol['object'] = _box(vol)
    a00 = helper.a0(power * helper.dnum('1 mJ'), t00, r00, w00)
    # at the box entrance, halfway from the lens to the focus: half the spot, converging with radius 20 cm
    w_in, R = 0.5 * r00, 0.5 * time_to_focus

    def incoming(w, x, y):
        rng = np.random.default_rng(seed)
        spec = np.exp(-(w - w00) ** 2 * t00 ** 2 / 4) * np.exp(2j * np.pi * 0.01 * (rng.random(w.shape) - 0.5))
        if coords == 'cartesian':
            r2 = np.add.outer(x ** 2, y ** 2)
        else:
            r2 = np.outer(x ** 2, np.ones(y.shape))
        env = np.exp(-r2 / w_in ** 2)[None] * np.exp(-1j * w[:, None, None] * r2[None] / (2 * R))
        # the envelope's time profile is ifft of the ifftshift-ordered spectrum: peak a00
        amp = a00 / np.max(np.abs(np.fft.ifft(np.fft.ifftshift(spec))))
        return amp * spec[:, None, None] * env
    vol['incoming wave'] = incoming
    if precision:
        vol['uppe precision'] = precision
    return vol


def argon_ionizer():
    '''Ar: Ip = 15.76 eV (photoionization_base.py:68), Z = 1, PPT tunnelling limit.'''
    helper = input_tools.InputHelper(MKS_LENGTH)
    return ionization.PPT_Tunneling(MKS_LENGTH, False, helper.dnum('15.76 eV'), 1.0, lstar=0, l=0, m=0, w0=1.0,
                                    terms=80)


def batched_radial(n=64, seed=1040, scale=1, steps=200, length='5 mm', precision=None):
    '''Independent radial UPPE instances for a batched run: Nr drawn uniformly from {64, 128, 256, 512}
    (each divided by `scale`, at least 8), Nw = 1025 over band (0, 4) (Nw - 1 and the band divided by `scale`),
    800 nm 40 fs pulses with w0 = 100 um and peak power log-uniform over 0.1-5 P_cr with 1 % spectral-phase
    noise, the gas O2 (12.07 eV, Z = 0.53) or Ar (15.76 eV, Z = 1) at 5.4e18 cm^-3 with Kerr, photoionisation
    (PPT tunnelling) and the plasma current; a 2 mm window and `steps` fixed steps over `length` (5 mm: the
    most powerful pulses self-focus within about 1 cm, where fixed 100 um steps no longer resolve the collapse
    and the field overflows, in FP64 as in FP32). Returns the list of vol_dicts and the drawn (Nr, P/P_cr, gas)
    of each.'''
    helper = input_tools.InputHelper(MKS_LENGTH)
    rng = np.random.default_rng(seed)
    nw = 1024 // scale + 1
    band = (0.0, 4.0 / scale)
    L = helper.dnum(length)
    t0 = helper.dnum('40 fs') / np.sqrt(2 * np.log(2))
    w0 = helper.dnum('100 um')
    vols, drawn = [], []
    for i in range(n):
        nr = max(8, int(rng.choice([64, 128, 256, 512])) // scale)
        power = 10 ** rng.uniform(np.log10(0.1), np.log10(5.0))
        gas = 'O2' if rng.integers(2) == 0 else 'Ar'
        air = dispersion.SimpleAir(MKS_LENGTH)
        vol = _base(helper, band, (helper.dnum('2 mm'), helper.dnum('2 mm'), L), (nw, nr, 1, 2), band, steps,
                    L / steps, 'cylindrical', nr)
        vol['dispersion inside'] = air
        vol['dispersion outside'] = dispersion.Vacuum()
        vol['window speed'] = air.GroupVelocityMagnitude(1.0)
        vol['chi3'] = helper.chi3(1.0, '%g m2/W' % N2_AIR)
        vol['ionizer'] = o2_ionizer('explicit') if gas == 'O2' else argon_ionizer()
        P0 = power * critical_power()
        energy = P0 * (t0 * helper.t1) / np.sqrt(2 / np.pi) / helper.u1
        a0 = helper.a0(energy, t0, w0, 1.0)
        phase_seed = seed * 1000 + i

        def incoming(w, r, phi, a0=a0, phase_seed=phase_seed):
            prng = np.random.default_rng(phase_seed)
            phase = 2 * np.pi * 0.01 * (prng.random(w.shape) - 0.5)
            spec = _gaussian_spectrum(w, t0)
            amp = _normalise(spec, a0)
            return amp * np.einsum('i,j,k', spec * np.exp(1j * phase), np.exp(-r ** 2 / w0 ** 2), np.ones(phi.shape))
        vol['incoming wave'] = incoming
        if precision:
            vol['uppe precision'] = precision
        vols.append(vol)
        drawn.append((nr, power, gas))
    return vols, drawn

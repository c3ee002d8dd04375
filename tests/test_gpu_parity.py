'''HIP library vs the FP64 oracle on identical seeded inputs, through the C ABI.

Tolerances: the FP64 paths (plain and fused) agree with the oracle to 1e-10 relative L2 (different FFT and
summation orders only); FP32 and BF16x3 agree to 1e-5 relative L2 on the field and 1e-4 on the electron
density, on the oracle's own step schedule.
'''
import sys

import numpy as np
import pytest

from conftest import rel_l2
from searay_amd import configs, paraxial_kernel, uppe_kernel

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_FIELD_TOL = 1e-5
FP32_NE_TOL = 1e-4


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _module(vol):
    return paraxial_kernel if vol.get('propagator', 'uppe') == 'paraxial' else uppe_kernel


def _run(vol, precision, record=None):
    return _module(vol).track_with(vol, uppe_kernel.hip_stepper_factory(precision), record)


@pytest.mark.parametrize('coords', ['cylindrical', 'cartesian'])
@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_gaussian_vacuum(oracle_mod, coords, precision):
    n = 128 if coords == 'cylindrical' else 64
    vol, p = configs.gaussian_vacuum(coords, n=n)
    _say('gaussian_vacuum %s %s n=%d' % (coords, precision, n))
    ref = oracle_mod.oracle_track(vol)
    got = _run(vol, precision)
    tol = FP64_TOL if precision.startswith('fp64') else FP32_FIELD_TOL
    for k in (1, 2):
        assert rel_l2(got[0][..., k], ref[0][..., k]) < tol
    assert got[5] == ref[5]


@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_radial_kerr(oracle_mod, precision):
    vol = configs.radial_kerr_air(scale=4, steps=6)
    _say('radial_kerr scale=4 steps=6 %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    got = _run(vol, precision)
    tol = FP64_TOL if precision.startswith('fp64') else FP32_FIELD_TOL
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < tol
    assert rel_l2(got[1][..., 1], ref[1][..., 1]) < (tol if precision.startswith('fp64') else 1e-4)
    assert rel_l2(got[2][..., 1], ref[2][..., 1]) < (tol if precision.startswith('fp64') else 1e-4)
    assert got[5] == ref[5]


@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_radial_raman(oracle_mod, precision):
    '''Kerr plus the 32-state N2 rotator (rotations.cl:80-129 inside the source term).'''
    vol = configs.radial_raman_air(scale=4, steps=6)
    _say('radial_raman scale=4 steps=6 %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    got = _run(vol, precision)
    kerr = _run(configs.radial_kerr_air(scale=4, steps=6, seed=1003), precision)
    fp64 = precision.startswith('fp64')
    tol = FP64_TOL if fp64 else FP32_FIELD_TOL
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < tol
    assert rel_l2(got[2][..., 1], ref[2][..., 1]) < (tol if fp64 else 1e-4)
    # the rotational term is a visible part of chiNL, not lost in the tolerance
    assert rel_l2(kerr[2][..., 1], ref[2][..., 1]) > 1e3 * (tol if fp64 else 1e-4)
    assert got[5] == ref[5]


@pytest.mark.parametrize('rate', ['explicit', 'series'])
@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_radial_ionization_adaptive(oracle_mod, rate, precision):
    vol = configs.radial_o2_ionization(scale=4, steps=4, rate=rate, length='1 mm', power=3.0)
    _say('radial_ionization scale=4 %s %s' % (rate, precision))
    rec_ref = []
    ref = oracle_mod.oracle_track(vol, record=rec_ref)
    rec = []
    got = _run(vol, precision, rec)
    if precision.startswith('fp64'):
        assert len(rec) == len(rec_ref)
        assert np.allclose(rec, rec_ref, rtol=1e-8, atol=0)
        tol, tol_ne = FP64_TOL, FP64_TOL
    else:
        tol, tol_ne = 1e-4, 1e-3   # adaptive schedules may differ slightly; exact replay is tested below
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < tol
    assert rel_l2(got[3][..., 1], ref[3][..., 1]) < tol_ne


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_radial_ionization_replayed_schedule(oracle_mod, precision):
    '''FP32 / BF16x3 on the oracle's own dz schedule: field within 1e-5, electron density within 1e-4.'''
    vol = configs.radial_o2_ionization(scale=4, steps=4, rate='explicit', length='1 mm', power=3.0)
    _say('radial_ionization replayed schedule %s' % precision)
    rec_ref = []
    ref = oracle_mod.oracle_track(vol, record=rec_ref)
    from searay_amd import uppe_kernel as uk
    rs = uk.RunSetup(vol)
    st = rs.make_stepper(uk.hip_stepper_factory(precision))
    zi, zf, L = rs.plane(0)
    st.set_field(rs.A0)
    n, _ = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, dz_schedule=np.array(rec_ref))
    a, J, chi_t, ne_t = st.finish(zf - zi)
    st.close()
    assert n == len(rec_ref)
    assert rel_l2(a, ref[0][..., 1]) < FP32_FIELD_TOL
    assert rel_l2(np.fft.rfft(ne_t, axis=0), ref[3][..., 1]) < FP32_NE_TOL


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_radial_ionization_simt_contraction(oracle_mod, precision, monkeypatch):
    '''The SIMT radial contraction (SEARAY_UPPE_SIMT_HANKEL=1, the fallback of the matrix-core path): the same
    adaptive run against the oracle, with its dz schedule (FP64) or within the adaptive FP32 tolerances.'''
    monkeypatch.setenv('SEARAY_UPPE_SIMT_HANKEL', '1')
    vol = configs.radial_o2_ionization(scale=4, steps=4, rate='explicit', length='1 mm', power=3.0)
    _say('radial_ionization simt contraction %s' % precision)
    rec_ref = []
    ref = oracle_mod.oracle_track(vol, record=rec_ref)
    rec = []
    got = _run(vol, precision, rec)
    if precision == 'fp64':
        assert len(rec) == len(rec_ref)
        assert np.allclose(rec, rec_ref, rtol=1e-8, atol=0)
        tol, tol_ne = FP64_TOL, FP64_TOL
    else:
        tol, tol_ne = 1e-4, 1e-3
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < tol
    assert rel_l2(got[3][..., 1], ref[3][..., 1]) < tol_ne


# ---- committed golden fixtures (tests/golden) --------------------------------------------------------------

import golden_cases  # noqa: E402


@pytest.mark.parametrize('name', golden_cases.NAMES)
@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64'])
def test_golden_fp64(name, precision):
    '''FP64 paths against the frozen oracle output: field and diagnostics 1e-10, same dz schedule.'''
    make, fx = golden_cases.load(name)
    _say('golden %s %s' % (name, precision))
    rec = []
    A, J, chi, ne, dom4d, evals = _run(make(), precision, rec)
    assert list(evals) == list(fx['evals'])
    assert np.allclose(rec, fx['dz'], rtol=1e-8, atol=0)
    assert rel_l2(A[..., -1], fx['A']) < FP64_TOL
    for key, arr in (('J', J), ('chi', chi), ('ne', ne)):
        if key in fx:
            assert rel_l2(arr[..., -1], fx[key]) < FP64_TOL


@pytest.mark.parametrize('name', golden_cases.NAMES)
@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_golden_fp32_replayed(name, precision):
    '''FP32 / BF16x3 on the fixture's dz schedule: field within 1e-5, electron density within 1e-4.'''
    make, fx = golden_cases.load(name)
    _say('golden replay %s %s' % (name, precision))
    vol = make()
    paraxial = vol.get('propagator', 'uppe') == 'paraxial'
    rs = _module(vol).RunSetup(vol)
    assert rs.diagnostic_steps == 1
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(precision))
    zi, zf, L = rs.plane(0)
    st.set_field(rs.shifted(rs.A0) if paraxial else rs.A0)
    n, _ = st.advance(zi, L, 0, rs.subcycles, 0.0 if paraxial else rs.dzmin, dz_schedule=fx['dz'])
    a, J, chi_t, ne_t = st.finish(zf - zi)
    st.close()
    assert n == len(fx['dz'])
    if paraxial:
        a = rs.sorted(a)
        ne_w = rs.sorted(np.fft.fft(ne_t, axis=0))
    else:
        ne_w = np.fft.rfft(ne_t, axis=0)
    assert rel_l2(a, fx['A']) < FP32_FIELD_TOL
    if 'ne' in fx:
        assert rel_l2(ne_w, fx['ne']) < FP32_NE_TOL


@pytest.mark.parametrize('propagator', ['uppe', 'paraxial'])
def test_volume_propagate_and_report(tmp_path, propagator):
    '''The volume's wave branches (volume.py:598-613, 627-641) through the HIP path: the files hold track's
    arrays.'''
    from searay_amd import volume
    if propagator == 'uppe':
        vol = configs.radial_kerr_air(scale=8, steps=2, precision='fp32')
        names = volume.UPPE_OUTPUTS
    else:
        vol = configs.paraxial_air_filament(scale=8, planes=3, length='2 mm', precision='fp32')
        names = volume.PARAXIAL_OUTPUTS
    _say('volume propagate + report (%s, fp32)' % propagator)
    box = vol['object']
    ret = box.Propagate(None, None, None)
    base = str(tmp_path / 'run')
    box.Report(base, configs.MKS_LENGTH)
    for suffix, ref in zip(names, ret):
        assert np.array_equal(np.load(base + '_' + box.name + '_' + suffix + '.npy'), ref)
    assert np.all(np.isfinite(ret[0]))


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_ray_boundary_track(oracle_mod, precision):
    '''track() driven by rays on the entrance plane (no 'incoming wave'), through the HIP stepper: the same run as
    the oracle's from the wave itself, and the relaunched rays carry the exit field (caustic_tools.py:237-254,
    309-338).'''
    from test_ray_coupling import radial_rays
    vol = configs.radial_kerr_air(scale=8, steps=3)
    _say('ray boundary radial_kerr scale=8 steps=3 %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    xp, eik, vg = radial_rays(vol)
    xp_o, eik_o, vg_o = xp.copy(), eik.copy(), vg.copy()
    vol_rays = dict(vol)
    del vol_rays['incoming wave']
    vol_rays['uppe precision'] = precision
    got = uppe_kernel.track(None, xp, eik, vg, vol_rays)
    oracle_mod.oracle_track(dict(vol_rays), None, xp_o, eik_o, vg_o)
    tol = FP64_TOL if precision == 'fp64' else FP32_FIELD_TOL
    assert rel_l2(got[0][..., 1:], ref[0][..., 1:]) < tol
    live = xp[:, 0, 4] > 0
    assert rel_l2(eik[live, 1], eik_o[live, 1]) < tol
    assert np.array_equal(xp[live], xp_o[live])


# ---- paraxial envelope (paraxial_kernel.py, paraxial.cl) ----------------------------------------------------

@pytest.mark.parametrize('coords', ['cartesian', 'cylindrical'])
@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_paraxial_gaussian_reference(coords, precision):
    """The reference's paraxial Gaussian-beam test (test/gaussian_test.py:63-108) through the HIP library: the
    lineout at the central frequency against a00 exp(-x^2 / (f r00^2)) / sqrt(f), f = 5 and 17, with the
    reference's tolerances (rtol 0.02, atol 0.01)."""
    vol, p = configs.gaussian_paraxial(coords, n=128, nw=16)
    _say('paraxial gaussian %s %s' % (coords, precision))
    A, _, _, _, dom4d, _ = _run(vol, precision)
    c = A.shape[0] // 2
    dx = (dom4d[3] - dom4d[2]) / A.shape[1]
    x = dom4d[2] + dx * (np.arange(A.shape[1]) + 0.5)
    if coords == 'cartesian':
        dy = (dom4d[5] - dom4d[4]) / A.shape[2]
        y = dom4d[4] + dy * 64.5
        lineout = lambda k: np.abs(A[c, :, 64, k])
    else:
        y = 0.0
        lineout = lambda k: np.abs(A[c, :, 0, k])
    for plane, f in ((1, 5), (2, 17)):
        expected = p['a00'] * np.exp(-y ** 2 / p['r00'] ** 2 / f) * np.exp(-x ** 2 / p['r00'] ** 2 / f) / f ** 0.5
        assert np.allclose(lineout(plane), expected, .02, .01)


@pytest.mark.parametrize('coords', ['cylindrical', 'cartesian'])
@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_paraxial_filament(oracle_mod, coords, precision):
    """The focused air filament (Kerr, cycle-averaged O2 rate, the 32-state rotator) against the oracle: FP64
    paths 1e-10 with the same step schedule; FP32 / BF16x3 on their own adaptive schedule 1e-4 (field) and
    1e-3 (electron density)."""
    scale = 4 if coords == 'cylindrical' else 16
    vol = configs.paraxial_air_filament(scale=scale, planes=2, length='4 mm', power=50.0 if scale == 4 else 200.0,
                                        coords=coords, rotator_states=32)
    _say('paraxial filament %s %s' % (coords, precision))
    rec_ref = []
    ref = oracle_mod.oracle_track(vol, record=rec_ref)
    rec = []
    got = _run(vol, precision, rec)
    if precision.startswith('fp64'):
        assert len(rec) == len(rec_ref)
        assert np.allclose(rec, rec_ref, rtol=1e-8, atol=0)
        tol, tol_ne = FP64_TOL, FP64_TOL
    else:
        tol, tol_ne = 1e-4, 1e-3
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < tol
    assert rel_l2(got[2][..., 1], ref[2][..., 1]) < tol_ne
    assert rel_l2(got[3][..., 1], ref[3][..., 1]) < tol_ne


@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_radial_ionization_8192_samples(oracle_mod, precision):
    '''Strips of Nt = 8192 samples (Nw = 4097: the photoionisation workload with its band doubled at the same
    frequency step), where the source kernel takes eight samples per thread and reads its twiddles from L2.'''
    vol = configs.radial_o2_ionization(nr=16, steps=2, length='1 mm')
    vol['wave grid'] = (4097,) + tuple(vol['wave grid'][1:])
    vol['frequency band'] = (0.0, 2 * vol['frequency band'][1])
    _say('radial_ionization Nt=8192 nr=16 %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    got = _run(vol, precision)
    fp64 = precision.startswith('fp64')
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < (FP64_TOL if fp64 else FP32_FIELD_TOL)
    assert rel_l2(got[3][..., 1], ref[3][..., 1]) < (FP64_TOL if fp64 else FP32_NE_TOL)
    assert got[5] == ref[5]


@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_radial_fewer_modes_than_cells(oracle_mod, precision):
    '''Nk = 40 radial modes on 64 cells (the reference zeroes the operator rows of the unused modes and gives them
    kr^2 = 0, grid_tools.py:268-272,313): contraction lengths and row counts that are not multiples of the MFMA
    tiles (40 = 2 x 16 + 8, padded to the 32-deep BF16 chunks).'''
    vol = configs.radial_o2_ionization(scale=4, nr=64, steps=2, length='2 mm')
    vol['radial modes'] = 40
    _say('radial_ionization nr=64 nk=40 %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    got = _run(vol, precision)
    fp64 = precision.startswith('fp64')
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < (FP64_TOL if fp64 else FP32_FIELD_TOL)
    assert rel_l2(got[3][..., 1], ref[3][..., 1]) < (FP64_TOL if fp64 else FP32_NE_TOL)
    assert got[5] == ref[5]


@pytest.mark.parametrize('max_steps', [0, 1])
@pytest.mark.parametrize('precision', ['fp64', 'bf16x3'])
def test_radial_fixed_then_adaptive(oracle_mod, precision, max_steps):
    '''A fixed-step advance call (dzmin >= the subcycle length: one step per subcycle), then an adaptive call on
    the same handle, on the HIP stepper and on the oracle. The adaptive step estimate takes amin from max|q|,
    which fixed steps form only on the step that ends their call: its last subcycle (max_steps = 0) or the step
    that reaches max_steps. A stale max|q| moves the adaptive schedule: FP64 must keep the oracle's to 1e-8 and
    its state to 1e-10; BF16x3 its state within the adaptive FP32 tolerance.'''
    vol = configs.radial_o2_ionization(scale=4, steps=4, rate='explicit', length='1 mm', power=3.0)
    _say('radial_ionization fixed then adaptive %s max_steps=%d' % (precision, max_steps))
    from searay_amd import uppe_kernel as uk
    rs = uk.RunSetup(vol)
    zi, zf, L = rs.plane(0)
    n_fixed = 1 if max_steps else 2
    runs = []
    for factory in (oracle_mod.oracle_factory, uk.hip_stepper_factory(precision)):
        st = rs.make_stepper(factory)
        st.set_field(rs.A0)
        n1, _ = st.advance(zi, L, 0, 2, L, max_steps=max_steps)
        n2, rec = st.advance(zi, L, n_fixed, rs.subcycles - n_fixed, rs.dzmin, record_cap=4096)
        runs.append((n1, n2, rec, st.get_state()))
        st.close()
    (n1_ref, n2_ref, rec_ref, q_ref), (n1, n2, rec, q) = runs
    assert n1 == n1_ref == n_fixed
    assert n2_ref > rs.subcycles - n_fixed  # the adaptive call subdivides its subcycles
    if precision == 'fp64':
        assert n2 == n2_ref
        assert np.allclose(rec, rec_ref, rtol=1e-8, atol=0)
        assert rel_l2(q, q_ref) < FP64_TOL
    else:
        assert rel_l2(q, q_ref) < 1e-4

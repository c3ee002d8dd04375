'''Full-size properties on the photoionisation workload (Nw = 2049, Nr = 256), where the oracle is too slow to
run to completion: agreement of the reduced-precision paths with the fused FP64 path on one step schedule,
FP64 fused against FP64 plain over the first steps, and conservation of energy in the linear regime; and the
radial Kerr workload (Nw = 513, Nr = 128) at full size against fused FP64 and the plain stepper.'''
import sys

import numpy as np
import pytest

from conftest import rel_l2
from searay_amd import configs, uppe_kernel

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope='module')
def c2():
    return uppe_kernel.RunSetup(configs.radial_o2_ionization())


def _steps(rs, precision, K, schedule=None, record=False):
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(precision))
    zi, zf, L = rs.plane(0)
    st.set_field(rs.A0)
    n, rec = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, max_steps=K, dz_schedule=schedule,
                        record_cap=K if record else 0)
    q = st.get_state()
    info = st.info()
    st.close()
    return n, np.array(rec), q, info


def test_fp64_fused_matches_plain(c2):
    _say('full size: fp64 fused vs fp64 plain, 6 steps')
    n0, rec0, q0, _ = _steps(c2, 'fp64-plain', 6, record=True)
    n1, rec1, q1, _ = _steps(c2, 'fp64', 6, record=True)
    assert n0 == n1 == 6
    assert np.allclose(rec1, rec0, rtol=1e-8, atol=0)
    assert rel_l2(q1, q0) < 1e-10


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_reduced_precision_on_fp64_schedule(c2, precision):
    '''200 steps: relative L2 of the k-space state against fused FP64 on its own schedule within 1e-5.'''
    _say('full size: %s vs fp64 on the fp64 schedule, 200 steps' % precision)
    n64, rec, q64, info64 = _steps(c2, 'fp64', 200, record=True)
    n, _, q, info = _steps(c2, precision, 200, schedule=rec)
    assert n == n64 == 200
    assert info['nan_steps'] == 0 and info64['nan_steps'] == 0
    assert rel_l2(q, q64) < 1e-5


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_radial_kerr_full_size(precision):
    '''The radial Kerr workload at full size (Nw = 513, Nr = 128, 100 fixed steps over 1 m): fused FP64 against
    the plain FP64 stepper over the first 10 steps (1e-10), and the whole run in the reduced precision against
    fused FP64 (1e-5; fixed steps give both the same schedule).'''
    rs = uppe_kernel.RunSetup(configs.radial_kerr_air())
    _say('full size radial Kerr: fp64 fused vs plain (10 steps), %s vs fp64 (100 steps)' % precision)
    n0, _, q0, _ = _steps(rs, 'fp64-plain', 10)
    n1, _, q1, _ = _steps(rs, 'fp64', 10)
    assert n0 == n1 == 10
    assert rel_l2(q1, q0) < 1e-10
    n64, _, q64, info64 = _steps(rs, 'fp64', 100)
    n, _, q, info = _steps(rs, precision, 100)
    assert n == n64 == 100
    assert info['nan_steps'] == 0 and info64['nan_steps'] == 0
    assert rel_l2(q, q64) < 1e-5


@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_linear_energy_conservation(precision):
    '''Kerr and ionisation off, lossless air: sum_w sum_r w^2 |A|^2 Lambda_r^2 after 20 steps and the finish
    (linear phase, inverse Hankel transform) equals its initial value within 1e-6 (1e-12 in FP64).'''
    vol = configs.radial_o2_ionization()
    vol['chi3'] = 0.0
    del vol['ionizer']
    rs = uppe_kernel.RunSetup(vol)
    _say('full size linear energy %s' % precision)
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(precision))
    zi, zf, L = rs.plane(0)
    st.set_field(rs.A0)
    n, _ = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, max_steps=20)
    z = st.info()['z_last']
    a, _, _, _ = st.finish(z, want_diagnostics=False)
    st.close()
    lam = rs.field_tool.GetTransverseTool().Lambda
    wgt = (rs.w[:, None, None] ** 2) * (lam[None, :, None] ** 2)
    e0 = np.sum(wgt * np.abs(rs.A0) ** 2)
    e1 = np.sum(wgt * np.abs(a) ** 2)
    assert n == 20
    assert abs(e1 / e0 - 1.0) < (1e-12 if precision == 'fp64' else 1e-6)


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
@pytest.mark.parametrize('scale,ionizing', [(1, False), (2, False), (4, False), (4, True), (8, True)])
def test_wave_source_matches_workgroup_source(precision, scale, ionizing, monkeypatch):
    '''The wave-per-strip source kernel (FP32 strips of Nt = 256 .. 1024, one wavefront each) against the workgroup
    source kernel on the same steps: the two differ only in FFT summation order and the depletion scan's tree, so
    the states agree to the FP32 field tolerance, 1e-5 relative L2, after 10 steps (Nt = 1024, 512, 256 Kerr-only;
    Nt = 1024 and 512 with photoionisation and the plasma current).'''
    if ionizing:
        vol = configs.radial_o2_ionization(scale=scale, steps=20, length='2 mm')
    else:
        vol = configs.radial_kerr_air(scale=scale, steps=20)
    rs = uppe_kernel.RunSetup(vol)
    _say('wave vs workgroup source %s grid %s ionizing=%s' % (precision, rs.N[:2], ionizing))
    monkeypatch.setenv('SEARAY_UPPE_WAVE_SOURCE', '0')
    n0, rec0, q0, i0 = _steps(rs, 'fp64', 10, record=True)
    _, _, qb, _ = _steps(rs, precision, 10, schedule=rec0)
    monkeypatch.setenv('SEARAY_UPPE_WAVE_SOURCE', '2')
    _, _, qw, iw = _steps(rs, precision, 10, schedule=rec0)
    assert iw['nan_steps'] == 0 and np.all(np.isfinite(qw))
    assert rel_l2(qw, qb) < 1e-5
    assert rel_l2(qw, q0) < 1e-5


@pytest.fixture(scope='module')
def c3():
    return uppe_kernel.RunSetup(configs.cartesian_filamentation())


def test_cartesian_full_size_fp64_fused_matches_plain(c3):
    '''The 3-D Cartesian workload at full size (Nw = 513, 256 x 256, Kerr + O2, 200 fixed steps over 10 cm): the
    fused FP64 stepper (batched LDS column FFTs) against the plain FP64 stepper (the reference's pass structure)
    after one step, 1e-10 relative L2.'''
    _say('full size Cartesian: fp64 fused vs plain, 1 step')
    n0, _, q0, _ = _steps(c3, 'fp64-plain', 1)
    n1, _, q1, _ = _steps(c3, 'fp64', 1)
    assert n0 == n1 == 1
    assert rel_l2(q1, q0) < 1e-10


def test_cartesian_full_size_fp32_matches_fp64(c3):
    '''FP32 against fused FP64 on the 3-D Cartesian workload after 3 fixed steps: 1e-5 relative L2.'''
    _say('full size Cartesian: fp32 vs fp64, 3 steps')
    n0, _, q0, i0 = _steps(c3, 'fp64', 3)
    n1, _, q1, i1 = _steps(c3, 'fp32', 3)
    assert n0 == n1 == 3
    assert i0['nan_steps'] == i1['nan_steps'] == 0
    assert rel_l2(q1, q0) < 1e-5


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_cartesian_full_size_linear_energy(precision):
    '''Kerr and ionisation off on the 3-D Cartesian grid: sum_w sum_xy w^2 |A|^2 after 20 fixed steps and the finish
    (linear phase, inverse transverse FFTs) equals its initial value within 1e-6 (1e-12 in FP64); the modes that
    the dry-air kz leaves evanescent hold no field.'''
    vol = configs.cartesian_filamentation()
    vol['chi3'] = 0.0
    del vol['ionizer']
    rs = uppe_kernel.RunSetup(vol)
    _say('full size Cartesian linear energy %s' % precision)
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(precision))
    zi, zf, L = rs.plane(0)
    st.set_field(rs.A0)
    n, _ = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, max_steps=20)
    z = st.info()['z_last']
    a, _, _, _ = st.finish(z, want_diagnostics=False)
    st.close()
    wgt = rs.w[:, None, None] ** 2
    a0 = rs.A0.copy()
    a0[:4] = 0.0  # the conditioner's bins
    e0 = np.sum(wgt * np.abs(a0) ** 2)
    e1 = np.sum(wgt * np.abs(a) ** 2)
    assert n == 20
    assert abs(e1 / e0 - 1.0) < (1e-12 if precision == 'fp64' else 1e-6)


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
@pytest.mark.parametrize('scale', [1, 2])
def test_long_strip_transforms(precision, scale):
    '''The FP32 source kernel's strip transforms at Nt = 4096 and 2048 on the photoionisation workload: against
    fused FP64 on its step schedule after 20 steps, 1e-5 relative L2, with the electron density growing (Kerr, rate,
    depletion and plasma current all on).'''
    vol = configs.radial_o2_ionization(scale=scale)
    rs = uppe_kernel.RunSetup(vol)
    _say('long strip transforms %s grid %s (Nt = %d)' % (precision, rs.N[:2], 2 * (rs.N[0] - 1)))
    n64, rec, q64, i64 = _steps(rs, 'fp64', 20, record=True)
    n, _, q, info = _steps(rs, precision, 20, schedule=rec)
    assert n == n64 == 20
    assert info['nan_steps'] == 0
    assert rel_l2(q, q64) < 1e-5

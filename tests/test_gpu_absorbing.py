'''Reduced-precision runs through an absorbing medium: the UV absorption band of the photoionisation workload
(air with an opacity region from 10 nm to 300 nm, Im kz up to 1.56e-4 per unit length) puts exp(Im kz z) of the
right-hand side past the FP32 range after ~6.5 cm. The FP32 state keeps its nonlinear-band columns relative to a
reference point that moves with the run.

Tolerances are those of the photoionisation workload: FP64 against the oracle 1e-10 relative L2, FP32 and BF16x3
1e-5 on the field and 1e-4 on the current, with the same number of right-hand-side evaluations as the oracle
(the adaptive schedule is formed from the interaction-picture state in FP64).
'''
import sys

import numpy as np
import pytest

from conftest import rel_l2
from searay_amd import configs, uppe_kernel

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _narrow(density_step=False):
    '''8 cm in 4 subcycles on a 2049 x 8 grid: max Im kz z = 98, Im kz L = 24.5.'''
    vol = configs.radial_o2_ionization(nr=8, steps=4, uv_opacity=True, length='8 cm')
    if density_step:
        zi, zf, _ = uppe_kernel.RunSetup(vol).plane(0)
        zmid = 0.5 * (zi + zf)
        # a quarter more gas behind the midpoint: the medium (and its Im kz) changes between subcycles 2 and 3
        vol['density lambda'] = lambda x, y, z, r2: np.where(z > zmid, 1.25, 1.0) * np.ones(r2.shape)
    return vol


@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_absorbing_medium_against_oracle(oracle_mod, precision):
    vol = _narrow()
    _say('absorbing medium 2049 x 8, 8 cm, %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    got = uppe_kernel.track_with(vol, uppe_kernel.hip_stepper_factory(precision))
    fp64 = precision == 'fp64'
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < (1e-10 if fp64 else 1e-5)
    assert rel_l2(got[1][..., 1], ref[1][..., 1]) < (1e-10 if fp64 else 1e-4)
    assert got[5] == ref[5]


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_absorbing_medium_density_step_against_oracle(oracle_mod, precision):
    '''The density steps up at the midpoint: the state, kept relative to its reference point under the first
    medium's Im kz, moves to the second medium's at the change (the reference keeps q across subcycles).'''
    vol = _narrow(density_step=True)
    rs = uppe_kernel.RunSetup(vol)
    assert not rs.medium_is_uniform(0)
    _say('absorbing medium with a density step 2049 x 8, 8 cm, %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    got = uppe_kernel.track_with(vol, uppe_kernel.hip_stepper_factory(precision))
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < 1e-5
    assert rel_l2(got[1][..., 1], ref[1][..., 1]) < 1e-4
    assert got[5] == ref[5]


def _interval(rs, precision, schedule=None):
    '''The whole interval in one advance (the density is uniform): steps, dz record, k-space state, field, current
    and counters.'''
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(precision))
    zi, zf, L = rs.plane(0)
    st.set_field(rs.A0)
    n, rec = st.advance(zi, L, 0, rs.subcycles, rs.dzmin, dz_schedule=schedule, record_cap=1 << 16)
    q = st.get_state()
    a, j, _, _ = st.finish(zf - zi)
    info = st.info()
    st.close()
    return n, rec, q, a, j, info


@pytest.mark.parametrize('precision', ['fp32', 'bf16x3'])
def test_absorbing_medium_full_size(precision):
    '''The photoionisation workload with the UV band at full size (2049 x 256, 20 cm in 400 subcycles, max Im kz
    z = 245) at 1 P_cr, against fused FP64: on the FP64 step schedule, the interaction-picture state (whose
    absorbing columns carry e^{Im kz z}, |q| up to 5e82), the field and the current at 20 cm; with its own adaptive
    steps, the FP64 schedule to 1e-4. (At the workload's 3 P_cr the collapse near 1.6 cm drives FP64 itself,
    fused and plain alike, to overflow with the UV band on this grid.)'''
    rs = uppe_kernel.RunSetup(configs.radial_o2_ionization(power=1.0, uv_opacity=True))
    _say('absorbing medium full size 2049 x 256, 20 cm: fp64, then %s on its schedule, then adaptive' % precision)
    n64, rec, q64, a64, j64, i64 = _interval(rs, 'fp64')
    assert i64['nan_steps'] == 0
    assert np.abs(q64).max() > 1e60  # far outside the FP32 range without the moving reference point
    n, _, q, a, j, info = _interval(rs, precision, schedule=rec)
    assert n == n64
    assert info['nan_steps'] == 0
    assert rel_l2(q, q64) < 1e-5
    assert rel_l2(a, a64) < 1e-5
    assert rel_l2(j, j64) < 1e-4
    na, reca, _, _, _, _ = _interval(rs, precision)
    _say('steps: fp64 %d, %s adaptive %d' % (n64, precision, na))
    assert na == n64
    assert np.allclose(reca, rec, rtol=1e-4, atol=0)


def test_fp32_refuses_subcycle_beyond_range():
    '''One 8 cm subcycle: Im kz L = 98 exceeds what the moving reference point covers (40); FP32 refuses.'''
    vol = configs.radial_o2_ionization(nr=8, steps=1, uv_opacity=True, length='8 cm')
    with pytest.raises(Exception, match='FP32 range'):
        uppe_kernel.track_with(vol, uppe_kernel.hip_stepper_factory('fp32'))

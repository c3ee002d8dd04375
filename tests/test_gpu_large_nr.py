'''Radial grids beyond 512 cells. The contraction then splits its rows into blocks of at most 512 (two 16-row
tiles on each of 16 waves), and stages its B slab in chunks of the contraction index where the whole slab would
not fit in LDS: FP64 at 1024 cells (chunks of 512 rows), every precision at 2048 (FP32 1152 or 1280 rows, FP64
512).

The workload is the photoionisation workload at a quarter of its band (Nw = 513, the carrier inside it), 1 mm in
two subcycles, against the oracle; tolerances as for the other parity tests (FP64 1e-10, FP32 and BF16x3 1e-5 on
the field and 1e-4 on the electron density, the same number of right-hand-side evaluations).
'''
import sys

import pytest

from conftest import rel_l2
from searay_amd import _abi, configs, uppe_kernel

pytestmark = pytest.mark.gpu

FP64_TOL = 1e-10
FP32_FIELD_TOL = 1e-5
FP32_NE_TOL = 1e-4

_refs = {}


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _vol(nr):
    return configs.radial_o2_ionization(scale=4, nr=nr, steps=2, length='1 mm')


def _oracle(oracle_mod, nr):
    '''The oracle's run of _vol(nr), once per module (20-80 s of host time at these sizes).'''
    if nr not in _refs:
        _say('oracle 513 x %d' % nr)
        _refs[nr] = oracle_mod.oracle_track(_vol(nr))
    return _refs[nr]


@pytest.mark.parametrize('nr', [1024, 2048])
@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_large_radial_grid_against_oracle(oracle_mod, nr, precision):
    ref = _oracle(oracle_mod, nr)
    _say('large radial grid 513 x %d, %s' % (nr, precision))
    got = uppe_kernel.track_with(_vol(nr), uppe_kernel.hip_stepper_factory(precision))
    fp64 = precision == 'fp64'
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < (FP64_TOL if fp64 else FP32_FIELD_TOL)
    assert rel_l2(got[3][..., 1], ref[3][..., 1]) < (FP64_TOL if fp64 else FP32_NE_TOL)
    assert got[5] == ref[5]


def test_radial_grid_beyond_limit_is_refused():
    '''More cells than the fused path takes: the handle is refused with UPPE_ERR_UNSUPPORTED, not mis-stepped.'''
    with pytest.raises(_abi.UppeError, match='create failed with status %d' % _abi.UPPE_ERR_UNSUPPORTED):
        uppe_kernel.track_with(_vol(4096), uppe_kernel.hip_stepper_factory('fp32'))

'''Cartesian lines of 1024 points: the transverse pass then takes 8 FP32 columns per line (a block of 1024
threads) instead of 16. The filamentation workload (Kerr, O2 photoionisation, plasma current) at Nw = 65 on a
1024 x 32 and a 32 x 1024 grid, 1 mm in two subcycles, against the oracle; tolerances as for the other parity
tests (FP64 1e-10, FP32 1e-5 on the field and 1e-4 on the electron density).
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


def _vol(nx, ny):
    vol = configs.cartesian_filamentation(scale=8, steps=2, length='1 mm')
    vol['wave grid'] = (vol['wave grid'][0], nx, ny) + tuple(vol['wave grid'][3:])
    return vol


def _oracle(oracle_mod, shape):
    if shape not in _refs:
        _say('oracle 65 x %d x %d' % shape)
        _refs[shape] = oracle_mod.oracle_track(_vol(*shape))
    return _refs[shape]


@pytest.mark.parametrize('shape', [(1024, 32), (32, 1024)])
@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_long_transverse_lines_against_oracle(oracle_mod, shape, precision):
    ref = _oracle(oracle_mod, shape)
    _say('Cartesian 65 x %d x %d, %s' % (shape + (precision,)))
    got = uppe_kernel.track_with(_vol(*shape), uppe_kernel.hip_stepper_factory(precision))
    fp64 = precision == 'fp64'
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < (FP64_TOL if fp64 else FP32_FIELD_TOL)
    assert rel_l2(got[3][..., 1], ref[3][..., 1]) < (FP64_TOL if fp64 else FP32_NE_TOL)
    assert got[5] == ref[5]


def test_transverse_line_beyond_limit_is_refused():
    with pytest.raises(_abi.UppeError, match='create failed with status %d' % _abi.UPPE_ERR_UNSUPPORTED):
        uppe_kernel.track_with(_vol(2048, 8), uppe_kernel.hip_stepper_factory('fp32'))

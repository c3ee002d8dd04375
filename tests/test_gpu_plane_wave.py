'''The HIP stepper's UPPE right-hand side against the closed-form plane-wave solutions of tests/plane_wave.py.

The same checks as the oracle's TestPlaneWaveRHS, through the C ABI: self-phase modulation of a monochromatic
plane wave in BK7 (kz differs from w by 50 %), transparent and absorbing, in the Cartesian geometry (transverse
FFT path) and in the radial geometry with an identity basis (the matrix-core contraction path). The third
harmonic is compared with the oracle run on the same input. Tolerances (relative to |q0|): FP64 1e-10;
FP32 and BF16x3 1e-5, the FP32 field tolerance (the phase of one radian is accumulated over 400 steps).
'''
import sys

import numpy as np
import pytest

import plane_wave as pw
from conftest import rel_l2
from searay_amd import uppe_kernel

pytestmark = pytest.mark.gpu

BAND = (6, 11)
TOL = {'fp64': 1e-10, 'fp32': 1e-5, 'bf16x3': 1e-5}
CASES = [('cartesian', 'fp64'), ('cartesian', 'fp32'), ('radial', 'fp64'), ('radial', 'fp32'), ('radial', 'bf16x3')]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize('absorb', [0.0, 3.8e-3])
@pytest.mark.parametrize('geometry,precision', CASES)
def test_self_phase(geometry, precision, absorb):
    _say('plane-wave self-phase %s %s absorb=%g' % (geometry, precision, absorb))
    kz = pw.bk7_kz(absorb)
    st, shape, nrm = pw.make(uppe_kernel.hip_stepper_factory(precision), geometry, kz, BAND)
    c = pw.spm_rate(kz[pw.M], nrm)
    L = 400.0
    q0 = np.sqrt(1.0 / (c * L)) * np.exp(0.3j)
    q = pw.run(st, pw.field(shape, q0 / nrm), L, 400)
    st.close()
    expect = pw.spm_solution(c, kz[pw.M], q0, L)
    tol = TOL[precision]
    cells = q[pw.M].reshape(-1)[:1] if geometry == 'cartesian' else q[pw.M].reshape(-1)
    for got in cells:
        assert abs(got - expect) < tol * abs(q0), (got, expect)
    others = np.delete(q.reshape(pw.NW, -1), pw.M, axis=0)
    assert np.abs(others).max() < (1e-12 if precision == 'fp64' else 1e-6) * abs(q0)


@pytest.mark.parametrize('precision', ['fp64', 'fp32'])
def test_third_harmonic_against_oracle(oracle_mod, precision):
    _say('plane-wave third harmonic %s' % precision)
    kz = pw.bk7_kz()
    out = []
    for factory in (oracle_mod.oracle_factory, uppe_kernel.hip_stepper_factory(precision)):
        st, shape, nrm = pw.make(factory, 'cartesian', kz, (0, pw.NW))
        c = pw.spm_rate(kz[pw.M], nrm)
        q0 = np.sqrt(0.1 / (c * 50.0)) * np.exp(0.3j)
        out.append(pw.run(st, pw.field(shape, q0 / nrm), 50.0, 400))
        st.close()
    ref, got = out
    h = 3 * pw.M
    assert abs(ref[h, 0, 0]) > 1e-3 * abs(ref[pw.M, 0, 0])
    assert rel_l2(got[h], ref[h]) < (1e-9 if precision == 'fp64' else 1e-4)
    assert rel_l2(got, ref) < (1e-10 if precision == 'fp64' else 1e-5)

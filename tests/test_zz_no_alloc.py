'''uppe_advance_static (the advance that allocates nothing) against uppe_advance, through the C ABI.

Both entry points launch the same kernels in the same order on the same inputs, so their states agree bit for
bit; the test asks for exact equality. Two calls in a row share the static workspace, the second on a smaller
grid than the first, so it also checks that a call does not read what the previous one left there. A request
larger than the workspace is refused with UPPE_ERR_ARG before any device work.
'''
import sys

import numpy as np
import pytest

from searay_amd import _abi, configs, uppe_kernel

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _run(vol, precision, static, plane_steps=None, record=False):
    rs = uppe_kernel.RunSetup(vol)
    st = rs.make_stepper(uppe_kernel.hip_stepper_factory(precision))
    st.set_field(rs.stepper_field())
    zi, zf, L = rs.plane(0)
    adv = st.advance_static if static else st.advance
    n, rec = adv(zi, L, 0, rs.subcycles, rs.dzmin, record_cap=64 if record else 0,
                 max_steps=plane_steps or 0)
    q = st.get_state()
    info = st.info()
    st.close()
    return n, rec, q, info


@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_static_advance_equals_advance(precision):
    # call 1: ionising radial workload at 257 x 32 with adaptive steps and a dz record
    big = configs.radial_o2_ionization(scale=8, steps=3, rate='explicit', length='0.5 mm')
    big['minimum step'] = big['minimum step'] / 20.0
    # call 2: a smaller Kerr workload (65 x 16) with fixed steps
    small = configs.radial_kerr_air(scale=8, steps=4)
    for name, vol, rec in (('ionising 257x32 adaptive', big, True), ('kerr 65x16 fixed', small, False)):
        _say('advance_static %s %s' % (precision, name))
        n0, r0, q0, i0 = _run(vol, precision, static=False, record=rec)
        n1, r1, q1, i1 = _run(vol, precision, static=True, record=rec)
        assert n0 == n1 and n0 > 0
        assert i0['steps'] == i1['steps'] and i0['nan_steps'] == i1['nan_steps'] == 0
        assert np.all(np.isfinite(q1))
        assert np.array_equal(q0, q1), (name, np.abs(q0 - q1).max())
        if rec:
            assert np.array_equal(r0, r1)


def test_static_advance_refuses_large_grid():
    # FP64, Nw = 2049, Nr = 512: five work arrays of 512 x 2080 x 16 bytes exceed the 48 MiB workspace
    from searay_amd import _native
    st = _native.HipStepper(_abi.UPPE_RADIAL, 2049, 512, 1, 512, _abi.PREC_FP64)
    try:
        with pytest.raises(_abi.UppeError, match='static workspace'):
            st.advance_static(0.0, 1.0, 0, 1, 1.0)
    finally:
        st.close()

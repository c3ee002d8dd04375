'''Radial grids with azimuthal points (N2 = 4, modes m = 0, 1, 2, -1) on the fused path against the FP64 oracle.

The field gets an azimuthal profile 1 + 0.3 cos(phi) + 0.2 sin(2 phi), so modes 0, +-1 and 2 all carry energy; each
mode has its own radial basis and kz, and the source acts pointwise in phi (grid_tools.py:266-325, fft.cl:401-477).
Tolerances as for the single-mode workloads: FP64 1e-10, FP32 and BF16x3 1e-5 on the field and 1e-4 on the current,
with the oracle's number of right-hand-side evaluations.
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


def with_azimuth(vol, n2=4):
    nw, nr, _, nz = vol['wave grid']
    vol['wave grid'] = (nw, nr, n2, nz)
    base = vol['incoming wave']

    def incoming(w, r, phi):
        return base(w, r, phi) * (1.0 + 0.3 * np.cos(phi) + 0.2 * np.sin(2 * phi))[None, None, :]
    vol['incoming wave'] = incoming
    return vol


def _check(ref, got, precision):
    fp64 = precision.startswith('fp64')
    assert rel_l2(got[0][..., 1], ref[0][..., 1]) < (1e-10 if fp64 else 1e-5)
    assert rel_l2(got[1][..., 1], ref[1][..., 1]) < (1e-10 if fp64 else 1e-4)
    assert got[5] == ref[5]


@pytest.mark.parametrize('precision', ['fp64-plain', 'fp64', 'fp32', 'bf16x3'])
def test_azimuthal_kerr(oracle_mod, precision):
    vol = with_azimuth(configs.radial_kerr_air(scale=4, steps=6))
    _say('radial Kerr with 4 azimuthal points %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    got = uppe_kernel.track_with(vol, uppe_kernel.hip_stepper_factory(precision))
    # the azimuthal profile survives: the phi-dependence of the output is not lost in the tolerance
    a = ref[0][..., 1]
    assert np.linalg.norm(a - a.mean(axis=2, keepdims=True)) > 0.1 * np.linalg.norm(a)
    _check(ref, got, precision)


@pytest.mark.parametrize('precision', ['fp64', 'fp32', 'bf16x3'])
def test_azimuthal_ionization(oracle_mod, precision):
    vol = with_azimuth(configs.radial_o2_ionization(scale=8, steps=4, length='2 mm'))
    _say('radial O2 ionisation with 4 azimuthal points %s' % precision)
    ref = oracle_mod.oracle_track(vol)
    got = uppe_kernel.track_with(vol, uppe_kernel.hip_stepper_factory(precision))
    _check(ref, got, precision)


def _coeffs(seed=5):
    return 0.05 * np.random.default_rng(seed).standard_normal(32)


def test_azimuthal_fp64_gradient_against_central_differences():
    '''The adjoint through the azimuthal modes (the transposed per-mode contractions and the adjoints of the phi
    transforms): FP64 against central differences of the FP64 forward, on the fluence objective (which sums over
    every radial cell of phi = 0) and the on-axis spectrum.'''
    from searay_amd import adjoint
    for objective in ('spectrum', 'fluence'):
        vol = with_azimuth(configs.radial_kerr_air(scale=8, steps=4))
        _say('adjoint with 4 azimuthal points, fp64 vs central FD, %s' % objective)
        p = adjoint.SpectralPhaseProblem(vol, configs.kerr_pulse_sigma(), precision='fp64', checkpoint_every=2,
                                         objective=objective)
        c = _coeffs()
        loss, g = p.value_and_grad(c)
        fd = np.array([(p.loss(c + 1e-4 * e) - p.loss(c - 1e-4 * e)) / 2e-4 for e in np.eye(32)])
        p.close()
        _say('  gradient vs FD %.2e' % rel_l2(g, fd))
        assert loss > 0 and np.linalg.norm(g) > 0
        assert rel_l2(g, fd) < 1e-6


@pytest.mark.parametrize('precision,tol', [('fp32', 1e-4), ('bf16x3', 1e-3)])
def test_azimuthal_reduced_precision_gradient(precision, tol):
    from searay_amd import adjoint
    vol = with_azimuth(configs.radial_kerr_air(scale=4, steps=10))
    _say('adjoint with 4 azimuthal points, %s vs fp64' % precision)
    sig = configs.kerr_pulse_sigma()
    c = _coeffs(6)
    p64 = adjoint.SpectralPhaseProblem(vol, sig, precision='fp64', checkpoint_every=5)
    _, g64 = p64.value_and_grad(c)
    p = adjoint.SpectralPhaseProblem(vol, sig, precision=precision, checkpoint_every=5, target=p64.target)
    _, g = p.value_and_grad(c)
    p64.close()
    p.close()
    _say('  %s gradient vs fp64 adjoint %.2e' % (precision, rel_l2(g, g64)))
    assert rel_l2(g, g64) < tol

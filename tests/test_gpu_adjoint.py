'''Adjoint gradients of the spectral-phase objective (searay_amd.adjoint) through the HIP library.

The FP64 adjoint is exact for the discrete forward map, so it must match central finite differences of the FP64
forward to their truncation error (h = 1e-4: ~1e-8 relative, tested at 1e-6), with Kerr only, with the rotational
Raman susceptibility, and through the photoionisation rate (explicit and series), depletion integral, soft window
and plasma current, for the on-axis
spectrum and the radial fluence objectives. FP32 and BF16x3 gradients are
compared with the FP64 adjoint (1e-4 and 1e-3 relative L2). Replays from checkpoints are deterministic, so the
gradient does not depend on the checkpoint interval.
'''
import sys

import numpy as np
import pytest

from conftest import rel_l2
from searay_amd import adjoint, configs

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _problem(scale, steps, precision, every, target=None):
    vol = configs.radial_kerr_air(scale=scale, steps=steps)
    return adjoint.SpectralPhaseProblem(vol, configs.kerr_pulse_sigma(), precision=precision,
                                        checkpoint_every=every, target=target)


def _coeffs(seed=3):
    return 0.05 * np.random.default_rng(seed).standard_normal(32)


def test_fp64_gradient_against_central_differences():
    _say('adjoint fp64 vs central FD, radial Kerr scale=8 steps=4')
    p = _problem(8, 4, 'fp64', 2)
    c = _coeffs()
    loss, g = p.value_and_grad(c)
    fd = np.array([(p.loss(c + 1e-4 * e) - p.loss(c - 1e-4 * e)) / 2e-4 for e in np.eye(32)])
    p.close()
    assert loss > 0 and np.linalg.norm(g) > 0
    assert rel_l2(g, fd) < 1e-6


def test_fp64_directional_derivative_with_segment_replay():
    '''20 steps in segments of 3 (the last one partial): every segment but the last is replayed.'''
    _say('adjoint fp64 directional derivative, scale=4 steps=20 every=3')
    p = _problem(4, 20, 'fp64', 3)
    c = _coeffs(4)
    _, g = p.value_and_grad(c)
    d = np.random.default_rng(5).standard_normal(32)
    d /= np.linalg.norm(d)
    fd = (p.loss(c + 1e-4 * d) - p.loss(c - 1e-4 * d)) / 2e-4
    p.close()
    assert abs(fd - g @ d) < 1e-6 * abs(fd)


def test_gradient_independent_of_checkpoint_interval():
    _say('adjoint fp32, checkpoint intervals 1, 7, 20')
    c = _coeffs(6)
    grads = []
    target = None
    for every in (1, 7, 20):
        p = _problem(4, 20, 'fp32', every, target)
        target = p.target
        grads.append(p.value_and_grad(c)[1])
        p.close()
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


@pytest.mark.parametrize('precision,tol', [('fp32', 1e-4), ('bf16x3', 1e-3)])
def test_reduced_precision_gradient_at_full_size(precision, tol):
    '''The radial Kerr workload at full size (Nw = 513, Nr = 128, 100 steps), checkpoints every 10 steps.'''
    _say('adjoint %s vs fp64 adjoint, radial Kerr 513 x 128, 100 steps' % precision)
    c = _coeffs(7)
    p64 = _problem(1, None, 'fp64', 10)
    _, g64 = p64.value_and_grad(c)
    p = _problem(1, None, precision, 10, p64.target)
    _, g = p.value_and_grad(c)
    p64.close()
    p.close()
    assert rel_l2(g, g64) < tol


def test_autograd_function():
    import torch
    _say('adjoint through torch.autograd.Function')
    p = _problem(8, 4, 'fp32', 2)
    c = _coeffs(8)
    loss_ref, g_ref = p.value_and_grad(c)
    fn = adjoint.spectral_phase_loss_function()
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    loss = fn.apply(ct, p)
    (3.0 * loss).backward()
    p.close()
    assert abs(float(loss) - loss_ref) <= 1e-12 * abs(loss_ref)
    assert np.allclose(ct.grad.numpy(), 3.0 * g_ref, rtol=1e-12, atol=0)


def _ionizing(precision, every, steps=4, scale=4, nr=16, rate='explicit', objective='spectrum', target=None,
              length='0.5 mm', rotator=False):
    '''The radial O2 photoionisation workload (Kerr, PPT rate, depletion, plasma current) at reduced size with fixed
    steps (the adjoint holds the step sizes constant), with the N2 rotator on request.'''
    vol = configs.radial_o2_ionization(scale=scale, steps=steps, rate=rate, length=length, nr=nr)
    if rotator:
        vol['rotator'] = configs.n2_rotator(32)
    vol['frequency band'] = (0.0, 2.0)
    vol['nonlinear band'] = (0.0, 2.0)
    L = vol['size'][2]
    vol['minimum step'] = L / steps
    return adjoint.SpectralPhaseProblem(vol, configs.ionizing_pulse_sigma(), precision=precision,
                                        checkpoint_every=every, objective=objective, target=target)


@pytest.mark.parametrize('rate', ['explicit', 'series'])
@pytest.mark.parametrize('objective', ['spectrum', 'fluence'])
def test_fp64_ionizing_gradient_against_central_differences(rate, objective):
    '''Through the photoionisation rate, the depletion integral, the soft window and the plasma current: the
    FP64 adjoint against central differences of the FP64 forward (h = 1e-4) on 8 random coefficients.'''
    _say('adjoint fp64 with ionization (%s rate, %s objective) vs central FD' % (rate, objective))
    p = _ionizing('fp64', 2, rate=rate, objective=objective)
    c = _coeffs(9)
    loss, g = p.value_and_grad(c)
    idx = np.random.default_rng(10).choice(32, 8, replace=False)
    fd = np.array([(p.loss(c + 1e-4 * e) - p.loss(c - 1e-4 * e)) / 2e-4 for e in np.eye(32)[idx]])
    # the plasma term contributes: the gradient of the Kerr-only problem differs
    ne_free = _ionizing('fp64', 2, rate=rate, objective=objective, target=p.target)
    ne_free.st.set_medium(ne_free.rs.w, ne_free.rs.mat.kz, ne_free.rs.mat.kg, ne_free.rs.mat.ngas,
                          ne_free.rs.nl_filter, ne_free.rs.cond_filter, ne_free.rs.mat.redchi3, ne_free.rs.dt,
                          dict(ne_free.rs.ion, kind=0))
    _, g_kerr = ne_free.value_and_grad(c)
    ne_free.close()
    p.close()
    _say('  gradient vs FD %.2e, Kerr-only gradient vs full %.2e' % (rel_l2(g[idx], fd), rel_l2(g_kerr, g)))
    assert loss > 0 and np.linalg.norm(g) > 0
    assert rel_l2(g[idx], fd) < 1e-6, rel_l2(g[idx], fd)
    assert rel_l2(g_kerr, g) > 1e-3, rel_l2(g_kerr, g)


@pytest.mark.parametrize('precision,tol', [('fp32', 1e-4), ('bf16x3', 1e-3)])
def test_reduced_precision_ionizing_gradient(precision, tol):
    _say('adjoint %s vs fp64 adjoint with ionization, 1025 x 32, 20 steps' % precision)
    c = _coeffs(11)
    p64 = _ionizing('fp64', 5, steps=20, scale=2, nr=32)
    _, g64 = p64.value_and_grad(c)
    p = _ionizing(precision, 5, steps=20, scale=2, nr=32, target=p64.target)
    _, g = p.value_and_grad(c)
    p64.close()
    p.close()
    _say('  %s gradient vs fp64 adjoint %.2e' % (precision, rel_l2(g, g64)))
    assert rel_l2(g, g64) < tol, rel_l2(g, g64)


def _raman(precision, every, steps=4, scale=8, target=None, rotator=True):
    vol = configs.radial_raman_air(scale=scale, steps=steps)
    if not rotator:
        del vol['rotator']
    return adjoint.SpectralPhaseProblem(vol, configs.kerr_pulse_sigma(), precision=precision,
                                        checkpoint_every=every, target=target)


def test_fp64_rotator_gradient_against_central_differences():
    '''Kerr plus the 32-state N2 rotator (rotations.cl:80-129): the adjoint of the Raman susceptibility's two
    running sums, against central differences of the FP64 forward; the Kerr-only gradient of the same problem
    differs by more than 1e-3, so the rotator's terms are exercised.'''
    _say('adjoint fp64 with the rotator vs central FD, radial Raman scale=8 steps=4')
    c = _coeffs(7)
    p = _raman('fp64', 2)
    loss, g = p.value_and_grad(c)
    fd = np.array([(p.loss(c + 1e-4 * e) - p.loss(c - 1e-4 * e)) / 2e-4 for e in np.eye(32)])
    pk = _raman('fp64', 2, target=p.target, rotator=False)
    _, gk = pk.value_and_grad(c)
    p.close()
    pk.close()
    _say('  rotator gradient vs FD %.2e, vs Kerr-only gradient %.2e' % (rel_l2(g, fd), rel_l2(gk, g)))
    assert loss > 0 and np.linalg.norm(g) > 0
    assert rel_l2(g, fd) < 1e-6
    assert rel_l2(gk, g) > 1e-3


def test_fp64_rotator_and_ionization_gradient_against_central_differences():
    '''The rotator and photoionisation together (one adjoint source kernel takes both).'''
    _say('adjoint fp64 with the rotator and ionisation vs central FD')
    c = _coeffs(9)
    p = _ionizing('fp64', 2, rotator=True)
    _, g = p.value_and_grad(c)
    fd = np.array([(p.loss(c + 1e-4 * e) - p.loss(c - 1e-4 * e)) / 2e-4 for e in np.eye(32)])
    p.close()
    _say('  gradient vs FD %.2e' % rel_l2(g, fd))
    assert rel_l2(g, fd) < 1e-6


@pytest.mark.parametrize('precision,tol', [('fp32', 1e-4), ('bf16x3', 1e-3)])
def test_reduced_precision_rotator_gradient(precision, tol):
    _say('adjoint %s with the rotator vs fp64 adjoint, radial Raman scale=4 steps=10' % precision)
    c = _coeffs(8)
    p64 = _raman('fp64', 5, steps=10, scale=4)
    _, g64 = p64.value_and_grad(c)
    p = _raman(precision, 5, steps=10, scale=4, target=p64.target)
    _, g = p.value_and_grad(c)
    p64.close()
    p.close()
    _say('  %s gradient vs fp64 adjoint %.2e' % (precision, rel_l2(g, g64)))
    assert rel_l2(g, g64) < tol


def test_config5_fp32_gradient_against_central_differences_full_size():
    '''The gradient check as the workload states it: the radial Kerr workload at full size (Nw = 513, Nr = 128,
    100 steps over 1 m), the FP32 adjoint with checkpoints every 10 steps against central finite differences
    (h = 1e-4) of the FP64 forward on all 32 Legendre coefficients: 1e-4 relative L2 (the FP32 rounding of the
    forward and adjoint sweeps; the FP64 adjoint matches the same differences to 1e-7).'''
    _say('adjoint full size: fp32 (checkpoints every 10) vs central FD of the fp64 forward, 32 coefficients')
    vol = configs.radial_kerr_air()
    sig = configs.kerr_pulse_sigma()
    c = _coeffs()
    p64 = adjoint.SpectralPhaseProblem(vol, sig, precision='fp64', checkpoint_every=10)
    _, g64 = p64.value_and_grad(c)
    fd = np.array([(p64.loss(c + 1e-4 * e) - p64.loss(c - 1e-4 * e)) / 2e-4 for e in np.eye(32)])
    p32 = adjoint.SpectralPhaseProblem(vol, sig, precision='fp32', checkpoint_every=10, target=p64.target)
    _, g32 = p32.value_and_grad(c)
    p64.close()
    p32.close()
    assert rel_l2(g64, fd) < 1e-7
    assert rel_l2(g32, fd) < 1e-4

'''Host side of the spectral-phase gradient (searay_amd.adjoint), on the CPU: the Legendre basis, and the
assembly dL/dc_n = -sum P_n Im(conj(G_0) A_0) from the cotangents of the on-axis spectrum and radial fluence
objectives, checked by central differences with a linear stand-in for the stepper (finish returns M a0 per radial
cell, the adjoint M^H g); the batched workload's seeded draws.'''
import numpy as np
import pytest

from conftest import rel_l2
from searay_amd import adjoint, configs


def test_legendre_basis():
    w = np.linspace(0.0, 2.0, 201)
    B = adjoint.legendre_basis(w, 1.0, 0.125, 4)  # 3 sigma = 0.375 exactly
    x = (w - 1.0) / 0.375
    inside = np.abs(x) <= 1
    assert np.all(B[:, ~inside] == 0.0)
    assert np.allclose(B[0, inside], 1.0) and np.allclose(B[1, inside], x[inside])
    assert np.allclose(B[2, inside], 0.5 * (3 * x[inside] ** 2 - 1))


class _LinearStepper:
    def __init__(self, nw, seed=9):
        rng = np.random.default_rng(seed)
        self.M = (rng.standard_normal((nw, nw)) + 1j * rng.standard_normal((nw, nw))) / np.sqrt(nw)
        self.a0 = None

    def set_hankel(self, *a):
        pass

    def set_medium(self, *a):
        pass

    def set_rotator(self, *a):
        pass

    def set_field(self, a0):
        self.a0 = np.array(a0)

    def advance(self, *a, **k):
        return 1, None

    def advance_checkpointed(self, *a, **k):
        return 1, None

    def finish(self, z, want_diagnostics=True):
        return np.einsum('fg,grm->frm', self.M, self.a0), None, None, None

    def adjoint(self, z, g):
        return np.einsum('gf,grm->frm', np.conj(self.M), g)

    def close(self):
        pass


@pytest.mark.parametrize('objective', ['spectrum', 'fluence'])
def test_gradient_assembly_against_central_differences(objective):
    vol = configs.radial_kerr_air(scale=16, steps=2)
    nw = vol['wave grid'][0]
    p = adjoint.SpectralPhaseProblem(vol, configs.kerr_pulse_sigma(), n_phase=8, objective=objective,
                                     stepper_factory=lambda *a, **k: _LinearStepper(nw))
    c = 0.3 * np.random.default_rng(2).standard_normal(8)
    loss, g = p.value_and_grad(c)
    fd = np.array([(p.loss(c + 1e-6 * e) - p.loss(c - 1e-6 * e)) / 2e-6 for e in np.eye(8)])
    assert loss > 0 and np.linalg.norm(g) > 0
    assert rel_l2(g, fd) < 1e-6


def test_batched_workload_draws():
    '''The 64-instance workload is seeded: its grids, powers and gases are the same on every call, within the
    stated ranges, and every instance is a radial vol_dict with fixed steps.'''
    vols, drawn = configs.batched_radial(n=64)
    vols2, drawn2 = configs.batched_radial(n=64)
    assert drawn == drawn2
    assert {d[0] for d in drawn} == {64, 128, 256, 512}
    assert all(0.1 <= d[1] <= 5.0 for d in drawn) and {d[2] for d in drawn} == {'O2', 'Ar'}
    for v, d in zip(vols, drawn):
        assert v['wave grid'][:3] == (1025, d[0], 1) and v['radial modes'] == d[0]
        assert v['minimum step'] == v['size'][2] / v['subcycles']


def _ramped(vol):
    '''The same workload with a gas density that rises along z.'''
    class Ramp:
        def GetRelDensity(self, xp):
            return 1.0 + 0.1 * xp[..., 3] / vol['size'][2] + 0.0 * xp[..., 1]
    vol = dict(vol)
    vol['object'] = Ramp()
    return vol


def test_uniform_medium_detection():
    from searay_amd import uppe_kernel
    vol = configs.radial_kerr_air(scale=16, steps=4)
    assert uppe_kernel.RunSetup(vol).medium_is_uniform(0)
    assert not uppe_kernel.RunSetup(_ramped(vol)).medium_is_uniform(0)


def test_batch_and_adjoint_refuse_z_varying_density():
    '''The batch and the adjoint load each instance's medium once (at the first subcycle's midpoint); an instance
    whose density changes along z is refused before any device work, instead of being stepped in the wrong medium.'''
    from searay_amd import batch, uppe_kernel
    vol = _ramped(configs.radial_kerr_air(scale=16, steps=4))

    class Stub:
        setups = [uppe_kernel.RunSetup(configs.radial_kerr_air(scale=16, steps=4)), uppe_kernel.RunSetup(vol)]
    with pytest.raises(ValueError, match='instance 1: the density varies along z'):
        batch.Batch.advance(Stub())
    with pytest.raises(ValueError, match='density varies along z'):
        adjoint.SpectralPhaseProblem(vol, configs.kerr_pulse_sigma(), stepper_factory=lambda *a, **k: None)

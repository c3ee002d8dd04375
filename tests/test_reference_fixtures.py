'''The host rate models, series fits and air susceptibilities against vectors the reference's own modules wrote.

tests/golden/reference_rates_chi.npz holds outputs of the reference's photoionization_base.py, ionization.py and
dispersion.py (tests/golden/make_reference_golden.py, run in the build container). These are the functions no
closed form or stored value pins: the PPT and PMPB rates between the tunnelling and multiphoton limits
(photoionization_base.py:337-412, the Fourier series over photon orders with Dawson's integral), the stitched
rate (:414-426), the quintic Ionizer fits with scipy.optimize.curve_fit (ionization.py:105-109), and chi(w) of
the air models (dispersion.py:371-429) with and without the UV opacity band of examples/uppe/air_fil.py:36.
Tolerance 1e-13 relative: the restatements evaluate the same expressions in the same order (they agree bit for
bit with this scipy and numpy); the fitted coefficients are compared at 1e-10 relative, the solver's own
reproducibility on identical inputs.
'''
import os

import numpy as np
import pytest

from searay_amd import configs, dispersion, ionization
from searay_amd import photoionization_base as pb

HERE = os.path.dirname(os.path.abspath(__file__))
G = np.load(os.path.join(HERE, 'golden', 'reference_rates_chi.npz'))
MK = configs.MKS_LENGTH
GASES = (('o2', 0.53), ('ar', 1.0))
MODELS = {'adk': pb.ADK, 'ppt_tunneling': pb.PPT_Tunneling, 'ppt': pb.PPT, 'pmpb': pb.PMPB,
          'stitched': pb.StitchedPPT}


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))


@pytest.mark.parametrize('avg', [True, False])
@pytest.mark.parametrize('model', sorted(MODELS))
@pytest.mark.parametrize('gas,z', GASES)
def test_rates_match_reference(gas, z, model, avg):
    E = G['E_au']
    uion = float(G['uion_au_' + gas])
    w0 = float(G['w0_au'])
    gam = np.sqrt(2 * uion) * w0 / E
    assert gam.min() < 0.1 and gam.max() > 20  # tunnelling and multiphoton limits and everything between
    got = MODELS[model](avg, uion, z, lstar=0, l=0, m=0, w0=w0, terms=80).Rate(E.copy())
    ref = G['rate_%s_%s_%s' % (model, gas, 'avg' if avg else 'static')]
    assert np.all(ref > 0)
    assert _rel(got, ref) < 1e-13


@pytest.mark.parametrize('model', ['ppt', 'pmpb', 'stitched'])
@pytest.mark.parametrize('gas,z', GASES)
def test_ionizer_coefficients_match_reference(gas, z, model):
    cls = {'ppt': ionization.PPT, 'pmpb': ionization.PMPB, 'stitched': ionization.StitchedPPT}[model]
    ion = ionization.Ionizer(cls(MK, False, float(G['uion_sim_' + gas]), z, lstar=0, l=0, m=0, w0=1.0, terms=80))
    ref_coeff = G['coeff_%s_%s' % (model, gas)]
    assert np.array_equal(np.array([ion.C1, ion.C2, ion.C3]), ref_coeff)
    key = 'clog_%s_%s' % (model, gas)
    if model == 'stitched':
        # StitchedPPT takes PPT_Tunneling's constructor: a closed form, no fit
        assert key not in G.files and ion.Clog is None and ref_coeff[0] > 0
        return
    assert _rel(ion.Clog, G[key]) < 1e-10
    # the fitted rate itself over the fit's field range
    E = np.logspace(np.log10(2e-3), 0, 64)
    lg = np.log(E)
    ours = np.polyval(np.asarray(ion.Clog)[::-1], lg)
    theirs = np.polyval(G[key][::-1], lg)
    assert np.max(np.abs(ours - theirs)) < 1e-10 * np.max(np.abs(theirs))


@pytest.mark.parametrize('grid', ['w513', 'w2049', 'w1025'])
def test_air_chi_matches_reference(grid):
    w = G[grid]
    media = {'simpleair': dispersion.SimpleAir(MK), 'dryair': dispersion.DryAir(MK),
             'humidair': dispersion.HumidAir(MK, 0.5)}
    uv = dispersion.SimpleAir(MK)
    cm, um = 100 * MK, 1e6 * MK
    uv.add_opacity_region(0.1 / cm, 0.01 / um, 0.3 / um)
    media['simpleair_uv'] = uv
    for name, med in media.items():
        ref = G['chi_%s_%s' % (name, grid)]
        with np.errstate(divide='ignore', invalid='ignore'):
            got = np.asarray(med.chi(w), dtype=complex)
        fin = np.isfinite(ref)
        assert np.array_equal(np.isfinite(got), fin), name
        scale = np.abs(ref[fin]).max()
        assert np.max(np.abs(got[fin] - ref[fin])) < 1e-13 * scale, name
    assert np.any(G['chi_simpleair_uv_' + grid].imag > 0) == (w.max() > 2 * np.pi / (0.3 / um))


def test_window_speed_matches_reference():
    assert dispersion.SimpleAir(MK).GroupVelocityMagnitude(1.0) == pytest.approx(float(G['vg_simpleair_w1']),
                                                                                 rel=1e-14)

'''The FP64 oracle against its committed golden fixtures (tests/golden): a change of the oracle's results shows
up here before it can move the GPU parity targets.'''
import numpy as np
import pytest

from conftest import rel_l2
import golden_cases


@pytest.mark.parametrize('name', golden_cases.NAMES)
def test_oracle_reproduces_fixture(oracle_mod, name):
    make, fx = golden_cases.load(name)
    rec = []
    A, J, chi, ne, dom4d, evals = oracle_mod.oracle_track(make(), record=rec)
    assert list(evals) == list(fx['evals'])
    assert np.allclose(rec, fx['dz'], rtol=1e-12, atol=0)
    assert rel_l2(A[..., -1], fx['A']) < 1e-12
    for key, arr in (('J', J), ('chi', chi), ('ne', ne)):
        if key in fx:
            assert rel_l2(arr[..., -1], fx[key]) < 1e-12


def test_fixtures_are_small():
    import os
    total = sum(os.path.getsize(os.path.join(golden_cases.GOLDEN, n + '.npz')) for n in golden_cases.NAMES)
    assert total < 1 << 20

'''Loader for the committed golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py).'''
import importlib.util
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _maker_module():
    spec = importlib.util.spec_from_file_location('make_golden', os.path.join(GOLDEN, 'make_golden.py'))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


CASES = _maker_module().CASES
NAMES = sorted(CASES)


def load(name):
    '''(vol_dict builder, fixture dict) of one case.'''
    make, _ = CASES[name]
    with np.load(os.path.join(GOLDEN, name + '.npz')) as z:
        fx = {k: z[k] for k in z.files}
    return make, fx

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP library)')


def _gpu_unavailable():
    '''Why gpu-marked tests cannot run here, or None: no /dev/kfd (no AMD GPU driver), or the HIP library is not
    built.'''
    if not os.path.exists('/dev/kfd'):
        return 'no AMD GPU here (/dev/kfd is absent)'
    try:
        from searay_amd import _native
        _native.library()
    except (ImportError, OSError) as e:
        return 'the HIP library does not load: %s' % e
    return None


def pytest_collection_modifyitems(config, items):
    gpu_items = [it for it in items if it.get_closest_marker('gpu') is not None]
    if not gpu_items:
        return
    reason = _gpu_unavailable()
    if reason is None:
        return
    skip = pytest.mark.skip(reason=reason)
    for it in gpu_items:
        it.add_marker(skip)


def rel_l2(a, b):
    '''Relative L2 distance ||a-b|| / ||b||.'''
    a = np.asarray(a)
    b = np.asarray(b)
    nb = np.linalg.norm(b.ravel())
    return np.linalg.norm((a - b).ravel()) / (nb if nb > 0 else 1.0)


@pytest.fixture(scope='session')
def oracle_mod():
    from oracle import oracle
    return oracle

import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP device) and the built library')


@pytest.fixture(scope='session')
def golden_small():
    return np.load(os.path.join(GOLDEN, 'golden_small.npz'))


@pytest.fixture(scope='session')
def golden_c1():
    return np.load(os.path.join(GOLDEN, 'golden_c1.npz'))


@pytest.fixture(scope='session')
def golden_c2():
    return np.load(os.path.join(GOLDEN, 'golden_c2.npz'))


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))

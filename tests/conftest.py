import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (HIP) device')


def golden(name):
    """Loads a golden fixture (plain arrays, no pickles)."""
    return np.load(os.path.join(GOLDEN, name + '.npz'), allow_pickle=False)


def golden_csr(g, key):
    from scipy.sparse import csr_matrix
    return csr_matrix((g[key + '_data'], g[key + '_indices'], g[key + '_indptr']),
                      shape=tuple(g[key + '_shape']))


def rel_err(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.fixture(scope='session')
def gpu():
    """Skips unless a HIP device and the built library are present."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no HIP device')
    import qspectra_amd
    qspectra_amd.require_native()
    return torch.device('cuda:0')

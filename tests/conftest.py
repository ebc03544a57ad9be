import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


def lattice(n, L=1.0):
    """examples/testHODLR3D.cpp:15-31: x_k = -L + 2L(k+1)/(n+1), z fastest, then x, then y."""
    t = -L + 2.0 * L * (np.arange(n) + 1.0) / (n + 1.0)
    j, k, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    return np.stack([t[k.ravel()], t[j.ravel()], t[i.ravel()]], axis=1)


def uniform(N, seed, lo=-1.0, hi=1.0):
    return np.random.default_rng(seed).uniform(lo, hi, (N, 3))


@pytest.fixture(scope="session")
def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False

"""Shared test setup: repository root on sys.path, the `gpu` marker, small helpers."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); run with -m gpu")


PARAMS = dict(pix_size=1.5, Cs=2.7, EkV=300.0, gaussEnv=np.inf, AmpContrast=0.1)
FILTER = dict(type="Butter", Qc=0.5, N=8)


def rel_offdiag(a, b):
    """Per-entry relative error of a against b over off-diagonal entries."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    off = ~np.eye(a.shape[0], dtype=bool)
    return np.abs(a - b)[off] / np.abs(b)[off]


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no ROCm GPU")
    return torch.device("cuda")


def curve_with_outlier(n=400, gap=0.13, seed=0):
    """Squared distances (x 1e4) of n points on a planar curve plus one point `gap` beyond its
    end: the diffusion map puts that point at ||psi[:, 0:3]|| ~ 6.6 and every curve point
    below 2.1, so trimming at rad = 5 removes exactly the outlier and then stops."""
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0, 1, n))
    X = np.stack([t, 0.1 * np.sin(6 * t)], 1) + 0.002 * rng.normal(size=(n, 2))
    out = np.array([[1 + gap, 0.1 * np.sin(6)]]) + 0.002 * rng.normal(size=(1, 2))
    X = np.concatenate([X, out])
    return np.sum((X[:, None] - X[None]) ** 2, -1) * 1e4

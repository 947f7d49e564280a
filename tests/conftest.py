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

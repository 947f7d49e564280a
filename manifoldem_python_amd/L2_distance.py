"""Drop-in for ``L2_distance`` (``L2_distance.py:8-41``): sqrt(aa + bb - 2 a^T b) with entries
below 1e-8 set to 0, on the GPU (float64)."""
import numpy as np
import torch

from . import pipeline as pl

eps = 1e-8


def op(a, b):
    if a.shape[0] != b.shape[0]:
        raise ValueError("A and B should be of same dimensionality")
    dev = pl.device()
    A = torch.as_tensor(np.asarray(a, dtype=np.float64), device=dev)
    B = torch.as_tensor(np.asarray(b, dtype=np.float64), device=dev)
    aa = (A * A).sum(0)
    bb = (B * B).sum(0)
    tmp = aa[:, None] + bb[None, :] - 2 * (A.T.contiguous() @ B)
    tmp = torch.where(tmp < eps, torch.zeros_like(tmp), tmp)
    return torch.sqrt(tmp).cpu().numpy()

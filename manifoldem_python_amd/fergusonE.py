"""Drop-in for ``fergusonE`` (``fergusonE.py:21-59``): the Ferguson bandwidth sweep over a value
list on the GPU, the tanh fit on the host (scipy curve_fit, unseeded restarts as in the
reference)."""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl

fun = pl.tanh_model


def find_thres(logEps, D2):
    """``:25-31`` (host form of the threshold the GPU sweep evaluates itself)."""
    eps = np.exp(logEps)
    d = 1.0 / (2.0 * np.max(eps)) * D2
    ss = np.sum(np.exp(-np.sort(d)))
    return max(-np.log(0.01 * ss / len(D2)), 10)


def sums(D, logEps):
    """(sum W, d log sum W / d log eps) for each eps, on the GPU."""
    dev = pl.device()
    v = torch.as_tensor(np.ascontiguousarray(np.ravel(D), dtype=np.float64), device=dev)
    le = torch.as_tensor(np.ascontiguousarray(logEps, dtype=np.float64), device=dev)
    out = torch.empty(3 * le.numel() + 1, dtype=torch.float64, device=dev)
    nat.ferguson_values(v, le, out)
    o = out.cpu().numpy()[:3 * le.numel()].reshape(-1, 3)
    return o[:, 0], o[:, 1] / o[:, 0]


def op(D, logEps, a0):
    S, _ = sums(D, logEps)
    with np.errstate(divide="ignore"):
        logSumWij = np.log(S)
    popt, resnorm, R_squared = pl.ferguson_fit(np.asarray(logEps), logSumWij, a0)
    return popt, logSumWij, resnorm, R_squared

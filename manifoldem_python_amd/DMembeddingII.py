"""Drop-in for ``DMembeddingII`` (Stage 2: diffusion-map embedding of one distance matrix).

``op(D, k, tune, prefsigma)`` keeps the reference signature and outputs
``(lamb, psi, sigma, mu, logEps, logSumWij, popt, R_squared)`` (``DMembeddingII.py:86-185``).
kNN selection, the Ferguson sweep, the Markov normalisation and the eigensolve run on the GPU;
the tanh fit runs on the host with numpy's global random restarts, as in the reference.
Like the reference (``:48``), the diagonal of the caller's D is set to -inf.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import p
from . import pipeline as pl


def initialize(nS, nN, D):
    """``:43-57``: kNN index/value lists (flattened column-major, nN x nS) from the GPU
    selection; sets D's diagonal to -inf like the reference."""
    dev = pl.device()
    Df = torch.as_tensor(np.asarray(D, dtype=np.float32).ravel(), device=dev)
    plan = nat.Plan([0], [nS], [nN], device=dev)
    mask = torch.empty(max(plan.info.mask_words, 1), dtype=torch.int32, device=dev)
    k = min(nN, nS)
    idx = torch.empty(k * nS, dtype=torch.int32, device=dev)
    val = torch.empty(k * nS, dtype=torch.float32, device=dev)
    nat.knn_select(plan, Df, mask, idx, val)
    np.fill_diagonal(D, -np.inf)
    yInd1 = idx.cpu().numpy().reshape(k, nS)
    yVal1 = val.cpu().numpy().astype(np.float64).reshape(k, nS)
    return yInd1.flatten("F"), yVal1.flatten("F")


def op(D, k, tune, prefsigma, logEps=None):
    dev = pl.device()
    D_arr = np.asarray(D)
    nS = D_arr.shape[0]
    Df = torch.as_tensor(np.ascontiguousarray(D_arr, dtype=np.float32).ravel(), device=dev)
    if isinstance(D, np.ndarray) and D.dtype.kind == "f":
        np.fill_diagonal(D, -np.inf)
    plan = nat.Plan([0], [nS], [int(k)], device=dev)
    le = pl.LOG_EPS if logEps is None else logEps
    r = pl.dm_embed(plan, Df, tune, p.num_eigs, le)[0]
    return r.as_tuple()

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
    selection; sets D's diagonal to -inf like the reference.  The reference sorts columns of D;
    the GPU selection sorts rows, so it is given D^T."""
    dev = pl.device()
    Df = torch.as_tensor(np.ascontiguousarray(np.asarray(D).T, dtype=np.float32).ravel(),
                         device=dev)
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


ZERO = 1e-6                                        # DMembeddingII.py:115
_F32_ZERO_FLOOR = np.nextafter(np.float32(ZERO), np.float32(np.inf))   # least float32 >= 1e-6


def kernel_distances(D):
    """The float32 distances the GPU kernel entries are built from, decided in float64 as the
    reference decides them.

    * Symmetrisation (``construct_matrix0``, ``:73-78``): the reference combines the two
      directions of a pair as a^2 + b^2 - a b with a = sqrt(D_ji), b = sqrt(D_ij) (a term is 0
      when its value is below 1e-6), which is D_ij itself for a symmetric D.  A D that is not
      bitwise symmetric (the stage-1 Gram computed by BLAS is symmetric only to rounding) is
      combined by that formula here, which is the reference's value wherever both directions
      are kNN entries (every entry when k = n, as in its callers).
    * The zero-entry test ``val < 1e-6`` (``:115``) is applied to the float64 values: those
      below become exact zeros and those at or above it stay at or above it in float32.
    kNN order is then decided on the float32 values (ties within float32 rounding go to the
    lower index)."""
    D = np.asarray(D, dtype=np.float64)
    off = ~np.eye(D.shape[0], dtype=bool)
    if not np.array_equal(D[off], D.T[off]):
        with np.errstate(invalid="ignore"):
            b = np.where(D >= ZERO, np.sqrt(np.abs(D)), 0.0)
        a = b.T
        D = a * a + b * b - a * b
    D32 = np.where(D >= ZERO, np.maximum(D, _F32_ZERO_FLOOR), 0.0).astype(np.float32)
    np.fill_diagonal(D32, 0.0)
    return D32


def op(D, k, tune, prefsigma, logEps=None):
    dev = pl.device()
    D_arr = np.asarray(D)
    nS = D_arr.shape[0]
    Df = torch.as_tensor(kernel_distances(D_arr).ravel(), device=dev)
    if isinstance(D, np.ndarray) and D.dtype.kind == "f":
        np.fill_diagonal(D, -np.inf)
    plan = nat.Plan([0], [nS], [int(k)], device=dev)
    le = pl.LOG_EPS if logEps is None else logEps
    r = pl.dm_embed(plan, Df, tune, p.num_eigs, le)[0]
    return r.as_tuple()

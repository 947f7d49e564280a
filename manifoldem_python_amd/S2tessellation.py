"""Drop-in for ``S2tessellation`` (``S2tessellation.py:36-147``): projection directions of the
(augmented) particles, reference directions spread over S^2 by ``distribute3Sphere``, each
particle assigned to its nearest reference direction on the GPU (``memi_nearest_bins``, float64
distances, in place of the reference's scikit-learn ball tree), and the admitted bins (the
half of S^2 whose occupancy list is longer, occupancy >= PDsizeTh, at most ``thres``
particles kept per bin).

``op(q, shAngWidth, PDsizeTh, visual, thres)`` returns the reference's
``(CG1, CG, nG, S2, S20_th, S20, NC)``.  ``visual`` (a matplotlib view) is not drawn.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import distribute3Sphere
from . import pipeline as pl


def get_S2(q):
    """Projection directions 3 x n of quaternions 4 x n (``:36-57``)."""
    q = np.asarray(q, dtype=np.float64)
    if q.shape[0] <= 3:
        raise ValueError("get_S2: q has wrong dimensions")
    return 2 * np.vstack((q[1] * q[3] - q[0] * q[2], q[0] * q[1] + q[2] * q[3],
                          q[0] ** 2 + q[3] ** 2 - 0.5))


def classS2(X, Q):
    """``:59-63``: IND (n x 1 int64) of the nearest row of X (m x 3) to each row of Q (n x 3),
    and NC = bincount(IND) (as long as the largest occupied bin + 1, like the reference's)."""
    dev = pl.device()
    pts = torch.as_tensor(np.ascontiguousarray(Q, dtype=np.float64), device=dev)
    ctr = torch.as_tensor(np.ascontiguousarray(X, dtype=np.float64), device=dev)
    idx = torch.empty(pts.shape[0], dtype=torch.int32, device=dev)
    nat.nearest_bins(pts, ctr, idx)
    NC = torch.bincount(idx.long()).cpu().numpy() if idx.numel() else np.zeros(0, np.int64)
    return idx.cpu().numpy().astype(np.int64).reshape(-1, 1), NC


def _members(IND, nbins):
    """Particles of every bin in increasing index order (``(IND == i).nonzero()[0]``)."""
    ind = IND[:, 0]
    order = np.argsort(ind, kind="stable")
    bounds = np.searchsorted(ind[order], np.arange(nbins + 1))
    return [order[bounds[i]:bounds[i + 1]] for i in range(nbins)]


def select_bins(NC, nbins, PDsizeTh):
    """``:85-108``: bins of the half of S^2 with the longer occupancy list whose occupancy is
    at least PDsizeTh (indices into S20)."""
    mid = int(np.floor(nbins / 2))
    NC1, NC2 = NC[:mid], NC[mid:]
    if len(NC1) >= len(NC2):
        return [int(i) for i in np.nonzero(NC1 >= PDsizeTh)[0]]
    return [mid + int(j) for j in np.nonzero(NC2 >= PDsizeTh)[0]]


def op(q, shAngWidth, PDsizeTh, visual, thres, *fig):
    nG = int(np.floor(4 * np.pi / (shAngWidth ** 2)))
    S20, _ = distribute3Sphere.op(nG)
    S20 = S20.T
    S2 = get_S2(q)
    IND, NC = classS2(S20.T, S2.T)
    members = _members(IND, S20.shape[1])
    CG1 = np.empty(len(members), dtype=object)
    CG1[:] = members
    NIND = select_bins(NC, S20.shape[1], PDsizeTh)
    S20_th = S20[:, NIND]
    CG = [members[b][:thres] if len(members[b]) > thres else members[b] for b in NIND]
    return CG1, CG, nG, S2, S20_th, S20, NC

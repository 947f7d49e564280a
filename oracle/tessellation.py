"""Oracle for the S^2 tessellation and conjugate augmentation of the alignment data (float64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``S2tessellation.py:36-147`` (get_S2, classS2, op), ``distribute3Sphere.py:19-49``,
``util.py:76-106`` (eul_to_quat, augment) and the quaternion product of
``read_alignfile.py:93-94``.  classS2 calls the same library routine as the reference
(scikit-learn's ball-tree 1-NN).

Pinned by (``tests/test_tessellation.py``):
  distribute3Sphere - unit norm, the requested count, ring structure (w1 rows at
                      (k + 1/2) delta, equal spacing along each ring);
  get_S2 / augment  - the augmented quaternion's direction is the antipode of the original's
                      (a closed form for unit quaternions), and S2 of a rotation about z is
                      (0, 0, 1);
  classS2           - brute-force argmin of the distances;
  op thresholds     - hand-built occupancies for both halving branches.
"""
from __future__ import annotations

import math

import numpy as np


def distribute3Sphere(numPts):
    """``distribute3Sphere.py:19-49``: rings at polar angle w1 = (k + 1/2) delta, points at
    azimuth (j + 1/2) delta / sin w1, delta rescaled until the count equals numPts (at most 100
    iterations).  Returns (points (numPts, 3), iterations)."""
    maxIter = 100
    K = numPts
    delta = math.exp(math.log(4 * np.pi / K) / 2.0)
    results = np.zeros((2 * K, 3))
    it = 0
    idx = 0
    while idx != K and it < maxIter:
        it += 1
        idx = 0
        dw1 = delta
        for w1 in np.arange(0.5 * dw1, np.pi, dw1):
            cw1, sw1 = math.cos(w1), math.sin(w1)
            dw2 = dw1 / sw1
            for w2 in np.arange(0.5 * dw2, 2 * np.pi, dw2):
                results[idx, :] = (cw1, sw1 * math.cos(w2), sw1 * math.sin(w2))
                idx += 1
        delta = delta * math.exp(math.log(float(idx) / K) / 2.0)
    return results[0:K, :], it


def get_S2(q):
    """``S2tessellation.py:36-57``: projection directions 3 x n of quaternions 4 x n."""
    return 2 * np.vstack((q[1] * q[3] - q[0] * q[2], q[0] * q[1] + q[2] * q[3],
                          q[0] ** 2 + q[3] ** 2 - 0.5))


def classS2(X, Q):
    """``S2tessellation.py:59-63``: nearest reference direction of each particle (1-NN)."""
    from sklearn.neighbors import NearestNeighbors
    nbrs = NearestNeighbors(n_neighbors=1, algorithm="ball_tree").fit(X)
    _, IND = nbrs.kneighbors(Q)
    NC = np.bincount(IND[:, 0].squeeze())
    return IND, NC


def select_bins(NC, nbins, PDsizeTh):
    """``S2tessellation.py:85-108``: the admitted bins of the half of S^2 with more occupancy
    entries (NC is bincount's output, as long as the largest occupied bin index + 1)."""
    mid = int(np.floor(nbins / 2))
    NC1, NC2 = NC[:mid], NC[mid:]
    NIND = []
    if len(NC1) >= len(NC2):
        for pd, c in enumerate(NC1):
            if c >= PDsizeTh:
                NIND.append(pd)
    else:
        for j, c in enumerate(NC2):
            if c >= PDsizeTh:
                NIND.append(mid + j)
    return NIND


def op(q, shAngWidth, PDsizeTh, visual, thres):
    """``S2tessellation.py:65-147`` -> (CG1, CG, nG, S2, S20_th, S20, NC)."""
    nG = int(np.floor(4 * np.pi / (shAngWidth ** 2)))
    S20, _ = distribute3Sphere(nG)
    S20 = S20.T
    S2 = get_S2(q)
    IND, NC = classS2(S20.T, S2.T)
    CG1 = [(IND == i).nonzero()[0] for i in range(S20.shape[1])]
    NIND = select_bins(NC, S20.shape[1], PDsizeTh)
    S20_th = S20[:, NIND]
    CG = []
    for b in NIND:
        a = (IND == b).nonzero()[0]
        if len(a) > thres:
            a = a[:thres]
        CG.append(a)
    return CG1, CG, nG, S2, S20_th, S20, NC


def eul_to_quat(phi, theta, psi, flip=True):
    """``util.py:76-92``: the three elementary quaternions of RELION's (rot, tilt, psi)."""
    z = np.zeros(phi.shape[0])
    qz = np.vstack((np.cos(phi / 2), z, z, -np.sin(phi / 2)))
    qy = np.vstack((np.cos(theta / 2), z, -np.sin(theta / 2), z))
    sp = np.sin(psi / 2)
    if flip:
        sp = -sp
    qzs = np.vstack((np.cos(psi / 2), z, z, sp))
    return qz, qy, qzs


def qmult(q1, q2):
    """Hamilton product of 4 x n quaternion arrays (``qMult_bsx.py:15-40``)."""
    a0, a1, a2, a3 = q1
    b0, b1, b2, b3 = q2
    return np.vstack((a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
                      a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                      a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1,
                      a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0))


def euler_q(phi, theta, psi, flip=True):
    """``read_alignfile.py:93-94``: q = qzs * (qy * qz)."""
    qz, qy, qzs = eul_to_quat(phi, theta, psi, flip)
    return qmult(qzs, qmult(qy, qz))


def augment(q):
    """``util.py:94-106``: append the conjugate orientations [-q1, q0, -q3, q2]."""
    qc = np.vstack((-q[1], q[0], -q[3], q[2]))
    return np.hstack((q, qc))

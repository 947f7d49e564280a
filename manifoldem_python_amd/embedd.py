"""Drop-in for ``embedd`` (``embedd.py:16-79``): re-embed one PD on the subset of points the
user kept in the 2-D manifold view, on the GPU.

``op(orig_zip, new_zip, PrD)`` keeps the reference signature and side effects: it reads D and
the trimmed posPath of the PD, keeps the points whose first coordinate appears among the new
points' coordinates, embeds D restricted to them with k = n, and rewrites ``eig_spec.txt`` and
the ``gC_trimmed_psi`` pickle (keys lamb, psi, sigma, mu, posPath, ind), then removes the PD's
psi-analysis pickles, their progress markers and ``class_avg.png`` so that they are rebuilt.
"""
from __future__ import annotations

import os

import numpy as np

from . import DMembeddingII, myio, p


def kept_points(orig_zip, new_zip):
    """``embedd.py:32-38``: indices r of the original points whose x coordinate, compared as a
    64-bit pattern, equals any coordinate (x or y) of a new point.

    The reference tests membership with ``np.in1d`` on the (n, 2) float64 arrays viewed as pairs
    of int32: each record is the bit pattern of one coordinate, and column 0 of the reshaped
    result is the x coordinate of each original point.  New points are copies of original
    points, so this selects exactly the points kept (and, as in the reference, any point whose
    x value happens to equal a kept coordinate)."""
    orig = np.asarray(list(orig_zip), dtype=np.float64).reshape(-1, 2)
    new = np.asarray(list(new_zip), dtype=np.float64).reshape(-1, 2)
    x_bits = np.ascontiguousarray(orig[:, 0]).view(np.int64)
    new_bits = np.ascontiguousarray(new).view(np.int64).ravel()
    return np.nonzero(np.isin(x_bits, new_bits))[0]


def op(orig_zip, new_zip, PrD):
    dist_file = "{}prD_{}".format(p.dist_file, PrD)
    psi_file = "{}prD_{}".format(p.psi_file, PrD)
    eig_file = "{}/topos/PrD_{}/eig_spec.txt".format(p.out_dir, PrD + 1)
    D = myio.fin1(dist_file)["D"]
    data = myio.fin1(psi_file)
    posPath, ind = data["posPath"], data["ind"]
    D = np.asarray(D, dtype=np.float64)[posPath][:, posPath]
    posPathInd = kept_points(orig_zip, new_zip)
    D1 = D[posPathInd][:, posPathInd]
    lamb, psi, sigma, mu, logEps, logSumWij, popt, R_squared = DMembeddingII.op(
        D1, D1.shape[0], p.tune, 60000)
    posPath = posPath[posPathInd]
    lamb = lamb[lamb > 0]
    os.makedirs(os.path.dirname(eig_file), exist_ok=True)
    if os.path.exists(eig_file):
        os.remove(eig_file)
    with open(eig_file, "a") as f:
        for i in range(len(lamb) - 1):
            f.write("%d\t%.5f\n" % (i + 1, lamb[i + 1]))
    myio.fout1(psi_file, ["lamb", "psi", "sigma", "mu", "posPath", "ind"],
               [lamb, psi, sigma, mu, posPath, ind])
    for psinum in range(p.num_psis):
        for f in ("{}prD_{}_psi_{}".format(p.psi2_file, PrD, psinum),
                  os.path.join(p.psi2_prog, "%s_%s" % (PrD, psinum))):
            if os.path.exists(f):
                os.remove(f)
    ca_file = "{}/topos/PrD_{}/class_avg.png".format(p.out_dir, PrD + 1)
    if os.path.exists(ca_file):
        os.remove(ca_file)

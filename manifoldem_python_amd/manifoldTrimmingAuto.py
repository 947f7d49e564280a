"""Drop-in for ``manifoldTrimmingAuto`` (``manifoldTrimmingAuto.py:18-95``): diffusion map with
k = n, then re-embedding while points fall outside radius ``rad`` on psi_1..psi_3; writes the
``gC_trimmed_psi`` pickle, the progress marker and ``eig_spec.txt``."""
from __future__ import annotations

import os

import numpy as np

from . import DMembeddingII, myio, p


def get_psiPath(psi, rad, plotNum):
    """``:18-28``."""
    d = np.sqrt(psi[:, plotNum] ** 2 + psi[:, plotNum + 1] ** 2 + psi[:, plotNum + 2] ** 2)
    return (d < rad).nonzero()[0]


def embed_trimmed(D, tune, rad, posPath=None):
    """The trimming loop of ``:48-77`` on an in-memory D; returns (posPath, DM tuple)."""
    nS = D.shape[1]
    if posPath is None:
        posPath = np.arange(nS)
    D = np.array(D[posPath][:, posPath], dtype=np.float64)
    nS = D.shape[1]
    res = DMembeddingII.op(D, nS, tune, 60000)
    posPath1 = get_psiPath(res[1], rad, 0)
    while len(posPath1) < nS:
        nS = len(posPath1)
        D1 = D[posPath1][:, posPath1]
        res = DMembeddingII.op(D1, D1.shape[0], tune, 600000)
        lamb = res[0]
        res = (lamb[lamb > 0],) + tuple(res[1:])
        posPathInt = get_psiPath(res[1], rad, 0)
        posPath1 = posPath1[posPathInt]
    return posPath[posPath1], res


def op(input_data, posPath, tune, rad, visual, doSave):
    dist_file, psi_file, eig_file, prD = input_data[:4]
    data = myio.fin1(dist_file)
    D, ind = data["D"], data["ind"]
    pp = None if (np.isscalar(posPath) and posPath == 0) else np.asarray(posPath)
    posPath, res = embed_trimmed(D, tune, rad, pp)
    lamb, psi, sigma, mu, logEps, logSumWij, popt, R_squared = res
    if doSave["Is"]:
        myio.fout1(psi_file, ["lamb", "psi", "sigma", "mu", "posPath", "ind", "logEps",
                              "logSumWij", "popt", "R_squared"],
                   [lamb, psi, sigma, mu, posPath, ind, logEps, logSumWij, popt, R_squared])
        open(os.path.join(p.psi_prog, "%s" % (prD,)), "a").close()
    if os.path.exists(eig_file):
        os.remove(eig_file)
    with open(eig_file, "a") as f:
        for i in range(len(lamb) - 1):
            f.write("%d\t%.5f\n" % (i + 1, lamb[i + 1]))
    return None

"""Drop-in for ``manifoldTrimmingAuto`` (``manifoldTrimmingAuto.py:18-95``): diffusion map with
k = n, then re-embedding while points fall outside radius ``rad`` on psi_1..psi_3; writes the
``gC_trimmed_psi`` pickle, the progress marker and ``eig_spec.txt``.  The rounds run on the
GPU with the PD's distances resident on the device (``trimming.trim_batch``)."""
from __future__ import annotations

import os

import numpy as np
import torch

from . import DMembeddingII, myio, p
from . import pipeline as pl
from . import trimming
from .trimming import get_psiPath  # noqa: F401  (reference name, manifoldTrimmingAuto.py:18)

OUT_KEYS = ["lamb", "psi", "sigma", "mu", "posPath", "ind", "logEps", "logSumWij", "popt",
            "R_squared"]


def upload_kernel_distances(Ds, dev):
    """Host distance matrices -> (plan, flat float32 device buffer) of their kernel distances
    (``DMembeddingII.kernel_distances``), k = n for every PD."""
    plan = trimming.packed_plan([d.shape[0] for d in Ds], dev)
    flat = torch.empty(max(plan.info.d_elems, 1), dtype=torch.float32, device=dev)
    for j, d in enumerate(Ds):
        plan.block(flat, j).copy_(torch.from_numpy(DMembeddingII.kernel_distances(d)))
    return plan, flat


def embed_trimmed_batch(Ds, tune, rad, logEps=None):
    """The loop of ``:48-77`` for several in-memory D (already restricted to their posPath);
    returns per PD (posPath1 into D's rows, DM tuple)."""
    dev = pl.device()
    plan, flat = upload_kernel_distances(Ds, dev)
    keep, res, _ = trimming.trim_batch(plan, flat, tune, rad, p.num_eigs,
                                       pl.LOG_EPS if logEps is None else logEps)
    return [(keep[j], res[j].as_tuple()) for j in range(len(Ds))]


def embed_trimmed(D, tune, rad, posPath=None):
    """The trimming loop of ``:48-77`` on an in-memory D; returns (posPath, DM tuple)."""
    nS = D.shape[1]
    if posPath is None:
        posPath = np.arange(nS)
    Dp = np.asarray(D, dtype=np.float64)[posPath][:, posPath]
    keep, res = embed_trimmed_batch([Dp], tune, rad)[0]
    return posPath[keep], res


def write_outputs(psi_file, eig_file, prD, posPath, ind, res, doSave):
    """``:79-95``: the gC_trimmed_psi pickle, its progress marker and eig_spec.txt."""
    lamb, psi, sigma, mu, logEps, logSumWij, popt, R_squared = res
    if doSave["Is"]:
        myio.fout1(psi_file, OUT_KEYS,
                   [lamb, psi, sigma, mu, posPath, ind, logEps, logSumWij, popt, R_squared])
        open(os.path.join(p.psi_prog, "%s" % (prD,)), "a").close()
    if os.path.exists(eig_file):
        os.remove(eig_file)
    with open(eig_file, "a") as f:
        for i in range(len(lamb) - 1):
            f.write("%d\t%.5f\n" % (i + 1, lamb[i + 1]))


def op(input_data, posPath, tune, rad, visual, doSave):
    dist_file, psi_file, eig_file, prD = input_data[:4]
    data = myio.fin1(dist_file)
    D, ind = data["D"], data["ind"]
    pp = None if (np.isscalar(posPath) and posPath == 0) else np.asarray(posPath)
    posPath, res = embed_trimmed(D, tune, rad, pp)
    write_outputs(psi_file, eig_file, prD, posPath, ind, res, doSave)
    return None

"""Drop-in for ``psiAnalysisParS2`` (``psiAnalysisParS2.py:45-170``): NLSA for each requested
diffusion coordinate of one PD on the GPU, tau binning into class frames, pickles and progress
markers.  The PD's distances, images and CTFs are uploaded once for all psi."""
from __future__ import annotations

import os

import numpy as np
import torch

from . import myio, p
from . import pipeline as pl
from . import spectral


def corr(a, b, n, m):
    A = a[:, n] - np.mean(a[:, n])
    B = b[:, m] - np.mean(b[:, m])
    with np.errstate(divide="ignore", invalid="ignore"):
        return np.dot(A, B) / (np.std(A) * np.std(B))


def diff_corr(a, b, mx):
    """``:189-192``."""
    return corr(a, b, 0, 0) + corr(a, b, mx, mx) - (corr(a, b, 0, mx) + corr(a, b, mx, 0))


def tau_bins(tau, IMGT, nClass):
    """``:100-124``: min-max normalised tau and the first frame of each of numclass bins."""
    nSrecon = min(IMGT.shape)
    numclass = int(min(nClass, np.floor(nSrecon / 2.0)))
    tau = (tau - min(tau)) / (max(tau) - min(tau))
    tauinds = []
    for i in range(numclass):
        ind1 = float(i) / numclass
        ind2 = ind1 + 1.0 / numclass
        if i == numclass - 1:
            tauind = ((tau >= ind1) & (tau <= ind2)).nonzero()[0]
        else:
            tauind = ((tau >= ind1) & (tau < ind2)).nonzero()[0]
        while tauind.size == 0:
            sc = 1.0 / (numclass * 2.0)
            ind1 = ind1 - sc * ind1
            ind2 = ind2 + sc * ind2
            tauind = ((tau >= ind1) & (tau < ind2)).nonzero()[0]
        tauinds.append(tauind[0])
    return tau, numclass, tauinds


def op(input_data, conOrderRange, traj_name, isFull, psiTrunc, *argv):
    dist_file, psi_file, psi2_file, EL_file, psinums, senses, prD = input_data[:7]
    psi_list = input_data[7] if len(input_data) == 8 else psinums
    data_IMG = myio.fin1(dist_file)
    data_psi = myio.fin1(psi_file)
    D = np.array(data_IMG["D"])
    imgAll = np.array(data_IMG["imgAll"])
    msk2 = np.array(data_IMG["msk2"])
    CTF = np.array(data_IMG["CTF"])
    psi = data_psi["psi"]
    posPath = np.squeeze(data_psi["posPath"])
    nS = len(posPath)
    ConOrder = int(np.floor(float(nS) / conOrderRange))
    dim = int(np.sqrt(imgAll.size / D.shape[0]))
    CTF = CTF.reshape(D.shape[0], dim, dim)
    D = D[posPath][:, posPath]
    dev = pl.device()
    Dd = torch.as_tensor(np.ascontiguousarray(D, dtype=np.float64), device=dev)
    imgs = torch.as_tensor(imgAll, dtype=torch.float32, device=dev)
    ctf_half = torch.as_tensor(np.ascontiguousarray(
        CTF[:, :, :dim // 2 + 1].reshape(CTF.shape[0], -1), dtype=np.float32), device=dev)
    m2 = None if msk2.ndim == 0 else torch.as_tensor(msk2.astype(np.float32), device=dev)
    for psinum in psi_list:
        if psinum == -1:
            continue
        PosPsi1 = np.argsort(psi[:, psinum])
        r = spectral.nlsa(Dd, torch.as_tensor(PosPsi1, device=dev),
                          torch.as_tensor(posPath[PosPsi1].astype(np.int64), device=dev), imgs,
                          ctf_half, dim, ConOrder, psiTrunc, nS=nS, tune=p.tune,
                          num_eigs=p.num_eigs, msk2=m2)
        tau, numclass, tauinds = tau_bins(r.tau.ravel(), r.IMGT_dev, p.nClass)
        IMG1 = r.IMGT_dev[:, torch.as_tensor(tauinds, device=dev)].cpu().numpy()
        tau = tau.reshape(r.tau.shape)
        if isFull:
            data = myio.fin1("{}_psi_{}".format(psi2_file, psinum))
            dc = diff_corr(IMG1, data["IMG1"], numclass - 1)
            if (senses[0] == -1 and dc > 0) or senses[0] == 1 and dc < 0:
                tau = 1 - tau
            IMGT = r.IMGT_dev.cpu().numpy()
            myio.fout1("{}_{}_1".format(EL_file, traj_name),
                       ["IMG1", "IMGT", "posPath", "PosPsi1", "psirec", "tau", "psiC1", "mu",
                        "VX", "sdiag", "Topo_mean", "tauinds"],
                       [IMG1, IMGT, posPath, PosPsi1, r.psirec, tau, r.psiC1, r.mu, r.VX,
                        r.sdiag, r.Topo_mean, tauinds])
            open(os.path.join(p.EL_prog, "%s" % (prD,)), "a").close()
        else:
            myio.fout1("{}_psi_{}".format(psi2_file, psinum),
                       ["IMG1", "psirec", "tau", "psiC1", "mu", "VX", "sdiag", "Topo_mean",
                        "tauinds"],
                       [IMG1, r.psirec, tau, r.psiC1, r.mu, r.VX, r.sdiag, r.Topo_mean,
                        tauinds])
            open(os.path.join(p.psi2_prog, "%s_%s" % (prD, psinum)), "a").close()
    return "ok"

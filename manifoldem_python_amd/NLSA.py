"""Drop-in for ``NLSA`` (``NLSA.py:23-158``): NLSA along one diffusion coordinate on the GPU.

``op(NLSAPar, DD, posPath, posPsi1, imgAll, msk2, CTF, ExtPar)`` keeps the reference signature
and returns ``(IMGT, Topo_mean, psirec, psiC1, sdiag, VX, mu, tau)``.  DD is the distance matrix
already sorted along psi; imgAll / CTF are the full per-PD stacks indexed by posPath[posPsi1].
With 'prD' in ExtPar the supervectors are the Wiener-filtered, masked images posPath[posPsi1]
(``NLSA.py:52-53``); with 'cuti' they are the raw images imgAll[posPsi1] (``NLSA.py:54-55``).
"""
from __future__ import annotations

import numpy as np
import torch

from . import myio, p
from . import pipeline as pl
from . import spectral


def _ctf_half(CTF, N):
    C = np.asarray(CTF).reshape(-1, N, N)
    return np.ascontiguousarray(C[:, :, :N // 2 + 1].reshape(C.shape[0], -1), dtype=np.float32)


def op(NLSAPar, DD, posPath, posPsi1, imgAll, msk2, CTF, ExtPar):
    if "prD" in ExtPar:
        wiener = True
    elif "cuti" in ExtPar:
        wiener = False
    else:
        raise ValueError("NLSA.op: ExtPar names neither 'prD' nor 'cuti'")
    dev = pl.device()
    num, ConOrder = NLSAPar["num"], NLSAPar["ConOrder"]
    nS, psiTrunc, tune = NLSAPar["nS"], NLSAPar["psiTrunc"], NLSAPar["tune"]
    imgAll = np.asarray(imgAll)
    N = imgAll.shape[1]
    idx = (np.asarray(posPath)[np.asarray(posPsi1)] if wiener else
           np.asarray(posPsi1)).astype(np.int64)
    m2 = None if not wiener or np.ndim(msk2) == 0 else torch.as_tensor(
        np.asarray(msk2, dtype=np.float32), device=dev)
    r = spectral.nlsa(torch.as_tensor(np.ascontiguousarray(DD, dtype=np.float64), device=dev),
                      torch.arange(num, dtype=torch.int64, device=dev),
                      torch.as_tensor(idx, device=dev),
                      torch.as_tensor(imgAll, dtype=torch.float32, device=dev),
                      torch.as_tensor(_ctf_half(CTF, N), device=dev), N, ConOrder, psiTrunc,
                      nS=nS, tune=tune, num_eigs=p.num_eigs, msk2=m2,
                      k=NLSAPar.get("k", num - ConOrder), wiener=wiener)
    r.IMGT = r.IMGT_dev.cpu().numpy()
    if NLSAPar.get("save", False):
        myio.fout1(ExtPar["filename"], ["psirec", "tau", "a", "b"], [r.psirec, r.tau, r.a, r.b])
    return r.as_tuple()

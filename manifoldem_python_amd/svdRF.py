"""Drop-in for ``svdRF`` (``svdRF.py:9-35``): SVD of a tall (or wide) matrix through the
eigen-decomposition of its small Gram matrix; Gram and back-projection on the GPU.

A short side of at most TALL_MAX (8, the NLSA case: psiTrunc columns) runs on the
one-pass tall-Gram kernels; a longer short side uses float64 library GEMMs on the device.
The small symmetric eigenproblem is solved on the host (LAPACK), as ``np.linalg.eigh`` is in
the reference."""
import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl

TALL_MAX = 8


def tidyUp(D, EV):
    order = np.argsort(D)[::-1]
    D = np.sort(D)[::-1]
    EV = EV[:, order]
    sqrtD = np.sqrt(D)
    return D, EV, np.diag(sqrtD), np.diag(1.0 / sqrtD)


def op(A):
    dev = pl.device()
    A = np.asarray(A, dtype=np.float64)
    D1, D2 = A.shape
    tall = D1 > D2
    M = A if tall else A.T
    c = M.shape[1]
    Md = torch.as_tensor(np.ascontiguousarray(M), device=dev)
    if c <= TALL_MAX:
        G = torch.empty((c, c), dtype=torch.float64, device=dev)
        nat.tall_gram(Md, G)
    else:
        G = Md.t().contiguous() @ Md
    D, E = np.linalg.eigh(G.cpu().numpy())
    D, E, S, invS = tidyUp(D, E)
    if c <= TALL_MAX:
        Y = torch.empty_like(Md)
        nat.tall_apply(Md, E @ invS, Y)
    else:
        Y = Md @ torch.as_tensor(E @ invS, device=dev)
    if tall:
        return Y.cpu().numpy(), S, E
    return E, S, Y.cpu().numpy()

"""Drop-in for ``svdRF`` (``svdRF.py:9-35``): SVD of a tall (or wide) matrix through the
eigen-decomposition of its small Gram matrix; Gram and back-projection on the GPU."""
import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl


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
    if M.shape[1] > 8:
        raise ValueError("svdRF on the GPU expects the short side <= 8 (got %d)" % M.shape[1])
    Md = torch.as_tensor(np.ascontiguousarray(M), device=dev)
    G = torch.empty((M.shape[1], M.shape[1]), dtype=torch.float64, device=dev)
    nat.tall_gram(Md, G)
    D, E = np.linalg.eigh(G.cpu().numpy())
    D, E, S, invS = tidyUp(D, E)
    Y = torch.empty_like(Md)
    nat.tall_apply(Md, E @ invS, Y)
    if tall:
        return Y.cpu().numpy(), S, E
    return E, S, Y.cpu().numpy()

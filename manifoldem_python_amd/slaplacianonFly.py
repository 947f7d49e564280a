"""Drop-in for ``slaplacianonFly`` (``slaplacianonFly.py:42-81``): Gaussian weights and the
alpha / symmetric normalisations of a COO kernel on the GPU; returns the same scipy csc matrix
|l + l^T| / 2 and sigmaTune."""
from __future__ import annotations

import numpy as np
import torch
from scipy.sparse import csc_matrix

from . import _native as nat
from . import pipeline as pl


def op(*arg):
    yVal, yCol, yRow, nS, options = arg[0], arg[1], arg[2], arg[3], arg[4]
    if getattr(options, "autotune", 0) > 0:
        print("Autotuning is not implemented in this version of slaplacian\n")
    sigmaTune = options.sigma
    dev = pl.device()
    row = np.ascontiguousarray(np.ravel(yRow), dtype=np.int64)
    col = np.ascontiguousarray(np.ravel(yCol), dtype=np.int64)
    val = np.ascontiguousarray(np.ravel(yVal), dtype=np.float64)
    out = torch.empty(val.shape[0], dtype=torch.float64, device=dev)
    deg = torch.empty(2 * int(nS), dtype=torch.float64, device=dev)
    nat.markov_coo(torch.as_tensor(row, device=dev), torch.as_tensor(col, device=dev),
                   torch.as_tensor(val, device=dev), int(nS), sigmaTune, options.alpha, out, deg)
    l = csc_matrix((out.cpu().numpy(), (row, col)), shape=(nS, nS))
    l = abs(l + l.T) / 2.0
    return l, sigmaTune

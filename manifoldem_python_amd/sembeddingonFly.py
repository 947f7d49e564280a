"""Drop-in for ``sembeddingonFly`` (``sembeddingonFly.py:17-39``): the normalised kernel
(slaplacianonFly) and its nEigs+1 largest-magnitude eigenpairs by GPU Lanczos, returned
sorted by descending eigenvalue."""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl
from . import slaplacianonFly


def op(yVal, yCol, yRow, nS, options1):
    class _O:
        pass
    options = _O()
    options.sigma, options.alpha, options.nEigs = options1.sigma, options1.alpha, options1.nEigs
    options.autotune = 0
    l, _ = slaplacianonFly.op(yVal, yCol, yRow, nS, options)
    dev = pl.device()
    L = torch.as_tensor(l.toarray().ravel(), device=dev)
    plan = nat.Plan([0], [int(nS)], [int(nS)], device=dev)
    lam, V, _, _ = pl.lanczos_eigs(plan, L, [options.nEigs + 1])[0]
    return lam, V

"""Manifold trimming on the GPU for a batch of PDs (``manifoldTrimmingAuto.py:48-77``).

Per PD the reference embeds D with k = n, keeps the points with ||psi[:, 0:3]|| < rad and
re-embeds D[posPath1][:, posPath1] until a round removes nothing, dropping non-positive
eigenvalues after each re-embedding (``:68``).  Here every PD's kernel distances stay in one
device buffer for all rounds.  A round builds a plan over the PDs that lost points in the
previous round, gathers all their sub-matrices from that buffer in one launch
(``memi_gather_blocks``) and re-embeds them together (``pipeline.dm_embed``); only the psi
coordinates go to the host, where the radius test picks the points kept.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl


def get_psiPath(psi, rad, plotNum=0):
    """``manifoldTrimmingAuto.py:18-28``: indices with ||psi[:, plotNum:plotNum+3]|| < rad."""
    d = np.sqrt(psi[:, plotNum] ** 2 + psi[:, plotNum + 1] ** 2 + psi[:, plotNum + 2] ** 2)
    return (d < rad).nonzero()[0]


def packed_plan(ns, dev, ks=None):
    """A plan of PDs of sizes ns packed from row 0 (k = n unless ks is given)."""
    ns = [int(n) for n in ns]
    rows0 = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.int64)
    return nat.Plan(rows0, ns, ns if ks is None else ks, device=dev)


def trim_batch(plan, D, tune, rad, num_eigs=15, logEps=pl.LOG_EPS, rand=None):
    """Trimming loop for every PD of ``plan``; D: float32 device kernel distances in the plan's
    blocks (each PD embedded with its plan k, n for the reference's callers).

    Returns (keep, results, rounds): per PD the int64 indices kept (into the PD's rows), the
    DMResult of its last embedding and the number of re-embeddings."""
    dev = D.device
    npd = plan.npd
    src_off = plan.pds["d_off"].astype(np.int64)
    src_n = plan.ns.astype(np.int32)
    results = list(pl.dm_embed(plan, D, tune, num_eigs, logEps, rand=rand))
    keep = [np.arange(int(n)) for n in plan.ns]
    rounds = [0] * npd
    active = []
    for p in range(npd):
        inner = get_psiPath(results[p].psi, rad)
        if len(inner) < len(keep[p]):
            keep[p] = keep[p][inner]
            active.append(p)
    while active:
        ns = [len(keep[p]) for p in active]
        if min(ns) < 4:
            raise RuntimeError("trim_batch: a PD was trimmed to %d points (PD %d); the "
                               "embedding needs at least 4" % (min(ns), active[int(np.argmin(ns))]))
        sub = packed_plan(ns, dev)
        Dsub = torch.empty(max(sub.info.d_elems, 1), dtype=torch.float32, device=dev)
        nat.gather_blocks(sub, D, src_off[active], src_n[active],
                          np.concatenate([keep[p] for p in active]), Dsub)
        res = pl.dm_embed(sub, Dsub, tune, num_eigs, logEps, rand=rand)
        nxt = []
        for j, p in enumerate(active):
            r = res[j]
            r.lamb = r.lamb[r.lamb > 0]
            results[p] = r
            rounds[p] += 1
            inner = get_psiPath(r.psi, rad)
            if len(inner) < ns[j]:
                keep[p] = keep[p][inner]
                nxt.append(p)
        active = nxt
    return keep, results, rounds

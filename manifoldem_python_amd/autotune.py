"""Pick the Gram kernel schedule for a batch by timing every variant on it.

The MFMA Gram (memi_ctf_distance) has schedule variants that trade prefetch depth against
occupancy (see include/memi.h).  Which one wins depends on the PD sizes and image side, so
``tune_gram`` times each on the batch's own operand with HIP events around the Gram kernel and
records the fastest in ``pipeline.GRAM_VARIANT``.  Results are cached per (precision, N,
size signature) in memory and, optionally, in a JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl

VARIANTS = (1, 2)
_cache = {}


def _key(precision, N, ns):
    ns = np.asarray(ns)
    return f"{int(precision)}:{int(N)}:{len(ns)}:{int(ns.sum())}:{int((ns.astype(np.int64) ** 2).sum())}"


def tune_gram(wsp, N, reps=3, cache_file=None):
    """Time each schedule variant on wsp's operand; set and return the fastest."""
    key = _key(wsp.precision, N, wsp.plan.ns)
    if cache_file and os.path.exists(cache_file) and not _cache:
        with open(cache_file) as f:
            _cache.update(json.load(f))
    if key in _cache:
        pl.GRAM_VARIANT[wsp.precision] = _cache[key]["best"]
        return _cache[key]
    ev = nat.HipEvents()
    times = {}
    for v in VARIANTS:
        pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, N, ws=wsp.ws, variant=v)
        pairs = [ev.new_pair() for _ in range(reps)]
        for p in pairs:
            pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, N, ws=wsp.ws, events=p,
                             variant=v)
        torch.cuda.synchronize()
        times[v] = float(np.median([ev.elapsed_ms(p) for p in pairs]))
    ev.close()
    best = min(times, key=times.get)
    pl.GRAM_VARIANT[wsp.precision] = best
    _cache[key] = {"best": best, "ms": {str(k): t for k, t in times.items()}}
    if cache_file:
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1)
    return _cache[key]

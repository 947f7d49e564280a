"""Pick the Gram kernel schedule and K split for a batch by timing them on it.

The MFMA Gram (memi_ctf_distance) runs one workgroup per (tile, segment, K chunk) work item.
Two things decide how well a launch fills the 256 CUs: the schedule variant, which trades
prefetch depth against occupancy (one or two workgroups per CU), and the number of K chunks,
which sets how many work items there are (a launch of 1.06 rounds of resident workgroups
runs almost as long as two full rounds).  ``tune_gram`` times every (variant, chunks) pair on
the batch's own operand with HIP events around the Gram kernel, installs the fastest plan in
the workspace and its variant in ``pipeline.GRAM_VARIANT``.  Results are cached per
(precision, N, size signature) in memory and, optionally, in a JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl

VARIANTS = (1, 2)
CHUNK_CANDIDATES = {nat.PREC_SPLIT_BF16: (1, 2, 3, 4, 6), nat.PREC_FP32: (2, 3, 4, 6, 8)}
WS_LIMIT_BYTES = 2 << 30          # skip splits whose partial tiles would exceed this
_cache = {}


def _key(precision, N, ns):
    ns = np.asarray(ns)
    return f"{int(precision)}:{int(N)}:{len(ns)}:{int(ns.sum())}:{int((ns.astype(np.int64) ** 2).sum())}"


def valid_chunks(precision, Khp, chunks):
    """True when `chunks` splits the Khp K range into non-empty chunks of whole 64-byte steps
    (the rule memi_ctf_distance enforces)."""
    steps = Khp // (16 if precision == nat.PREC_FP32 else 32)
    spc = -(-steps // chunks)
    return chunks >= 1 and (chunks - 1) * spc < steps


def _install(wsp, chunks):
    p = wsp.plan
    if p.chunks != chunks:
        wsp.plan = nat.Plan(p.rows0, p.ns, p.ks, chunks=chunks, device=wsp.operand.hi.device)
    need = max(wsp.plan.info.ws_floats, 1)
    if wsp.ws.numel() < need:
        wsp.ws = torch.empty(need, dtype=torch.float32, device=wsp.ws.device)


def tune_gram(wsp, N, reps=3, cache_file=None, chunk_candidates=None):
    """Time each (variant, chunks) on wsp's operand; install and return the fastest."""
    prec = wsp.precision
    key = _key(prec, N, wsp.plan.ns)
    if cache_file and os.path.exists(cache_file) and not _cache:
        with open(cache_file) as f:
            _cache.update(json.load(f))
    if key in _cache:
        _install(wsp, _cache[key]["chunks"])
        pl.GRAM_VARIANT[prec] = _cache[key]["best"]
        return _cache[key]
    Khp = wsp.operand.Khp
    cands = chunk_candidates or CHUNK_CANDIDATES[prec]
    ut = int(wsp.plan.info.n_utiles)
    cands = [c for c in cands if valid_chunks(prec, Khp, c)
             and ut * 4 * c * 128 * 128 * 4 <= WS_LIMIT_BYTES] or [wsp.plan.chunks]
    ev = nat.HipEvents()
    times = {}
    for c in cands:
        _install(wsp, c)
        for v in VARIANTS:
            pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, N, ws=wsp.ws, variant=v)
            pairs = [ev.new_pair() for _ in range(reps)]
            for p in pairs:
                pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, N, ws=wsp.ws,
                                 events=p, variant=v)
            torch.cuda.synchronize()
            times[(v, c)] = float(np.median([ev.elapsed_ms(p) for p in pairs]))
    ev.close()
    best_v, best_c = min(times, key=times.get)
    _install(wsp, best_c)
    pl.GRAM_VARIANT[prec] = best_v
    _cache[key] = {"best": best_v, "chunks": best_c,
                   "ms": {f"v{v}c{c}": t for (v, c), t in times.items()}}
    if cache_file:
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1)
    return _cache[key]

"""Pick the Gram kernel's tile shape, schedule and K split for a batch by timing them on it.

The MFMA Gram (memi_ctf_distance) runs one workgroup per (tile, half, K chunk) work item.
Three things decide how well a launch fills the 256 CUs: the tile shape (128 x 128 tiles at
two workgroups per CU, or 256 x 256 tiles at one, which stage half the operand bytes per
MFMA but leave fewer, larger work items and more padding on small PDs), the schedule variant
(prefetch one or two K steps through registers, or load straight into LDS) and the number of
K chunks, which sets how many work items there are (a launch of 1.06 rounds of resident workgroups runs almost as
long as two full rounds).  ``tune_gram`` times every (tile, variant, chunks) triple on the
batch's own operand with HIP events around the Gram kernel, installs the fastest plan in the
workspace and its variant in ``pipeline.GRAM_VARIANT``.  Results are cached per (precision,
N, size signature) in memory and, optionally, in a JSON file.
"""
from __future__ import annotations

import json
import os

import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl

# (tile, variant) pairs: variants 1 and 2 prefetch one or two K steps through registers (the
# 256-wide kernel one only), variant 3 loads each K step straight into LDS
TILE_VARIANTS = ((128, 1), (128, 2), (128, 3), (256, 1), (256, 3))
# K chunks per half (two operand segments).  Split-BF16 takes two or more, so that no float32
# accumulator chain is longer than one segment's K range (its D error model assumes that
# length); FP32's products carry no split error and may run a half in one chain.
CHUNK_CANDIDATES = {
    (nat.PREC_SPLIT_BF16, 128): (2, 4, 6, 8, 12),
    (nat.PREC_SPLIT_BF16, 256): (2, 3, 4, 6, 7, 8),
    (nat.PREC_FP32, 128): (4, 6, 8, 12, 16),
    (nat.PREC_FP32, 256): (1, 2, 3, 4, 6, 7, 8),
}
WS_LIMIT_BYTES = 4 << 30          # skip splits whose partial tiles would exceed this
_cache = {}


def _key(precision, N, ns):
    ns = np.asarray(ns)
    return f"{int(precision)}:{int(N)}:{len(ns)}:{int(ns.sum())}:{int((ns.astype(np.int64) ** 2).sum())}"


def valid_chunks(precision, Khp, chunks):
    """True when `chunks` splits a half's K range (two segments of Khp elements) into non-empty
    chunks of whole 64-byte steps (the rule memi_ctf_distance enforces)."""
    steps = 2 * (Khp // (16 if precision == nat.PREC_FP32 else 32))
    spc = -(-steps // chunks)
    return chunks >= 1 and (chunks - 1) * spc < steps


def _install(wsp, chunks, tile):
    p = wsp.plan
    if p.chunks != chunks or p.gram_tile != tile:
        wsp.plan = nat.Plan(p.rows0, p.ns, p.ks, chunks=chunks, device=wsp.operand.hi.device,
                            gram_tile=tile)
    need = max(wsp.plan.info.ws_floats, 1)
    if wsp.ws.numel() < need:
        wsp.ws = torch.empty(need, dtype=torch.float32, device=wsp.ws.device)


def _utiles(ns, tile):
    T = (np.asarray(ns, dtype=np.int64) + tile - 1) // tile
    return int(np.sum(T * (T + 1) // 2))


def tune_gram(wsp, N, reps=3, cache_file=None, chunk_candidates=None, tile_variants=None):
    """Time each (tile, variant, chunks) on wsp's operand; install and return the fastest."""
    prec = wsp.precision
    key = _key(prec, N, wsp.plan.ns)
    if cache_file and os.path.exists(cache_file) and not _cache:
        with open(cache_file) as f:
            _cache.update(json.load(f))
    if key in _cache:
        _install(wsp, _cache[key]["chunks"], _cache[key]["tile"])
        pl.GRAM_VARIANT[prec] = _cache[key]["best"]
        return _cache[key]
    Khp = wsp.operand.Khp
    ev = nat.HipEvents()
    times = {}
    for tile, v in (tile_variants or TILE_VARIANTS):
        ut = _utiles(wsp.plan.ns, tile)
        cands = chunk_candidates or CHUNK_CANDIDATES[(prec, tile)]
        cands = [c for c in cands if valid_chunks(prec, Khp, c)
                 and ut * 2 * c * tile * tile * 4 <= WS_LIMIT_BYTES]
        for c in cands:
            _install(wsp, c, tile)
            pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, N, ws=wsp.ws, variant=v)
            pairs = [ev.new_pair() for _ in range(reps)]
            for p in pairs:
                pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, N, ws=wsp.ws,
                                 events=p, variant=v)
            torch.cuda.synchronize()
            times[(tile, v, c)] = float(np.median([ev.elapsed_ms(p) for p in pairs]))
    ev.close()
    if not times:
        raise RuntimeError("tune_gram: no valid (tile, variant, chunks) candidate")
    best_t, best_v, best_c = min(times, key=times.get)
    _install(wsp, best_c, best_t)
    pl.GRAM_VARIANT[prec] = best_v
    _cache[key] = {"best": best_v, "chunks": best_c, "tile": best_t,
                   "ms": {f"t{t}v{v}c{c}": ms for (t, v, c), ms in times.items()}}
    if cache_file:
        with open(cache_file, "w") as f:
            json.dump(_cache, f, indent=1)
    return _cache[key]

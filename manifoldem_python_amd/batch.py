"""Batched per-PD pipeline: many projection directions in one set of launches.

Replaces the reference's per-PD process pool (``GetDistancesS2.py:53-124``,
``manifoldAnalysis.py:57-116``): all images of all PDs are prepared in one chunked pass, all
PD distance matrices come from one grouped MFMA launch over a (PD, tile, segment) work list,
and the kNN / Ferguson / normalisation kernels each run once over the whole batch.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import dataclass

import numpy as np
import torch

from . import _native as nat
from . import pipeline as pl


@dataclass
class Batch:
    """Device-resident inputs of a batch of PDs (images packed PD after PD)."""
    N: int
    stack: torch.Tensor        # (n_img, N, N) float32 SPIDER-order images
    src: torch.Tensor          # (rows,) int64 stack index per particle
    conj: torch.Tensor         # (rows,) int32 conjugate flag per particle
    rot1: torch.Tensor         # (rows,) float64 first rotation (deg) = -(180/pi) Psi
    rot2: torch.Tensor         # (rows,) float64 second rotation (deg) = -psi_p of its PD
    df: torch.Tensor           # (rows,) float64 defocus (A)
    ns: np.ndarray             # particles per PD
    ks: np.ndarray             # kNN per PD
    geometry: list             # per-PD host geometry (PD, Psi, Nom, Dnom, psi_p)
    ctf: object                # nat.CtfParams
    filterPar: dict

    @property
    def rows0(self):
        return np.concatenate(([0], np.cumsum(self.ns)[:-1])).astype(np.int64)

    @property
    def rows(self):
        return int(np.sum(self.ns))


def make_batch(stack, pds, N, ctf, filterPar, device):
    """pds: list of dicts {src, conj, q, df, k} (numpy) for stack images on `device`."""
    src, conj, rot1, rot2, df, ns, ks, geo = [], [], [], [], [], [], [], []
    for d in pds:
        g = pl.pd_geometry(d["q"])
        n = len(d["src"])
        src.append(np.asarray(d["src"], np.int64))
        conj.append(np.asarray(d["conj"], np.int32))
        rot1.append(-(180 / math.pi) * g["Psi"])
        rot2.append(np.full(n, -g["psi_p"]))
        df.append(np.asarray(d["df"], np.float64))
        ns.append(n)
        ks.append(min(int(d.get("k", n)), n))
        geo.append(g)
    all_src = np.concatenate(src)
    if all_src.size and (all_src.min() < 0 or all_src.max() >= stack.shape[0]):
        raise ValueError("make_batch: particle index outside the image stack")
    cat = lambda xs, dt: torch.as_tensor(np.concatenate(xs).astype(dt), device=device)
    return Batch(N=N, stack=stack, src=cat(src, np.int64), conj=cat(conj, np.int32),
                 rot1=cat(rot1, np.float64), rot2=cat(rot2, np.float64), df=cat(df, np.float64),
                 ns=np.asarray(ns), ks=np.asarray(ks), geometry=geo, ctf=ctf,
                 filterPar=filterPar)


class Workspace:
    """Buffers reused across steps of the same batch and precision."""

    def __init__(self, batch, precision, chunk_rows=16384, gram_tile=None, chunks=None):
        dev = batch.stack.device
        self.precision = precision
        self.plan = nat.Plan(batch.rows0, batch.ns, batch.ks,
                             chunks=pl.CHUNKS[precision] if chunks is None else chunks,
                             device=dev, gram_tile=gram_tile or pl.GRAM_TILE)
        self.operand = pl.Operand(batch.rows, batch.N, precision, dev)
        Nh = batch.N // 2 + 1
        self.fy = torch.empty((batch.rows, batch.N, Nh), dtype=torch.complex64, device=dev)
        self.ctf = torch.empty((batch.rows, batch.N * Nh), dtype=torch.float32, device=dev)
        self.ws = torch.empty(max(self.plan.info.ws_floats, 1), dtype=torch.float32, device=dev)
        self.flags = torch.empty(3 * self.plan.info.flag_cap, dtype=torch.int32, device=dev)
        self.chunk_rows = chunk_rows


def prepare(batch, wsp):
    """a1-a6 for every particle of the batch, in chunks of rows (bounded scratch)."""
    R = batch.rows
    for r0 in range(0, R, wsp.chunk_rows):
        r1 = min(R, r0 + wsp.chunk_rows)
        sl = slice(r0, r1)
        pl.prepare_images(batch.stack, batch.src[sl], batch.conj[sl], batch.rot1[sl],
                          batch.rot2[sl], batch.df[sl], batch.N, batch.filterPar, batch.ctf,
                          operand=wsp.operand, row0=r0, fy_out=wsp.fy[sl], ctf_out=wsp.ctf[sl],
                          want_images=False, src_max=batch.stack.shape[0] - 1)


def distances(batch, wsp, events=None, D=None):
    """a1-a8: prepared operand then the grouped Gram; returns pl.Distances."""
    prepare(batch, wsp)
    return pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, batch.N, ws=wsp.ws,
                            events=events, flags=wsp.flags, D=D)


class GraphedDistances:
    """a1-a8 of a batch captured once into a HIP graph (torch.cuda.CUDAGraph records a
    hipGraph on ROCm) and replayed per step: one graph launch instead of the ~20 kernel and
    FFT launches per row chunk plus the three Gram launches.

    The captured work clears its own counters with kernels (no memset nodes).  Every replay
    overwrites the same device buffers: ``run()`` returns the same Distances object each time,
    so a caller that needs an earlier step's D must copy it before the next replay."""

    def __init__(self, batch, wsp):
        self.batch, self.wsp = batch, wsp
        self.D = torch.empty(max(wsp.plan.info.d_elems, 1), dtype=torch.float32,
                             device=wsp.operand.hi.device)
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):           # eager run first: FFT plans and workspaces
            distances(batch, wsp, D=self.D)
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            self.dist = distances(batch, wsp, D=self.D)

    def run(self):
        self.graph.replay()
        return self.dist


class OverlappedSteps:
    """Steps of the same batch (a stream of PD batches) with step k+1's a1-a8 overlapping step
    k's a10-a16: two buffer sets, each with its captured distance graph; the graph replays run
    on a side stream while the embeddings (small, launch-bound kernels that leave most CUs
    idle) run on the caller's stream.  Events order each set's reuse: step k's replay waits
    until step k-2's embedding is done with the set, and step k's embedding waits for its
    replay.  Every step does its whole batch."""

    def __init__(self, batch, wsp_a, wsp_b, graphs=None, split=0):
        """split: CUs of every 32 given to the embeddings alone, the rest to the distances
        alone (0: both streams on every CU).  Latency-bound embedding kernels sharing SIMDs
        with the prep and Gram kernels ran 5-10 times slower than alone.  The embeddings keep
        mask bits 32 - split .. 31 of every 32.  Each stream needs CUs on every XCD, and the
        mask enumerates CUs interleaved over the eight XCDs (a mask of bits 28-31 of every 32
        was refused), so 8 <= split <= 24."""
        self.batch = batch
        ga, gb = graphs if graphs is not None else (GraphedDistances(batch, wsp_a),
                                                     GraphedDistances(batch, wsp_b))
        self.sets = [(wsp_a, ga), (wsp_b, gb)]
        self.split = int(split)
        if self.split and not 8 <= self.split <= 24:
            raise ValueError("OverlappedSteps: split must be in 8..24 (each stream keeps CUs "
                             "on all eight XCDs)")
        if self.split:
            # disjoint CU masks; both masked streams are blocking, so the embeddings do not
            # run on a caller's null stream.  A device that refuses the masks gets the shared
            # streams (split 0), with a warning
            cut = 32 - self.split
            try:
                self.main = nat.cu_masked_stream(lambda b: b >= cut)
                self.side = nat.cu_masked_stream(lambda b: b < cut)
            except RuntimeError as e:
                warnings.warn(f"OverlappedSteps: CU-masked streams unavailable ({e}); the two "
                              "streams share all CUs", RuntimeWarning, stacklevel=2)
                self.split = 0
        if not self.split:
            self.main = torch.cuda.current_stream()
            self.side = torch.cuda.Stream()
        self.done_dist = [torch.cuda.Event(), torch.cuda.Event()]
        self.done_embed = [torch.cuda.Event(), torch.cuda.Event()]
        for e in self.done_embed:
            e.record(self.main)

    def run(self, n, tune=3, num_eigs=15, logEps=pl.LOG_EPS, rand=None, record=None,
            on_result=None):
        """n steps; returns the last step's embedding results.  record: optional list of n
        (start, end) torch events put around each step's replay on the side stream;
        on_result(k, results) sees every step's results."""
        dists = [None] * n
        caller = torch.cuda.current_stream()
        self.main.wait_stream(caller)
        self.side.wait_stream(caller)

        def launch(k):
            wsp, g = self.sets[k % 2]
            with torch.cuda.stream(self.side):
                self.side.wait_event(self.done_embed[k % 2])
                if record is not None:
                    record[k][0].record(self.side)
                dists[k] = g.run()
                if record is not None:
                    record[k][1].record(self.side)
                self.done_dist[k % 2].record(self.side)

        launch(0)
        res = None
        with torch.cuda.stream(self.main):
            for k in range(n):
                if k + 1 < n:
                    launch(k + 1)
                self.main.wait_event(self.done_dist[k % 2])
                res = embed(self.sets[k % 2][0], dists[k], tune, num_eigs, logEps, rand=rand)
                self.done_embed[k % 2].record(self.main)
                if on_result is not None:
                    on_result(k, res)
        caller.wait_stream(self.main)
        caller.wait_stream(self.side)
        return res


def embed(wsp, dist, tune=3, num_eigs=15, logEps=pl.LOG_EPS, rand=None, timer=None):
    """a10-a16 for every PD of the batch."""
    return pl.dm_embed(wsp.plan, dist.D, tune, num_eigs, logEps, rand=rand, timer=timer)


def breakdown(batch, wsp, num_eigs, rand, logEps=pl.LOG_EPS):
    """Wall time per stage of one step (synchronising between stages), in ms."""
    tm = pl.StageTimer()
    tm.start()
    prepare(batch, wsp)
    tm.mark("prep")
    dist = pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, batch.N, ws=wsp.ws,
                            flags=wsp.flags)
    tm.mark("gram")
    embed(wsp, dist, num_eigs=num_eigs, rand=rand, timer=tm, logEps=logEps)
    return {k: 1e3 * v for k, v in tm.t.items()}

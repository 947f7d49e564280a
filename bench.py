#!/usr/bin/env python
"""Throughput of the per-PD hot path on one MI355X.

A step is one pass of the hot path over the workload already resident in HBM: image
preparation (load/normalise, Butterworth, two spline rotations, CTF, FFTs), the CTF-weighted
distance Gram, and the diffusion-map embedding (kNN, 1501-point Ferguson sweep, tanh fit,
Markov normalisation, float64 eigensolve).  Prints ONE JSON line.

Workload (default, --workload c2): one PD of n=2000 synthetic 128x128 particles, defocus
U(0.5, 3.5) um, kNN k=200, 10 eigenvectors; the same line carries a nested "c3" record of the
batched workload (500 PDs, lognormal sizes (median 400, sigma 0.6, clipped to [100, 3000]),
96x96, k=min(n,150), 8 eigenvectors, prep + Gram replayed from a HIP graph), split-BF16 against
FP32.  --workload c3 measures that workload alone; --workload c1: the plain method on one PD
(n=200, 70x70, full kernel, 101 eps, 8 eigenvectors), whose CPU baseline runs whole.

value = image-pair distances per second over the whole step (split-BF16 Gram);
fp32 = the same with the plain FP32 MFMA Gram.  cpu_baseline: the float64 oracle on this
host's cores in the reference's two modes (one process with all BLAS threads; one
single-threaded process per core on separate PDs).

Secondary workloads (their own JSON line, same contract):
  --workload c4: the differentiable distance -> Gaussian kernel -> log-sum op at n=1000,
    128x128, 41 eps values 2^-20..2^20 x median(D); a step is forward + hand-written backward
    (dL/d eps and dL/dX).
  --workload c5: the full per-PD path with NLSA: PDs of n=1500 160x160 particles (shift
    8 px), stage 1 + diffusion map (8 psi), then NLSA (ConOrder 50, 8 topos) along each of the
    8 psi.  --c5-pds PDs are timed end to end; the 200-PD figure is extrapolated from them
    and labelled so.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

FILTER = dict(type="Butter", Qc=0.5, N=8)
PEAK_BF16_TFLOPS = 2500.0
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--no-c3", action="store_true",
                    help="c2 only: skip the nested 500-PD batched measurement")
    ap.add_argument("--gram-config", default=None,
                    help="tile,variant,chunks: use this Gram configuration for the split-BF16 "
                         "path instead of autotuning (for counter runs of one configuration)")
    ap.add_argument("--c5-pds", type=int, default=2,
                    help="PDs of the NLSA workload timed end to end (c5 only)")
    ap.add_argument("--lanczos-graph", type=int, default=1,
                    help="replay blocks of Lanczos steps from a captured HIP graph (default 1)")
    ap.add_argument("--lanczos-split", type=int, default=None,
                    help="1/0: force the grid-parallel / per-PD orthogonalisation")
    ap.add_argument("--cu-split", type=int, default=None,
                    help="overlapped steps: CUs of every 32 kept for the embedding alone, the "
                         "rest for the prep + Gram (0: shared CUs; default: the measured best)")
    ap.add_argument("--no-pipeline", action="store_true",
                    help="time steps one after another instead of overlapping step k+1's "
                         "prep + Gram with step k's embedding on a second stream")
    ap.add_argument("--graph", type=int, default=None,
                    help="1 (default): replay prep + Gram as a captured HIP graph; 0: eager")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-fp32", action="store_true", help="skip the plain FP32 path")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 error check")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 device copies of 512 MiB (PMC byte-counter calibration)")
    ap.add_argument("--prep-images", default=None, metavar="N,R",
                    help="only time the fused image prep on R random N x N images (split-BF16 "
                         "operand; for counter runs of that kernel)")
    ap.add_argument("--optflow", default=None, metavar="M,F,DIM",
                    help="only time the optical flow of M synthetic F-frame DIM x DIM movies "
                         "(Horn-Schunck, 200 iterations per frame pair)")
    return ap.parse_args()


def prep_images_only(spec, steps):
    """steps launches of the fused image prep (memi_prep_images) on R random N x N images with
    random angles and defoci; prints the per-launch times (HIP events) as one JSON line."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl, synth
    N, R = (int(x) for x in spec.split(","))
    dev = torch.device("cuda")
    nat.load()
    g = torch.Generator(device=dev).manual_seed(1)
    stack = torch.randn((R, N, N), generator=g, device=dev)
    rng = np.random.default_rng(0)
    src = torch.arange(R, device=dev)
    conj = torch.zeros(R, dtype=torch.int32, device=dev)
    rot1, rot2 = (torch.as_tensor(rng.uniform(0, 360, R), device=dev) for _ in range(2))
    df = torch.as_tensor(rng.uniform(5000, 30000, R), device=dev)
    ctf = nat.ctf_params(synth.PIX, synth.CS, synth.EKV, synth.AMPC, np.inf)
    op = pl.Operand(R, N, nat.PREC_SPLIT_BF16, dev)
    fy = torch.empty((R, N, N // 2 + 1), dtype=torch.complex64, device=dev)
    ct = torch.empty((R, N * (N // 2 + 1)), dtype=torch.float32, device=dev)
    ms = []
    for _ in range(steps):
        # the prep kernel runs on torch's current stream, which these events time
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        pl.prepare_images(stack, src, conj, rot1, rot2, df, N, FILTER, ctf, operand=op, row0=0,
                          fy_out=fy, ctf_out=ct, want_images=False, src_max=R - 1,
                          prep="fused")
        b.record()
        torch.cuda.synchronize()
        ms.append(a.elapsed_time(b))
    print(json.dumps({"kernel": "k_prep_images", "image": N, "images": R, "launch_ms": ms,
                      "us_per_image": 1e3 * min(ms[1:] or ms) / R}), flush=True)


def optflow_only(spec, steps, warmup, with_cpu):
    """steps batched optical-flow passes (OpticalFlowMovie.op, 'HS', blockSize_avg 5) over M
    movies of F frames of DIM x DIM; the Horn-Schunck kernel is also timed alone with HIP events
    on the stream it runs on.  One JSON line."""
    import torch
    from manifoldem_python_amd import _native as nat, optflow
    M, F, dim = (int(x) for x in spec.split(","))
    dev = torch.device("cuda")
    nat.load()
    g = torch.Generator(device=dev).manual_seed(3)
    movies = 50.0 * torch.randn((M, F, dim, dim), generator=g, device=dev)
    movies = optflow.gaussian(movies.reshape(M * F, dim, dim), 2.0,
                              torch.float32).reshape(M, F, dim, dim)
    for _ in range(max(1, warmup)):
        optflow.movie_flows(movies, 5, "HS")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        optflow.movie_flows(movies, 5, "HS")
    torch.cuda.synchronize()
    t_step = (time.perf_counter() - t0) / steps
    # the Horn-Schunck kernel alone: 200 iterations = 50 launches of 4 fused iterations
    A = -(-F // 5)
    nP = M * A
    fr = movies[:, ::5].reshape(nP, dim, dim).contiguous()
    pairs = optflow.frame_pairs(M, A, dev)
    fx, fy, ft = optflow.derivatives(fr, pairs)
    U = torch.zeros((nP, dim, dim), dtype=torch.float64, device=dev)
    V = torch.zeros_like(U)
    tu, tv = torch.empty_like(U), torch.empty_like(V)
    nat.hs_iterate(fx, fy, ft, 2.0, 200, U, V, tu, tv)
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    a.record()  # hs_iterate launches on torch's current stream, which these events time
    for _ in range(steps):
        nat.hs_iterate(fx, fy, ft, 2.0, 200, U, V, tu, tv)
    b.record()
    torch.cuda.synchronize()
    launches = steps * 50
    us_launch = 1e3 * a.elapsed_time(b) / launches
    # algorithmic bytes of one launch (4 fused iterations): U, V read and written once in
    # float64 and fx, fy, ft read once in float32 -- 44 B per pixel of every frame pair
    alg = nP * dim * dim * 44.0
    ach = alg / (us_launch * 1e-6) / 1e9
    line = {"metric": "optical_flow_movies_per_s", "value": M / t_step, "unit": "movies/s",
            "n_gpus": 1, "steps": steps, "warmup": warmup, "ms_per_step": 1e3 * t_step,
            "higher_is_better": True, "dtype": "f64 (flow), f32 (images)", "data": "synthetic",
            "config": {"workload": "optflow_hs", "movies": M, "frames": F, "image": dim,
                       "frame_pairs": nP, "hs_iterations": 200, "block_size_avg": 5},
            "roofline": {"kernel": "k_hs_iterate<double,32,4,256>", "bound": "hbm",
                         "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": ach / PEAK_HBM_GBS, "traffic": None,
                         "us_per_launch": us_launch, "alg_bytes_per_launch": alg}}
    if with_cpu:
        from oracle import optflow as oof
        mv = movies[0].reshape(F, dim * dim).cpu().numpy()
        t0 = time.perf_counter()
        oof.movie_flow(mv, 5, "HS")
        tc = time.perf_counter() - t0
        line["cpu_baseline"] = {"value": 1.0 / tc, "unit": "movies/s", "cores": 1,
                                "kind": "port", "sample": "1 movie, oracle.optflow.movie_flow"}
    print(json.dumps(line), flush=True)


def calibrate_copy():
    """Three 512 MiB device-to-device copies: the reference for FETCH_SIZE / WRITE_SIZE."""
    import torch
    a = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}), flush=True)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def build_workload(name, dev):
    import torch
    from manifoldem_python_amd import synth, batch as B, _native as nat
    ctf = nat.ctf_params(synth.PIX, synth.CS, synth.EKV, synth.AMPC, np.inf)
    if name == "c1":
        n, N = 200, 70
        s = synth.make_pd(synth.seeds(6)[0], n, N, device=dev, as_torch=True)
        pds = [dict(src=np.arange(n), conj=np.zeros(n, np.int32), q=s.q, df=s.df, k=n)]
        stack = s.raw
        meta = dict(workload="c1_plain_one_pd", n_pd=1, particles=n, image=N, knn="n (full)",
                    n_eigs=8, n_eps=101)
        num_eigs = 8
    elif name == "c2":
        n, N = 2000, 128
        s = synth.make_pd(synth.seeds(3)[1], n, N, df_range_um=(0.5, 3.5), device=dev,
                          as_torch=True)
        pds = [dict(src=np.arange(n), conj=np.zeros(n, np.int32), q=s.q, df=s.df, k=200)]
        stack = s.raw
        meta = dict(workload="c2_single_large_pd", n_pd=1, particles=n, image=N, knn=200,
                    n_eigs=10, n_eps=1501)
        num_eigs = 10
    else:
        N = 96
        sizes = synth.lognormal_sizes(synth.seeds(4)[2], 500)
        seeds = np.random.SeedSequence(synth.SEED + 3).spawn(len(sizes))
        stacks, pds, off = [], [], 0
        for p, n in enumerate(sizes):
            rng = np.random.default_rng(seeds[p])
            axis = synth.random_axis(rng)
            s = synth.make_pd(seeds[p], int(n), N, axis=axis, device=dev, as_torch=True)
            stacks.append(s.raw)
            pds.append(dict(src=np.arange(n) + off, conj=np.zeros(n, np.int32), q=s.q, df=s.df,
                            k=min(int(n), 150)))
            off += int(n)
        stack = torch.cat(stacks)
        meta = dict(workload="c3_batched_pds", n_pd=len(sizes), particles=int(sizes.sum()),
                    median_particles=int(np.median(sizes)), image=N, knn="min(n,150)",
                    n_eigs=8, n_eps=1501)
        num_eigs = 8
    batch = B.make_batch(stack, pds, N, ctf, FILTER, dev)
    batch.q_host = [d["q"] for d in pds]            # per-PD quaternions (CPU baseline)
    return batch, meta, num_eigs


# overlapped steps: CUs of every 32 kept for the embedding alone, the rest for the prep + Gram
# (0: shared CUs).  Measured on one PD of 2000 images: 0 -> 3.13-3.15 ms, 8 -> 3.18,
# 12 -> 3.29, 16 -> 2.90-2.92, 20 -> 3.98, 24 -> 3.92 ms per step
CU_SPLIT = 16


def run_precision(batch, precision, steps, warmup, num_eigs, events, graph=False,
                  logEps=None, gram_config=None, pipeline=True, cu_split=0):
    """Tune the Gram on the prepared operand, warm up, then time exactly `steps` full steps;
    returns timings and the last results.  graph: the prep + Gram part of each step is one
    replay of a captured HIP graph; the Gram kernel is then timed with HIP events in `steps`
    eager launches right after the timed region (events cannot sit inside the graph).

    pipeline (with graph): consecutive steps overlap on two streams -- step k+1's prep + Gram
    (one graph replay into the second buffer set, on a side stream) runs while step k's
    embedding (launch-bound small kernels that leave most CUs idle) runs on the main stream.
    Every step still does its whole PD batch; events order the reuse of each buffer set.  The
    same number of steps is then timed one after another (``t_serial``) and the last pipelined
    step's eigenvalues are compared with a serial step's."""
    import torch
    from manifoldem_python_amd import batch as B
    from manifoldem_python_amd import autotune
    from manifoldem_python_amd import pipeline as pl
    wsp = B.Workspace(batch, precision)
    le = pl.LOG_EPS if logEps is None else logEps
    rng = np.random.default_rng(0)
    rand = lambda *sh: rng.random(sh)
    B.distances(batch, wsp)
    if gram_config is not None:
        t, v, c = (int(x) for x in gram_config.split(","))
        autotune._install(wsp, c, t)
        pl.GRAM_VARIANT[precision] = v
        tuned = {"best": v, "chunks": c, "tile": t, "ms": {}, "forced": True}
    else:
        tuned = autotune.tune_gram(wsp, batch.N)
    pipeline = bool(pipeline and graph)
    g = B.GraphedDistances(batch, wsp) if graph else None
    dist_fn = (lambda ev=None: g.run()) if graph else (
        lambda ev=None: B.distances(batch, wsp, events=ev))

    def seeded(seed):
        r = np.random.default_rng(seed)
        return lambda *sh: r.random(sh)

    def serial_steps(n, record=None, rand=rand):
        res = None
        for k in range(n):
            if record is not None:
                record[k][0].record()
            dist = dist_fn(pairs[k] if record is not None else None)
            if record is not None:
                record[k][1].record()
            res = B.embed(wsp, dist, num_eigs=num_eigs, rand=rand, logEps=le)
        return res

    pairs = [events.new_pair() for _ in range(steps)]
    dev_ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
              for _ in range(steps)]
    if pipeline:
        wsp2 = B.Workspace(batch, precision)
        autotune._install(wsp2, wsp.plan.chunks, wsp.plan.gram_tile)
        B.distances(batch, wsp2)
        ov = B.OverlappedSteps(batch, wsp, wsp2, graphs=(g, B.GraphedDistances(batch, wsp2)),
                               split=cu_split)
        cu_split = ov.split                     # 0 when the device refused the CU masks

        def pipelined_steps(n, record=None, rand=rand):
            return ov.run(n, num_eigs=num_eigs, logEps=le, rand=rand, record=record)

        run_steps = pipelined_steps
    else:
        run_steps = serial_steps
    run_steps(max(warmup, 2))
    torch.cuda.synchronize()
    # the prep + Gram part of each step is timed by events on its stream, so that nothing
    # waits between it and the embedding (the host queues the next stage meanwhile)
    t0 = time.perf_counter()
    res = run_steps(steps, dev_ev, rand=seeded(1))
    torch.cuda.synchronize()
    t = time.perf_counter() - t0
    t_dist = sum(a.elapsed_time(b) for a, b in dev_ev) / 1e3
    out_extra = {}
    if pipeline:
        # the same steps one after another, from the same random fit starts
        t1 = time.perf_counter()
        res_s = serial_steps(steps, rand=seeded(1))
        torch.cuda.synchronize()
        out_extra["t_serial"] = time.perf_counter() - t1
        out_extra["pipelined_vs_serial_max_dlambda"] = float(max(
            np.max(np.abs(np.asarray(a.lamb) - np.asarray(b.lamb))) for a, b in zip(res, res_s)))
    if graph:
        for k in range(steps):
            pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, batch.N, ws=wsp.ws,
                             events=pairs[k], flags=wsp.flags)
        torch.cuda.synchronize()
    gram_ms = [events.elapsed_ms(p) for p in pairs]
    stages = B.breakdown(batch, wsp, num_eigs, rand, logEps=le)
    J = np.array([getattr(r, "lanczos_steps", 0) for r in res])
    lanczos = dict(mean_steps=float(J.mean()), max_steps=int(J.max()),
                   p90_steps=float(np.percentile(J, 90)),
                   max_residual=float(max(getattr(r, "lanczos_residual", 0.0) for r in res)))
    dist = g.dist if graph else B.distances(batch, wsp)
    return dict(t=t, t_dist=t_dist, gram_ms=float(np.mean(gram_ms)), wsp=wsp, dist=dist,
                res=res, flags=int(dist.flag_count.item()), stages=stages, tuned=tuned,
                graph=graph, lanczos=lanczos, pipeline=pipeline,
                cu_split=cu_split if pipeline else None, **out_extra)


def issued_gram_flops(plan, Khp, products):
    """MFMA flops the Gram kernel issues: for every upper tile of the plan's tile side, the
    16 x 16 blocks that hold data (edge blocks beyond a PD's size and, on diagonal tiles,
    blocks wholly below the diagonal are skipped), 4 segments of Khp, `products` MFMA products
    per block (3 for split-BF16, 1 for FP32)."""
    bt = int(plan.gram_tile)
    tot = 0
    for n in plan.ns:
        nb = -(-int(n) // 16)                    # 16-blocks per side holding data
        per = bt // 16
        T = -(-int(n) // bt)
        for a in range(T):
            ra = range(a * per, min((a + 1) * per, nb))
            for c in range(a, T):
                rc = range(c * per, min((c + 1) * per, nb))
                if a == c:
                    tot += sum(1 for i in ra for j in rc if i <= j)
                else:
                    tot += len(ra) * len(rc)
    return float(tot) * 256 * 2.0 * 4 * Khp * products


def fp64_error(batch, wsp, dist):
    """Per-entry relative error of D against a float64 Gram of the same spectra/CTFs."""
    import torch
    N = batch.N
    Nh = N // 2 + 1
    w = torch.full((Nh,), 2.0, dtype=torch.float64, device=wsp.fy.device)
    w[0] = 1.0
    if N % 2 == 0:
        w[-1] = 1.0
    errs = []
    for p in range(wsp.plan.npd if wsp.plan.npd <= 8 else 8):
        r0, n = int(wsp.plan.rows0[p]), int(wsp.plan.ns[p])
        X = wsp.fy[r0:r0 + n].to(torch.complex128)
        C = wsp.ctf[r0:r0 + n].to(torch.float64).view(n, N, Nh)
        A = (X.abs() ** 2 * w).reshape(n, -1)
        Bm = (C ** 2).reshape(n, -1)
        Y = (C * X * w.sqrt()).reshape(n, -1)
        P = A @ Bm.T
        D64 = P + P.T - 2 * (Y @ Y.conj().T).real
        Dg = dist.block(p).double()
        off = ~torch.eye(n, dtype=torch.bool, device=Dg.device)
        errs.append(((Dg - D64).abs()[off] / D64.abs()[off]).cpu().numpy())
    e = np.concatenate(errs)
    return dict(max=float(e.max()), p99=float(np.percentile(e, 99)), median=float(np.median(e)))


def _blas_limit(threads):
    from threadpoolctl import threadpool_limits
    return threadpool_limits(limits=threads)


def _cpu_unit_costs(stack, df, N, n_img, n_g, nnz, n_ref, k_ref, threads, seed=0):
    """One bounded sample of the oracle's per-part costs on this host (float64 numpy/scipy, the
    reference's operations): per-image prep (load, filter, two 3x3-tile spline rotations, CTF,
    FFTs), the float64 Gram per image pair, the Ferguson sweep per (eps, kernel value) with
    Python's builtin sum as the reference sums, and one eigsh on a kernel-sized sparse matrix.
    Runs with `threads` BLAS threads."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import eigsh
    from oracle import distance as od, dm as odm
    with _blas_limit(threads):
        p = dict(nPix=N, pix_size=1.5, Cs=2.7, EkV=300.0, gaussEnv=np.inf, AmpContrast=0.1)
        ind = np.arange(n_img)
        qq = np.tile(np.array([[1.0], [0], [0], [0]]), (1, n_img))
        t0 = time.perf_counter()
        od.op([ind, qq, df[:n_img], None, 0], FILTER, stack[:n_img], 2 * n_img,
              dict(avgOnly=True), p)
        t_img = (time.perf_counter() - t0) / n_img
        rng = np.random.default_rng(seed)
        fy = rng.normal(size=(n_g, N * N)) + 1j * rng.normal(size=(n_g, N * N))
        C = rng.normal(size=(n_g, N * N))
        t0 = time.perf_counter()
        od.gram_distance(fy, C)
        t_pair = (time.perf_counter() - t0) / (n_g * n_g)
        del fy, C
        vals = rng.uniform(1e6, 2e6, nnz)
        le = np.linspace(-150, 150, 1501)[700:705]       # eps where every value is summed
        t0 = time.perf_counter()
        odm.ferguson_sums(np.sqrt(vals), le, builtin=True)
        t_eps = (time.perf_counter() - t0) / (len(le) * nnz)
        A = sp.random(n_ref, n_ref, density=min(1.0, 2 * k_ref / n_ref), random_state=1)
        A = (A + A.T) / 2
        t0 = time.perf_counter()
        eigsh(A, k=11, maxiter=300)
        t_eig = time.perf_counter() - t0
    return dict(t_img=t_img, t_pair=t_pair, t_eps=t_eps, t_eig=t_eig)


def _cpu_unit_costs_job(args):
    return _cpu_unit_costs(*args)


def _cpu_full_pd(raw, q, df, ind, nStot, N, k, n_eps, num_eigs, threads):
    """The oracle's whole per-PD path (getDistanceCTF...op, then DMembeddingII.op with the
    builtin-sum Ferguson sweep) on one PD with `threads` BLAS threads; returns seconds."""
    from oracle import distance as od, dm as odm
    with _blas_limit(threads):
        p = dict(nPix=N, pix_size=1.5, Cs=2.7, EkV=300.0, gaussEnv=np.inf, AmpContrast=0.1)
        t0 = time.perf_counter()
        out = od.op([ind, q, df, None, 0], FILTER, raw, nStot, dict(avgOnly=False), p)
        odm.op(out["D"], k, 3, 60000, num_eigs=num_eigs, rng=np.random.default_rng(0),
               logEps=np.linspace(-150, 150, n_eps), builtin=True)
        return time.perf_counter() - t0


def _cpu_full_pd_job(args):
    return _cpu_full_pd(*args)


def _pool_run(fn, jobs):
    """Run jobs in spawned worker processes (one per job) and return (results, wall seconds)."""
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    t0 = time.perf_counter()
    with ProcessPoolExecutor(len(jobs), mp_context=mp.get_context("spawn")) as ex:
        res = list(ex.map(fn, jobs))
    return res, time.perf_counter() - t0


def cpu_baseline(meta, batch, workload):
    """The float64 oracle (numpy/scipy restatement of the reference path) timed on this host's
    cores in the reference's two execution modes: A, one process with every BLAS thread
    (GetDistancesS2.py:102-108, ncpu = 1), and B, one single-threaded process per core working
    on separate PDs (GetDistancesS2.py:110-120, the multiprocessing pool).  The one-PD
    plain-method workload runs whole; the larger ones time a bounded sample of each part and
    extrapolate with the cost model stated in `sample`."""
    from threadpoolctl import threadpool_info
    N = batch.N
    ns = np.asarray(batch.ns)
    pairs = float(np.sum(ns * (ns - 1) / 2))
    cores = os.cpu_count() or 1
    threads = sum(x.get("num_threads", 0) for x in threadpool_info()) or 1
    procs = max(1, min(16, cores))
    stack = batch.stack.cpu().numpy()
    df = batch.df.cpu().numpy()
    if workload == "c1":
        n = int(ns[0])
        ind = np.arange(n)
        job = (stack[:n], batch.q_host[0], df[:n], ind, 2 * n, N, n, 101, 8)
        tA = _cpu_full_pd(*job, threads)
        tB, _ = _pool_run(_cpu_full_pd_job, [job + (1,)] * procs)
        wB = max(tB)           # the processes run together; each reports its own PD's time
        per_pd = pairs
        return dict(value=per_pd / tA, unit="pairs/s", cores=int(threads), kind="port",
                    sample=f"the whole one-PD path on the host, float64 oracle: stage 1 "
                           f"(n={n}, {N}x{N}) and DMembeddingII with 101 eps (builtin-sum "
                           f"Ferguson sweep), mode A ({threads} BLAS threads, {tA:.2f} s)",
                    seconds_per_step=tA,
                    mode_b=dict(value=procs * per_pd / wB, unit="pairs/s", processes=procs,
                                wall_s=wB, note=f"{procs} single-threaded processes, one PD "
                                                f"each, run together (no extrapolation; "
                                                f"the slowest process's time)"))
    n_ref, k_ref = int(ns.max()), int(max(batch.ks))
    nnz = int(2 * n_ref * min(k_ref, n_ref))
    n_img, n_g = 16, 600
    args = (stack, df, N, n_img, n_g, nnz, n_ref, k_ref)
    uA = _cpu_unit_costs(*args, threads)
    resB, wB = _pool_run(_cpu_unit_costs_job, [args + (1, p) for p in range(procs)])
    uB = {k: float(np.mean([r[k] for r in resB])) for k in resB[0]}

    def per_step(u):
        t = 0.0
        for n, k in zip(ns, batch.ks):
            n = int(n)
            nnz_p = 2 * n * min(int(k), n)
            t += (n * u["t_img"] + n * n * u["t_pair"] + 1501 * nnz_p * u["t_eps"] +
                  u["t_eig"] * (n / n_ref) ** 2)
        return t

    tA, tB1 = per_step(uA), per_step(uB)
    sample = (f"float64 oracle on this host, bounded sample extrapolated per PD as "
              f"n t_img + n^2 t_pair + 1501 nnz t_eps + t_eig (n/{n_ref})^2: t_img from "
              f"{n_img} images at {N}x{N} (load, filter, two 3x3-tile spline rotations, CTF, "
              f"FFTs), t_pair from a {n_g}^2 float64 Gram, t_eps from 5 eps x {nnz} values "
              f"summed with builtin sum (fergusonE.py:43), t_eig from eigsh k=11 at n={n_ref}; "
              f"mode A {threads} BLAS threads")
    return dict(value=pairs / tA, unit="pairs/s", cores=int(threads), kind="port",
                sample=sample, seconds_per_step=tA, unit_costs_mode_a=uA,
                mode_b=dict(value=procs * pairs / tB1, unit="pairs/s", processes=procs,
                            seconds_per_step_one_process=tB1, unit_costs=uB,
                            note=f"{procs} single-threaded processes timed together on the "
                                 f"same sample, each working on its own PDs (throughput "
                                 f"= {procs} x one process's, extrapolated as mode A)"))


def _ctf():
    from manifoldem_python_amd import synth, _native as nat
    return nat.ctf_params(synth.PIX, synth.CS, synth.EKV, synth.AMPC, np.inf)


def run_c4(args, dev):
    """Forward + backward of the fused distance -> Gaussian kernel -> log-sum op."""
    import torch
    from manifoldem_python_amd import synth, autograd as ag, pipeline as pl, _native as nat
    n, N = 1000, 128
    s = synth.make_pd(synth.seeds(5)[3], n, N, device=dev, as_torch=True)
    ctf = _ctf()
    g = pl.pd_geometry(s.q)
    out = pl.prepare_images(s.raw, torch.arange(n, device=dev),
                            torch.zeros(n, dtype=torch.int32, device=dev),
                            torch.as_tensor(-(180 / math.pi) * g["Psi"], device=dev),
                            torch.full((n,), -g["psi_p"], dtype=torch.float64, device=dev),
                            torch.as_tensor(s.df, device=dev), N, FILTER, ctf, want_images=False)
    fy0 = out["fy"].contiguous()
    df = torch.as_tensor(s.df, device=dev)
    P = n * (n - 1) / 2
    Kh = pl.half_len(N)
    res = {}
    for name, prec in (("fp32", nat.PREC_FP32), ("split-bf16", nat.PREC_SPLIT_BF16)):
        fy = fy0.clone().requires_grad_(True)
        plan = nat.Plan([0], [n], [n], chunks=pl.CHUNKS[prec], device=dev)
        op = pl.Operand(n, N, prec, dev)
        ch = torch.empty((n, Kh), dtype=torch.float32, device=dev)
        nat.ctf_pack(fy0, df, N, ctf, prec, op.Khp, 0, op.hi, op.lo if op.lo is not None else op.hi,
                     ch)
        Dm = pl.ctf_distances(plan, op, fy0, ch, N).block(0)
        med = float(Dm[~torch.eye(n, dtype=torch.bool, device=dev)].median())
        eps = (2.0 ** torch.linspace(-20, 20, 41, dtype=torch.float64, device=dev)) * med
        eps = eps.requires_grad_(True)

        def step():
            fy.grad = None
            eps.grad = None
            L = ag.kernel_logsum(fy, df, eps, ctf, prec)
            L.sum().backward()
            return L

        for _ in range(args.warmup):
            step()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t = (time.perf_counter() - t0) / args.steps
        t0 = time.perf_counter()
        for _ in range(args.steps):
            with torch.no_grad():
                ag.kernel_logsum(fy.detach(), df, eps.detach(), ctf, prec)
        torch.cuda.synchronize()
        t_fwd = (time.perf_counter() - t0) / args.steps
        res[name] = dict(ms_per_step=1e3 * t, forward_ms=1e3 * t_fwd, value=P / t,
                         grad_finite=bool(torch.isfinite(fy.grad).all() and
                                          torch.isfinite(eps.grad).all()))
    out = {
        "metric": "image-pair distances/s through forward + backward of the kernel log-sum op",
        "value": res["fp32"]["value"], "unit": "pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": res["fp32"]["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32 Gram and backward GEMMs, fp64 row sums", "data": "synthetic",
        "config": dict(workload="c4_autograd", particles=n, image=N, n_eps=41,
                       eps="2^-20..2^20 x median(D)", precision="fp32"),
        "gram_flop_per_step": 8.0 * Kh * P,
        "backward_gemm_flop_per_step": 3 * 2.0 * n * n * N * (N // 2 + 1),
        "split_bf16": res["split-bf16"], "fp32": res["fp32"],
    }
    return out


def run_c5(args, dev):
    """PDs of n=1500 160x160 particles: stage 1 + diffusion map, then NLSA along 8 psi."""
    import torch
    from manifoldem_python_amd import synth, batch as B, pipeline as pl, spectral
    from manifoldem_python_amd import _native as nat, autotune
    n, N, con, n_psi = 1500, 160, 50, 8
    npd = max(1, args.c5_pds)
    ss = np.random.SeedSequence(synth.SEED + 5).spawn(npd)
    stacks, pds = [], []
    for p in range(npd):
        rng = np.random.default_rng(ss[p])
        s = synth.make_pd(ss[p], n, N, shift_px=8.0, axis=synth.random_axis(rng), device=dev,
                          as_torch=True)
        stacks.append(s.raw)
        pds.append(dict(src=np.arange(n) + p * n, conj=np.zeros(n, np.int32), q=s.q, df=s.df,
                        k=n))
    batch = B.make_batch(torch.cat(stacks), pds, N, _ctf(), FILTER, dev)
    del stacks
    wsp = B.Workspace(batch, nat.PREC_SPLIT_BF16)
    imgs = torch.empty((batch.rows, N, N), dtype=torch.float32, device=dev)
    rng = np.random.default_rng(0)
    rand = lambda *sh: rng.random(sh)
    ev = nat.HipEvents()
    rows0 = batch.rows0

    def stage1(events=None):
        for p in range(npd):
            sl = slice(int(rows0[p]), int(rows0[p]) + n)
            o = pl.prepare_images(batch.stack, batch.src[sl], batch.conj[sl], batch.rot1[sl],
                                  batch.rot2[sl], batch.df[sl], N, FILTER, batch.ctf,
                                  operand=wsp.operand, row0=int(rows0[p]), fy_out=wsp.fy[sl],
                                  ctf_out=wsp.ctf[sl], want_images=True,
                                  src_max=batch.stack.shape[0] - 1)
            imgs[sl].copy_(o["imgAll"])
        dist = pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, N, ws=wsp.ws,
                                events=events, flags=wsp.flags)
        return dist, B.embed(wsp, dist, num_eigs=n_psi, rand=rand)

    def nlsa_pd(p, dist, res, psis):
        sl = slice(int(rows0[p]), int(rows0[p]) + n)
        D = dist.block(p)
        out = []
        for k in psis:
            od = torch.as_tensor(np.argsort(res[p].psi[:, k]), device=dev)
            out.append(spectral.nlsa(D, od, od, imgs[sl], wsp.ctf[sl], N, con, 8, nS=n, tune=3,
                                     num_eigs=15, rand=rand))
        return out

    dist, res = stage1()
    autotune.tune_gram(wsp, N)
    log("c5 warmup")
    for _ in range(max(1, args.warmup)):
        dist, res = stage1()
        nlsa_pd(0, dist, res, [0])
    torch.cuda.synchronize()
    steps = max(1, args.steps)
    pair = [ev.new_pair() for _ in range(steps)]
    t1 = 0.0
    t_psi = []
    t_start = time.perf_counter()
    for k in range(steps):
        a = time.perf_counter()
        dist, res = stage1(pair[k])
        torch.cuda.synchronize()
        t1 += time.perf_counter() - a
        log(f"c5 step {k}: stage 1 + DM {1e3 * (time.perf_counter() - a):.1f} ms")
        for p in range(npd):
            for j in range(n_psi):
                a = time.perf_counter()
                nlsa_pd(p, dist, res, [j])
                torch.cuda.synchronize()
                t_psi.append(time.perf_counter() - a)
    t_all = time.perf_counter() - t_start
    gram_ms = float(np.mean([ev.elapsed_ms(q) for q in pair]))
    ev.close()
    # every NLSA call of the last step again, split by stage (device synchronised between
    # stages, so these sum to a little more than the timed calls)
    per_call = []
    for p in range(npd):
        sl = slice(int(rows0[p]), int(rows0[p]) + n)
        for j in range(n_psi):
            tm = pl.StageTimer()
            od = torch.as_tensor(np.argsort(res[p].psi[:, j]), device=dev)
            r = spectral.nlsa(dist.block(p), od, od, imgs[sl], wsp.ctf[sl], N, con, 8, nS=n,
                              tune=3, num_eigs=15, rand=rand, timer=tm, keep=True)
            tm.t["dm1.lanczos_steps"] = 1e-3 * r.dm1.lanczos_steps
            tm.t["dm2.lanczos_steps"] = 1e-3 * r.dm2.lanczos_steps
            per_call.append(tm.t)
    stage_names = list(per_call[0])
    t_pd = t_all / steps / npd
    P = n * (n - 1) / 2
    Kh = pl.half_len(N)
    F = 8.0 * Kh * P * npd
    ach = F / (gram_ms / 1e3) / 1e12
    out = {
        "metric": "PDs/s through the full per-PD path with NLSA (8 psi per PD)",
        "value": 1.0 / t_pd, "unit": "PDs/s", "n_gpus": 1, "steps": steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * t_all / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 (split) Gram, fp32 prep/NLSA transforms, fp64 NLSA sums/SVD/eigensolve",
        "data": "synthetic",
        "config": dict(workload="c5_nlsa", n_pd_timed=npd, particles=n, image=N,
                       con_order=con, n_psi=n_psi, topos=8, knn="n (full kernel)"),
        "stage1_dm_ms_per_step": 1e3 * t1 / steps,
        "nlsa_ms_per_psi": dict(mean=1e3 * float(np.mean(t_psi)),
                                min=1e3 * float(np.min(t_psi)), max=1e3 * float(np.max(t_psi))),
        "ms_per_pd": 1e3 * t_pd,
        "extrapolated_200_pd_s": 200 * t_pd,
        "extrapolation_note": "200-PD time = 200 x measured per-PD time of the timed PDs "
                              "(stage 1 batched over them, NLSA run per psi)",
        "nlsa_stage_ms": {k: dict(mean=1e3 * float(np.mean([c[k] for c in per_call])),
                                  max=1e3 * float(np.max([c[k] for c in per_call])))
                          for k in stage_names},
        "gram_kernel_ms": gram_ms,
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": ach / PEAK_BF16_TFLOPS, "traffic": None,
                     "kernel": "k_gram_tiles<split-bf16>"},
    }
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); PDs are "
              "independent, so a multi-GPU run would give each GPU a share of the PD list, "
              "balanced by cost (n^2 K_h + n N^2 log N), with no collective.", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        calibrate_copy()
        return
    if args.prep_images:
        prep_images_only(args.prep_images, max(args.steps, 2))
        return
    if args.optflow:
        optflow_only(args.optflow, max(args.steps, 1), args.warmup, not args.no_cpu)
        return
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    dev = torch.device("cuda")
    nat.load()
    pl.LANCZOS.update(graph=bool(args.lanczos_graph),
                      split=None if args.lanczos_split is None else bool(args.lanczos_split))
    if args.workload in ("c4", "c5"):
        log("timing", args.workload)
        print(json.dumps((run_c4 if args.workload == "c4" else run_c5)(args, dev)), flush=True)
        return
    ev = nat.HipEvents()
    out = measure(args.workload, args, dev, ev, args.steps, args.warmup,
                  with_fp32=not args.no_fp32, with_check=not args.no_check,
                  with_cpu=not args.no_cpu)
    if args.workload == "c2" and not args.no_c3:
        # the batched 500-PD workload, measured in the same run (split-BF16 against FP32)
        torch.cuda.empty_cache()
        c3 = measure("c3", args, dev, ev, max(2, min(args.steps, 3)), 1, with_fp32=True,
                     with_check=not args.no_check, with_cpu=False)
        keep = ("value", "unit", "pds_per_s", "ms_per_step", "steps", "config", "gram_kernel_ms",
                "stage_ms", "gram_tuned", "kappa_flagged", "roofline", "d_rel_err_vs_fp64",
                "speedup_vs_fp32", "lanczos", "ms_per_step_serial",
                "pipelined_vs_serial_max_dlambda")
        out["c3"] = {k: c3[k] for k in keep if k in c3}
        if "fp32" in c3:
            out["c3"]["fp32"] = {k: c3["fp32"][k] for k in ("value", "pds_per_s", "ms_per_step",
                                                            "ms_per_step_serial",
                                                            "gram_kernel_ms", "roofline")}
    ev.close()
    print(json.dumps(out), flush=True)


def measure(workload, args, dev, ev, steps, warmup, with_fp32=True, with_check=True,
            with_cpu=True):
    """One JSON record for a distance + diffusion-map workload (c1, c2, c3)."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    log("building workload", workload)
    batch, meta, num_eigs = build_workload(workload, dev)
    torch.cuda.synchronize()
    logEps = np.linspace(-150, 150, 101) if workload == "c1" else None
    log("timing split-BF16", workload)
    graph = bool(args.graph) if args.graph is not None else True
    # overlapping steps pays where the embedding is launch-bound (one PD per step); on the
    # 500-PD batch both stages are heavy and overlap gained nothing (93.5 vs 93.6 ms)
    pipeline = not args.no_pipeline and workload in ("c1", "c2")
    cu_split = CU_SPLIT if args.cu_split is None else args.cu_split
    bf = run_precision(batch, nat.PREC_SPLIT_BF16, steps, warmup, num_eigs, ev, graph=graph,
                       logEps=logEps, gram_config=args.gram_config, pipeline=pipeline,
                       cu_split=cu_split)
    ns = batch.ns
    pairs = float(np.sum(ns * (ns - 1) / 2))
    Kh = pl.half_len(batch.N)
    flops = 8.0 * Kh * pairs                       # algorithmic: 8 K_h per unordered pair
    exec_flops = issued_gram_flops(bf["wsp"].plan, bf["wsp"].operand.Khp, 3)
    traffic, traffic_note = gram_traffic_lookup(workload, bf["tuned"])
    gram_s = bf["gram_ms"] / 1e3
    achieved = flops / gram_s / 1e12
    out = {
        "metric": "image-pair distances/s & PDs/s, split-BF16 vs FP32 at <=1e-5 rel err vs FP64",
        "value": pairs * steps / bf["t"],
        "unit": "pairs/s",
        "n_gpus": 1, "steps": steps, "warmup": warmup,
        "ms_per_step": 1e3 * bf["t"] / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "bf16x3 (split) Gram, fp32 prep, fp64 kernel + eigensolve",
        "data": "synthetic",
        "config": dict(meta, precision="split-bf16", hip_graph=graph,
                       pipelined_steps=bf["pipeline"],
                       cu_split_per_32=bf["cu_split"],
                       lanczos=dict(pl.LANCZOS), kernel_dtype=str(pl.KERNEL_DTYPE)),
        "pds_per_s": batch.ns.size * steps / bf["t"],
        "distance_ms_per_step": 1e3 * bf["t_dist"] / steps,
        "gram_kernel_ms": bf["gram_ms"],
        "stage_ms": bf["stages"],
        "gram_tuned": bf["tuned"],
        "lanczos": bf["lanczos"],
        "kappa_flagged": bf["flags"],
        "ms_per_step_serial": (1e3 * bf["t_serial"] / steps) if "t_serial" in bf else None,
        "pipelined_vs_serial_max_dlambda": bf.get("pipelined_vs_serial_max_dlambda"),
        "pipelined_note": ("steps overlap on two streams: step k+1's prep + Gram (second buffer "
                           "set) runs during step k's embedding; every step does its PD's whole "
                           "pipeline.  pipelined_vs_serial_max_dlambda compares every step's "
                           "eigenvalues with the same steps run one after another from the same "
                           "random fit starts") if bf["pipeline"] else None,
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic,
                     "traffic_note": traffic_note,
                     "kernel": "k_gram_tiles<split-bf16>",
                     "flop_per_launch": flops,
                     "mfma_util": exec_flops / gram_s / 1e12 / PEAK_BF16_TFLOPS},
    }
    if with_check:
        out["d_rel_err_vs_fp64"] = fp64_error(batch, bf["wsp"], bf["dist"])
    d_sample = None
    if with_cpu:
        # a block of the timed run's D for the end-to-end check in the CPU baseline leg
        n0 = int(batch.ns[0])
        rows = np.unique(np.linspace(0, n0 - 1, min(n0, 24)).astype(np.int64))
        blk = bf["dist"].block(0).double().cpu().numpy()
        d_sample = (rows, blk[np.ix_(rows, rows)])
    del bf
    if with_fp32:
        log("timing FP32", workload)
        f32 = run_precision(batch, nat.PREC_FP32, steps, warmup, num_eigs, ev, graph=graph,
                            logEps=logEps, pipeline=pipeline, cu_split=cu_split)
        a32 = flops / (f32["gram_ms"] / 1e3) / 1e12
        out["fp32"] = {"value": pairs * steps / f32["t"],
                       "ms_per_step": 1e3 * f32["t"] / steps,
                       "ms_per_step_serial": ((1e3 * f32["t_serial"] / steps)
                                              if "t_serial" in f32 else None),
                       "pds_per_s": batch.ns.size * steps / f32["t"],
                       "gram_kernel_ms": f32["gram_ms"],
                       "stage_ms": f32["stages"],
                       "gram_tuned": f32["tuned"],
                       "roofline": {"bound": "mfma", "achieved": a32, "peak": PEAK_FP32_TFLOPS,
                                    "unit": "TFLOP/s", "frac": a32 / PEAK_FP32_TFLOPS,
                                    "mfma_util": issued_gram_flops(
                                        f32["wsp"].plan, f32["wsp"].operand.Khp, 1)
                                    / (f32["gram_ms"] / 1e3) / 1e12 / PEAK_FP32_TFLOPS}}
        if with_check:
            out["fp32"]["d_rel_err_vs_fp64"] = fp64_error(batch, f32["wsp"], f32["dist"])
        out["speedup_vs_fp32"] = out["value"] / out["fp32"]["value"]
        del f32
    if with_cpu:
        log("timing CPU baseline", workload)
        out["cpu_baseline"] = cpu_baseline(meta, batch, workload)
        if d_sample is not None:
            out["cpu_baseline"]["e2e_check"] = e2e_check(batch, *d_sample)
    return out


def e2e_check(batch, rows, D_gpu):
    """Relative error of the timed run's D on a sample of PD 0's images against the float64
    oracle run from the same raw images (its prep a1-a6 with the PD geometry of all of PD 0's
    particles, then its Gram, getDistanceCTF...:244-397): the whole path, not the Gram alone."""
    from oracle import distance as od
    N = batch.N
    n0 = int(batch.ns[0])
    src = batch.src[:n0].cpu().numpy()
    conj = batch.conj[:n0].cpu().numpy().astype(bool)
    nStot = 2 * int(batch.stack.shape[0])
    ind = np.where(conj, src + nStot // 2, src)
    stack = batch.stack[:int(src.max()) + 1].cpu().numpy()
    p = dict(nPix=N, pix_size=1.5, Cs=2.7, EkV=300.0, gaussEnv=np.inf, AmpContrast=0.1)
    pr = od.prep_rows([ind, batch.q_host[0], batch.df[:n0].cpu().numpy()], FILTER, stack,
                      nStot, p, rows)
    D64 = od.gram_distance(pr["fy"], pr["CTF"])
    off = ~np.eye(len(rows), dtype=bool)
    e = np.abs(D_gpu - D64)[off] / np.abs(D64)[off]
    return {"images": int(len(rows)), "rel_err_max": float(e.max()),
            "rel_err_median": float(np.median(e)),
            "note": "D of a sample of PD 0's images: this run's split-BF16 path from raw images "
                    "vs the float64 oracle's prep + Gram of the same raw images"}


def gram_traffic_lookup(workload, tuned):
    """HBM bytes per Gram launch for the (tile, variant, chunks) this run tuned to, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE table (a lookup, not a counter read of this
    run; null when the table has no row for the configuration)."""
    tfile = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                         "k_gram_tiles_traffic.json")
    if not os.path.exists(tfile):
        return None, None
    with open(tfile) as f:
        tr = json.load(f)
    key = f"{workload}:t{tuned.get('tile', 128)}v{tuned['best']}c{tuned['chunks']}"
    cfg = tr["configs"].get(key)
    if cfg is None:
        return None, f"no counter row for {key} in profiles/k_gram_tiles_traffic.json"
    return cfg["bytes_per_launch"], (
        f"lookup: rocprofv3 FETCH_SIZE x {tr['fetch_scale']:.3f} + WRITE_SIZE x "
        f"{tr['write_scale']:.3f} (copy-calibrated) per launch of {key}, measured by "
        f"{tr['source']}")


if __name__ == "__main__":
    main()

"""GPU parity of Stage 2 (kNN, Ferguson sweep, Markov normalisation, eigensolve)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _distance_matrix(n, seed):
    """A realistic symmetric distance matrix: squared distances of noisy points on a curve."""
    rng = np.random.default_rng(seed)
    t = np.sort(rng.uniform(0, 1, n))
    X = np.stack([np.cos(3 * t), np.sin(3 * t), t], axis=1) + 0.05 * rng.normal(size=(n, 3))
    D = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1) * 1.0e4
    return D.astype(np.float32).astype(np.float64)   # representable in float32 exactly


def _plan(n, k):
    from manifoldem_python_amd import _native as nat
    return nat.Plan([0], [n], [k], chunks=1)


@pytest.mark.parametrize("n,k", [(200, 40), (333, 150), (129, 129)])
def test_knn_lists_match_sort(cuda, n, k):
    import torch
    from manifoldem_python_amd import _native as nat
    from oracle import dm
    D = _distance_matrix(n, 3)
    yInd1, yVal1 = dm.initialize(n, k, D.copy())
    plan = _plan(n, k)
    Dg = torch.as_tensor(D.astype(np.float32).ravel(), device=cuda)
    mask = torch.empty(plan.info.mask_words, dtype=torch.int32, device=cuda)
    idx = torch.empty(plan.info.list_elems, dtype=torch.int32, device=cuda)
    val = torch.empty(plan.info.list_elems, dtype=torch.float32, device=cuda)
    nat.knn_select(plan, Dg, mask, idx, val)
    gi = idx.cpu().numpy().reshape(k, n)
    gv = val.cpu().numpy().reshape(k, n)
    ri = yInd1.reshape(n, k).T
    rv = yVal1.reshape(n, k).T
    np.testing.assert_array_equal(gi, ri)
    np.testing.assert_allclose(gv, rv, rtol=1e-7)
    # bitmask rows agree with the lists
    W = (n + 31) // 32
    m = mask.cpu().numpy().view(np.uint32).reshape(n, W)
    bits = np.unpackbits(m.view(np.uint8).reshape(n, W, 4)[:, :, ::-1], axis=2,
                         bitorder="big").reshape(n, W * 32)
    # big-endian unpack of little-endian words: rebuild bit j of word w
    member = np.zeros((n, n), bool)
    for i in range(n):
        for w in range(W):
            word = int(m[i, w])
            for b in range(32):
                j = 32 * w + b
                if j < n and (word >> b) & 1:
                    member[i, j] = True
    ref = np.zeros((n, n), bool)
    for i in range(n):
        ref[i, ri[:, i]] = True
    np.testing.assert_array_equal(member, ref)
    del bits


@pytest.mark.parametrize("n,k,n_eps", [(200, 200, 101), (300, 60, 1501)])
def test_dm_embed_matches_oracle(cuda, n, k, n_eps):
    import torch
    from manifoldem_python_amd import pipeline as pl
    from oracle import dm
    D = _distance_matrix(n, 11)
    logEps = np.linspace(-150, 150, n_eps) if n_eps != 1501 else dm.LOG_EPS
    rng_o = np.random.default_rng(7)
    ref = dm.op(D, k, 3, 60000, num_eigs=8, rng=rng_o, logEps=logEps)
    rng_g = np.random.default_rng(7)
    plan = _plan(n, k)
    Dg = torch.as_tensor(D.astype(np.float32).ravel(), device=cuda)
    res = pl.dm_embed(plan, Dg, tune=3, num_eigs=8, logEps=logEps,
                      rand=lambda *sh: rng_g.random(sh))[0]
    lamb, psi, sigma, mu, _, lsw, popt, R2 = ref
    assert np.max(np.abs(res.logSumWij - lsw)) < 1e-6
    assert abs(res.sigma - sigma) / sigma < 1e-5
    assert np.max(np.abs(res.lamb - lamb)) < 1e-6
    # eigenvectors up to sign: compare psi via the subspace of the leading vectors
    Vr = np.column_stack([np.sqrt(mu)] + [psi[:, j] * np.sqrt(mu) for j in range(psi.shape[1])])
    Vg = res.V
    for j in range(Vg.shape[1]):
        c = abs(np.dot(Vr[:, j] / np.linalg.norm(Vr[:, j]), Vg[:, j]))
        assert c > 1 - 1e-6, (j, c)
    np.testing.assert_allclose(res.mu, mu, rtol=1e-5, atol=1e-12)


def test_ferguson_slope_matches_finite_differences(cuda):
    import torch
    from manifoldem_python_amd import pipeline as pl
    n = 250
    D = _distance_matrix(n, 5)
    # keep every x = D/(2 eps) below the threshold 10 so log S is smooth in eps
    logEps = np.linspace(np.log(D.max() / 20) + 0.5, 25, 301)
    plan = _plan(n, n)
    Dg = torch.as_tensor(D.astype(np.float32).ravel(), device=cuda)
    S, slope, _ = pl.ferguson_sums(plan, Dg, pl.knn_mask(plan, Dg), logEps)
    # float64 finite differences of log S(eps) on the same kernel entries
    Df = D.astype(np.float32).astype(np.float64)
    v = Df[~np.eye(n, dtype=bool)]
    v = v[v >= 1e-6]
    nz = n  # diagonal zero entries

    def logS(le):
        x = v / (2 * np.exp(le))
        return np.log(nz + np.sum(np.exp(-x[x < 10])))

    h = 1e-4
    fd = np.array([(logS(le + h) - logS(le - h)) / (2 * h) for le in logEps])
    assert np.max(np.abs(slope[0] - fd)) < 1e-6
    assert np.max(np.abs(np.log(S[0]) - np.array([logS(le) for le in logEps]))) < 1e-6


@pytest.mark.parametrize("split,check,ns,kk", [
    (False, "gpu", [150, 400, 257], [12, 16, 9]),
    (True, "gpu", [150, 400, 257], [12, 16, 9]),
    (False, "host", [150, 400, 257], [12, 16, 9]),
    # two PDs exhaust their n < m steps before the first check
    (True, "gpu", [20, 333, 16], [8, 14, 6]),
    (False, "gpu", [20, 333, 16], [8, 14, 6]),
])
def test_batched_lanczos_matches_dense(cuda, split, check, ns, kk):
    """Three PDs of ragged sizes in one plan: GPU Lanczos eigenpairs (largest magnitude)
    against numpy eigh of each normalised kernel block."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    ks = [n if n < 300 else 120 for n in ns]
    D = [_distance_matrix(n, 20 + i) for i, n in enumerate(ns)]
    rows0 = np.concatenate(([0], np.cumsum(ns)[:-1]))
    plan = nat.Plan(rows0, ns, ks, chunks=1)
    flat = torch.as_tensor(np.concatenate([d.astype(np.float32).ravel() for d in D]),
                           device=cuda)
    mask = pl.knn_mask(plan, flat)
    sigma = torch.as_tensor([np.sqrt(np.median(d)) for d in D], device=cuda)
    L = torch.empty(plan.info.d_elems, dtype=torch.float32, device=cuda)
    deg = torch.empty(2 * plan.info.n_rows, dtype=torch.float64, device=cuda)
    nat.markov_normalize(plan, flat, mask, sigma, L, deg)
    res = pl.lanczos_eigs(plan, L, kk, split=split, check=check)
    for p in range(3):
        A = plan.block(L, p).double().cpu().numpy()
        assert np.allclose(A, A.T)
        w, U = np.linalg.eigh(A)
        sel = np.argsort(-np.abs(w), kind="stable")[:kk[p]]
        sel = sel[np.argsort(-w[sel], kind="stable")]
        lam, Y, J, r = res[p]
        np.testing.assert_allclose(lam, w[sel], atol=1e-9)
        assert Y.shape == (ns[p], kk[p])
        for c in range(kk[p]):
            assert abs(abs(np.dot(Y[:, c], U[:, sel[c]])) - 1) < 1e-6, (p, c)


@pytest.mark.parametrize("split", [True, False])
def test_sparse_kernel_lanczos_matches_dense(cuda, split):
    """The sparse rows of a kNN kernel (memi_kernel_ell) rebuild the dense float64 kernel
    exactly, and Lanczos over them (memi_lanczos_steps_sparse) gives the dense path's
    eigenpairs, for PDs with k << n (one of 2000 rows, as the single-large-PD workload)."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    ns, ks, kk = [2000, 600, 333], [200, 60, 40], [11, 9, 7]
    D = [_distance_matrix(n, 60 + i) for i, n in enumerate(ns)]
    rows0 = np.concatenate(([0], np.cumsum(ns)[:-1]))
    plan = nat.Plan(rows0, ns, ks, chunks=1)
    flat = torch.as_tensor(np.concatenate([d.astype(np.float32).ravel() for d in D]),
                           device=cuda)
    mask = pl.knn_mask(plan, flat)
    sigma = torch.as_tensor([np.sqrt(np.median(d)) for d in D], device=cuda)
    L = torch.empty(plan.info.d_elems, dtype=torch.float64, device=cuda)
    deg = torch.empty(2 * plan.info.n_rows, dtype=torch.float64, device=cuda)
    nat.markov_normalize(plan, flat, mask, sigma, L, deg)
    col = torch.empty(plan.info.d_elems, dtype=torch.int32, device=cuda)
    val = torch.empty(plan.info.d_elems, dtype=torch.float64, device=cuda)
    cnt = torch.empty(plan.info.n_rows, dtype=torch.int32, device=cuda)
    nat.kernel_ell(plan, L, col, val, cnt)
    cnt_h = cnt.cpu().numpy()
    for p in range(3):
        n = ns[p]
        A = plan.block(L, p).cpu().numpy()
        c = plan.block(col, p).cpu().numpy()
        v = plan.block(val, p).cpu().numpy()
        R = np.zeros((n, n))
        for i in range(n):
            k = cnt_h[rows0[p] + i]
            assert k == np.count_nonzero(A[i])
            assert np.all(np.diff(c[i, :k]) > 0)
            R[i, c[i, :k]] = v[i, :k]
        np.testing.assert_array_equal(R, A)
        assert cnt_h[rows0[p]:rows0[p] + n].max() < n // 2      # the sparse case
    dense = pl.lanczos_eigs(plan, L, kk, split=split)
    sparse = pl.lanczos_eigs(plan, (col, val, cnt), kk, split=split)
    for p in range(3):
        ld, Yd, Jd, _ = dense[p]
        ls, Ys, Js, _ = sparse[p]
        np.testing.assert_allclose(ls, ld, rtol=0, atol=1e-12)
        assert abs(Js - Jd) <= 8
        for q in range(kk[p]):
            assert abs(abs(Ys[:, q] @ Yd[:, q]) - 1) < 1e-8, (p, q)


@pytest.mark.parametrize("split", [True, False])
def test_mixed_sparse_dense_lanczos_matches_dense(cuda, split):
    """A batch whose PDs are partly sparse (k <= n/4) and partly dense: memi_kernel_ell with a
    selection writes the selected PDs' rows only, and Lanczos with the selection (sparse matvec
    for those, dense for the rest) gives the all-dense eigenpairs."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    ns, ks, kk = [900, 300, 640, 150], [150, 150, 150, 120], [9, 9, 9, 9]
    D = [_distance_matrix(n, 80 + i) for i, n in enumerate(ns)]
    rows0 = np.concatenate(([0], np.cumsum(ns)[:-1]))
    plan = nat.Plan(rows0, ns, ks, chunks=1)
    flat = torch.as_tensor(np.concatenate([d.astype(np.float32).ravel() for d in D]),
                           device=cuda)
    mask = pl.knn_mask(plan, flat)
    sigma = torch.as_tensor([np.sqrt(np.median(d)) for d in D], device=cuda)
    L = torch.empty(plan.info.d_elems, dtype=torch.float64, device=cuda)
    deg = torch.empty(2 * plan.info.n_rows, dtype=torch.float64, device=cuda)
    nat.markov_normalize(plan, flat, mask, sigma, L, deg)
    sel_h = np.array([k * pl.SPARSE_RATIO <= n for k, n in zip(ks, ns)], dtype=np.int32)
    assert sel_h.tolist() == [1, 0, 1, 0]
    sel = torch.as_tensor(sel_h, device=cuda)
    col = torch.full((plan.info.d_elems,), -7, dtype=torch.int32, device=cuda)
    val = torch.empty(plan.info.d_elems, dtype=torch.float64, device=cuda)
    cnt = torch.full((plan.info.n_rows,), -7, dtype=torch.int32, device=cuda)
    nat.kernel_ell(plan, L, col, val, cnt, sel=sel)
    cnt_h = cnt.cpu().numpy()
    for p in range(len(ns)):
        n = ns[p]
        c = plan.block(col, p).cpu().numpy()
        if sel_h[p]:
            A = plan.block(L, p).cpu().numpy()
            v = plan.block(val, p).cpu().numpy()
            for i in range(n):
                k = cnt_h[rows0[p] + i]
                assert k == np.count_nonzero(A[i])
                np.testing.assert_array_equal(v[i, :k], A[i, c[i, :k]])
        else:                                          # untouched
            assert np.all(cnt_h[rows0[p]:rows0[p] + n] == -7) and np.all(c == -7)
    dense = pl.lanczos_eigs(plan, L, kk, split=split)
    mixed = pl.lanczos_eigs(plan, (col, val, cnt, sel, L), kk, split=split)
    for p in range(len(ns)):
        ld, Yd, Jd, _ = dense[p]
        lm, Ym, Jm, _ = mixed[p]
        np.testing.assert_allclose(lm, ld, rtol=0, atol=1e-12)
        assert abs(Jm - Jd) <= 8
        for q in range(kk[p]):
            assert abs(abs(Ym[:, q] @ Yd[:, q]) - 1) < 1e-8, (p, q)


def test_gpu_tanh_fit_matches_curve_fit(cuda):
    """Batched LM fit of the Ferguson curve against scipy curve_fit (the reference's fitter)
    on sweeps of several realistic kernels, from the same start points."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    from oracle import dm
    curves, a0s = [], []
    rng = np.random.default_rng(12)
    for seed in range(6):
        D = _distance_matrix(150 + 20 * seed, 40 + seed)
        yRow, yCol, yVal = dm.assemble(D.copy(), 60)
        curves.append(dm.ferguson_sums(np.sqrt(yVal), dm.LOG_EPS)[0])
        a0s.append(rng.random((4, 1)) - 0.5)
    logS = np.stack(curves)
    r1 = np.random.default_rng(3)
    popt, R2 = pl.ferguson_fit_batch(dm.LOG_EPS, logS, a0s, rand=lambda *sh: r1.random(sh))
    r2 = np.random.default_rng(3)
    for p in range(len(curves)):
        ref, _, R2r = dm.ferguson_fit(dm.LOG_EPS, logS[p], a0s[p], r2)
        s_ref = np.sqrt(2 * np.exp(-ref[1] / ref[0]))
        s_gpu = np.sqrt(2 * np.exp(-popt[p][1] / popt[p][0]))
        assert abs(s_gpu - s_ref) / s_ref < 1e-6, (p, s_gpu, s_ref)
        assert abs(R2[p] - R2r) < 1e-9


def test_fit_bandwidths_restart_path_matches_curve_fit(cuda):
    """fit_bandwidths when a PD's own start fails the restart rule: the host restart loop takes
    over from the device's first fit (not fitting it again) and gives the sequential
    reference's bandwidth and R^2 from the same random draws.  One PD per call, so that the
    reference's global draw order (a draw after every fit) and the batch's agree."""
    import torch
    from manifoldem_python_amd import pipeline as pl
    from oracle import dm
    rng = np.random.default_rng(21)
    restarted = 0
    for seed in range(4):
        D = _distance_matrix(140 + 25 * seed, 70 + seed)
        yRow, yCol, yVal = dm.assemble(D.copy(), 60)
        logS = dm.ferguson_sums(np.sqrt(yVal), dm.LOG_EPS)[0][None, :]
        a0 = (rng.random((4, 1)) - 0.5) * 40.0          # far start: huge pcov, the rule refits
        class Counted:                                  # the oracle draws rng.random(shape)
            def __init__(self, seed):
                self.r, self.n = np.random.default_rng(seed), 0

            def random(self, sh):
                self.n += 1
                return self.r.random(sh)

        r2 = Counted(5 + seed)
        pr, _, R2r = dm.ferguson_fit(dm.LOG_EPS, logS[0], a0, r2)
        restarted += r2.n > 1                           # a draw after every fit: > 1 fit
        r1 = np.random.default_rng(5 + seed)
        sigma, finish = pl.fit_bandwidths(dm.LOG_EPS, torch.as_tensor(logS, device=cuda), [a0],
                                          1.0, rand=lambda *sh: r1.random(sh))
        popt, R2 = finish(logS)
        s_ref = np.sqrt(2 * np.exp(-pr[1] / pr[0]))
        s_gpu = float(sigma.cpu().numpy()[0])
        assert abs(s_gpu - s_ref) / s_ref < 1e-6, (seed, s_gpu, s_ref)
        assert abs(R2[0] - R2r) < 1e-9, seed
    assert restarted >= 2, "too few far starts failed: the test does not reach the fallback"


def test_tanh_select_matches_host_rule(cuda):
    """memi_tanh_select against the host's form of the restart rule (numpy inv of J^T J,
    sum sqrt|diag pcov| <= 100, the first passing start) and sigma = tune sqrt(2 exp(-b/a)),
    on fits from good and bad starts (several per PD), plus an exactly singular J^T J."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    from oracle import dm
    curves = []
    for seed in range(4):
        D = _distance_matrix(120 + 30 * seed, 50 + seed)
        yRow, yCol, yVal = dm.assemble(D.copy(), 50)
        curves.append(dm.ferguson_sums(np.sqrt(yVal), dm.LOG_EPS)[0])
    logS = np.stack(curves)
    npd, m = logS.shape
    R = 6
    rng = np.random.default_rng(7)
    starts = rng.random((npd, R, 4)) - 0.5
    starts[:, 1] *= 40.0                     # far starts: some end with huge pcov
    x = torch.as_tensor(dm.LOG_EPS, device=cuda)
    y = torch.as_tensor(np.repeat(logS, R, axis=0), device=cuda)
    out = torch.empty(npd * R * 16, dtype=torch.float64, device=cuda)
    nat.tanh_fit(x, y, torch.as_tensor(starts.reshape(-1, 4), device=cuda), out)
    o = out.cpu().numpy().reshape(npd, R, 16)
    o[3, 0, 4:14] = 0.0                      # an exactly singular J^T J fails the test
    o[3, 0, 4] = 1.0
    out = torch.as_tensor(o.reshape(-1), device=cuda)
    ok = torch.empty(npd, dtype=torch.int32, device=cuda)
    sig = torch.full((npd,), -1.0, dtype=torch.float64, device=cuda)
    res = torch.empty(npd * R, dtype=torch.float64, device=cuda)
    nat.tanh_select(out, npd, R, m, 0.7, 100.0, ok, sig, res)
    iu = np.triu_indices(4)
    J = np.zeros((npd, R, 4, 4))
    J[:, :, iu[0], iu[1]] = o[:, :, 4:14]
    J = J + np.triu(J, 1).transpose(0, 1, 3, 2)
    with np.errstate(all="ignore"):
        diag = np.diagonal(pl.batched_inv(J.reshape(-1, 4, 4)), axis1=1, axis2=2)
        want = np.sum(np.sqrt(np.fabs(diag.reshape(npd, R, 4) *
                                      (o[:, :, 14] / (m - 4))[:, :, None])), axis=2)
    want[~np.isfinite(want)] = np.inf
    got = res.cpu().numpy().reshape(npd, R)
    fin = np.isfinite(want)
    assert np.array_equal(fin, np.isfinite(got))
    # well-conditioned fits agree closely; from nearly singular ones (values 1e7 and up) any
    # two inverses differ widely, and only the side of the threshold matters
    wc = fin & (want < 1e3)
    np.testing.assert_allclose(got[wc], want[wc], rtol=1e-7)
    np.testing.assert_array_equal(got <= 100, want <= 100)
    assert not fin[3, 0]
    okw = want <= 100
    assert okw.any() and not okw.all()       # both outcomes occur
    pick = np.where(okw.any(axis=1), np.argmax(okw, axis=1) + 1, 0)
    np.testing.assert_array_equal(ok.cpu().numpy(), pick)
    sg = sig.cpu().numpy()
    for p in range(npd):
        if pick[p]:
            a = o[p, pick[p] - 1]
            assert abs(sg[p] - 0.7 * np.sqrt(2 * np.exp(-a[1] / a[0]))) <= 1e-14 * sg[p]
        else:
            assert sg[p] == -1.0


def test_ferguson_moments_match_direct_sweep(cuda):
    """Bucket-moment sums (memi_ferguson_moments) against the direct sweep over every entry at
    every eps, for a batch of PDs with kNN and full kernels, duplicate points (zero entries)
    and value spreads from one to eight decades, on the default 1501-point grid, a coarser
    one, a fine one (dle = 0.01: eight decades need more buckets than the moments carry) and a
    very coarse one (dle = 2: too coarse for the series); the last two go to the direct sweep
    inside memi_ferguson_moments."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    ns, ks = [150, 400, 257, 90], [150, 60, 257, 30]
    D = [_distance_matrix(n, 40 + i) for i, n in enumerate(ns)]
    D[1][5, 6] = D[1][6, 5] = 0.0                     # a duplicate pair: zero entries
    D[2] = D[2] ** 4 / 1e12                           # eight decades of values
    rows0 = np.concatenate(([0], np.cumsum(ns)[:-1]))
    plan = nat.Plan(rows0, ns, ks, chunks=1)
    flat = torch.as_tensor(np.concatenate([d.astype(np.float32).ravel() for d in D]),
                           device=cuda)
    mask = pl.knn_mask(plan, flat)
    for le in (pl.LOG_EPS, np.arange(-40, 40.5, 0.5), np.arange(-60, 60.005, 0.01),
               np.arange(-150, 151, 2.0)):
        S1, sl1, c1 = pl.ferguson_sums(plan, flat, mask, le, method="moments")
        S0, sl0, c0 = pl.ferguson_sums(plan, flat, mask, le, method="direct")
        np.testing.assert_array_equal(c1, c0)
        ok = S0 > 0
        # the direct sweep takes exp(-x) through a float32 exp2 of the fractional part
        # (~1e-7 per term); the moment sums are float64 throughout
        assert np.max(np.abs(np.log(S1[ok]) - np.log(S0[ok]))) < 1e-7
        assert np.max(np.abs(sl1[ok] - sl0[ok])) < 1e-6


@pytest.mark.parametrize("split", [True, False])
def test_batched_lanczos_many_pds(cuda, split):
    """80 PDs of 40-160 points: the many-PD orthogonalisation (waves walking the live
    (PD, column) pairs, row-list updates) against dense eigh per PD."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    rng = np.random.default_rng(9)
    ns = rng.integers(40, 160, size=80)
    ks = np.minimum(ns, 30)
    D = [_distance_matrix(int(n), 100 + i) for i, n in enumerate(ns)]
    rows0 = np.concatenate(([0], np.cumsum(ns)[:-1]))
    plan = nat.Plan(rows0, ns, ks, chunks=1)
    flat = torch.as_tensor(np.concatenate([d.astype(np.float32).ravel() for d in D]),
                           device=cuda)
    mask = pl.knn_mask(plan, flat)
    sigma = torch.as_tensor([np.sqrt(np.median(d)) for d in D], device=cuda)
    L = torch.empty(plan.info.d_elems, dtype=torch.float32, device=cuda)
    deg = torch.empty(2 * plan.info.n_rows, dtype=torch.float64, device=cuda)
    nat.markov_normalize(plan, flat, mask, sigma, L, deg)
    kk = [9] * len(ns)
    res = pl.lanczos_eigs(plan, L, kk, split=split)
    for p in range(0, len(ns), 7):
        A = plan.block(L, p).double().cpu().numpy()
        w, U = np.linalg.eigh(A)
        sel = np.argsort(-np.abs(w), kind="stable")[:9]
        sel = sel[np.argsort(-w[sel], kind="stable")]
        lam, Y, J, r = res[p]
        np.testing.assert_allclose(lam, w[sel], atol=1e-9)
        for c in range(9):
            assert abs(abs(np.dot(Y[:, c], U[:, sel[c]])) - 1) < 1e-6, (p, c)

"""Pins of the Stage-2 oracle (oracle/dm.py): brute force, closed forms, invariants.  CPU only."""
import os

import numpy as np
import pytest

from oracle import dm


def _D(n, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, 3))
    D = ((X[:, None] - X[None]) ** 2).sum(-1) * 100.0
    np.fill_diagonal(D, 0.0)
    return D


def test_initialize_is_columnwise_sort_with_self_first():
    D = _D(8, 1)
    n, k = 8, 5
    yInd1, yVal1 = dm.initialize(n, k, D.copy())
    I = yInd1.reshape(n, k)       # row i = point i (F-order flatten of k x n)
    V = yVal1.reshape(n, k)
    for i in range(n):
        col = D[:, i].copy()
        col[i] = -np.inf
        order = np.argsort(col)[:k]
        assert I[i, 0] == i and V[i, 0] == 0.0
        np.testing.assert_array_equal(I[i], order)
        np.testing.assert_allclose(V[i, 1:], D[order[1:], i])


def test_assemble_symmetrised_entries_brute_force():
    n, k = 8, 3
    D = _D(n, 2)
    D[0, 5] = D[5, 0] = 1e-9   # a "zero" entry off the diagonal
    yRow, yCol, yVal = dm.assemble(D.copy(), k)
    got = {}
    for r, c, v in zip(yRow, yCol, yVal):
        assert (r, c) not in got
        got[(r, c)] = v
    # brute force: kNN lists (self first), zero entries (val < 1e-6) on the lists,
    # non-zero entries symmetrised: d^2 where j on i's list or i on j's list
    lists = []
    for i in range(n):
        col = D[:, i].copy()
        col[i] = -np.inf
        lists.append(list(np.argsort(col)[:k]))
    expect = {}
    for i in range(n):
        for j in lists[i]:
            if j == i or D[j, i] < 1e-6:
                expect[(i, j)] = 0.0
    for i in range(n):
        for j in lists[i]:
            if j != i and D[j, i] >= 1e-6:
                expect[(i, j)] = D[i, j]
                expect[(j, i)] = D[i, j]
    assert set(got) == set(expect)
    for key, v in expect.items():
        assert got[key] == pytest.approx(v, rel=1e-12)


def test_ferguson_limits_and_slope():
    rng = np.random.default_rng(3)
    v = np.concatenate([np.zeros(7), rng.uniform(1.0, 5.0, 200)])
    le = np.array([-150.0, -140.0, 140.0, 150.0])
    s, thr = dm.ferguson_sums(np.sqrt(v), le)
    assert thr == 10.0
    # eps -> 0: only the zero entries survive; eps -> inf: every entry counts exp(0) = 1
    np.testing.assert_allclose(s[:2], np.log(7.0))
    np.testing.assert_allclose(s[2:], np.log(207.0))
    # d log S / d log eps = sum(W x) / sum(W) (all x < thr for this range)
    le = np.linspace(2.0, 4.0, 11)
    s, _ = dm.ferguson_sums(np.sqrt(v), le)
    h = 1e-5
    sp, _ = dm.ferguson_sums(np.sqrt(v), le + h)
    sm, _ = dm.ferguson_sums(np.sqrt(v), le - h)
    fd = (sp - sm) / (2 * h)
    x = v[None, :] / (2 * np.exp(le)[:, None])
    an = np.sum(np.exp(-x) * x, 1) / np.sum(np.exp(-x), 1)
    np.testing.assert_allclose(fd, an, atol=1e-8)


def test_normalisation_invariants():
    n = 30
    D = _D(n, 4)
    yRow, yCol, yVal = dm.assemble(D.copy(), n)
    sigma = np.sqrt(np.median(D))
    l = dm.slaplacian(yVal, yCol, yRow, n, sigma).toarray()
    assert np.allclose(l, l.T)
    # alpha = 1: the normalised kernel is similar to a Markov matrix: top eigenvalue 1,
    # eigenvector proportional to sqrt of the column sums of W/(d d)
    W = np.exp(-D / sigma ** 2)
    d = W.sum(0)
    W1 = W / np.outer(d, d)
    d1 = W1.sum(0)
    w, U = np.linalg.eigh(l)
    assert w[-1] == pytest.approx(1.0, abs=1e-12)
    u = np.sqrt(d1) / np.linalg.norm(np.sqrt(d1))
    assert abs(abs(u @ U[:, -1]) - 1) < 1e-12
    lam, V = dm.top_eigs(dm.slaplacian(yVal, yCol, yRow, n, sigma), 6)
    lam_d, V_d = dm.top_eigs(l, 6, dense=True)
    np.testing.assert_allclose(lam, lam_d, atol=1e-10)
    mu = V[:, 0] ** 2
    assert mu.sum() == pytest.approx(1.0)


def test_golden_dm_fixture():
    path = os.path.join(os.path.dirname(__file__), "golden", "dm_n40_k25.npz")
    g = np.load(path, allow_pickle=False)
    lamb, psi, sigma, mu, _, lsw, popt, R2 = dm.op(g["D"], 25, 3, 60000, num_eigs=8,
                                                  rng=np.random.default_rng(6), dense=True,
                                                  logEps=g["logEps"])
    np.testing.assert_allclose(lsw, g["logSumWij"], rtol=1e-12)
    assert sigma == pytest.approx(float(g["sigma"]), rel=1e-8)
    np.testing.assert_allclose(lamb, g["lamb"], atol=1e-10)
    np.testing.assert_allclose(mu, g["mu"], atol=1e-10)


def test_trim_keeps_points_inside_radius():
    n = 60
    D = _D(n, 5)
    D[0] *= 50.0
    D[:, 0] *= 50.0
    pos, res = dm.trim(D, 3, 5, num_eigs=8, rng=np.random.default_rng(1), dense=True)
    psi = res[1]
    assert np.all(np.sqrt(np.sum(psi[:, :3] ** 2, 1)) < 5)
    assert len(pos) == psi.shape[0]


def test_trim_removes_the_outlier_at_rad5():
    """manifoldTrimmingAuto.py:48-77 at the default rad = 5: the point off the end of the
    curve lies outside the radius (||psi_1..3|| ~ 6.6, curve points < 2.1), so one re-embedding
    runs on the curve alone and removes nothing; its eigenvalues are the positive ones."""
    from conftest import curve_with_outlier
    D = curve_with_outlier()
    n = D.shape[0]
    res0 = dm.op(D, n, 3, 60000, 15, np.random.default_rng(1))
    r0 = np.sqrt(np.sum(res0[1][:, :3] ** 2, 1))
    assert r0[-1] > 6 and np.max(r0[:-1]) < 2.5
    pos, res = dm.trim(D, 3, 5, num_eigs=15, rng=np.random.default_rng(1))
    np.testing.assert_array_equal(pos, np.arange(n - 1))
    assert res[1].shape[0] == n - 1 and np.all(res[0] > 0)
    ref = dm.op(D[:n - 1, :n - 1], n - 1, 3, 600000, 15, np.random.default_rng(1))
    assert res[2] == pytest.approx(ref[2], rel=1e-6)            # sigma (fit restarts differ)
    np.testing.assert_allclose(res[0], ref[0][ref[0] > 0], atol=1e-6)


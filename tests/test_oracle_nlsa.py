"""Pins of the Stage-3 oracle (oracle/nlsa.py, oracle/fit1d.py).  CPU only."""
import numpy as np
import pytest
from scipy.fftpack import fft2, ifft2

from oracle import fit1d, nlsa


def test_condist_matches_triple_loop():
    rng = np.random.default_rng(1)
    n, C = 17, 4
    DD = rng.normal(size=(n, n))
    ConD = nlsa.condist(DD, n, C)
    m = n - C
    brute = np.zeros((m, m))
    for a in range(m):                      # NLSA.py:36-41 (commented reference form)
        for b in range(m):
            for c in range(C):
                brute[a, b] += DD[a + c, b + c]
    np.testing.assert_allclose(ConD, brute, rtol=1e-14, atol=1e-14)


def test_wiener_window_offsets():
    rng = np.random.default_rng(2)
    n, C, N = 9, 3, 4
    CTF = rng.normal(size=(n, N, N))
    W, CTF1 = nlsa.get_wiener(CTF, np.arange(n), np.arange(n), C, n)
    for i in range(n - C):
        expect = sum(CTF[C - ii + i] ** 2 for ii in range(C)) + 0.2
        np.testing.assert_allclose(W[i], expect)


def test_supervectors_by_linearity():
    """The per-(ii, i) FFT loop of NLSA.py:66-86 equals one inverse FFT per (ii, column) of the
    mu_psi-weighted spectral sum -- the identity the GPU kernel is built on."""
    rng = np.random.default_rng(3)
    n, C, N, ell = 14, 3, 6, 4
    m = n - C
    IMG = rng.normal(size=(n, N, N))
    CTF = rng.normal(size=(n, N, N))
    W, CTF1 = nlsa.get_wiener(CTF, np.arange(n), np.arange(n), C, n)
    mu_psi = rng.normal(size=(m, ell))
    msk2 = (rng.uniform(size=(N, N)) > 0.3).astype(float)
    A = nlsa.supervectors(IMG, CTF1, W, mu_psi, C, n, msk2)
    B = np.zeros_like(A)
    for ii in range(C):
        for c in range(ell):
            spec = sum(mu_psi[i, c] * fft2(IMG[C - ii + i - 1]) * CTF1[C - ii + i - 1] / W[i]
                       for i in range(m))
            img = ifft2(spec).real * msk2
            B[ii * N * N:(ii + 1) * N * N, c] = img.T.reshape(-1)
    np.testing.assert_allclose(A, B, rtol=1e-10, atol=1e-12)


def test_raw_supervectors_window():
    """'cuti' branch: with image j constant at j + 1, block ii column c is the constant
    sum_i (ConOrder - ii + i) mu_psi[i, c] (image index ConOrder - ii + i - 1); and the raw sum
    equals the Wiener form with CTF = W = 1 (an FFT round trip of each image)."""
    rng = np.random.default_rng(13)
    n, C, N, ell = 12, 4, 5, 3
    m = n - C
    mu_psi = rng.normal(size=(m, ell))
    IMG = np.arange(1, n + 1, dtype=float)[:, None, None] * np.ones((n, N, N))
    A = nlsa.supervectors_raw(IMG, mu_psi, C, n)
    for ii in range(C):
        expect = (C - ii + np.arange(m)) @ mu_psi
        np.testing.assert_allclose(A[ii * N * N:(ii + 1) * N * N], np.tile(expect, (N * N, 1)),
                                   rtol=1e-13)
    IMG = rng.normal(size=(n, N, N))
    ones = np.ones((n, N, N))
    np.testing.assert_allclose(nlsa.supervectors_raw(IMG, mu_psi, C, n),
                               nlsa.supervectors(IMG, ones, ones[:m], mu_psi, C, n),
                               rtol=1e-10, atol=1e-12)


def test_svdrf_and_l2():
    rng = np.random.default_rng(4)
    A = rng.normal(size=(200, 6))
    U, S, V = nlsa.svdRF(A)
    s = np.linalg.svd(A, compute_uv=False)
    np.testing.assert_allclose(np.diag(S), s, rtol=1e-12)
    np.testing.assert_allclose(U @ S @ V.T, A, atol=1e-10)
    for shape in ((7, 40), (12, 12)):               # wide and square: eigh(A A^T) branch
        A = rng.normal(size=shape)
        U, S, V = nlsa.svdRF(A)
        np.testing.assert_allclose(np.diag(S), np.linalg.svd(A, compute_uv=False), rtol=1e-10)
        np.testing.assert_allclose(U @ S @ V.T, A, atol=1e-9)
    a = rng.normal(size=(5, 7))
    d = nlsa.l2_distance(a, a)
    brute = np.array([[np.linalg.norm(a[:, i] - a[:, j]) for j in range(7)] for i in range(7)])
    np.testing.assert_allclose(d, brute, atol=1e-7)
    assert np.all(np.diag(d) == 0)


def test_fit1d_recovers_tau_on_exact_curve():
    rng = np.random.default_rng(5)
    tau = np.sort(rng.uniform(0.02, 0.98, 120))
    a = np.array([1.0, 0.35, 0.12])
    b = np.array([0.1, -0.2, 0.05])
    psi = np.stack([a[j] * np.cos((j + 1) * np.pi * tau) + b[j] for j in range(3)], axis=1)
    af, bf, tf = fit1d.op(psi)
    t = tf.ravel()
    err = min(np.max(np.abs(t - tau)), np.max(np.abs((1 - t) - tau)))
    assert err < 2e-2
    # each returned tau minimises the residual over a dense grid (up to the grid step)
    grid = np.linspace(0, 1, 2001)
    for p in range(0, 120, 17):
        rg = fit1d.residual(psi[p, :3], af, bf, grid)
        rt = fit1d.residual(psi[p, :3], af, bf, np.array([t[p]]))[0]
        assert rt <= rg.min() + 1e-6


def test_tau_bins_pick_frame_in_bin():
    rng = np.random.default_rng(6)
    tau = rng.uniform(0, 1, 300)
    IMGT = rng.normal(size=(400, 300))      # pixels x frames: numclass = min(50, 300 // 2)
    tn, IMG1, inds = nlsa.tau_bins(tau, IMGT, nClass=50)
    assert tn.min() == 0 and tn.max() == 1 and len(inds) == 50
    for i, ix in enumerate(inds):
        lo, hi = i / 50, (i + 1) / 50
        assert lo - 0.05 <= tn[ix] <= hi + 0.05
        np.testing.assert_array_equal(IMG1[:, i], IMGT[:, ix])

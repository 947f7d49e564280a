"""Pins of the Stage-1 oracle (oracle/distance.py) by closed forms, invariants, brute force and
library special cases.  CPU only."""
import numpy as np
import pytest
from scipy.fftpack import fft2

from oracle import distance as od


def test_annular_mask_geometry():
    # N = 4: centre row N/2-1 = 1, centre column N/2 = 2, radius 2 (strict)
    m = od.annular_mask(0, 2.0, 4, 4)
    # row r: (r-1)^2 + (c-2)^2 < 4
    expect = np.array([[0, 1, 1, 1],
                       [0, 1, 1, 1],
                       [0, 1, 1, 1],
                       [0, 0, 0, 0]], dtype=float)
    np.testing.assert_array_equal(m, expect)
    # large N: area approaches pi r^2, disk symmetric about its centre
    N = 200
    m = od.annular_mask(0, N / 2.0, N, N)
    assert abs(m.sum() - np.pi * (N / 2) ** 2) < 2 * np.pi * N / 2
    c_r, c_c = N // 2 - 1, N // 2
    assert np.array_equal(m[c_r - 50:c_r + 51, c_c], m[c_r - 50:c_r + 51, c_c][::-1])


def test_grid_and_butterworth_closed_forms():
    for N in (8, 9):
        Q = od.create_grid(N, N / 2.0)
        c = N // 2 if N % 2 == 0 else (N - 1) // 2
        assert Q[c, c] == 0.0
        # one grid step = 1/(N/2)
        assert np.isclose(Q[c, c + 1], 2.0 / N)
    G = od.create_filter("Butter", 8, 0.5, np.array([0.0, 0.5, 1.0]))
    np.testing.assert_allclose(G, [1.0, np.sqrt(0.5), np.sqrt(1 / (1 + 2.0 ** 16))])
    Gg = od.create_filter("Gauss", 8, 0.5, np.array([0.0, 0.5]))
    np.testing.assert_allclose(Gg, [1.0, np.sqrt(0.5)])
    with pytest.raises(ValueError):
        od.create_filter("Box", 8, 0.5, np.zeros(2))


def test_rotatefill_identity_and_quarter_turn():
    rng = np.random.default_rng(1)
    img = rng.normal(size=(16, 16))
    np.testing.assert_allclose(od.rotatefill(img, 0.0), img, atol=1e-12)
    # +90 deg about ((N-1)/2, (N-1)/2) maps the grid onto itself: an exact permutation
    r = od.rotatefill(img, 90.0)
    cands = [np.rot90(img, 1), np.rot90(img, -1)]
    assert min(np.max(np.abs(r - c)) for c in cands) < 1e-9
    # four quarter turns give the identity
    x = img
    for _ in range(4):
        x = od.rotatefill(x, 90.0)
    np.testing.assert_allclose(x, img, atol=1e-9)


def test_ctf_closed_forms():
    p = [2.7, 15000.0, 300.0, np.inf, 0.1]
    assert od.ctf(np.array([0.0]), p)[0] == pytest.approx(-0.1)
    # zeros where sin(wr) = ampc cos(wr): wr = arctan(0.1) + j pi; invert wr(k) = a k^4 - c k^2
    Cs = 2.7e7
    wav = 12.3986 / np.sqrt((2 * 511.0 + 300.0) * 300.0)
    a = 0.5 * np.pi * Cs * wav ** 3
    c = np.pi * wav * 15000.0
    for j in range(-6, -1):
        wr = np.arctan(0.1) + j * np.pi
        k2 = (c - np.sqrt(c * c + 4 * a * wr)) / (2 * a)   # small root of a u^2 - c u = wr
        assert abs(od.ctf(np.array([np.sqrt(k2)]), p)[0]) < 1e-9
    # Gaussian envelope halves the amplitude at k = B
    pe = [2.7, 15000.0, 300.0, 0.2, 0.1]
    k = np.array([0.2])
    assert np.isclose(od.ctf(k, pe)[0] / od.ctf(k, p)[0], 0.5)


def test_q2spider_rebuilds_quaternion():
    rng = np.random.default_rng(2)
    for _ in range(5):
        PD = rng.normal(size=3)
        PD /= np.linalg.norm(PD)
        Qr = np.array([1 + PD[2], PD[1], -PD[0], 0])
        Qr /= np.linalg.norm(Qr)
        a = od.q2spider(Qr)
        q1 = np.array([np.cos(a[0] / 2), 0, 0, -np.sin(a[0] / 2)])
        q2 = np.array([np.cos(a[1] / 2), 0, -np.sin(a[1] / 2), 0])
        q3 = np.array([np.cos(a[2] / 2), 0, 0, -np.sin(a[2] / 2)])
        q = od.qmult(q3, od.qmult(q2, q1)).ravel()
        assert min(np.max(np.abs(q - Qr)), np.max(np.abs(q + Qr))) < 1e-8


def test_gram_equals_direct_form_and_parseval():
    rng = np.random.default_rng(3)
    n, N = 7, 10
    imgs = rng.normal(size=(n, N, N))
    fy = np.stack([fft2(x) for x in imgs])
    C = rng.normal(size=(n, N, N))
    D = od.gram_distance(fy, C)
    Dd = od.direct_distance(fy, C)
    off = ~np.eye(n, dtype=bool)
    np.testing.assert_allclose(D[off], Dd[off], rtol=1e-10)
    assert np.max(np.abs(np.diag(D))) < 1e-8 * np.max(D)
    # C == 1: D_ij = N^2 ||img_i - img_j||^2 (unnormalised FFT, Parseval)
    D1 = od.gram_distance(fy, np.ones_like(C))
    brute = np.array([[N * N * np.sum((imgs[i] - imgs[j]) ** 2) for j in range(n)]
                      for i in range(n)])
    np.testing.assert_allclose(D1[off], brute[off], rtol=1e-10)


def test_half_spectrum_weights_reproduce_full_sums():
    """The identity the GPU operand packing relies on: for real images and even CTFs the
    weighted rfft2 half spectrum gives the same Gram as the full spectrum."""
    rng = np.random.default_rng(4)
    for N in (8, 9):
        n = 5
        imgs = rng.normal(size=(n, N, N))
        F = np.fft.fft2(imgs)
        H = np.fft.rfft2(imgs)
        k = np.fft.fftfreq(N) * N
        K = np.sqrt(k[:, None] ** 2 + k[None, :] ** 2)
        C = np.cos(0.3 * K)[None] * (1 + 0.1 * np.arange(n))[:, None, None]
        Nh = N // 2 + 1
        w = np.full(Nh, 2.0)
        w[0] = 1.0
        if N % 2 == 0:
            w[-1] = 1.0
        full = od.gram_distance(F, C)
        Ch = C[:, :, :Nh]
        A = (np.abs(H) ** 2 * w).reshape(n, -1)
        B = (Ch ** 2).reshape(n, -1)
        Y = (Ch * H * np.sqrt(w)).reshape(n, -1)
        half = A @ B.T + B @ A.T - 2 * np.real(Y @ Y.conj().T)
        np.testing.assert_allclose(half, full, rtol=1e-10, atol=1e-9 * np.max(full))


def test_stage1_invariants_and_conjugates():
    from manifoldem_python_amd import synth
    s = synth.make_pd(synth.seeds(1)[0], 16, 24, conj_frac=0.5)
    p = dict(nPix=24, pix_size=1.5, Cs=2.7, EkV=300.0, gaussEnv=np.inf, AmpContrast=0.1)
    out = od.op([s.ind, s.q, s.df, None, 0], dict(type="Butter", Qc=0.5, N=8), s.raw, s.nStot,
                dict(avgOnly=False), p)
    D = out["D"]
    assert np.array_equal(D, D.T)
    assert np.min(D) > -1e-8 * np.max(D)
    assert np.max(np.abs(np.diag(D))) < 1e-8 * np.max(D)
    assert set(out["imgLabels"]) <= {1, -1} and (out["imgLabels"] == -1).sum() == 8
    # a conjugate is stored vertically flipped; loading flips it back
    msk = od.annular_mask(0, 12.0, 24, 24)
    i = int(np.nonzero(s.ind >= s.nStot / 2)[0][0])
    src = int(s.ind[i] - s.nStot / 2)
    a = od.load_normalized(s.raw, src, True, msk)
    b = od.load_normalized(np.flip(s.raw, axis=2), src, False, msk)
    np.testing.assert_allclose(a, b, atol=1e-12)
    assert len(out) >= 19


def test_golden_stage1_fixture():
    import os
    path = os.path.join(os.path.dirname(__file__), "golden", "stage1_n12_N24.npz")
    g = np.load(path, allow_pickle=False)
    p = dict(nPix=int(g["N"]), pix_size=1.5, Cs=2.7, EkV=300.0, gaussEnv=np.inf, AmpContrast=0.1)
    out = od.op([g["ind"], g["q"], g["df"], None, 0], dict(type="Butter", Qc=0.5, N=8), g["raw"],
                int(g["nStot"]), dict(avgOnly=False), p)
    np.testing.assert_allclose(out["D"], g["D"], rtol=1e-10, atol=1e-6)
    np.testing.assert_allclose(out["imgAll"], g["imgAll"], atol=1e-10)
    np.testing.assert_allclose(out["CTF"], g["CTF"], atol=1e-12)

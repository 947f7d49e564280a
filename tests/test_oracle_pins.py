"""Independent pins of oracle functions that no library routine reproduces: closed forms,
brute-force index loops written from the reference's index rules, and invariants of synthetic
geometry.  Each check is built so that a plausible slip in the oracle (a sign, a window offset,
a masked-only mean, a transposed index) fails it.  CPU only."""
import math

import numpy as np
import pytest

from oracle import distance as od, nlsa


# ------------------------------------------------------------------ a1 load + normalise ----
def _brute_normalize(raw_img, conj):
    """getDistanceCTF...:254-281 by explicit loops: t = raw.T (flipud for a conjugate); mask =
    annularMask(0, N/2, N, N), a disk of radius N/2 centred on row N/2 - 1, column N/2
    (annularMask.py:24-30); mean and ddof-0 std of t * (1 - mask) over all N^2 pixels."""
    N = raw_img.shape[0]
    t = [[0.0] * N for _ in range(N)]
    for r in range(N):
        for c in range(N):
            rr = N - 1 - r if conj else r
            t[r][c] = float(raw_img[c][rr])             # (raw.T)[rr][c] = raw[c][rr]
    b = [[0.0] * N for _ in range(N)]
    for r in range(N):
        for c in range(N):
            x = r - N / 2 + 1
            y = c - N / 2
            inside = x * x + y * y < (N / 2) ** 2
            b[r][c] = 0.0 if inside else t[r][c]
    s = sum(b[r][c] for r in range(N) for c in range(N))
    mean = s / (N * N)
    var = sum((b[r][c] - mean) ** 2 for r in range(N) for c in range(N)) / (N * N)
    sd = math.sqrt(var)
    return np.array([[(t[r][c] - mean) / sd for c in range(N)] for r in range(N)])


@pytest.mark.parametrize("N", [6, 7])
@pytest.mark.parametrize("conj", [False, True])
def test_load_normalized_against_explicit_loops(N, conj):
    rng = np.random.default_rng(N + 10 * conj)
    raw = rng.normal(size=(3, N, N)) + np.arange(N)[None, :, None] * 0.3   # not symmetric
    msk = od.annular_mask(0, N / 2.0, N, N)
    got = od.load_normalized(raw, 1, conj, msk)
    np.testing.assert_allclose(got, _brute_normalize(raw[1], conj), rtol=1e-12, atol=1e-12)


def test_load_normalized_rejects_masked_only_statistics():
    """The mean over the background pixels only (a plausible reading of the reference) gives
    a different result: the statistics include the zeroed interior."""
    N = 8
    rng = np.random.default_rng(3)
    raw = rng.normal(size=(1, N, N)) + 2.0
    msk = od.annular_mask(0, N / 2.0, N, N)
    got = od.load_normalized(raw, 0, False, msk)
    t = raw[0].T
    bg = t[msk == 0]
    wrong = (t - bg.mean()) / bg.std()
    assert np.max(np.abs(got - wrong)) > 0.1
    np.testing.assert_allclose(got, _brute_normalize(raw[0], False), rtol=1e-12, atol=1e-12)


# ------------------------------------------------------------------ a3 PD geometry --------
def _q_from_euler(phi, theta, psi):
    """util.py:76-92 (flip=True) composed as read_alignfile.py:93-94, written out from the
    half-angle quaternions: q = qzs * (qy * qz)."""
    def mul(a, b):
        a0, a1, a2, a3 = a
        b0, b1, b2, b3 = b
        return np.array([a0 * b0 - a1 * b1 - a2 * b2 - a3 * b3,
                         a0 * b1 + a1 * b0 + a2 * b3 - a3 * b2,
                         a0 * b2 - a1 * b3 + a2 * b0 + a3 * b1,
                         a0 * b3 + a1 * b2 - a2 * b1 + a3 * b0])
    out = []
    for f, t, p in zip(phi, theta, psi):
        qz = np.array([math.cos(f / 2), 0, 0, -math.sin(f / 2)])
        qy = np.array([math.cos(t / 2), 0, -math.sin(t / 2), 0])
        qzs = np.array([math.cos(p / 2), 0, 0, -math.sin(p / 2)])
        out.append(mul(qzs, mul(qy, qz)))
    return np.array(out).T


@pytest.mark.parametrize("phi,theta", [(0.3, 0.7), (2.0, 1.2), (-1.0, 2.5), (0.0, 0.0),
                                       (3.0, 0.05)])
def test_pd_axis_and_inplane_angle_of_synthetic_particles(phi, theta):
    """Particles seen along one axis (rot phi, tilt theta) with in-plane angles psi_i:
    calc_avg_pd gives the axis (cos phi sin theta, sin phi sin theta, cos theta) for every
    particle, and get_psi gives Psi_i = psi_i + phi (mod 2 pi) -- the sign convention of the
    in-plane alignment (getDistanceCTF...:169-202)."""
    rng = np.random.default_rng(int(100 * phi) % 97)
    n = 12
    psi = rng.uniform(0, 2 * np.pi, n)
    q = _q_from_euler(np.full(n, phi), np.full(n, theta), psi)
    PDs = od.calc_avg_pd(q, n)
    axis = np.array([math.cos(phi) * math.sin(theta), math.sin(phi) * math.sin(theta),
                     math.cos(theta)])
    np.testing.assert_allclose(PDs, np.repeat(axis[:, None], n, 1), atol=1e-12)
    PD = PDs.sum(1) / np.linalg.norm(PDs.sum(1))
    for i in range(n):
        Psi, s, c = od.get_psi(q, PD, i)
        d = np.mod(Psi - psi[i] - phi + np.pi, 2 * np.pi) - np.pi
        assert abs(d) < 1e-9, (i, Psi, psi[i])


def test_pd_geometry_of_a_cap_is_its_mean_direction():
    """Orientations spread in a cap: PD is the normalised sum of the per-particle view
    directions (the third row of each rotation matrix)."""
    rng = np.random.default_rng(7)
    n = 50
    phi = 0.4 + rng.normal(scale=0.02, size=n)
    theta = 1.1 + rng.normal(scale=0.02, size=n)
    q = _q_from_euler(phi, theta, rng.uniform(0, 2 * np.pi, n))
    dirs = np.stack([np.cos(phi) * np.sin(theta), np.sin(phi) * np.sin(theta), np.cos(theta)])
    np.testing.assert_allclose(od.calc_avg_pd(q, n), dirs, atol=1e-12)
    PDs, PD, _ = od.pd_geometry(q)
    np.testing.assert_allclose(PD, dirs.sum(1) / np.linalg.norm(dirs.sum(1)), atol=1e-12)


# ------------------------------------------------------------------ a6 / a7 averages ------
def _stack(n, N, seed):
    rng = np.random.default_rng(seed)
    return rng.normal(size=(n, N, N))


@pytest.mark.parametrize("c", [1.0, -1.0, 0.5])
def test_wiener_average_constant_ctf(c):
    """C == c on every image: W = -(n c^2 + 1/5) (negative: getDistanceCTF...:423) and
    imgAvg = c sum_i img_i / (W n), so c = 1 gives sum img / (-(n + 0.2) n)."""
    n, N = 7, 8
    img = _stack(n, N, 1)
    C = np.full((n, N, N), c)
    imgAvg, _, _ = od.averages(img, img, C)
    expect = c * img.sum(0) / (-(n * c * c + 0.2) * n)
    np.testing.assert_allclose(imgAvg, expect, rtol=1e-12, atol=1e-14)


def test_phase_flip_averages_with_signed_ctf():
    """sign(C) == +1: imgAllFlip = img, imgAvgFlip = mean(img), imgAllIntensity =
    mean(img^2); sign(C) == -1 flips the first and keeps the second (getDistanceCTF...:346,
    :367, :400)."""
    n, N = 6, 8
    img = _stack(n, N, 2)
    fy = np.fft.fft2(img)
    for sgn in (1.0, -1.0):
        C = sgn * np.random.default_rng(3).uniform(0.1, 1.0, size=(n, N, N))
        flip = np.stack([od.phase_flip(fy[i], C[i]) for i in range(n)])
        np.testing.assert_allclose(flip, sgn * img, atol=1e-12)
        _, avgf, inten = od.averages(img, flip, C)
        np.testing.assert_allclose(avgf, sgn * img.mean(0), atol=1e-12)
        np.testing.assert_allclose(inten, (img ** 2).mean(0), atol=1e-12)


# ------------------------------------------------------------------ a21 Wiener window -----
def test_get_wiener_constant_and_ramp_ctf():
    """Constant C = c: W_i = ConOrder c^2 + 0.2.  Per-image constants C_j = j: W_i = sum of j^2
    over j = i+1 .. i+ConOrder (get_wiener.py:15-16, CTF index ConOrder - ii + i for
    ii < ConOrder), the closed form S(i + ConOrder) - S(i), S(m) = m(m+1)(2m+1)/6."""
    n, C, N = 12, 4, 3
    W, _ = nlsa.get_wiener(np.full((n, N, N), 0.7), np.arange(n), np.arange(n), C, n)
    np.testing.assert_allclose(W, np.full((n - C, N, N), C * 0.49 + 0.2), rtol=1e-14)
    CTF = np.arange(n, dtype=float)[:, None, None] * np.ones((n, N, N))
    W, CTF1 = nlsa.get_wiener(CTF, np.arange(n), np.arange(n), C, n)
    S = lambda m: m * (m + 1) * (2 * m + 1) / 6.0
    for i in range(n - C):
        np.testing.assert_allclose(W[i], S(i + C) - S(i) + 0.2, rtol=1e-14)
    # the sorted order is applied before the window: CTF1 = CTF[posPath[posPsi1]]
    perm = np.random.default_rng(1).permutation(n)
    _, CTF1 = nlsa.get_wiener(CTF, np.arange(n), perm, C, n)
    np.testing.assert_array_equal(CTF1[:, 0, 0], perm.astype(float))


def test_supervector_image_window_is_one_behind_the_ctf_window():
    """NLSA.py:66-70: block ii, column i uses image ConOrder - ii + i - 1 (one before the first
    CTF of W_i's window).  With images that are constant planes img_j = j and C == 1, block ii
    column i of tmp is the constant (ConOrder - ii + i - 1) / W_i, W_i = ConOrder + 0.2."""
    n, C, N = 10, 3, 4
    m = n - C
    IMG = np.arange(n, dtype=float)[:, None, None] * np.ones((n, N, N))
    CTF = np.ones((n, N, N))
    W, CTF1 = nlsa.get_wiener(CTF, np.arange(n), np.arange(n), C, n)
    mu_psi = np.eye(m)[:, :m]                      # picks tmp's columns out one by one
    A = nlsa.supervectors(IMG, CTF1, W, mu_psi, C, n)
    for ii in range(C):
        for i in range(m):
            col = A[ii * N * N:(ii + 1) * N * N, i]
            np.testing.assert_allclose(col, (C - ii + i - 1) / (C + 0.2), rtol=1e-12)


# ------------------------------------------------------------------ a23 reconstruction ----
def _reconstruct_brute(U, S, V, psiC, ConOrder, num, nS, Npix):
    """NLSA.py:93-138 scalar by scalar: Topo_mean[p, j] = mean over k of U[k Npix + p, j];
    frame ii = sum_{i < ConOrder} sum_{r in {0, 1}} U[i Npix + p, r] s_r (V[:, r] . psiC[i+ii])
    for ii < num - 2 ConOrder (zero beyond, up to nS - 2 ConOrder), then per-frame
    (x - mean) / std (ddof 0)."""
    s = np.diag(S)
    psiTrunc = psiC.shape[1]
    topo = np.zeros((Npix, psiTrunc))
    for j in range(psiTrunc):
        for p in range(Npix):
            topo[p, j] = sum(U[k * Npix + p, j] for k in range(ConOrder)) / ConOrder
    nfr = nS - 2 * ConOrder
    IMGT = np.zeros((Npix, nfr))
    for ii in range(num - 2 * ConOrder):
        for i in range(ConOrder):
            for r in (0, 1):
                z = s[r] * float(np.dot(V[:, r], psiC[i + ii]))
                for p in range(Npix):
                    IMGT[p, ii] += U[i * Npix + p, r] * z
    for ii in range(nfr):
        col = IMGT[:, ii]
        mu = sum(col) / Npix
        sd = math.sqrt(sum((x - mu) ** 2 for x in col) / Npix)
        IMGT[:, ii] = (col - mu) / sd
    return IMGT, topo


def test_reconstruct_matches_scalar_loops():
    rng = np.random.default_rng(11)
    ConOrder, num, Npix, psiTrunc = 3, 14, 10, 4
    m = num - ConOrder
    U = rng.normal(size=(ConOrder * Npix, psiTrunc))
    V = np.linalg.qr(rng.normal(size=(psiTrunc, psiTrunc)))[0]
    S = np.diag(np.array([3.0, 2.0, 1.0, 0.5]))
    psiC = rng.normal(size=(m, psiTrunc))
    IMGT, topo, sdiag = nlsa.reconstruct(U, S, V, psiC, ConOrder, num, num, Npix)
    IMGT_b, topo_b = _reconstruct_brute(U, S, V, psiC, ConOrder, num, num, Npix)
    np.testing.assert_allclose(topo, topo_b, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(IMGT, IMGT_b, rtol=1e-10, atol=1e-12)
    np.testing.assert_array_equal(sdiag, np.diag(S))


def test_reconstruct_block_constant_closed_form():
    """U with the same two pixel patterns u0, u1 (zero mean, orthogonal) in every delay block
    and chronos z_0(j) = j, z_1(j) = 1 (psiC V = [z_0, z_1, 0, ...]): frame ii is
    s_0 Z_0(ii) u0 + s_1 Z_1(ii) u1 with the window sums Z_0(ii) = ConOrder ii +
    ConOrder (ConOrder - 1) / 2 and Z_1 = ConOrder, so the z-scored frame is known exactly."""
    ConOrder, num, Npix, psiTrunc = 4, 15, 12, 3
    m = num - ConOrder
    x = np.arange(Npix)
    u0 = np.cos(2 * np.pi * x / Npix)
    u1 = np.sin(4 * np.pi * x / Npix)
    U = np.zeros((ConOrder * Npix, psiTrunc))
    for k in range(ConOrder):
        U[k * Npix:(k + 1) * Npix, 0] = u0
        U[k * Npix:(k + 1) * Npix, 1] = u1
        U[k * Npix:(k + 1) * Npix, 2] = 1.0
    V = np.linalg.qr(np.random.default_rng(2).normal(size=(psiTrunc, psiTrunc)))[0]
    Z = np.zeros((m, psiTrunc))
    Z[:, 0] = np.arange(m)
    Z[:, 1] = 1.0
    psiC = Z @ V.T                                   # psiC V = Z
    s = np.array([2.0, 5.0, 7.0])
    IMGT, _, _ = nlsa.reconstruct(U, np.diag(s), V, psiC, ConOrder, num, num, Npix)
    for ii in range(num - 2 * ConOrder):
        a = s[0] * (ConOrder * ii + ConOrder * (ConOrder - 1) / 2.0)
        b = s[1] * ConOrder
        frame = a * u0 + b * u1
        sd = math.sqrt((a * a * np.sum(u0 ** 2) + b * b * np.sum(u1 ** 2)) / Npix)
        np.testing.assert_allclose(IMGT[:, ii], frame / sd, rtol=1e-10, atol=1e-12)

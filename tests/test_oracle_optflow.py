"""Pins of the optical-flow oracle (CPU): brute-force filters, closed-form flows, invariants of
the diffusion (CC/hornschunck_simple.py, CC/OpticalFlowMovie.py)."""
import numpy as np
import pytest

from oracle import optflow as of


def _refl(i, n):
    """Half-sample reflection (scipy mode 'reflect'), repeated for long extensions."""
    i = i % (2 * n)
    return i if i < n else 2 * n - 1 - i


def _conv_bf(im, k):
    """The convolution sum out[i, j] = sum k[a, b] in[i - a + c0, j - b + c1], centre
    c = size // 2, accumulated in float64 and stored in the input's dtype."""
    H, W = im.shape
    c0, c1 = k.shape[0] // 2, k.shape[1] // 2
    out = np.zeros((H, W))
    for i in range(H):
        for j in range(W):
            s = 0.0
            for a in range(k.shape[0]):
                for b in range(k.shape[1]):
                    s += k[a, b] * float(im[_refl(i - a + c0, H), _refl(j - b + c1, W)])
            out[i, j] = s
    return out.astype(im.dtype)


def _gauss_bf(im, sig):
    r = int(4.0 * sig + 0.5)
    x = np.arange(-r, r + 1)
    w = np.exp(-0.5 / (sig * sig) * x ** 2)
    w /= w.sum()
    H, W = im.shape
    t = np.zeros((H, W))
    for i in range(H):
        for j in range(W):
            t[i, j] = sum(w[k + r] * float(im[_refl(i + k, H), j]) for k in range(-r, r + 1))
    t = t.astype(im.dtype)
    out = np.zeros((H, W))
    for i in range(H):
        for j in range(W):
            out[i, j] = sum(w[k + r] * float(t[i, _refl(j + k, W)]) for k in range(-r, r + 1))
    return out.astype(im.dtype)


def test_derivatives_match_convolution_definition():
    rng = np.random.default_rng(0)
    im1 = rng.normal(size=(7, 9)).astype(np.float32)
    im2 = rng.normal(size=(7, 9)).astype(np.float32)
    fx, fy, ft = of.compute_derivatives(im1, im2)
    bx = _conv_bf(im1, of.KX) + _conv_bf(im2, of.KX)
    by = _conv_bf(im1, of.KY) + _conv_bf(im2, of.KY)
    bt = _conv_bf(im1, of.KT) + _conv_bf(im2, -of.KT)
    for a, b in ((fx, bx), (fy, by), (ft, bt)):
        assert a.dtype == np.float32
        np.testing.assert_array_equal(a, b)
    # the convolution flips the kernel: fx is minus the forward x difference (summed over
    # the 2 x 2 block and both frames), ft is frame 1 minus frame 2
    ramp = np.tile(np.arange(9, dtype=np.float32), (7, 1))
    fx, fy, ft = of.compute_derivatives(ramp, ramp - 0.5)
    np.testing.assert_array_equal(fx[:, :-1], -1.0)
    np.testing.assert_array_equal(fy, 0.0)
    np.testing.assert_array_equal(ft, 0.5)
    u = rng.normal(size=(6, 5))
    np.testing.assert_allclose(of.convolve(u, of.HSKERN), _conv_bf(u, of.HSKERN), rtol=1e-13)


@pytest.mark.parametrize("shape,sig", [((11, 13), 1.5), ((11, 13), 2.0), ((5, 4), 1.5)])
def test_lowpass_matches_separable_gaussian(shape, sig):
    rng = np.random.default_rng(1)
    for dt in (np.float64, np.float32):
        im = rng.normal(size=shape).astype(dt)
        got = of.lowpassfilt(im, sig)
        assert got.dtype == dt
        np.testing.assert_allclose(got, _gauss_bf(im, sig),
                                   rtol=1e-12 if dt == np.float64 else 1e-6, atol=1e-7)


def test_hs_identical_frames_zero_flow():
    im = np.random.default_rng(2).normal(size=(16, 20))
    U, V = of.hornschunck(im, im, np.zeros(1), np.zeros(1), 2.0, 2.0, 30)
    assert U.dtype == np.float64
    np.testing.assert_array_equal(U, 0.0)
    np.testing.assert_array_equal(V, 0.0)


@pytest.mark.parametrize("niter", [1, 20])
def test_hs_translated_ramp(niter):
    """Content moved by d along x: U = d, V = 0 away from the edges (small alpha)."""
    d = 0.3
    j = np.arange(64, dtype=np.float64)
    im1 = np.tile(2.0 * j, (64, 1))
    im2 = np.tile(2.0 * (j - d), (64, 1))
    U, V = of.hornschunck(im1, im2, np.zeros(1), np.zeros(1), 1.0, 0.001, niter)
    np.testing.assert_array_equal(V, 0.0)
    c = slice(24, 40)
    np.testing.assert_allclose(U[c, c], d, rtol=1e-5)
    # and with the content moved along y the roles swap
    U2, V2 = of.hornschunck(im1.T.copy(), im2.T.copy(), np.zeros(1), np.zeros(1), 1.0, 0.001,
                            niter)
    np.testing.assert_array_equal(U2, 0.0)
    np.testing.assert_allclose(V2[c, c], d, rtol=1e-5)


def test_hs_transpose_swaps_components():
    rng = np.random.default_rng(3)
    im1 = rng.normal(size=(20, 24))
    im2 = np.roll(im1, 1, axis=1) + 0.1 * rng.normal(size=(20, 24))
    U, V = of.hornschunck(im1, im2, np.zeros(1), np.zeros(1), 1.5, 2.0, 25)
    Ut, Vt = of.hornschunck(im1.T.copy(), im2.T.copy(), np.zeros(1), np.zeros(1), 1.5, 2.0, 25)
    # equal up to the float32 rounding of the Gaussian passes, which run in the other order
    tol = 1e-6 * max(np.abs(U).max(), np.abs(V).max())
    np.testing.assert_allclose(Ut, V.T, rtol=0, atol=tol)
    np.testing.assert_allclose(Vt, U.T, rtol=0, atol=tol)
    # a float32 initial flow keeps the iteration in float32
    U32, _ = of.hornschunck(im1, im2, np.zeros((20, 24), np.float32),
                            np.zeros((20, 24), np.float32), 1.5, 2.0, 25)
    assert U32.dtype == np.float32
    np.testing.assert_allclose(U32, U, rtol=1e-4, atol=1e-6 * np.abs(U).max())


def _pm_step_flux(s, kappa, gamma, step):
    """One Perona-Malik step in flux form, per pixel: each of the 6 neighbours contributes
    g(delta) * delta / step_axis, with no flux through the stack's faces."""
    F, H, W = s.shape
    s = s.astype(np.float64)
    out = s.copy()
    for z in range(F):
        for y in range(H):
            for x in range(W):
                acc = 0.0
                for ax, (dz, dy, dx) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    for sgn in (1, -1):
                        zz, yy, xx = z + sgn * dz, y + sgn * dy, x + sgn * dx
                        if 0 <= zz < F and 0 <= yy < H and 0 <= xx < W:
                            dlt = s[zz, yy, xx] - s[z, y, x]
                            acc += np.exp(-(dlt / kappa) ** 2) * dlt / step[ax]
                out[z, y, x] += gamma * acc
    return out


def test_anisodiff3_invariants():
    rng = np.random.default_rng(4)
    st = (50 * rng.normal(size=(6, 7, 8))).astype(np.float32)
    np.testing.assert_array_equal(of.anisodiff3(st, niter=1), st)
    one = of.anisodiff3(st, niter=2, kappa=50, gamma=0.1, step=(5., 3., 3.))
    np.testing.assert_allclose(one, _pm_step_flux(st, 50, 0.1, (5., 3., 3.)), rtol=2e-5,
                               atol=2e-4)
    out = of.anisodiff3(st, niter=5, kappa=50, gamma=0.1, step=(5., 3., 3.))
    assert out.dtype == np.float32
    np.testing.assert_allclose(out.astype(np.float64).sum(), st.astype(np.float64).sum(),
                               rtol=1e-5)
    assert out.var() < st.var()


def test_block_average_and_static_movie():
    rng = np.random.default_rng(5)
    M = rng.normal(size=(12, 3, 4)).astype(np.float32)
    A = of.block_average(M, 5)
    assert A.shape == (3, 3, 4)
    for b, (lo, hi) in enumerate(((0, 5), (5, 10), (10, 12))):
        s = np.zeros((3, 4))
        for f in range(lo, hi):
            s += M[f].astype(np.float64)
        np.testing.assert_array_equal(A[b], s / (hi - lo))
    frame = rng.normal(size=(16 * 16,))
    mov = np.tile(frame, (10, 1))
    VxM, VyM = of.movie_flow(mov, 5, Niter=10)
    np.testing.assert_array_equal(VxM, 0.0)
    np.testing.assert_array_equal(VyM, 0.0)


def test_orient_mag_precedence():
    X = np.array([1.0, -1.0, 0.0, 3.0])
    Y = np.array([1.0, 0.0, -1.0, 4.0])
    o, m = of.get_orient_mag(X, Y)
    np.testing.assert_allclose(o, [np.pi / 4 * np.pi, (np.pi - 2) * np.pi,
                                   (2 - np.pi / 2) * np.pi, np.arctan2(4, 3) * np.pi],
                               rtol=1e-14)
    np.testing.assert_allclose(m, [np.sqrt(2), 1, 1, 5], rtol=1e-15)

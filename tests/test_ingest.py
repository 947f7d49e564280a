"""Ingest of RELION particle stacks on the host side: the MRC reader (against headers packed
by hand here, not by the library's writer) and the semantics of the reference's recentring
shift, scipy.ndimage.shift(order=3, mode='wrap') (getDistanceCTF_local_Conj9combinedS2.py:264),
which the GPU kernel restates.  CPU only."""
import math
import struct

import numpy as np
import pytest
from scipy import ndimage


def _pack_mrc(path, data, order="<", nsymbt=0, mode=2):
    nz, ny, nx = data.shape
    hdr = bytearray(1024)
    struct.pack_into(order + "4i", hdr, 0, nx, ny, nz, mode)
    struct.pack_into(order + "i", hdr, 92, nsymbt)
    hdr[208:212] = b"MAP "
    hdr[212:214] = bytes([0x11, 0x11]) if order == ">" else bytes([0x44, 0x44])
    dt = np.dtype({2: np.float32, 1: np.int16}[mode]).newbyteorder(order)
    with open(path, "wb") as f:
        f.write(bytes(hdr))
        f.write(b"\x07" * nsymbt)
        f.write(np.ascontiguousarray(data, dtype=dt).tobytes())


@pytest.mark.parametrize("order,nsymbt", [("<", 0), ("<", 96), (">", 0)])
def test_mrc_reader_against_hand_packed_files(tmp_path, order, nsymbt):
    from manifoldem_python_amd import mrc
    rng = np.random.default_rng(len(order) + nsymbt)
    data = rng.normal(size=(5, 6, 7)).astype(np.float32)          # nz, ny, nx: x fastest
    path = str(tmp_path / "s.mrcs")
    _pack_mrc(path, data, order, nsymbt)
    h = mrc.header(path)
    assert (h["nx"], h["ny"], h["nz"], h["mode"], h["offset"]) == (7, 6, 5, 2, 1024 + nsymbt)
    np.testing.assert_array_equal(mrc.read_images(path), data)
    np.testing.assert_array_equal(mrc.read_images(path, [4, 0, 4]), data[[4, 0, 4]])
    with pytest.raises(IndexError):
        mrc.read_images(path, [5])


def test_mrc_writer_round_trip_and_mode_check(tmp_path):
    from manifoldem_python_amd import mrc
    data = np.random.default_rng(1).normal(size=(3, 8, 8)).astype(np.float32)
    path = str(tmp_path / "w.mrcs")
    mrc.write_stack(path, data)
    np.testing.assert_array_equal(mrc.read_images(path), data)
    p16 = str(tmp_path / "i.mrcs")
    _pack_mrc(p16, (data * 100).astype(np.int16), mode=1)
    with pytest.raises(NotImplementedError, match="mode 1"):
        mrc.read_images(p16)


def _b3(t):
    t = abs(t)
    if t < 1:
        return 2 / 3 - t * t + t * t * t / 2
    if t < 2:
        return (2 - t) ** 3 / 6
    return 0.0


def _shift_restated(img, sh):
    """The rule memi_shift_images implements: mirror-boundary cubic B-spline coefficients,
    coordinates outside [0, L-1] wrapped with period L-1, mirrored taps."""
    N0, N1 = img.shape

    def wrap(c, L):
        P = L - 1
        if c < 0:
            r = math.fmod(-c, P)
            return P - r if r != 0 else P
        return math.fmod(c, P) if c > P else c

    def mir(k, L):
        return -k if k < 0 else (2 * (L - 1) - k if k > L - 1 else k)

    coef = ndimage.spline_filter(img, order=3, mode="mirror")
    out = np.zeros_like(img)
    for i in range(N0):
        for j in range(N1):
            y, x = wrap(i - sh[0], N0), wrap(j - sh[1], N1)
            fy, fx = math.floor(y), math.floor(x)
            out[i, j] = sum(_b3(y - a) * _b3(x - b) * coef[mir(a, N0), mir(b, N1)]
                            for a in range(fy - 1, fy + 3) for b in range(fx - 1, fx + 3))
    return out


@pytest.mark.parametrize("N", [9, 12])
def test_wrap_shift_semantics(N):
    """scipy's shift(order=3, mode='wrap') equals the restated rule, and an integer shift by 1
    is the sample permutation that rule predicts: out[0] = in[N-2], out[i] = in[i-1] (the
    coordinate -1 wraps with period N-1), on both axes."""
    rng = np.random.default_rng(N)
    img = rng.normal(size=(N, N))
    for sh in [(0.3, -1.7), (2.5, 3.25), (-0.5, -0.5), (5.9, -7.2)]:
        np.testing.assert_allclose(ndimage.shift(img, sh, order=3, mode="wrap"),
                                   _shift_restated(img, sh), rtol=0, atol=1e-12)
    out = ndimage.shift(img, (1.0, 0.0), order=3, mode="wrap")
    perm = np.concatenate(([N - 2], np.arange(N - 1)))
    np.testing.assert_allclose(out, img[perm], atol=1e-12)
    out = ndimage.shift(img, (0.0, 1.0), order=3, mode="wrap")
    np.testing.assert_allclose(out, img[:, perm], atol=1e-12)


def test_mirror_spline_lu_solves_the_system():
    """The host LU factors memi_shift_images uses: forward/back substitution with them gives
    the coefficients of scipy's mirror-mode spline_filter1d."""
    from manifoldem_python_amd._native import spline_lu
    for L in (3, 4, 9, 64, 131):
        x = np.random.default_rng(L).normal(size=L)
        lu = spline_lu(L)
        m, ib = lu[:L], lu[L:]
        d = 6 * x.copy()
        for k in range(1, L):
            d[k] = 6 * x[k] - m[k] * d[k - 1]
        c = d.copy()
        c[L - 1] = d[L - 1] * ib[L - 1]
        for k in range(L - 2, -1, -1):
            c[k] = (d[k] - (2.0 if k == 0 else 1.0) * c[k + 1]) * ib[k]
        np.testing.assert_allclose(c, ndimage.spline_filter1d(x, order=3, mode="mirror"),
                                   rtol=0, atol=1e-12)


def test_nearest_mode_interpolation_semantics():
    """scipy's map_coordinates / affine_transform with order 3 and mode 'nearest' (projectMask,
    CC/projectMask.py:51) equals: pad 12 edge samples per side, half-sample-reflection spline
    coefficients of the padded array, interpolate at coordinate + 12 with taps clamped to the
    padded array -- the rule memi_project_mask implements."""
    rng = np.random.default_rng(0)
    for L in (9, 14):
        x = rng.normal(size=L)
        coords = np.array([-15.0, -3.2, -1.0, -0.6, 0.0, 0.4, 4.5, L - 1.3, L - 1.0, L + 0.1,
                           L + 3.0, 3 * L + 0.5])
        ref = ndimage.map_coordinates(x, [coords], order=3, mode="nearest")
        xp = np.pad(x, 12, mode="edge")
        coef = ndimage.spline_filter1d(xp, 3, mode="reflect")
        got = []
        for c in coords:
            y = c + 12
            f = math.floor(y)
            got.append(sum(_b3(y - a) * coef[min(max(a, 0), len(xp) - 1)]
                           for a in range(f - 1, f + 3)))
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12)


def test_reflect_spline_lu_solves_the_system():
    from manifoldem_python_amd._native import spline_lu
    # lengths of padded mask volumes (side + 24): scipy's truncated start is accurate there
    for L in (26, 40, 64):
        x = np.random.default_rng(L).normal(size=L)
        lu = spline_lu(L, "reflect")
        m, ib = lu[:L], lu[L:]
        d = 6 * x.copy()
        for k in range(1, L):
            d[k] = 6 * x[k] - m[k] * d[k - 1]
        c = d.copy()
        c[L - 1] = d[L - 1] * ib[L - 1]
        for k in range(L - 2, -1, -1):
            c[k] = (d[k] - c[k + 1]) * ib[k]
        # scipy starts its reflect-mode recursion from a truncated sum (accurate to ~1e-12 at
        # these lengths, far worse for very short lines); the LU solve is exact
        np.testing.assert_allclose(c, ndimage.spline_filter1d(x, order=3, mode="reflect"),
                                   rtol=0, atol=1e-11)
        np.testing.assert_allclose((5 * c[0] + c[1]) / 6, x[0], atol=1e-14)
        np.testing.assert_allclose((c[:-2] + 4 * c[1:-1] + c[2:]) / 6, x[1:-1], atol=1e-14)


def test_project_mask_along_z_is_the_z_projection():
    """PD = +z: the Euler angles are 0, the rotation is the identity and the order-3
    interpolation reproduces the samples, so the mask is sum_z vol[z] > 1 in the (y, x) layout
    of mrcfile's data (CC/projectMask.py:118-140)."""
    from oracle import distance as od
    rng = np.random.default_rng(4)
    vol = (rng.uniform(size=(12, 12, 12)) > 0.6).astype(float) * rng.uniform(0.2, 1.0)
    proj, msk = od.project_mask(vol, np.array([0.0, 0.0, 1.0]))
    np.testing.assert_allclose(proj, vol.sum(axis=0), atol=1e-12)
    np.testing.assert_array_equal(msk, vol.sum(axis=0) > 1)


def test_mask_transform_matches_oracle_rotation():
    """The host affine map of the GPU path against the oracle's restatement of
    eulerRotMatrix3DSpider and the centre offset (CC/projectMask.py:24-53)."""
    from manifoldem_python_amd import pipeline as pl
    from oracle import distance as od
    rng = np.random.default_rng(5)
    for _ in range(4):
        PD = rng.normal(size=3)
        PD /= np.linalg.norm(PD)
        T = pl.mask_transform(PD, 20)
        Qr = np.array([1 + PD[2], PD[1], -PD[0], 0])
        phi, theta, _ = od.q2spider(Qr / np.sqrt(np.sum(Qr ** 2)))
        R = od.spider_rotation(0.0, -theta, -phi)
        c = np.full(3, 10.0)
        np.testing.assert_allclose(T[:9], R.ravel(), atol=1e-12)
        np.testing.assert_allclose(T[9:], c - R @ c, atol=1e-12)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-12)

"""GPU parity of Stage 1 (image preparation + distance Gram) against the float64 oracle."""
import math
import sys

import numpy as np
import pytest
from scipy.fftpack import fft2, ifft2, ifftshift

from conftest import FILTER, PARAMS, rel_offdiag

pytestmark = pytest.mark.gpu


def _setup(n, N, seed=0, conj_frac=0.0):
    from manifoldem_python_amd import synth
    s = synth.make_pd(synth.seeds(seed + 1)[seed], n, N, conj_frac=conj_frac)
    return s


def _oracle(s):
    from oracle import distance as od
    p = dict(PARAMS, nPix=s.N)
    return od.op([s.ind, s.q, s.df, None, 0], FILTER, s.raw, s.nStot, dict(avgOnly=False), p)


def _gpu_prep(s, cuda, precision, prep=None):
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    N = s.N
    n = len(s.ind)
    conj = (s.ind >= s.nStot / 2)
    src = np.where(conj, s.ind - s.nStot // 2, s.ind)
    geo = pl.pd_geometry(s.q)
    stack = torch.as_tensor(s.raw, device=cuda)
    t_src = torch.as_tensor(src.astype(np.int64), device=cuda)
    t_conj = torch.as_tensor(conj.astype(np.int32), device=cuda)
    rot1 = torch.as_tensor(-(180 / math.pi) * geo["Psi"], device=cuda)
    rot2 = torch.full((n,), -geo["psi_p"], dtype=torch.float64, device=cuda)
    df = torch.as_tensor(s.df, device=cuda)
    ctf = nat.ctf_params(PARAMS["pix_size"], PARAMS["Cs"], PARAMS["EkV"], PARAMS["AmpContrast"],
                         PARAMS["gaussEnv"])
    op = pl.Operand(n, N, precision, cuda)
    out = pl.prepare_images(stack, t_src, t_conj, rot1, rot2, df, N, FILTER, ctf, operand=op,
                            prep=prep)
    return geo, out, op


@pytest.mark.parametrize("N,conj_frac", [(48, 0.0), (45, 0.25)])
def test_load_normalize_and_filter(cuda, N, conj_frac):
    import torch
    from manifoldem_python_amd import _native as nat
    from oracle import distance as od
    s = _setup(24, N, conj_frac=conj_frac)
    conj = (s.ind >= s.nStot / 2)
    src = np.where(conj, s.ind - s.nStot // 2, s.ind)
    msk = od.annular_mask(0, N / 2.0, N, N)
    ref = np.stack([od.load_normalized(s.raw, src[i], conj[i], msk) for i in range(len(src))])
    y = torch.empty((len(src), N, N), dtype=torch.float32, device=cuda)
    nat.load_normalize(torch.as_tensor(s.raw, device=cuda),
                       torch.as_tensor(src.astype(np.int64), device=cuda),
                       torch.as_tensor(conj.astype(np.int32), device=cuda), N, y)
    got = y.cpu().numpy()
    assert np.max(np.abs(got - ref)) < 2e-6 * np.max(np.abs(ref))
    # Butterworth filter (create_grid / create_filter, :285-293)
    G = ifftshift(od.create_filter("Butter", 8, 0.5, od.create_grid(N, N / 2.0)))
    reff = np.stack([ifft2(fft2(r) * G).real for r in ref])
    Y = torch.fft.rfft2(y)
    nat.spectrum_scale(Y, N, nat.FILTER_BUTTER, 0.5, 8)
    gotf = torch.fft.irfft2(Y, s=(N, N)).cpu().numpy()
    assert np.max(np.abs(gotf - reff)) < 2e-6 * np.max(np.abs(reff))


@pytest.mark.parametrize("N", [48, 45])
def test_spline_rotate_matches_ndimage(cuda, N):
    import torch
    from manifoldem_python_amd import _native as nat
    from oracle import distance as od
    rng = np.random.default_rng(5)
    n = 12
    img = rng.normal(size=(n, N, N)) * od.annular_mask(0, N / 2.0, N, N)
    ang = np.concatenate(([0.0, 90.0, -90.0, 180.0], rng.uniform(-360, 360, n - 4)))
    ref = np.stack([od.rotatefill(img[i], ang[i]) for i in range(n)])
    y = torch.as_tensor(img, dtype=torch.float32, device=cuda)
    Y = torch.fft.rfft2(y)
    nat.spectrum_scale(Y, N, nat.FILTER_BSPLINE)
    coef = torch.fft.irfft2(Y, s=(N, N)).contiguous()
    out = torch.empty_like(coef)
    nat.spline_rotate(coef, torch.as_tensor(ang, device=cuda), N, out)
    got = out.cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < 2e-6, err


@pytest.mark.parametrize("N", [40, 45, 96, 128, 160])
def test_fused_alignment_matches_ndimage(cuda, N):
    """memi_align_images (disk mask, then two rotations, recursive B-spline prefilter in LDS)
    against the oracle's rotatefill(rotatefill(img * mask, a1), a2) -- scipy.ndimage.rotate
    on the 3x3-tiled image -- and against the FFT-prefilter path on the same input.  N = 40 /
    45 / 128 / 160 cover the 64- and 141-wide single-launch kernels and the two-launch one."""
    import torch
    from manifoldem_python_amd import _native as nat
    from oracle import distance as od
    rng = np.random.default_rng(11)
    n = 8 if N > 100 else 12
    img = rng.normal(size=(n, N, N))
    a1 = np.concatenate(([0.0, 90.0, -90.0, 180.0], rng.uniform(-360, 360, n - 4)))
    a2 = np.concatenate(([0.0, 33.0, 180.0, -270.0], rng.uniform(0, 360, n - 4)))
    msk = od.annular_mask(0, N / 2.0, N, N)
    ref = np.stack([od.rotatefill(od.rotatefill(img[i] * msk, a1[i]), a2[i])
                    for i in range(n)])
    y = torch.as_tensor(img, dtype=torch.float32, device=cuda)
    t1 = torch.as_tensor(a1, device=cuda)
    t2 = torch.as_tensor(a2, device=cuda)
    out = torch.empty_like(y)
    nat.align_images(y, t1, t2, N, out, disk=True)
    got = out.cpu().numpy()
    err = np.max(np.abs(got - ref)) / np.max(np.abs(ref))
    assert err < 3e-6, err
    # in place gives the same result
    y2 = y.clone()
    nat.align_images(y2, t1, t2, N, y2, disk=True)
    np.testing.assert_array_equal(y2.cpu().numpy(), got)
    # FFT-prefilter path on the same input
    z = y.clone()
    nat.mask_multiply(z, N, disk=True)
    for t in (t1, t2):
        Y = torch.fft.rfft2(z)
        nat.spectrum_scale(Y, N, nat.FILTER_BSPLINE)
        coef = torch.fft.irfft2(Y, s=(N, N)).contiguous()
        nat.spline_rotate(coef, t, N, z)
    assert np.max(np.abs(z.cpu().numpy() - got)) / np.max(np.abs(ref)) < 3e-6


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("n,N,conj_frac", [(64, 48, 0.0), (150, 70, 0.2), (40, 45, 0.0)])
def test_stage1_matches_oracle(cuda, precision, n, N, conj_frac):
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    s = _setup(n, N, seed=1, conj_frac=conj_frac)
    ref = _oracle(s)
    geo, out, op = _gpu_prep(s, cuda, precision)
    assert abs(geo["psi_p"] - ref["psi_p"]) < 1e-9
    np.testing.assert_allclose(geo["Psi"], ref["Psis"].ravel(), atol=1e-12)
    # aligned images
    img = out["imgAll"].cpu().numpy()
    assert np.max(np.abs(img - ref["imgAll"])) < 5e-6 * np.max(np.abs(ref["imgAll"]))
    # CTF (full layout) and half spectra
    Cfull = pl.full_ctf(out["ctf"], N).cpu().numpy()
    assert np.max(np.abs(Cfull - ref["CTF"])) < 1e-6
    fy = out["fy"].cpu().numpy()
    fref = ref["fy"][:, :, :N // 2 + 1]
    assert np.max(np.abs(fy - fref)) < 5e-6 * np.max(np.abs(fref))
    # distances
    plan = nat.Plan([0], [n], [n], chunks=pl.CHUNKS[precision])
    dist = pl.ctf_distances(plan, op, out["fy"], out["ctf"], N)
    D = dist.block(0).cpu().numpy()
    assert np.array_equal(D, D.T)
    assert np.all(np.diag(D) == 0)
    err = rel_offdiag(D, ref["D"])
    assert err.max() < 1e-5, (err.max(), np.percentile(err, 99))


@pytest.mark.parametrize("precision", [0, 1])
def test_gram_alone_vs_float64(cuda, precision):
    """The Gram on the GPU's own spectra against the float64 Gram of the same spectra, and a
    duplicate pair (D = 0 in exact arithmetic) resolved by the direct-form recompute.  rocFFT
    is not bitwise identical across batch positions, so the duplicate's spectra differ in the
    last bits and its D is ~1e-15 of a typical D rather than exactly 0."""
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    from oracle import distance as od
    n, N = 300, 64
    s = _setup(n, N, seed=2)
    # make images 7 and 8 identical (a duplicate pair: D = 0 exactly in exact arithmetic)
    s.raw[8] = s.raw[7]
    s.q[:, 8] = s.q[:, 7]
    s.df[8] = s.df[7]
    geo, out, op = _gpu_prep(s, cuda, precision)
    plan = nat.Plan([0], [n], [n], chunks=pl.CHUNKS[precision])
    dist = pl.ctf_distances(plan, op, out["fy"], out["ctf"], N)
    D = dist.block(0).cpu().numpy()
    Nh = N // 2 + 1
    fy = out["fy"].cpu().numpy().astype(np.complex128)
    full = np.empty((n, N, N), dtype=np.complex128)
    full[:, :, :Nh] = fy
    rows = (-np.arange(N)) % N
    full[:, :, Nh:] = np.conj(fy[:, rows][:, :, N - np.arange(Nh, N)])
    C = pl.full_ctf(out["ctf"], N).cpu().numpy().astype(np.float64).reshape(n, N, N)
    ref = od.gram_distance(full, C)
    typical = np.median(D[~np.eye(n, dtype=bool)])
    assert D[7, 8] == D[8, 7] and abs(D[7, 8]) < 1e-12 * typical
    mask = np.ones((n, n), bool)
    mask[7, 8] = mask[8, 7] = False
    err = rel_offdiag(D, ref)[mask[~np.eye(n, dtype=bool)]]
    assert err.max() < 1e-5, (err.max(), np.median(err))
    assert int(dist.flag_count.item()) >= 1


@pytest.mark.parametrize("precision", [0, 1])
def test_gram_schedule_variants_agree(cuda, precision):
    """Both Gram schedules issue the same MFMAs in the same order: identical D."""
    from manifoldem_python_amd import _native as nat, pipeline as pl
    n, N = 333, 48
    s = _setup(n, N, seed=3)
    geo, out, op = _gpu_prep(s, cuda, precision)
    plan = nat.Plan([0], [n], [n], chunks=pl.CHUNKS[precision])
    D1 = pl.ctf_distances(plan, op, out["fy"], out["ctf"], N, variant=1).block(0).cpu().numpy()
    D2 = pl.ctf_distances(plan, op, out["fy"], out["ctf"], N, variant=2).block(0).cpu().numpy()
    np.testing.assert_array_equal(D1, D2)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("chunks,tile", [(5, 128), (7, 128), (1, 256), (3, 256), (7, 256)])
def test_gram_uneven_chunks_vs_float64(cuda, precision, chunks, tile):
    """K splits whose last chunk is shorter (N = 64: 132 split-BF16 / 264 FP32 steps per half)
    with 128- and 256-wide tiles (n = 300: full, edge and diagonal tiles of both) against the
    float64 Gram of the same spectra, and against the default split."""
    from manifoldem_python_amd import _native as nat, pipeline as pl
    from oracle import distance as od
    n, N = 300, 64
    s = _setup(n, N, seed=4)
    geo, out, op = _gpu_prep(s, cuda, precision)
    plan = nat.Plan([0], [n], [n], chunks=chunks, gram_tile=tile)
    D = pl.ctf_distances(plan, op, out["fy"], out["ctf"], N).block(0).cpu().numpy()
    base = nat.Plan([0], [n], [n], chunks=pl.CHUNKS[precision])
    D0 = pl.ctf_distances(base, op, out["fy"], out["ctf"], N).block(0).cpu().numpy()
    Nh = N // 2 + 1
    fy = out["fy"].cpu().numpy().astype(np.complex128)
    full = np.empty((n, N, N), dtype=np.complex128)
    full[:, :, :Nh] = fy
    rows = (-np.arange(N)) % N
    full[:, :, Nh:] = np.conj(fy[:, rows][:, :, N - np.arange(Nh, N)])
    C = pl.full_ctf(out["ctf"], N).cpu().numpy().astype(np.float64).reshape(n, N, N)
    ref = od.gram_distance(full, C)
    assert rel_offdiag(D, ref).max() < 1e-5
    assert rel_offdiag(D, D0).max() < 2e-5
    np.testing.assert_array_equal(D, D.T)
    assert np.all(np.diag(D) == 0)


def _operand_values(op):
    import torch
    if op.lo is None:
        return op.hi.cpu().numpy().astype(np.float64)
    f = lambda t: (t.cpu().numpy().astype(np.int32) & 0xFFFF).astype(np.uint32) << 16
    return (f(op.hi).view(np.float32).astype(np.float64) +
            f(op.lo).view(np.float32).astype(np.float64))


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("N", [40, 45, 70, 96, 128, 150, 160])
def test_fused_prep_matches_staged(cuda, precision, N):
    """memi_prep_images (load, filter, rotations, FFTs and CTF packing in one kernel with LDS
    FFTs) against the staged path (rocFFT + separate kernels) on the same inputs: aligned
    images, half spectra, CTF and operand rows."""
    n = 37
    s = _setup(n, N, seed=6, conj_frac=0.3)
    _, a, opa = _gpu_prep(s, cuda, precision, prep="fused")
    _, b, opb = _gpu_prep(s, cuda, precision, prep="staged")
    ia, ib = a["imgAll"].cpu().numpy(), b["imgAll"].cpu().numpy()
    assert np.max(np.abs(ia - ib)) < 2e-6 * np.max(np.abs(ib))
    fa, fb = a["fy"].cpu().numpy(), b["fy"].cpu().numpy()
    assert np.max(np.abs(fa - fb)) < 2e-6 * np.max(np.abs(fb))
    np.testing.assert_allclose(a["ctf"].cpu().numpy(), b["ctf"].cpu().numpy(), atol=1e-6)
    va, vb = _operand_values(opa), _operand_values(opb)
    Kp = opa.Khp
    # per segment, relative to the segment's scale: the split-BF16 pair hi + lo holds a value to
    # about 2^-17 of itself, so values that differ by float32 rounding between the two paths can
    # land up to ~2 x 2^-17 apart after the split
    tol = 1e-5 if precision == 0 else 2e-5
    for sg in range(4):
        xa = va.reshape(n, 4, Kp)[:, sg]
        xb = vb.reshape(n, 4, Kp)[:, sg]
        assert np.max(np.abs(xa - xb)) < tol * np.max(np.abs(xb)), sg

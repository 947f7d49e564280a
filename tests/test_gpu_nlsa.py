"""GPU parity of Stage 3 (NLSA along one diffusion coordinate) against the float64 oracle."""
import numpy as np
import pytest

from conftest import FILTER, PARAMS

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pd_case():
    from manifoldem_python_amd import synth
    from oracle import distance as od, dm
    n, N = 120, 32
    s = synth.make_pd(synth.seeds(2)[1], n, N)
    out = od.op([s.ind, s.q, s.df, None, 0], FILTER, s.raw, s.nStot, dict(avgOnly=False),
                dict(PARAMS, nPix=N))
    res = dm.op(out["D"], n, 3, 60000, num_eigs=15, rng=np.random.default_rng(4))
    return dict(n=n, N=N, out=out, psi=res[1])


def _align(a, b):
    """Flip columns of a to match the signs of b."""
    s = np.sign(np.sum(a * b, axis=0))
    s[s == 0] = 1
    return a * s


def test_condist_f64_bitwise(cuda, pd_case):
    import torch
    from manifoldem_python_amd import _native as nat
    from oracle import nlsa
    D = pd_case["out"]["D"]
    n = pd_case["n"]
    C = 6
    order = np.argsort(pd_case["psi"][:, 0])
    ref = nlsa.condist(D[order][:, order], n, C)
    got = torch.empty((n - C) ** 2, dtype=torch.float64, device=cuda)
    nat.nlsa_condist(torch.as_tensor(D, device=cuda), torch.as_tensor(order, device=cuda), n, C,
                     got)
    np.testing.assert_array_equal(got.cpu().numpy().reshape(n - C, n - C), ref)


def test_supervectors_match_fft_loop(cuda, pd_case):
    import torch
    from manifoldem_python_amd import _native as nat
    from oracle import nlsa
    n, N = pd_case["n"], pd_case["N"]
    C, ell = 6, 8
    out = pd_case["out"]
    order = np.argsort(pd_case["psi"][:, 0])
    CTF = out["CTF"].reshape(n, N, N)
    W, CTF1 = nlsa.get_wiener(CTF, np.arange(n), order, C, n)
    IMG1 = out["imgAll"][order]
    rng = np.random.default_rng(1)
    mu_psi = rng.normal(size=(n - C, ell))
    ref = nlsa.supervectors(IMG1, CTF1, W, mu_psi, C, n)
    spec = torch.fft.rfft2(torch.as_tensor(IMG1, dtype=torch.float32, device=cuda)).contiguous()
    ch = torch.as_tensor(CTF1[:, :, :N // 2 + 1].reshape(n, -1), dtype=torch.float32,
                         device=cuda).contiguous()
    sv = torch.empty((C, ell, N, N // 2 + 1), dtype=torch.complex64, device=cuda)
    nat.nlsa_supervectors(spec, ch, n, N, C, torch.as_tensor(mu_psi, device=cuda), 5.0, sv)
    imgs = torch.fft.irfft2(sv, s=(N, N))
    A = imgs.permute(0, 3, 2, 1).reshape(C * N * N, ell).double().cpu().numpy()
    err = np.max(np.abs(A - ref)) / np.max(np.abs(ref))
    assert err < 2e-6, err


def test_nlsa_matches_oracle(cuda, pd_case):
    import torch
    from manifoldem_python_amd import spectral, fit1d
    from oracle import nlsa
    n, N = pd_case["n"], pd_case["N"]
    out = pd_case["out"]
    C = 6
    order = np.argsort(pd_case["psi"][:, 0])
    CTF = out["CTF"].reshape(n, N, N)
    DD = out["D"][order][:, order]
    par = dict(num=n, ConOrder=C, k=n - C, tune=3, nS=n, save=False, psiTrunc=8)
    ref = nlsa.op(par, DD, np.arange(n), order, out["imgAll"], 1, CTF,
                  rng=np.random.default_rng(9))
    IMGT_r, Topo_r, psirec_r, psiC1_r, sdiag_r, VX_r, mu_r, tau_r, extra = ref
    ctf_half = torch.as_tensor(CTF[:, :, :N // 2 + 1].reshape(n, -1), dtype=torch.float32,
                               device=cuda).contiguous()
    imgAll = torch.as_tensor(out["imgAll"], dtype=torch.float32, device=cuda)
    rng = np.random.default_rng(9)
    r = spectral.nlsa(torch.as_tensor(out["D"], device=cuda), torch.as_tensor(order, device=cuda),
                      torch.as_tensor(order, device=cuda), imgAll, ctf_half, N, C, 8, nS=n,
                      tune=3, num_eigs=15, rand=lambda *sh: rng.random(sh), keep=True)
    # first DM on ConD
    psiC = _align(r.psiC1, psiC1_r)
    assert np.max(np.abs(psiC - psiC1_r)) < 1e-4 * np.max(np.abs(psiC1_r))
    # singular values of the supervector matrix
    np.testing.assert_allclose(r.sdiag, sdiag_r, rtol=1e-5)
    # reconstructed, z-scored frames (invariant to eigenvector signs)
    IMGT = r.IMGT_dev.cpu().numpy()
    assert IMGT.shape == IMGT_r.shape
    assert np.max(np.abs(IMGT - IMGT_r)) < 1e-4 * np.max(np.abs(IMGT_r))
    # frame distances and the second embedding
    from oracle.nlsa import l2_distance
    D2r = l2_distance(IMGT_r, IMGT_r) ** 2
    off = ~np.eye(D2r.shape[0], dtype=bool)
    D2 = r.Drecon2.cpu().numpy().astype(np.float64)
    assert np.max(np.abs(D2 - D2r)[off] / D2r[off]) < 1e-4
    psirec = _align(r.psirec, psirec_r)
    assert np.max(np.abs(psirec - psirec_r)) < 1e-3 * np.max(np.abs(psirec_r))
    # tau from the sign-aligned coordinates
    _, _, tau = fit1d.fit(psirec)
    assert np.max(np.abs(tau - tau_r)) < 1e-3


@pytest.mark.parametrize("q,ncols", [(12, 70), (100, 300)])
def test_recon_dist_matches_l2_of_frames(cuda, q, ncols):
    """memi_nlsa_recon_dist from (Gq, H) against the oracle's L2_distance of the explicit
    z-scored frames IMGT = Ubar H diag(1/sd) (NLSA.py:112-146), float64 on the host."""
    import torch
    from manifoldem_python_amd import _native as nat
    from oracle.nlsa import l2_distance
    rng = np.random.default_rng(q)
    npix = 400
    Ubar = rng.normal(size=(npix, q))
    Ubar -= Ubar.mean(axis=0)
    H = rng.normal(size=(q, ncols))
    H[:, 5] = H[:, 3]                       # two identical frames: distance exactly 0
    frames = Ubar @ H
    sd = np.sqrt(np.sum(frames ** 2, axis=0) / npix)
    IMGT = frames / sd
    ref = l2_distance(IMGT, IMGT) ** 2
    Gq = torch.as_tensor(Ubar.T @ Ubar, device=cuda).contiguous()
    inv_sd = torch.empty(ncols, dtype=torch.float64, device=cuda)
    D2 = torch.empty((ncols, ncols), dtype=torch.float32, device=cuda)
    ws = torch.empty((q + 1) * ncols, dtype=torch.float64, device=cuda)
    nat.nlsa_recon_dist(Gq, torch.as_tensor(H, device=cuda).contiguous(), npix, inv_sd, D2, ws)
    np.testing.assert_allclose(inv_sd.cpu().numpy(), 1.0 / sd, rtol=1e-12)
    got = D2.cpu().numpy().astype(np.float64)
    assert np.all(np.diag(got) == 0) and got[3, 5] == 0 and got[5, 3] == 0
    off = ref > 1e-6
    assert np.max(np.abs(got - ref)[off] / ref[off]) < 1e-6
    np.testing.assert_allclose(got, got.T, rtol=1e-6, atol=1e-6)


def _curve_points(rng, n, A, B, noise):
    t = rng.uniform(0, 1, n)
    return B + A * np.cos(np.pi * t[:, None] * np.arange(1, 4)) + noise * rng.normal(size=(n, 3))


@pytest.mark.parametrize("noise,a3", [(0.0, 0.2), (0.05, 0.2), (0.5, 0.7), (0.05, 1e-3),
                                      (0.05, 1e-6)])
def test_fit1d_tau_update_matches_np_roots(cuda, noise, a3):
    """memi_fit1d_tau against the oracle's per-point np.roots solve (solve_d_R_d_tau_p_3D.py
    + R_p.py as restated in oracle/fit1d.py) on unflagged points, and the flagged points are
    the ones the host re-solves; the sums equal the fit loop's (fit_1D_open_manifold_3D.py:85-93)
    for the returned tau."""
    import torch
    from manifoldem_python_amd import _native as nat
    from oracle import fit1d as ofit
    rng = np.random.default_rng(int(noise * 100) + 7)
    n = 700                                  # three blocks, the last one partial
    A = np.array([1.0, -0.5, a3])
    B = rng.normal(size=3)
    x = _curve_points(rng, n, A, B, noise)
    a = A * (1 + 0.05 * rng.normal(size=3))
    b = B + 0.05 * rng.normal(size=3)
    xd = torch.as_tensor(x, device=cuda).contiguous()
    tau = torch.empty(n, dtype=torch.float64, device=cuda)
    flag = torch.empty(n, dtype=torch.int32, device=cuda)
    sums = torch.empty(13, dtype=torch.float64, device=cuda)
    ws = torch.empty(nat.fit1d_ws_len(n), dtype=torch.float64, device=cuda)
    nat.fit1d_tau(xd, a, b, tau, flag, sums, ws)
    t = tau.cpu().numpy()
    fl = flag.cpu().numpy().astype(bool)
    ref = np.array([ofit.solve_tau(x[p], a, b)[0] for p in range(n)]).ravel()
    assert np.max(np.abs(t - ref)[~fl], initial=0.0) < 1e-10
    if a3 < 1e-5:
        assert fl.all()                      # leading coefficient negligible: all re-solved
    else:
        assert fl.mean() < 0.02
    cj = np.cos(np.dot(t[:, None], np.pi * np.array([[1, 2, 3]])))
    want = np.concatenate([np.sum(cj ** 2, 0), np.sum(cj, 0), np.sum(x * cj, 0), np.sum(x, 0),
                           [fl.sum()]])
    np.testing.assert_allclose(sums.cpu().numpy(), want, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("noise", [0.01, 0.2])
def test_fit1d_matches_oracle(cuda, noise):
    """The whole fit (GPU tau updates, host 2x2 solves) against oracle/fit1d.op."""
    from manifoldem_python_amd import fit1d
    from oracle import fit1d as ofit
    rng = np.random.default_rng(3)
    psi = np.hstack([_curve_points(rng, 900, np.array([1.0, 0.4, 0.25]), np.zeros(3), noise),
                     rng.normal(size=(900, 2))])
    a, b, tau = fit1d.fit(psi, device=cuda)
    ar, br, tr = ofit.op(psi)
    np.testing.assert_allclose(a, ar, rtol=1e-9, atol=1e-12)
    np.testing.assert_allclose(b, br, rtol=1e-9, atol=1e-12)
    assert tau.shape == tr.shape
    assert np.max(np.abs(tau - tr)) < 1e-9


def test_nlsa_cuti_branch_matches_oracle(cuda, pd_case):
    """'cuti' supervectors (NLSA.py:54-55): the raw images imgAll[posPsi1], no Wiener filter,
    no mask; singular values and the reconstructed frames against the oracle."""
    import torch
    from manifoldem_python_amd import spectral
    from oracle import nlsa
    n, N = pd_case["n"], pd_case["N"]
    out = pd_case["out"]
    C = 6
    order = np.argsort(pd_case["psi"][:, 0])
    CTF = out["CTF"].reshape(n, N, N)
    DD = out["D"][order][:, order]
    par = dict(num=n, ConOrder=C, k=n - C, tune=3, nS=n, save=False, psiTrunc=8)
    ref = nlsa.op(par, DD, np.arange(n), order, out["imgAll"], 1, CTF,
                  rng=np.random.default_rng(9), branch="cuti")
    IMGT_r, sdiag_r, extra = ref[0], ref[4], ref[8]
    imgAll = torch.as_tensor(out["imgAll"], dtype=torch.float32, device=cuda)
    ctf_half = torch.zeros((n, N * (N // 2 + 1)), dtype=torch.float32, device=cuda)
    rng = np.random.default_rng(9)
    r = spectral.nlsa(torch.as_tensor(out["D"], device=cuda), torch.as_tensor(order, device=cuda),
                      torch.as_tensor(order, device=cuda), imgAll, ctf_half, N, C, 8, nS=n,
                      tune=3, num_eigs=15, rand=lambda *sh: rng.random(sh), keep=True,
                      wiener=False)
    A = r.A.cpu().numpy()
    Ar = extra["A"]
    # columns 1.. carry eigenvector signs: compare column by column up to sign
    for c in range(A.shape[1]):
        s = np.sign(A[:, c] @ Ar[:, c]) or 1.0
        assert np.max(np.abs(s * A[:, c] - Ar[:, c])) < 1e-4 * np.max(np.abs(Ar[:, c])), c
    np.testing.assert_allclose(r.sdiag, sdiag_r, rtol=1e-5)
    IMGT = r.IMGT_dev.cpu().numpy()
    assert np.max(np.abs(IMGT - IMGT_r)) < 1e-4 * np.max(np.abs(IMGT_r))


def test_nlsa_dropin_cuti_indexes_imgall_by_pospsi1(cuda, pd_case):
    """NLSA.op with 'cuti' reads imgAll[posPsi1] (not posPath[posPsi1]): a posPath that is a
    permutation leaves the result unchanged."""
    from manifoldem_python_amd import NLSA
    n, N = pd_case["n"], pd_case["N"]
    out = pd_case["out"]
    C = 6
    order = np.argsort(pd_case["psi"][:, 0])
    CTF = out["CTF"].reshape(n, N, N)
    DD = out["D"][order][:, order]
    par = dict(num=n, ConOrder=C, k=n - C, tune=3, nS=n, save=False, psiTrunc=8)
    perm = np.random.default_rng(3).permutation(n)
    s1 = NLSA.op(par, DD, np.arange(n), order, out["imgAll"], 1, CTF, {"cuti": 1})[4]
    s2 = NLSA.op(par, DD, perm, order, out["imgAll"], 1, CTF, {"cuti": 1})[4]
    np.testing.assert_allclose(s1, s2, rtol=1e-4)

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
    assert np.max(np.abs(r.Drecon2 - D2r)[off] / D2r[off]) < 1e-4
    psirec = _align(r.psirec, psirec_r)
    assert np.max(np.abs(psirec - psirec_r)) < 1e-3 * np.max(np.abs(psirec_r))
    # tau from the sign-aligned coordinates
    _, _, tau = fit1d.fit(psirec)
    assert np.max(np.abs(tau - tau_r)) < 1e-3

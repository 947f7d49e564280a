"""Host-side pieces of the GPU pipeline that need no GPU."""
import numpy as np


def test_lm_ritz_matches_full_decomposition():
    from manifoldem_python_amd import pipeline as pl
    rng = np.random.default_rng(0)
    for J, k in ((10, 4), (80, 9), (200, 16)):
        a = rng.normal(size=J)
        b = rng.uniform(0.1, 1.0, J - 1)
        T = np.diag(a) + np.diag(b, 1) + np.diag(b, -1)
        w, U = np.linalg.eigh(T)
        sel = np.argsort(-np.abs(w), kind="stable")[:k]
        th, S = pl.lm_ritz(a, b, k)
        np.testing.assert_allclose(np.sort(th), np.sort(w[sel]), atol=1e-12)
        for c in range(k):
            j = int(np.argmin(np.abs(w - th[c])))
            assert abs(abs(S[:, c] @ U[:, j]) - 1) < 1e-9


def test_geometry_matches_oracle():
    from manifoldem_python_amd import pipeline as pl, synth
    from oracle import distance as od
    s = synth.make_pd(synth.seeds(1)[0], 30, 24)
    g = pl.pd_geometry(s.q)
    PD = od.calc_avg_pd(s.q, 30).sum(1)
    PD = PD / np.linalg.norm(PD)
    np.testing.assert_allclose(g["PD"], PD, atol=1e-15)
    assert abs(g["psi_p"] - od.psi_ang(PD)) < 1e-12
    for i in range(30):
        Psi, sn, c = od.get_psi(s.q, PD, i)
        assert abs(g["Psi"][i] - (0 if np.isnan(Psi) else Psi)) < 1e-14


def test_fit1d_matches_oracle():
    from manifoldem_python_amd import fit1d
    from oracle import fit1d as ofit
    rng = np.random.default_rng(4)
    t = rng.uniform(0, 1, 150)
    psi = np.stack([np.cos(np.pi * t), 0.3 * np.cos(2 * np.pi * t), 0.1 * np.cos(3 * np.pi * t)],
                   1) + 0.05 * rng.normal(size=(150, 3))
    a1, b1, t1 = ofit.op(psi)
    a2, b2, t2 = fit1d.fit(psi)
    np.testing.assert_allclose(a1, a2, atol=1e-12)
    np.testing.assert_allclose(t1, t2, atol=1e-12)


def test_tau_bins_match_oracle():
    from manifoldem_python_amd import psiAnalysisParS2
    from oracle import nlsa
    rng = np.random.default_rng(6)
    tau = rng.uniform(0, 1, 300)
    IMGT = rng.normal(size=(400, 300))
    tn, numclass, inds = psiAnalysisParS2.tau_bins(tau, IMGT, 50)
    tr, IMG1, ir = nlsa.tau_bins(tau, IMGT, 50)
    np.testing.assert_array_equal(inds, ir)
    np.testing.assert_allclose(tn, tr)

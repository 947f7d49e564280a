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


def test_fit1d_host_resolve_matches_oracle():
    """The host re-solve of flagged points (batched companion eigenvalues) against the oracle's
    per-point np.roots, including a degenerate quintic (zero trailing coefficient)."""
    from manifoldem_python_amd import fit1d
    from oracle import fit1d as ofit
    rng = np.random.default_rng(4)
    t = rng.uniform(0, 1, 150)
    psi = np.stack([np.cos(np.pi * t), 0.3 * np.cos(2 * np.pi * t), 0.1 * np.cos(3 * np.pi * t)],
                   1) + 0.05 * rng.normal(size=(150, 3))
    a, b = ofit.initial_guess(psi)
    x = psi.copy()
    x[0, 0] = b[0] + 3 * a[2] * (x[0, 2] - b[2]) / a[0]     # trailing coefficient ~ 0
    got = fit1d.solve_tau(x, a, b)
    ref = np.array([ofit.solve_tau(x[p], a, b)[0] for p in range(len(x))]).ravel()
    np.testing.assert_allclose(got, ref, atol=1e-12)


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


def test_recursive_bspline_prefilter_equals_fft_division():
    """The fused alignment kernel prefilters each image line with the exact periodic recursive
    filter (pole z = sqrt(3) - 2, initial values from 24-term periodic sums, gain 6).  Restated
    here in float64, it must equal the frequency-domain division by (4 + 2 cos w) / 6 that the
    FFT path uses, including lines shorter than the 24-term sums."""
    z = np.sqrt(3.0) - 2.0

    def recursive(s, nt=24):
        N = len(s)
        zn = z ** N
        cp = np.empty(N)
        cp[0] = sum(z ** j * s[(-j) % N] for j in range(min(N, nt))) / (1 - zn)
        for k in range(1, N):
            cp[k] = s[k] + z * cp[k - 1]
        cm = np.empty(N)
        cm[N - 1] = -z / (1 - zn) * sum(z ** j * cp[(N - 1 + j) % N] for j in range(min(N, nt)))
        for k in range(N - 2, -1, -1):
            cm[k] = z * (cm[k + 1] - cp[k])
        return 6 * cm

    rng = np.random.default_rng(3)
    for N in (8, 9, 23, 24, 25, 70, 128, 160, 200):
        s = rng.normal(size=N)
        k = np.arange(N)
        ref = np.real(np.fft.ifft(np.fft.fft(s) / ((4 + 2 * np.cos(2 * np.pi * k / N)) / 6)))
        assert np.max(np.abs(recursive(s) - ref)) < 1e-12 * np.max(np.abs(ref)), N


def test_uniform_grid_detection():
    """The moment-based Ferguson sweep is used only for log-uniform eps grids."""
    from manifoldem_python_amd import pipeline as pl
    g = pl.uniform_grid(pl.LOG_EPS)
    assert g is not None and g[0] == -150.0 and abs(g[1] - 0.2) < 1e-12
    assert pl.uniform_grid(np.array([0.0, 1.0, 3.0])) is None
    assert pl.uniform_grid(np.array([1.0])) is None
    assert pl.uniform_grid(np.array([2.0, 1.0, 0.0])) is None


def test_check_pacing_predicts_last_convergence():
    """The Lanczos check pacing on a measured residual history (one PD, n = 2000, log10 of
    residual / target = 7.21, 4.42, 0.94, -2.72 at 32, 40, 48, 56 steps): after two checks the
    next check is one block on, and the predicted convergence step is the log-linear
    extrapolation of the last two checks."""
    from manifoldem_python_amd import pipeline as pl
    tol, th = 1e-10, np.array([1.0])
    res = lambda l10: np.array([tol * 10.0 ** l10])
    act = np.array([1], dtype=np.int32)
    gap, hist, last = pl._next_check_gap(32, res(7.21), th, act, tol, 8, None)
    assert gap == 8 and last is None                  # no trend from one check
    gap, hist, last = pl._next_check_gap(40, res(4.42), th, act, tol, 8, hist)
    assert gap == 8 and abs(last - (40 + 4.42 / ((7.21 - 4.42) / 8))) < 1e-9
    assert last > 48                                  # the check at 48 is predicted not final
    gap, hist, last = pl._next_check_gap(48, res(0.94), th, act, tol, 8, hist)
    assert gap == 8 and last < 56                     # ... and the one at 56 is
    # a PD whose residual stopped falling leaves the prediction open
    two = np.array([1, 1], dtype=np.int32)
    _, h2, _ = pl._next_check_gap(32, np.r_[res(7.0), res(6.0)], np.r_[th, th], two, tol, 8, None)
    _, _, last = pl._next_check_gap(40, np.r_[res(5.0), res(6.5)], np.r_[th, th], two, tol, 8, h2)
    assert last is None


def test_batched_inv_matches_per_matrix_inv():
    """The fit's batched inverse equals np.linalg.inv matrix by matrix, with NaN exactly where
    inv rejects the matrix (singular J^T J of a saturated fit), and NaN in, NaN out."""
    from manifoldem_python_amd import pipeline as pl
    rng = np.random.default_rng(9)
    J = rng.normal(size=(40, 6, 4))
    A = np.einsum("bij,bik->bjk", J, J)
    A[3] = 0.0                                       # all-zero
    A[7][:, 2] = A[7][2, :] = 0.0                    # a zero column and row
    A[11][0, 0] = np.nan
    got = pl.batched_inv(A)
    for q in range(len(A)):
        try:
            ref = np.linalg.inv(A[q])
        except np.linalg.LinAlgError:
            assert np.all(np.isnan(got[q])), q
            continue
        np.testing.assert_array_equal(np.isnan(got[q]), np.isnan(ref))
        np.testing.assert_allclose(got[q][~np.isnan(ref)], ref[~np.isnan(ref)], rtol=1e-12)


def test_fit1d_curve_solve_matches_lstsq():
    """The 1-D fit's batched 2x2 solve equals the reference's per-coordinate lstsq, including
    a singular system (all tau equal), where lstsq's minimum-norm solution is kept."""
    from manifoldem_python_amd import fit1d
    rng = np.random.default_rng(5)
    nS = 300
    for singular in (False, True):
        tau = np.full(nS, 0.3) if singular else rng.uniform(0, 1, nS)
        cj = np.cos(np.pi * tau[:, None] * np.array([[1, 2, 3]]))
        x = rng.normal(size=(nS, 3))
        s = np.concatenate([np.sum(cj ** 2, 0), np.sum(cj, 0), np.sum(x * cj, 0), np.sum(x, 0)])
        a, b = fit1d._lstsq_curve(s, nS)
        for q in range(3):
            M = np.array([[s[q], s[3 + q]], [s[3 + q], nS]])
            ref = np.linalg.lstsq(M, np.array([s[6 + q], s[9 + q]]), rcond=None)[0]
            np.testing.assert_allclose([a[q], b[q]], ref, rtol=1e-9, atol=1e-12)


def test_embedd_kept_points_selects_the_surviving_points():
    """embedd.py:32-38: the GUI's new points are copies of the original points outside the
    user's polygon, so the selection is exactly their indices (in the original order)."""
    from manifoldem_python_amd import embedd
    rng = np.random.default_rng(7)
    x, y = rng.normal(size=50), rng.normal(size=50)
    keep = np.sort(rng.choice(50, 31, replace=False))
    got = embedd.kept_points(zip(x, y), zip(x[keep], y[keep]))
    np.testing.assert_array_equal(got, keep)
    # the reference compares bit patterns: x equal to a kept point's y coordinate also counts
    x2 = x.copy()
    drop = np.setdiff1d(np.arange(50), keep)[0]
    x2[drop] = y[keep[0]]
    got = embedd.kept_points(zip(x2, y), zip(x2[keep], y[keep]))
    np.testing.assert_array_equal(got, np.sort(np.append(keep, drop)))
    # -0.0 and 0.0 differ as bit patterns
    got = embedd.kept_points([(-0.0, 1.0), (0.5, 2.0)], [(0.0, 1.0), (0.5, 2.0)])
    np.testing.assert_array_equal(got, [1])

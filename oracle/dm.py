"""Oracle for Stage 2: diffusion-map embedding of one distance matrix (float64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``DMembeddingII.py:16-185``, ``fergusonE.py:21-59``, ``slaplacianonFly.py:42-81``,
``sembeddingonFly.py:17-39`` and the trimming loop of ``manifoldTrimmingAuto.py:38-95``.

Deviations (numerical only, documented):
  * ``fergusonE`` sums with ``np.sum`` (pairwise) by default where the reference uses the
    Python builtin ``sum`` (sequential); both are float64, difference <= 1e-13 relative.
    ``builtin=True`` sums as the reference does (the CPU baseline times that form).
  * the unseeded ``np.random.rand`` restarts of the reference (``DMembeddingII.py:142``,
    ``fergusonE.py:52``) take an explicit ``rng`` here so tests are reproducible.
  * ``eigsh`` may be replaced by a dense ``eigh`` (``dense=True``) with the same
    largest-magnitude selection, for checking the ARPACK path itself.

How each part is pinned (``tests/test_oracle_dm.py``):
  knn/assembly  - brute force on n <= 8 with distinct values, and the symmetrisation identity;
  ferguson      - closed forms: eps -> inf gives log(#terms), eps -> 0 gives log(#zeros),
                  log-log slope equals sum(W d)/sum(W) (central differences);
  normalisation - row sums of D^-1 W equal 1 after alpha=1 normalisation, top eigenvalue 1
                  with eigenvector proportional to sqrt(degree), sum(mu) = 1;
  eigensolve    - ARPACK against dense eigh.
"""
from __future__ import annotations

import numpy as np
from scipy.optimize import curve_fit, OptimizeWarning
from scipy.sparse import csr_matrix, csc_matrix
from scipy.sparse.linalg import eigsh, ArpackNoConvergence
import warnings

warnings.simplefilter("ignore", OptimizeWarning)

LOG_EPS = np.arange(-150, 150.2, 0.2)          # DMembeddingII.py:146 (1501 points)


def initialize(nS, nN, D):
    """``DMembeddingII.py:43-57``: per-column sort; mutates the diagonal of D to -inf."""
    yInd1 = np.zeros((nN, nS), dtype="int32")
    yVal1 = np.zeros((nN, nS), dtype="float64")
    for iS in range(nS):
        D[iS, iS] = -np.inf
        B = np.sort(D[:, iS])
        IX = np.argsort(D[:, iS])
        yInd1[:, iS] = IX[:nN]
        yVal1[:, iS] = B[:nN]
        yVal1[0, iS] = 0
    return yInd1.flatten("F"), yVal1.flatten("F")


def assemble(D, k):
    """``DMembeddingII.py:86-140``: kNN lists -> COO (yRow, yCol, yVal) of squared distances,
    zero entries first (value 0), then the symmetrised non-zero entries."""
    nS = D.shape[0]
    nN = k
    yInd1, yVal1 = initialize(nS, nN, D)
    # get_yColVal with a single batch (nB = nS)
    yVal = yVal1.reshape(nS, nN).T[:nN, :].copy()
    yVal[0, :] = 0
    yVal = yVal.reshape(nN * nS, 1)
    yCol = yInd1.reshape(nS, nN).T[:nN, :].reshape(nN * nS, 1).astype(float)
    yRow = (np.ones((nN, 1)) * range(nS)).reshape(nS * nN, 1)
    ifZero = yVal < 1e-6
    yRowNZ = yRow[~ifZero]
    yColNZ = yCol[~ifZero]
    yValNZ = np.sqrt(yVal[~ifZero])
    yRowZ = yRow[ifZero]
    yColZ = yCol[ifZero]
    nZ = len(yRowZ)
    # construct_matrix0 (DMembeddingII.py:73-78)
    y = csr_matrix((yValNZ, (yRowNZ, yColNZ)), shape=(nS, nS)).toarray()
    y2 = y * y.T
    y = y ** 2
    y = y + y.T - y2
    nzr, nzc = y.nonzero()
    yValNZ = y[nzr, nzc]
    z = csr_matrix((np.ones(nZ), (yRowZ, yColZ)), shape=(nS, nS)).toarray()
    zr, zc = z.nonzero()
    yRow = np.hstack((zr, nzr)).astype(int)
    yCol = np.hstack((zc, nzc)).astype(int)
    yVal = np.hstack((np.zeros(len(zr)), yValNZ))
    return yRow, yCol, yVal


def fun(xx, aa0, aa1, aa2, aa3):
    """``fergusonE.py:21-23``."""
    return aa3 + aa2 * np.tanh(aa0 * xx + aa1)


def find_thres(logEps, D2):
    """``fergusonE.py:25-31``."""
    eps = np.exp(logEps)
    d = 1.0 / (2.0 * np.max(eps)) * D2
    ss = np.sum(np.exp(-np.sort(d)))
    return max(-np.log(0.01 * ss / len(D2)), 10)


def ferguson_sums(Dv, logEps, builtin=False):
    """``fergusonE.py:33-43``: logSumWij[k] = log sum_{d < thr} exp(-d), d = D^2/(2 eps_k).
    builtin: sum with Python's builtin ``sum`` over the array as the reference does (``:43``;
    sequential float64 adds, slower) instead of ``np.sum`` (pairwise)."""
    D2 = Dv * Dv
    thr = find_thres(logEps, D2)
    out = np.zeros(len(logEps))
    for k in range(len(logEps)):
        d = 1.0 / (2.0 * np.exp(logEps[k])) * D2
        w = np.exp(-d[d < thr])
        out[k] = np.log(sum(w) if builtin else np.sum(w))
    return out, thr


def ferguson_fit(logEps, logSumWij, a0, rng):
    """``fergusonE.py:45-57``: tanh fit restarted while sum(sqrt|diag pcov|) > 100."""
    resnorm = np.inf
    while resnorm > 100:
        popt, pcov = curve_fit(fun, logEps, logSumWij, p0=np.ravel(a0))
        resnorm = np.sum(np.sqrt(np.fabs(np.diag(pcov))))
        a0 = 1 * (rng.random((4, 1)) - 0.5)
        residuals = logSumWij - fun(logEps, *popt)
        ss_res = np.sum(residuals ** 2)
        ss_tot = np.sum((logSumWij - np.mean(logSumWij)) ** 2)
        R_squared = 1 - ss_res / ss_tot
    return popt, resnorm, R_squared


def fergusonE(Dv, logEps, a0, rng, builtin=False):
    """``fergusonE.op`` -> (popt, logSumWij, resnorm, R_squared)."""
    logSumWij, _ = ferguson_sums(Dv, logEps, builtin)
    popt, resnorm, R2 = ferguson_fit(logEps, logSumWij, a0, rng)
    return popt, logSumWij, resnorm, R2


def slaplacian(yVal, yCol, yRow, nS, sigma, alpha=1.0):
    """``slaplacianonFly.py:42-81``: Gaussian weights, alpha normalisation, symmetric
    normalisation by sqrt of column sums, l = |l + l^T| / 2."""
    yVal = np.exp(-(yVal / sigma ** 2))
    l = csc_matrix((yVal, (yRow, yCol)), shape=(nS, nS))
    d = np.asarray(l.sum(axis=0)).ravel()
    if alpha != 1:
        d = d ** alpha
    yVal = yVal / (d[yRow] * d[yCol])
    l = csc_matrix((yVal, (yRow, yCol)), shape=(nS, nS))
    d = np.sqrt(np.asarray(l.sum(axis=0)).ravel())
    yVal = yVal / (d[yRow] * d[yCol])
    l = csc_matrix((yVal, (yRow, yCol)), shape=(nS, nS))
    return abs(l + l.T) / 2.0


def top_eigs(l, k, dense=False):
    """``sembeddingonFly.py:17-39``: eigsh(which='LM', k), sorted descending."""
    if dense:
        lam, V = np.linalg.eigh(l.toarray() if hasattr(l, "toarray") else l)
        sel = np.argsort(-np.abs(lam), kind="stable")[:k]
        lam, V = lam[sel], V[:, sel]
    else:
        try:
            lam, V = eigsh(l, k=k, maxiter=300)
        except ArpackNoConvergence as e:
            lam, V = e.eigenvalues, e.eigenvectors
    ix = np.argsort(lam)[::-1]
    return lam[ix], V[:, ix]


def op(D, k, tune, prefsigma, num_eigs=15, rng=None, dense=False, logEps=None, builtin=False):
    """``DMembeddingII.op`` -> (lamb, psi, sigma, mu, logEps, logSumWij, popt, R_squared).

    D is copied (the reference mutates the caller's diagonal in place)."""
    rng = np.random.default_rng(0) if rng is None else rng
    D = np.array(D, dtype=np.float64, copy=True)
    nS = D.shape[0]
    yRow, yCol, yVal = assemble(D, k)
    a0 = rng.random((4, 1)) - 0.5
    logEps = LOG_EPS if logEps is None else logEps
    popt, logSumWij, resnorm, R2 = fergusonE(np.sqrt(yVal), logEps, a0, rng, builtin)
    nEigs = min(num_eigs, nS - 3)
    sigma = tune * np.sqrt(2 * np.exp(-popt[1] / popt[0]))
    l = slaplacian(yVal, yCol, yRow, nS, sigma)
    lamb, v = top_eigs(l, nEigs + 1, dense=dense)
    true_shape = v.shape[1] - 1
    psi = np.zeros((v.shape[0], nEigs))
    psi[:, :true_shape] = v[:, 1:] / v[:, 0:1]
    mu = v[:, 0] ** 2
    return lamb, psi, sigma, mu, logEps, logSumWij, popt, R2


def kernel_dense(yRow, yCol, yVal, nS, sigma):
    """Dense symmetric normalised kernel from the COO lists (helper for checks)."""
    return slaplacian(yVal, yCol, yRow, nS, sigma).toarray()


def trim(D, tune, rad, num_eigs=15, rng=None, dense=False):
    """``manifoldTrimmingAuto.py:48-77``: embed, keep ||psi[:, 0:3]|| < rad, re-embed until
    no point is removed.  Returns (posPath, last DM tuple)."""
    nS = D.shape[0]
    posPath = np.arange(nS)
    res = op(D, nS, tune, 60000, num_eigs, rng, dense)
    psi = res[1]
    posPath1 = np.nonzero(np.sqrt(np.sum(psi[:, 0:3] ** 2, axis=1)) < rad)[0]
    while len(posPath1) < nS:
        nS = len(posPath1)
        D1 = D[posPath1][:, posPath1]
        res = op(D1, D1.shape[0], tune, 600000, num_eigs, rng, dense)
        res = (res[0][res[0] > 0],) + tuple(res[1:])          # manifoldTrimmingAuto.py:68
        psi = res[1]
        posPathInt = np.nonzero(np.sqrt(np.sum(psi[:, 0:3] ** 2, axis=1)) < rad)[0]
        posPath1 = posPath1[posPathInt]
    return posPath[posPath1], res

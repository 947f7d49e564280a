"""Oracle for Stage 3: NLSA along one diffusion coordinate (float64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``NLSA.py:23-158``, ``get_wiener.py:9-21``, ``svdRF.py:9-35``, ``L2_distance.py:8-41``
and the per-psi body of ``psiAnalysisParS2.py:45-170`` (sort along psi, NLSA, tau binning),
keeping the reference's index conventions, including the +1 offset between the Wiener
window (``get_wiener.py:16``: CTF index ConOrder-ii+i) and the image it divides
(``NLSA.py:70``: image index ConOrder-ii+i-1).

Pinned by (``tests/test_oracle_nlsa.py``):
  condist   - brute-force triple loop of the commented-out form at ``NLSA.py:36-41``;
  supervec  - the per-(ii, i) FFT loop equals the linearity-restructured spectral sum, and
              with CTF = W = 1 it equals the raw ('cuti') window sum;
  svd       - ``np.linalg.svd`` (singular values; vectors up to sign);
  l2        - brute-force pairwise norms;
  binning   - every class index lands in its tau bin.
"""
from __future__ import annotations

import numpy as np
from scipy.fftpack import fft2, ifft2

from . import dm as odm
from . import fit1d as ofit


def condist(DD, num, ConOrder):
    """``NLSA.py:31-34``: ConD = sum_{i<ConOrder} DD[i:i+m, i:i+m] (sequential order)."""
    m = num - ConOrder
    ConD = np.zeros((m, m))
    for i in range(ConOrder):
        ConD += DD[i:i + m, i:i + m]
    return ConD


def get_wiener(CTF, posPath, posPsi1, ConOrder, num):
    """``get_wiener.py:9-21``: W_i = sum_{ii<ConOrder} CTF1[ConOrder-ii+i]^2 + 1/5."""
    dim = CTF.shape[1]
    CTF1 = CTF[posPath[posPsi1], :, :]
    wiener_dom = np.zeros((num - ConOrder, dim, dim))
    for i in range(num - ConOrder):
        for ii in range(ConOrder):
            wiener_dom[i] += CTF1[ConOrder - ii + i] ** 2
    return wiener_dom + 1.0 / 5, CTF1


def supervectors(IMG1, CTF1, wiener_dom, mu_psi, ConOrder, num, msk2=1):
    """``NLSA.py:63-86``: A[ii*N^2:(ii+1)*N^2, :] = tmp_ii @ mu_psi with tmp columns the
    Wiener-filtered images flattened in column-major order."""
    dim = IMG1.shape[1]
    ell1 = mu_psi.shape[1]
    A = np.zeros((ConOrder * dim * dim, ell1))
    tmp = np.zeros((dim * dim, num - ConOrder))
    for ii in range(ConOrder):
        for i in range(num - ConOrder):
            ind3 = ConOrder - ii + i - 1
            img_f = fft2(IMG1[ind3])
            img = ifft2(img_f * (CTF1[ind3] / wiener_dom[i])).real * msk2
            tmp[:, i] = img.T.reshape(-1)
        A[ii * dim * dim:(ii + 1) * dim * dim, :] = np.matmul(tmp, mu_psi)
    return A


def supervectors_raw(IMG1, mu_psi, ConOrder, num):
    """``NLSA.py:54-55,63-86`` 'cuti' branch: the same window sums as ``supervectors`` over the
    images themselves (no Wiener filter, no mask), flattened in column-major order."""
    dim = IMG1.shape[1]
    A = np.zeros((ConOrder * dim * dim, mu_psi.shape[1]))
    tmp = np.zeros((dim * dim, num - ConOrder))
    for ii in range(ConOrder):
        for i in range(num - ConOrder):
            tmp[:, i] = IMG1[ConOrder - ii + i - 1].T.reshape(-1)
        A[ii * dim * dim:(ii + 1) * dim * dim, :] = np.matmul(tmp, mu_psi)
    return A


def svdRF(A):
    """``svdRF.py:9-35``."""
    D1, D2 = A.shape

    def tidy(D, EV):
        order = np.argsort(D)[::-1]
        D = np.sort(D)[::-1]
        EV = EV[:, order]
        sqrtD = np.sqrt(D)
        return D, EV, np.diag(sqrtD), np.diag(1.0 / sqrtD)

    if D1 > D2:
        D, V = np.linalg.eigh(np.matmul(A.T, A))
        D, V, S, invS = tidy(D, V)
        U = np.matmul(A, np.matmul(V, invS))
    else:
        D, U = np.linalg.eigh(np.matmul(A, A.T))
        D, U, S, invS = tidy(D, U)
        V = np.matmul(A.T, np.matmul(U, invS))
    return U, S, V


def l2_distance(a, b):
    """``L2_distance.py:8-41``: sqrt(aa + bb - 2 a^T b) with entries < 1e-8 set to 0."""
    aa = np.sum(a ** 2, axis=0)
    bb = np.sum(b ** 2, axis=0)
    tmp = aa[:, None] + bb[None, :] - 2 * np.matmul(a.T, b)
    tmp[tmp < 1e-8] = 0
    return np.sqrt(tmp)


def reconstruct(U, S, V, psiC, ConOrder, num, nS, Npixel):
    """``NLSA.py:96-138``: Topo_mean, rank-2 ConImgT, IMGT accumulation, per-frame z-score."""
    psiTrunc = psiC.shape[1]
    sdiag = np.diag(S)
    Topo_mean = np.zeros((Npixel, psiTrunc))
    for ii in range(psiTrunc):
        Topo = np.stack([U[k * Npixel:(k + 1) * Npixel, ii] for k in range(ConOrder)], axis=1)
        Topo_mean[:, ii] = np.mean(Topo, axis=1)
    ConImgT = np.zeros((U.shape[0], psiTrunc))
    for i in range(0, 2):
        ConImgT = ConImgT + np.matmul(U[:, i].reshape(-1, 1), sdiag[i] * V[:, i].reshape(1, -1))
    IMGT = np.zeros((Npixel, nS - 2 * ConOrder))
    for i in range(ConOrder):
        tmp = np.matmul(ConImgT[i * Npixel:(i + 1) * Npixel, :], psiC.T)
        for ii in range(num - 2 * ConOrder):
            IMGT[:, ii] = IMGT[:, ii] + tmp[:, i + ii]
    for i in range(IMGT.shape[1]):
        t = IMGT[:, i]
        IMGT[:, i] = (t - np.mean(t)) / np.std(t)
    return IMGT, Topo_mean, sdiag


def op(NLSAPar, DD, posPath, posPsi1, imgAll, msk2, CTF, tune=3, num_eigs=15, rng=None,
       dense=False, logEps=None, branch="prD"):
    """``NLSA.op`` -> (IMGT, Topo_mean, psirec, psiC1, sdiag, VX, mu, tau); branch 'prD'
    (Wiener-filtered images posPath[posPsi1]) or 'cuti' (raw images posPsi1, ``NLSA.py:54-55``)."""
    rng = np.random.default_rng(0) if rng is None else rng
    num, ConOrder = NLSAPar["num"], NLSAPar["ConOrder"]
    k, nS, psiTrunc = NLSAPar["k"], NLSAPar["nS"], NLSAPar["psiTrunc"]
    ConD = condist(DD, num, ConOrder)
    lambdaC, psiC, sigmaC, mu, _, _, _, _ = odm.op(ConD, k, tune, 600000, num_eigs, rng, dense,
                                                   logEps)
    psiC1 = np.copy(psiC)
    if branch == "prD":
        IMG1 = imgAll[posPath[posPsi1], :, :]
        wiener_dom, CTF1 = get_wiener(CTF, posPath, posPsi1, ConOrder, num)
    else:
        IMG1 = imgAll[posPsi1, :, :]
    dim = CTF.shape[1]
    ell = psiTrunc - 1
    m = psiC.shape[0]
    psiC = np.hstack((np.ones((m, 1)), psiC[:, 0:ell]))
    mu_psi = mu.reshape((-1, 1)) * psiC
    if branch == "prD":
        A = supervectors(IMG1, CTF1, wiener_dom, mu_psi, ConOrder, num, msk2)
    else:
        A = supervectors_raw(IMG1, mu_psi, ConOrder, num)
    U, S, V = svdRF(A)
    VX = np.matmul(V.T, psiC.T)
    IMGT, Topo_mean, sdiag = reconstruct(U, S, V, psiC, ConOrder, num, nS, dim * dim)
    Drecon = l2_distance(IMGT, IMGT)
    nSrecon = min(IMGT.shape)
    _, psirec, _, mu2, _, _, _, _ = odm.op(Drecon ** 2, nSrecon, tune, 30, num_eigs, rng, dense,
                                           logEps)
    a, b, tau = ofit.op(psirec)
    return IMGT, Topo_mean, psirec, psiC1, sdiag, VX, mu2, tau, dict(ConD=ConD, A=A, U=U, S=S,
                                                                       V=V, psiC=psiC, mu=mu)


def tau_bins(tau, IMGT, nClass=50):
    """``psiAnalysisParS2.py:100-124``: min-max normalised tau and one frame per class."""
    nSrecon = min(IMGT.shape)
    numclass = int(min(nClass, np.floor(nSrecon / 2.0)))
    tau = (tau - min(tau)) / (max(tau) - min(tau))
    tauinds = []
    IMG1 = np.zeros((IMGT.shape[0], numclass))
    for i in range(numclass):
        ind1 = float(i) / numclass
        ind2 = ind1 + 1.0 / numclass
        if i == numclass - 1:
            tauind = ((tau >= ind1) & (tau <= ind2)).nonzero()[0]
        else:
            tauind = ((tau >= ind1) & (tau < ind2)).nonzero()[0]
        while tauind.size == 0:
            sc = 1.0 / (numclass * 2.0)
            ind1 = ind1 - sc * ind1
            ind2 = ind2 + sc * ind2
            tauind = ((tau >= ind1) & (tau < ind2)).nonzero()[0]
        IMG1[:, i] = IMGT[:, tauind[0]]
        tauinds.append(tauind[0])
    return tau, IMG1, tauinds

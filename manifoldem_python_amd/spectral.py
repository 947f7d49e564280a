"""NLSA along one diffusion coordinate on the GPU (``NLSA.op``, ``NLSA.py:23-158``).

Stages and where they run:
  a19 delay-embedding distances ConD          HIP (memi_nlsa_condist), float64 sums
  a20 diffusion map of ConD                   the Stage-2 GPU path (pipeline.dm_embed)
  a21 Wiener-filtered supervectors A          HIP spectral contraction (memi_nlsa_supervectors)
                                              + rocFFT inverse transforms, by linearity: one
                                              inverse FFT per (delay, eigenvector) instead of
                                              one FFT pair per (delay, image)
  a22 svdRF                                   HIP tall-skinny A^T A and A V S^-1; 8x8 eigh on host
  a23 rank-2 reconstruction IMGT + z-score    factored form: IMGT = Ubar H diag(1/sd), Ubar the
                                              pixel-centred 2*ConOrder columns of U, H the Hankel
                                              matrix of the rank-2 chronos (HIP GEMM)
  a24 L2 distances of IMGT frames             from the factored Gram H^T (Ubar^T Ubar) H with the
                                              L2_distance epilogue fused (memi_nlsa_recon_dist)
  a20 second diffusion map                    Stage-2 GPU path
  a25 1-D open-manifold fit                   HIP tau update + sums per iteration
                                              (memi_fit1d_tau), 2x2 solves on host
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat
from . import fit1d
from . import pipeline as pl


_plans = {}


def _plan(n, k, dev):
    """One-PD plan of size n with k neighbours, reused by every call of that size: the plan
    keeps the kernel-matrix buffer and the captured Lanczos graph, so the eight psi of a PD
    (and PDs of equal size) replay one graph instead of capturing their own."""
    key = (int(n), int(k), str(dev))
    p = _plans.get(key)
    if p is None:
        p = _plans[key] = nat.Plan([0], [n], [k], device=dev)
    return p


class NLSAResult:
    __slots__ = ("IMGT", "Topo_mean", "psirec", "psiC1", "sdiag", "VX", "mu", "tau", "a", "b",
                 "IMGT_dev", "Drecon2", "dm1", "dm2", "A", "U")

    def as_tuple(self):
        return (self.IMGT, self.Topo_mean, self.psirec, self.psiC1, self.sdiag, self.VX, self.mu,
                self.tau)


def nlsa(D, ord_d, img_idx, imgAll, ctf_half, N, ConOrder, psiTrunc, nS=None, tune=3,
         num_eigs=15, msk2=None, logEps=pl.LOG_EPS, rand=None, snr=5.0, want_imgt=True,
         keep=False, timer=None, k=None, wiener=True):
    """One psi of psiAnalysisParS2 / NLSA.op.

    D: (n, n) float32/float64 device distances of the PD (posPath order); ord_d: int64 device
    sort order along psi (PosPsi1), so DD = D[ord][:, ord]; img_idx: int64 device indices of
    the images in psi order into imgAll / ctf_half (posPath[PosPsi1]).  timer: an optional
    pipeline.StageTimer that gets one mark per stage (it synchronises the device at each).
    k: neighbours kept in the embedding of ConD (NLSAPar['k'], NLSA.py:45); None = all m.
    wiener: True for the 'prD' supervectors (Wiener-filtered, masked images, NLSA.py:52-53,
    66-78); False for the 'cuti' branch (NLSA.py:54-55), which contracts the images themselves
    (ctf_half and msk2 unused).
    """
    mark = timer.mark if timer is not None else (lambda name: None)
    if timer is not None:
        timer.start()
    dev = D.device
    num = int(ord_d.shape[0])
    nS = num if nS is None else nS
    m = num - ConOrder
    Npix = N * N
    Nh = N // 2 + 1
    r = NLSAResult()
    # a19
    ConD = torch.empty(m * m, dtype=torch.float64, device=dev)
    nat.nlsa_condist(D.reshape(-1), ord_d, num, ConOrder, ConD)
    mark("condist")
    # a20
    plan = _plan(m, m if k is None else max(1, min(int(k), m)), dev)
    sub = pl.StageTimer() if timer is not None else None
    dm1 = pl.dm_embed(plan, ConD.float(), tune, num_eigs, logEps, rand=rand, timer=sub)[0]
    psiC, mu = dm1.psi, dm1.mu
    r.psiC1 = psiC.copy()
    mark("dm1")
    if timer is not None:
        timer.t.update({"dm1." + k: v for k, v in sub.t.items()})
    # a21
    IMG1 = imgAll.index_select(0, img_idx)
    ell = psiTrunc - 1
    psiCx = np.hstack((np.ones((m, 1)), psiC[:, 0:ell]))
    mu_psi = torch.as_tensor(mu.reshape(-1, 1) * psiCx, device=dev).contiguous()
    if wiener:
        CTF1 = ctf_half.index_select(0, img_idx).contiguous()
        spec = torch.fft.rfft2(IMG1).contiguous()
        sv = torch.empty((ConOrder, psiTrunc, N, Nh), dtype=torch.complex64, device=dev)
        nat.nlsa_supervectors(spec, CTF1, num, N, ConOrder, mu_psi, snr, sv)
        imgs = torch.fft.irfft2(sv, s=(N, N))
        if msk2 is not None:
            imgs = imgs * msk2
        # column-major flatten of each image (NLSA.py:80: img.T.reshape(-1))
        A = imgs.permute(0, 3, 2, 1).reshape(ConOrder * Npix, psiTrunc).double().contiguous()
        del sv, imgs, spec
    else:
        # block ii = X[ConOrder-ii-1 : ConOrder-ii-1+m]^T @ mu_psi, X the column-major
        # flattened images in psi order (NLSA.py:63-86 without the filter)
        X = IMG1.double().transpose(1, 2).reshape(num, Npix)
        A = torch.empty((ConOrder * Npix, psiTrunc), dtype=torch.float64, device=dev)
        for ii in range(ConOrder):
            s0 = ConOrder - ii - 1
            torch.mm(X[s0:s0 + m].t(), mu_psi, out=A[ii * Npix:(ii + 1) * Npix])
        del X
    del IMG1
    mark("supervectors")
    # a22
    G = torch.empty((psiTrunc, psiTrunc), dtype=torch.float64, device=dev)
    nat.tall_gram(A, G)
    Dv, Vv = np.linalg.eigh(G.cpu().numpy())
    order = np.argsort(Dv)[::-1]
    Dv = np.sort(Dv)[::-1]
    Vv = Vv[:, order]
    S = np.sqrt(Dv)
    U = torch.empty_like(A)
    nat.tall_apply(A, Vv @ np.diag(1.0 / S), U)
    r.VX = Vv.T @ psiCx.T
    r.sdiag = S
    mark("svd")
    # a23
    U3 = U.view(ConOrder, Npix, psiTrunc)
    r.Topo_mean = U3.mean(dim=0).cpu().numpy()
    n_rec = nS - 2 * ConOrder
    n_fill = num - 2 * ConOrder
    Ubar = U3[:, :, 0:2].permute(1, 2, 0).reshape(Npix, 2 * ConOrder).contiguous()
    z = (S[0:2, None] * (Vv[:, 0:2].T @ psiCx.T))          # (2, m): z_r(j)
    H = np.zeros((2 * ConOrder, n_rec))
    jj = np.arange(ConOrder)[:, None] + np.arange(n_fill)[None, :]
    H[:ConOrder, :n_fill] = z[0][jj]
    H[ConOrder:, :n_fill] = z[1][jj]
    Gq = torch.empty((2 * ConOrder, 2 * ConOrder), dtype=torch.float64, device=dev)
    nat.center_gram(Ubar, Gq)
    Hd = torch.as_tensor(H, device=dev).contiguous()
    inv_sd = torch.empty(n_rec, dtype=torch.float64, device=dev)
    # a24: L2_distance on the z-scored frames, from the factored Gram H^T Gq H
    Drecon2 = torch.empty((n_rec, n_rec), dtype=torch.float32, device=dev)
    ws = torch.empty((2 * ConOrder + 1) * n_rec, dtype=torch.float64, device=dev)
    nat.nlsa_recon_dist(Gq, Hd, Npix, inv_sd, Drecon2, ws)
    del ws
    if want_imgt:
        IMGT = torch.empty((Npix, n_rec), dtype=torch.float64, device=dev)
        nat.nlsa_imgt(Ubar, Hd, inv_sd, IMGT)
        r.IMGT_dev = IMGT
    r.Drecon2 = Drecon2
    mark("reconstruct")
    # second DM (NLSA.py:141-146)
    nSrecon = min(Npix, n_rec)
    plan2 = _plan(n_rec, nSrecon, dev)
    sub = pl.StageTimer() if timer is not None else None
    dm2 = pl.dm_embed(plan2, Drecon2.reshape(-1), tune, num_eigs, logEps, rand=rand, timer=sub)[0]
    r.psirec, r.mu = dm2.psi, dm2.mu
    mark("dm2")
    if timer is not None:
        timer.t.update({"dm2." + k: v for k, v in sub.t.items()})
    # a25
    r.a, r.b, r.tau = fit1d.fit(r.psirec, device=dev)
    mark("fit1d")
    if keep:
        r.dm1, r.dm2, r.A, r.U = dm1, dm2, A, U
    return r

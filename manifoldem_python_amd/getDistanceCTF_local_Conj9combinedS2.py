"""Drop-in for ``getDistanceCTF_local_Conj9combinedS2`` (Stage 1: CTF-aware distances of one PD).

``op(input_data, filterPar, imgFileName, sh, nStot, options)`` keeps the reference signature
(``getDistanceCTF_local_Conj9combinedS2.py:216``) and writes the same 19-key pickle plus the
progress marker (``:409-419``).  The work runs on the GPU: image preparation as rocFFT
transforms with HIP kernels between them, the distance Gram on the MFMA units
(``p.precision``: split-BF16 or FP32), averages as HIP reductions.

RELION input (``options['relion_data']``, ``:259-264``) is read from the MRCS stack with numpy
(``mrc.py``) and shifted on the GPU (``memi_shift_images``); a volume mask ``p.mask_vol_file``
(``:304-310``) is projected along the PD on the GPU (``memi_project_mask``, ``projectMask``).
"""
from __future__ import annotations

import math
import os

import numpy as np
import torch

from . import _native as nat
from . import mrc, myio, p
from . import pipeline as pl

version = "getDistanceCTF_local9, V 1.0"


def create_grid(N, N2):
    """``:139-154`` (host helper kept for callers)."""
    if N % 2 == 1:
        a = np.arange(-(N - 1) / 2, (N - 1) / 2 + 1)
    else:
        a = np.arange(-N2, N / 2)
    X, Y = np.meshgrid(a, a)
    return (1.0 / (N / 2.0)) * np.sqrt(X ** 2 + Y ** 2)


calc_avg_pd = pl.calc_avg_pd
psi_ang = pl.psi_ang


def get_psi(q, PD, iS):
    """``:187-202`` for one image."""
    s = -(1 + PD[2]) * q[3, iS] - PD[0] * q[1, iS] - PD[1] * q[2, iS]
    c = (1 + PD[2]) * q[0, iS] + PD[1] * q[1, iS] - PD[0] * q[2, iS]
    with np.errstate(divide="ignore", invalid="ignore"):
        return 2 * np.arctan(s / c), s, c


def precision_code(name=None):
    name = p.precision if name is None else name
    if name in ("split_bf16", "bf16", nat.PREC_SPLIT_BF16):
        return nat.PREC_SPLIT_BF16
    if name in ("fp32", nat.PREC_FP32):
        return nat.PREC_FP32
    raise ValueError("unknown distance precision %r" % (name,))


def read_spider(imgFileName, src, N):
    """Raw float32 SPIDER images by index (np.memmap, ``:254-258``), as one array."""
    mm = np.memmap(imgFileName, dtype="float32", mode="r")
    out = np.empty((len(src), N, N), dtype=np.float32)
    for j, s in enumerate(src):
        out[j] = mm[N * N * int(s):N * N * (int(s) + 1)].reshape(N, N)
    return out


def compute(ind, q, df, filterPar, stack, nStot, options, N=None, precision=None):
    """All Stage-1 outputs of one PD as a dict (numpy), from a SPIDER-order stack (numpy or
    device tensor) indexed by the reference's conjugate convention."""
    dev = pl.device()
    N = p.nPix if N is None else N
    ind = np.asarray(ind)
    nS = ind.shape[0]
    conj = ~(ind < nStot / 2)
    src = np.where(conj, ind - nStot // 2, ind).astype(np.int64)
    if isinstance(stack, torch.Tensor):
        stack_d = stack.to(dev)
    else:
        uniq, inv = np.unique(src, return_inverse=True)
        stack_d = torch.as_tensor(np.ascontiguousarray(stack[uniq]), device=dev)
        src = inv.astype(np.int64)
    geo = pl.pd_geometry(q)
    # volume mask projected along PD (:304-310), or none
    msk2_d, msk2 = None, 1
    if p.mask_vol_file:
        msk2_d, _ = pl.volume_mask(mrc.read_volume(p.mask_vol_file), geo["PD"])
        msk2 = msk2_d.cpu().numpy() > 0
    ctf = nat.ctf_params(p.pix_size, p.Cs, p.EkV, p.AmpContrast, p.gaussEnv)
    prec = precision_code(precision)
    op = pl.Operand(nS, N, prec, dev)
    out = pl.prepare_images(stack_d, torch.as_tensor(src, device=dev),
                            torch.as_tensor(conj.astype(np.int32), device=dev),
                            torch.as_tensor(-(180 / math.pi) * geo["Psi"], device=dev),
                            torch.full((nS,), -geo["psi_p"], dtype=torch.float64, device=dev),
                            torch.as_tensor(np.asarray(df, dtype=np.float64), device=dev), N,
                            filterPar, ctf, operand=op, msk2=msk2_d)
    avg = pl.averages(out["imgAll"], out["fy"], out["ctf"], N, msk2=msk2_d)
    if options.get("avgOnly", False):
        D = np.zeros((nS, nS))
    else:
        plan = nat.Plan([0], [nS], [nS], chunks=pl.CHUNKS[prec], device=dev)
        D = pl.ctf_distances(plan, op, out["fy"], out["ctf"], N, kappa=p.kappa).block(0)
        D = D.double().cpu().numpy()
    return dict(D=D, ind=ind, q=q, df=df,
                CTF=pl.full_ctf(out["ctf"], N).double().cpu().numpy(),
                imgAll=out["imgAll"].double().cpu().numpy(), msk2=msk2, PD=geo["PD"],
                PDs=geo["PDs"], Psis=geo["Psi"].reshape(-1, 1),
                imgAvg=avg["imgAvg"].cpu().numpy(), imgAvgFlip=avg["imgAvgFlip"].cpu().numpy(),
                imgAllFlip=avg["imgAllFlip"].double().cpu().numpy(),
                imgLabels=np.where(conj, -1, 1), Dnom=geo["Dnom"].reshape(-1, 1),
                Nom=geo["Nom"].reshape(-1, 1),
                imgAllIntensity=avg["imgAllIntensity"].cpu().numpy(), version=version,
                options=options)


KEYS = ["D", "ind", "q", "df", "CTF", "imgAll", "msk2", "PD", "PDs", "Psis", "imgAvg",
        "imgAvgFlip", "imgAllFlip", "imgLabels", "Dnom", "Nom", "imgAllIntensity", "version",
        "options"]


def write(outFile, res, prD):
    """The stage contract: pickle + empty progress marker (``:409-419``)."""
    myio.fout1(outFile, KEYS, [res[k] for k in KEYS])
    open(os.path.join(p.dist_prog, "%s" % (prD,)), "a").close()


def read_relion(imgFileName, src, sh):
    """RELION particles ``src`` of an MRCS stack (``:259-264``): read, then shifted by
    (sh[1] - 0.5, sh[0] - 0.5) with the order-3 'wrap' spline on the GPU; returned as a
    SPIDER-layout device stack."""
    img = mrc.read_images(imgFileName, src)
    s0 = np.asarray(sh[1], dtype=np.float64)[src] - 0.5
    s1 = np.asarray(sh[0], dtype=np.float64)[src] - 0.5
    return pl.relion_stack(img, s0, s1)


def op(input_data, filterPar, imgFileName, sh, nStot, options):
    ind, q, df, outFile, prD = input_data[:5]
    N = p.nPix
    ind = np.asarray(ind)
    conj = ~(ind < nStot / 2)
    src = np.where(conj, ind - nStot // 2, ind).astype(np.int64)
    uniq, inv = np.unique(src, return_inverse=True)
    if options.get("relion_data", False):
        stack = read_relion(imgFileName, uniq, sh)
    else:
        stack = read_spider(imgFileName, uniq, N)
    # compute() re-derives src from ind; give it the compact stack with remapped indices
    ind_c = np.where(conj, inv + nStot // 2, inv)
    res = compute(ind_c, q, df, filterPar, stack, nStot, options, N=N)
    res["ind"] = ind
    write(outFile, res, prD)


def get_wiener1(CTF1):
    """``:422-430``."""
    return np.sum(CTF1 ** 2, axis=0) + 1.0 / 5

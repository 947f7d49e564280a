"""Oracle for Stage 1: CTF-aware image distances of one PD (float64 / complex128).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``getDistanceCTF_local_Conj9combinedS2.py:139-430`` with its leaves
``annularMask.py:19-32``, ``rotatefill.py:21-28``, ``ctemh_cryoFrank.py:24-44``,
``q2Spider.py:19-60`` and ``qMult_bsx.py:15-40``.  Images are taken from an in-memory SPIDER
raw stack instead of ``np.memmap`` (same bytes, same ``.T``).

How each part is pinned (``tests/test_oracle_distance.py``):
  annular_mask  - closed-form disk with the off-centre row (brute force);
  create_grid/filter - closed forms at known points;
  rotatefill    - identity at 0 deg, exact pixel permutation at 90 deg (even N);
  ctf           - closed form at k=0 (-ampc) and at the zeros tan(wr)=ampc;
  q2spider      - quaternion rebuilt from the fitted angles equals the input up to sign;
  gram          - brute-force direct form sum_k |C_j x_i - C_i x_j|^2, Parseval with C=1;
  op            - flipped-conjugate and normalisation invariants, golden fixture.
"""
from __future__ import annotations

import math

import numpy as np
from scipy import optimize
from scipy.fftpack import fft2, ifft2, ifftshift
from scipy.ndimage import rotate


def annular_mask(a, b, N, M):
    """``annularMask.py:19-32``: ring a <= r < b centred on row N/2-1, column M/2."""
    xx = np.arange(N)[:, None] - N / 2 + 1
    yy = np.arange(M)[None, :] - M / 2
    r2 = xx * xx + yy * yy
    return ((r2 >= a * a) & (r2 < b * b)).astype(float)


def create_grid(N, N2):
    """``getDistanceCTF...:139-154``."""
    assert N > 0
    if N % 2 == 1:
        a = np.arange(-(N - 1) / 2, (N - 1) / 2 + 1)
    else:
        a = np.arange(-N2, N / 2)
    X, Y = np.meshgrid(a, a)
    return (1.0 / (N / 2.0)) * np.sqrt(X ** 2 + Y ** 2)


def create_filter(filter_type, NN, Qc, Q):
    """``getDistanceCTF...:156-167``."""
    if filter_type == "Gauss":
        return np.exp(-(np.log(2) / 2.0) * (Q / Qc) ** 2)
    if filter_type == "Butter":
        return np.sqrt(1.0 / (1 + (Q / Qc) ** (2 * NN)))
    raise ValueError("%s filter is unsupported" % filter_type)


def calc_avg_pd(q, nS):
    """``getDistanceCTF...:169-185``."""
    return 2 * np.vstack((q[1] * q[3] - q[0] * q[2],
                          q[0] * q[1] + q[2] * q[3],
                          q[0] ** 2 + q[3] ** 2 - np.ones((1, nS)) / 2.0))


def get_psi(q, PD, iS):
    """``getDistanceCTF...:187-202`` (arctan, not arctan2)."""
    s = -(1 + PD[2]) * q[3, iS] - PD[0] * q[1, iS] - PD[1] * q[2, iS]
    c = (1 + PD[2]) * q[0, iS] + PD[1] * q[1, iS] - PD[0] * q[2, iS]
    with np.errstate(divide="ignore", invalid="ignore"):
        Psi = 2 * np.arctan(s / c)
    return Psi, s, c


def qmult(q, s):
    """``qMult_bsx.py:15-40``."""
    q = q.reshape(4, -1)
    s = s.reshape(4, -1)
    q0, qv, s0, sv = q[0], q[1:4], s[0], s[1:4]
    c = np.vstack((qv[1] * sv[2] - qv[2] * sv[1],
                   qv[2] * sv[0] - qv[0] * sv[2],
                   qv[0] * sv[1] - qv[1] * sv[0]))
    return np.vstack((q0 * s0 - np.sum(qv * sv, axis=0), q0 * sv + s0 * qv + c))


def q2spider(q):
    """``q2Spider.py:19-60``: ZYZ angles by TRF least squares from x0 = 0, ftol 1e-12."""
    q = q / np.sqrt(np.sum(q ** 2))

    def dev1(a):
        q1 = np.array([np.cos(a[0] / 2.0), 0.0, 0.0, -np.sin(a[0] / 2.0)])
        q2 = np.array([np.cos(a[1] / 2.0), 0, -np.sin(a[1] / 2.0), 0.0])
        q3 = np.array([np.cos(a[2] / 2.0), 0.0, 0.0, -np.sin(a[2] / 2)])
        return q - qmult(q3, qmult(q2, q1)).flatten()

    res = optimize.least_squares(dev1, np.array([0, 0, 0]), bounds=(-np.inf, np.inf), ftol=1e-12)
    return res.x[0], res.x[1], res.x[2]


def psi_ang(PD):
    """``getDistanceCTF...:206-213``: in-plane reference angle in degrees."""
    Qr = np.array([1 + PD[2], PD[1], -PD[0], 0])
    Qr = Qr / np.sqrt(np.sum(Qr ** 2))
    _, _, psi = q2spider(Qr)
    return np.mod(psi, 2 * np.pi) * (180 / np.pi)


def spider_rotation(Phi, Theta, Psi):
    """``CC/projectMask.py:24-37`` (eulerRotMatrix3DSpider, radians)."""
    from math import cos, sin
    return np.array([[cos(Phi) * cos(Psi) * cos(Theta) + (-1) * sin(Phi) * sin(Psi),
                      cos(Psi) * cos(Theta) * sin(Phi) + cos(Phi) * sin(Psi),
                      (-1) * cos(Psi) * sin(Theta)],
                     [(-1) * cos(Psi) * sin(Phi) + (-1) * cos(Phi) * cos(Theta) * sin(Psi),
                      cos(Phi) * cos(Psi) + (-1) * cos(Theta) * sin(Phi) * sin(Psi),
                      sin(Psi) * sin(Theta)],
                     [cos(Phi) * sin(Theta), sin(Phi) * sin(Theta), cos(Theta)]])


def project_mask(vol, PD):
    """``CC/projectMask.py:118-140`` (projectMask.op) in float64: swap axes 0 and 2, rotate by
    the inverse of PD's Euler angles with psi = 0 (scipy affine_transform, order 3, mode
    'nearest', about the centre dims/2), sum along the last axis, transpose, threshold > 1.
    Returns (projection sums, mask)."""
    from scipy.ndimage import affine_transform
    vol = np.swapaxes(np.asarray(vol, dtype=np.float64), 0, 2)
    nPix = vol.shape[0]
    Qr = np.array([1 + PD[2], PD[1], -PD[0], 0])
    phi, theta, psi = q2spider(Qr / np.sqrt(np.sum(Qr ** 2)))
    sym = -np.array([phi, theta, 0.0])
    T = spider_rotation(sym[2], sym[1], sym[0])
    c = 0.5 * np.array(vol.shape)
    rho = affine_transform(input=vol, matrix=T, offset=c - T @ c, output_shape=vol.shape,
                           mode="nearest")
    proj = np.sum(rho, axis=2).reshape(nPix, nPix).T
    return proj, proj > 1


def rotatefill(img, angle):
    """``rotatefill.py:21-28``: order-3 spline rotation of the 3x3-tiled image, centre crop."""
    n = img.shape[0]
    rep = np.tile(img, (3, 3))
    out = rotate(rep, angle, reshape=False)
    return out[n:2 * n, n:2 * n]


def ctf(k, params):
    """``ctemh_cryoFrank.py:24-44``; params = [Cs mm, df A, kV, gaussEnv, ampc]."""
    Cs = params[0] * 1.0e7
    df, kev, B, ampc = params[1], params[2], params[3], params[4]
    wav = 12.3986 / np.sqrt((2 * 511.0 + kev) * kev)
    w1 = np.pi * Cs * wav ** 3
    w2 = np.pi * wav * df
    k2 = k * k
    sigm = B / math.sqrt(2 * math.log(2))
    wi = np.exp(-k2 / (2 * sigm ** 2))
    wr = (0.5 * w1 * k2 - w2) * k2
    return (np.sin(wr) - ampc * np.cos(wr)) * wi


def load_normalized(raw, src, conj, msk):
    """``getDistanceCTF...:246-281`` for one SPIDER image: ``.T``, flipud, background norm
    with mean/std over the full N^2 array (ddof 0)."""
    tmp = np.asarray(raw[src], dtype=np.float64).T
    if conj:
        tmp = np.flipud(tmp)
    backg = tmp * (1 - msk)
    return (tmp - backg.mean()) / backg.std()


def load_normalized_relion(data, src, conj, msk, sh):
    """``getDistanceCTF...:259-281`` for one RELION image: ``mrcfile`` data[src] (no transpose),
    ``ndimage.shift`` by (sh[1][src] - 0.5, sh[0][src] - 0.5) with order 3 and mode 'wrap',
    flipud for a conjugate, background normalisation as ``load_normalized``."""
    from scipy.ndimage import shift
    tmp = np.asarray(data[src])
    tmp = shift(tmp, (sh[1][src] - 0.5, sh[0][src] - 0.5), order=3, mode="wrap")
    tmp = np.asarray(tmp, dtype=np.float64)
    if conj:
        tmp = np.flipud(tmp)
    backg = tmp * (1 - msk)
    return (tmp - backg.mean()) / backg.std()


def gram_distance(fy, CTF):
    """``getDistanceCTF...:392-397``: D = |C|^2 |fy|^2T + (.)^T - 2 Re(CTFfy CTFfy^H)."""
    nS = fy.shape[0]
    fy = fy.reshape(nS, -1)
    CTF = CTF.reshape(nS, -1)
    CTFfy = CTF.conj() * fy
    D = np.dot(np.abs(CTF) ** 2, (np.abs(fy) ** 2).T)
    return D + D.T - 2 * np.real(np.dot(CTFfy, CTFfy.conj().T))


def direct_distance(fy, CTF):
    """Dead direct form of ``getDistanceCTF...:106-122``: ||C_x fy_y - C_y fy_x||^2."""
    nS = fy.shape[0]
    D = np.zeros((nS, nS))
    for x in range(nS):
        for y in range(nS):
            if x != y:
                D[x, y] = np.linalg.norm(CTF[x] * fy[y] - CTF[y] * fy[x]) ** 2
    return D


def pd_geometry(q):
    """``getDistanceCTF...:297-301`` and ``:206-213``: per-image PD vectors, the normalised
    PD sum and the in-plane reference angle psi_p (degrees)."""
    nS = q.shape[1]
    PDs = calc_avg_pd(q, nS)
    PD = np.sum(PDs, 1)
    PD = PD / np.linalg.norm(PD)
    return PDs, PD, psi_ang(PD)


def prep_rows(input_data, filterPar, raw, nStot, params, rows, geometry=None, msk2=1,
              relion_sh=None):
    """a1-a6 of ``getDistanceCTF...:244-349`` for the images ``rows`` of one PD only (every
    image is processed on its own in the reference; only PD and psi_p use all of q).

    raw: the SPIDER stack, or with relion_sh = (shx, shy) the RELION stack in mrcfile's layout
    (``:259-264``).  Returns a dict of per-row arrays: imgAll, fy, CTF (N x N, fftshift-free),
    imgAllFlip, Psis, Nom, Dnom, imgLabels."""
    ind, q, df = input_data[0], input_data[1], input_data[2]
    N = params["nPix"]
    N2 = N / 2.0
    PDs, PD, psi_p = pd_geometry(q) if geometry is None else geometry
    msk = annular_mask(0, N2, N, N)
    Q = create_grid(N, N2)
    G = ifftshift(create_filter(filterPar["type"], filterPar["N"], filterPar["Qc"], Q))
    rows = np.asarray(rows, dtype=np.int64)
    m = len(rows)
    out = dict(imgAll=np.zeros((m, N, N)), fy=np.zeros((m, N, N), dtype=complex),
               CTF=np.zeros((m, N, N)), imgAllFlip=np.zeros((m, N, N)),
               Psis=np.zeros(m), Nom=np.zeros(m), Dnom=np.zeros(m),
               imgLabels=np.zeros(m, dtype=int))
    for r, iS in enumerate(rows):
        conj = not (ind[iS] < nStot / 2)
        src = int(ind[iS] - nStot / 2) if conj else int(ind[iS])
        out["imgLabels"][r] = -1 if conj else 1
        if relion_sh is None:
            y = load_normalized(raw, src, conj, msk)
        else:
            y = load_normalized_relion(raw, src, conj, msk, relion_sh)
        y = ifft2(fft2(y) * G).real
        Psi, s, c = get_psi(q, PD, iS)
        if np.isnan(Psi):
            Psi = 0
        out["Psis"][r], out["Nom"][r], out["Dnom"][r] = Psi, s, c
        img = y * msk
        img = rotatefill(img, -(180 / math.pi) * Psi)
        img = rotatefill(img, -psi_p)
        C = ifftshift(ctf(Q / (2 * params["pix_size"]),
                          [params["Cs"], df[iS], params["EkV"], params["gaussEnv"],
                           params["AmpContrast"]]))
        out["CTF"][r] = C
        out["fy"][r] = fft2(img * msk2)
        out["imgAllFlip"][r] = phase_flip(out["fy"][r], C)
        out["imgAll"][r] = img
    return out


def prep_rows_parallel(input_data, filterPar, raw, nStot, params, workers=8, msk2=1):
    """``prep_rows`` over all images of one PD, split across `workers` spawned processes, each
    given only the stack images its rows read.  Same values as the serial loop."""
    import multiprocessing as mp
    ind, q, df = input_data[0], input_data[1], input_data[2]
    n = len(ind)
    geometry = pd_geometry(q)
    chunks = [c for c in np.array_split(np.arange(n), max(1, workers)) if len(c)]
    jobs = []
    for c in chunks:
        conj = ~(ind[c] < nStot / 2)
        src = np.where(conj, ind[c] - nStot // 2, ind[c]).astype(np.int64)
        uniq, inv = np.unique(src, return_inverse=True)
        sub = np.ascontiguousarray(raw[uniq])
        m = len(uniq)
        # local stack of m images, local nStot = 2m: conjugates index m + local
        loc = np.where(conj, inv + m, inv)
        ind_l = np.zeros(n, dtype=np.int64)
        ind_l[c] = loc
        jobs.append((ind_l, q, df, filterPar, sub, 2 * m, params, c, geometry, msk2))
    # spawned workers (the caller may hold a GPU context, which a fork would copy); an
    # executor, not a Pool, so a worker that dies raises instead of being replaced forever
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(len(jobs), mp_context=mp.get_context("spawn")) as ex:
        parts = list(ex.map(_prep_job, jobs))
    out = {}
    for k in parts[0]:
        out[k] = np.concatenate([p[k] for p in parts])
    return out, geometry


def _prep_job(args):
    ind_l, q, df, filterPar, sub, nStot_l, params, rows, geometry, msk2 = args
    return prep_rows([ind_l, q, df], filterPar, sub, nStot_l, params, rows, geometry, msk2)


def averages(imgAll, imgAllFlip, CTF, msk2=1, snr=5.0):
    """a6-a7 outputs (``getDistanceCTF...:348``, ``:352-367``, ``:400``, ``:422-430``):
    the Wiener average imgAvg with W = -(sum_i C_i^2 + 1/SNR) (SNR = 5, note the sign), the
    phase-flipped average imgAvgFlip and imgAllIntensity = mean(imgAllFlip^2)."""
    nS, N = imgAll.shape[0], imgAll.shape[1]
    imgAvgFlip = np.zeros((N, N))
    for iS in range(nS):
        imgAvgFlip = imgAvgFlip + imgAllFlip[iS]
    wiener_dom = -(np.sum(CTF ** 2, axis=0) + 1.0 / snr)
    imgAvg = 0
    for iS in range(nS):
        imgAvg = imgAvg + ifft2(fft2(imgAll[iS]) * (CTF[iS] / wiener_dom)).real
    imgAvg = imgAvg * msk2 / nS
    imgAvgFlip = imgAvgFlip.real * msk2 / nS
    imgAllIntensity = np.mean(imgAllFlip ** 2, axis=0)
    return imgAvg, imgAvgFlip, imgAllIntensity


def phase_flip(fy, C):
    """``getDistanceCTF...:346``: real(ifft2(sign(C) fy))."""
    return ifft2(np.sign(C) * fy).real


def op(input_data, filterPar, raw, nStot, options, params, msk2=1, sh=None):
    """``getDistanceCTF_local_Conj9combinedS2.op`` (``:216-420``) returning the 19-key dict
    it pickles instead of writing it.

    input_data = [ind, q (4 x n), df (n), outFile, prD]; ``raw`` is the SPIDER image stack
    (n_img, N, N); ``params`` carries nPix, pix_size, Cs, EkV, gaussEnv, AmpContrast.
    """
    ind, q, df = input_data[0], input_data[1], input_data[2]
    nS = ind.shape[0]
    N = params["nPix"]
    PDs, PD, psi_p = geometry = pd_geometry(q)
    relion = sh if options.get("relion_data", False) else None
    pr = prep_rows(input_data, filterPar, raw, nStot, params, np.arange(nS), geometry, msk2,
                   relion)
    imgAll, fy, CTF, imgAllFlip = pr["imgAll"], pr["fy"], pr["CTF"], pr["imgAllFlip"]
    imgAvg, imgAvgFlip, imgAllIntensity = averages(imgAll, imgAllFlip, CTF, msk2)
    D = np.zeros((nS, nS))
    if not options.get("avgOnly", False):
        D = gram_distance(fy, CTF)
    return dict(D=D, ind=ind, q=q, df=df, CTF=CTF.reshape(nS, N * N), imgAll=imgAll, msk2=msk2,
                PD=PD, PDs=PDs, Psis=pr["Psis"].reshape(nS, 1), imgAvg=imgAvg,
                imgAvgFlip=imgAvgFlip, imgAllFlip=imgAllFlip, imgLabels=pr["imgLabels"],
                Dnom=pr["Dnom"].reshape(nS, 1), Nom=pr["Nom"].reshape(nS, 1),
                imgAllIntensity=imgAllIntensity, version="getDistanceCTF_local9, V 1.0",
                options=options, fy=fy, psi_p=psi_p)

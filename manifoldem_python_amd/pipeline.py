"""GPU orchestration of the per-PD hot path on one MI355X.

Stage 1 (``getDistanceCTF_local_Conj9combinedS2``): image preparation as batched rocFFT
transforms (``torch.fft``) with the elementwise work between them in HIP kernels, then the MFMA
distance Gram.  Stage 2 (``DMembeddingII``): kNN selection, Ferguson sweep, the batched
Levenberg-Marquardt tanh fit and Markov normalisation in HIP kernels, then the largest-magnitude
eigenpairs of every PD's float64 kernel by batched Lanczos on the GPU (a dense float64 eigh is
kept as an option).

Host-side work kept from the reference: per-image in-plane angles and the per-PD reference
angle (``get_psi``, ``calc_avg_pd``, ``q2Spider``), the fit's restart rule (scipy's curve_fit
remains available as ``ferguson_fit``), the small tridiagonal Ritz problems, and the final
psi/mu arithmetic on the (n x nEigs+1) eigenvectors.
"""
from __future__ import annotations

import warnings

import numpy as np
import torch
from scipy import optimize
from scipy.optimize import OptimizeWarning, curve_fit

from . import _native as nat

warnings.simplefilter("ignore", OptimizeWarning)

LOG_EPS = np.arange(-150, 150.2, 0.2)        # DMembeddingII.py:146 (1501 points)
LOG_EPS.setflags(write=False)                # a constant; read-only arrays are cached by identity
KAPPA = {nat.PREC_FP32: 1.0e4, nat.PREC_SPLIT_BF16: 8.0}
CHUNKS = {nat.PREC_FP32: 8, nat.PREC_SPLIT_BF16: 2}     # default K chunks per Gram half
GRAM_TILE = 128                                          # default Gram output tile side


def round_up(x, m):
    return (x + m - 1) // m * m


def half_len(N):
    return N * (N // 2 + 1)


def khp_of(N):
    return round_up(half_len(N), 64)


def device():
    if not torch.cuda.is_available():
        raise RuntimeError("manifoldem_python_amd needs a ROCm GPU (torch.cuda is unavailable); "
                           "there is no CPU path")
    return torch.device("cuda")


# ------------------------------------------------------------------ host geometry ---------
def calc_avg_pd(q):
    """``getDistanceCTF_local_Conj9combinedS2.py:169-185``."""
    nS = q.shape[1]
    return 2 * np.vstack((q[1] * q[3] - q[0] * q[2],
                          q[0] * q[1] + q[2] * q[3],
                          q[0] ** 2 + q[3] ** 2 - np.ones((1, nS)) / 2.0))


def get_psi(q, PD):
    """``getDistanceCTF...:187-202`` for all images: Psi = 2 arctan(s/c), NaN -> 0 (:319-320)."""
    s = -(1 + PD[2]) * q[3] - PD[0] * q[1] - PD[1] * q[2]
    c = (1 + PD[2]) * q[0] + PD[1] * q[1] - PD[0] * q[2]
    with np.errstate(divide="ignore", invalid="ignore"):
        Psi = 2 * np.arctan(s / c)
    Psi = np.where(np.isnan(Psi), 0.0, Psi)
    return Psi, s, c


def _qmult(q, s):
    q0, qv, s0, sv = q[0], q[1:4], s[0], s[1:4]
    c = np.array([qv[1] * sv[2] - qv[2] * sv[1], qv[2] * sv[0] - qv[0] * sv[2],
                  qv[0] * sv[1] - qv[1] * sv[0]])
    return np.concatenate(([q0 * s0 - np.dot(qv, sv)], q0 * sv + s0 * qv + c))


def psi_ang(PD):
    """``getDistanceCTF...:206-213`` with ``q2Spider.py:19-60``: ZYZ angles by the same TRF
    least-squares fit from x0 = 0 (ftol 1e-12); returns psi in degrees, mod 360."""
    return np.mod(spider_angles(PD)[2], 2 * np.pi) * (180 / np.pi)


def spider_angles(PD):
    """(phi, theta, psi) in radians of the rotation taking the z axis to PD: ``q2Spider.op`` of
    the normalised Qr = [1 + PD_2, PD_1, -PD_0, 0] (``getDistanceCTF...:206-213``,
    ``CC/projectMask.py:76-79``), the TRF least-squares fit from x0 = 0 (ftol 1e-12)."""
    Qr = np.array([1 + PD[2], PD[1], -PD[0], 0.0])
    q = Qr / np.sqrt(np.sum(Qr ** 2))

    def dev1(a):
        q1 = np.array([np.cos(a[0] / 2.0), 0.0, 0.0, -np.sin(a[0] / 2.0)])
        q2 = np.array([np.cos(a[1] / 2.0), 0.0, -np.sin(a[1] / 2.0), 0.0])
        q3 = np.array([np.cos(a[2] / 2.0), 0.0, 0.0, -np.sin(a[2] / 2.0)])
        return q - _qmult(q3, _qmult(q2, q1))

    res = optimize.least_squares(dev1, np.array([0, 0, 0]), bounds=(-np.inf, np.inf), ftol=1e-12)
    return res.x[0], res.x[1], res.x[2]


def spider_rotation(Phi, Theta, Psi):
    """``CC/projectMask.py:24-37`` (eulerRotMatrix3DSpider, radians)."""
    cf, sf, ct, st, cp, sp = (np.cos(Phi), np.sin(Phi), np.cos(Theta), np.sin(Theta),
                              np.cos(Psi), np.sin(Psi))
    return np.array([[cf * cp * ct - sf * sp, cp * ct * sf + cf * sp, -cp * st],
                     [-cp * sf - cf * ct * sp, cf * cp - ct * sf * sp, sp * st],
                     [cf * st, sf * st, ct]])


def mask_transform(PD, L):
    """The affine map of ``CC/projectMask.py:40-53,116-127`` for PD: the Euler angles of PD with
    psi set to 0 and negated, R = eulerRotMatrix3DSpider(0, -theta, -phi), offset = c - R c with
    c = L/2 on every axis; returned as float64 [12] = {R row-major, offset}."""
    phi, theta, _ = spider_angles(PD)
    R = spider_rotation(0.0, -theta, -phi)
    c = 0.5 * np.array([L, L, L], dtype=np.float64)
    return np.concatenate([R.ravel(), c - R @ c])


def volume_mask(vol, PD):
    """msk2 of ``getDistanceCTF...:304-310`` on the GPU: ``projectMask.op(vol, PD)`` for a
    cubic mask volume vol (numpy, mrcfile's (z, y, x) orientation).  Returns (mask float32
    (L, L) device tensor of 1.0 / 0.0, projection sums float64 device tensor)."""
    L = vol.shape[0]
    if vol.shape != (L, L, L):
        raise ValueError(f"volume_mask: the mask volume must be cubic, got {vol.shape}")
    sw = np.ascontiguousarray(np.swapaxes(np.asarray(vol, dtype=np.float32), 0, 2))
    proj, mask = nat.project_mask(torch.as_tensor(sw, device=device()), mask_transform(PD, L))
    return mask, proj


def pd_geometry(q):
    """PD axis, per-image (Psi, Nom, Dnom) and psi_p of one PD."""
    PDs = calc_avg_pd(q)
    PD = np.sum(PDs, 1)
    PD = PD / np.linalg.norm(PD)
    Psi, s, c = get_psi(q, PD)
    return dict(PDs=PDs, PD=PD, Psi=Psi, Nom=s, Dnom=c, psi_p=psi_ang(PD))


# ----------------------------------------------------------------------- Stage 1 ----------
class Operand:
    """Packed distance operand for `rows` images of side N in one precision."""

    def __init__(self, rows, N, precision, dev):
        self.N = N
        self.Khp = khp_of(N)
        self.precision = precision
        shape = (rows, 4 * self.Khp)
        if precision == nat.PREC_FP32:
            self.hi = torch.empty(shape, dtype=torch.float32, device=dev)
            self.lo = None
        else:
            self.hi = torch.empty(shape, dtype=torch.int16, device=dev)
            self.lo = torch.empty(shape, dtype=torch.int16, device=dev)


def filter_kind(filterPar):
    t = filterPar.get("type", "Butter")
    if t == "Butter":
        return nat.FILTER_BUTTER
    if t == "Gauss":
        return nat.FILTER_GAUSS
    raise ValueError("%s filter is unsupported" % t)


ALIGN = "auto"      # "fused": memi_align_images; "fft": FFT prefilter + memi_spline_rotate
PREP = "auto"       # "fused": memi_prep_images (one kernel); "staged": the kernel sequence
PREP_FUSED_MAX_N = 160  # "auto" uses the fused kernel up to this image side


def prepare_images(stack, src, conj, rot1_deg, rot2_deg, df, N, filterPar, ctf, operand=None,
                   row0=0, msk2=None, fy_out=None, ctf_out=None, want_images=True, src_max=None,
                   align=None, prep=None):
    """a1-a6 for one set of images (all device tensors): load + normalise, Butterworth,
    disk mask, two cubic-spline rotations, fy = rfft2(img * msk2), CTF and operand rows.
    prep: "fused" (memi_prep_images: the whole chain in one kernel, images in LDS), "staged"
    (rocFFT transforms with the elementwise kernels between them) or "auto"/None (fused when N
    and the inputs allow).  align (staged only): "fused" (memi_align_images, 8 <= N <= 196),
    "fft" (B-spline prefilter by FFT division, any N) or "auto"/None.

    Returns a dict of device tensors: imgAll (n,N,N) f32, fy (n,N,Nh) c64, ctf (n,Kh) f32.
    """
    n = int(src.shape[0])
    dev = stack.device
    prep = prep or PREP
    if prep == "auto":
        # measured on MI355X with one LDS image per block (in-place FFTs): 3.94 vs 6.99 ms for
        # 16384 96x96 images, 0.86 vs 1.38 ms for 2000 128x128, 2.60 vs 4.23 ms for 3000
        # 160x160 (fused vs staged)
        prep = "fused" if (nat.prep_supported(N) and N <= PREP_FUSED_MAX_N and msk2 is None and
                           filterPar.get("type", "Butter") in ("Butter", "Gauss")) else "staged"
    if prep == "fused":
        if msk2 is not None:
            raise ValueError("prepare_images: the fused kernel has no mask-volume input")
        op = operand if operand is not None else Operand(n, N, nat.PREC_FP32, dev)
        r0 = row0 if operand is not None else 0
        fy = (fy_out.view(n, N, N // 2 + 1) if fy_out is not None else
              torch.empty((n, N, N // 2 + 1), dtype=torch.complex64, device=dev))
        ctf_half = ctf_out if ctf_out is not None else torch.empty((n, half_len(N)),
                                                                  dtype=torch.float32, device=dev)
        img = torch.empty((n, N, N), dtype=torch.float32, device=dev) if want_images else None
        ws = torch.empty(6 * half_len(N), dtype=torch.float64, device=dev)
        nat.prep_images(stack, src, conj, rot1_deg, rot2_deg, df, N, filter_kind(filterPar),
                        filterPar["Qc"], filterPar["N"], ctf, op.precision, op.Khp, r0, op.hi,
                        op.lo if op.lo is not None else op.hi, fy, ctf_half, ws, img_out=img,
                        src_max=src_max)
        return dict(imgAll=img, fy=fy, ctf=ctf_half)
    y = torch.empty((n, N, N), dtype=torch.float32, device=dev)
    nat.load_normalize(stack, src, conj, N, y, src_max=src_max)
    Y = torch.fft.rfft2(y)
    nat.spectrum_scale(Y, N, filter_kind(filterPar), filterPar["Qc"], filterPar["N"])
    y = torch.fft.irfft2(Y, s=(N, N)).contiguous()
    del Y
    align = align or ALIGN
    if align == "auto":
        align = "fused" if nat.ALIGN_MIN_N <= N <= nat.ALIGN_MAX_N else "fft"
    if align == "fused":
        nat.align_images(y, rot1_deg, rot2_deg, N, y, disk=True)
    else:
        nat.mask_multiply(y, N, disk=True)
        for ang in (rot1_deg, rot2_deg):
            Y = torch.fft.rfft2(y)
            nat.spectrum_scale(Y, N, nat.FILTER_BSPLINE)
            coef = torch.fft.irfft2(Y, s=(N, N)).contiguous()
            nat.spline_rotate(coef, ang, N, y)
        del Y, coef
    img = y
    if msk2 is not None:
        img = y.clone()
        nat.mask_multiply(img, N, disk=False, msk2=msk2)
    if fy_out is not None:
        fy = fy_out.view(n, N, N // 2 + 1)
        torch.fft.rfft2(img, out=fy)
    else:
        fy = torch.fft.rfft2(img)
    ctf_half = ctf_out if ctf_out is not None else torch.empty((n, half_len(N)),
                                                              dtype=torch.float32, device=dev)
    if operand is not None:
        nat.ctf_pack(fy, df, N, ctf, operand.precision, operand.Khp, row0, operand.hi,
                     operand.lo if operand.lo is not None else operand.hi, ctf_half)
    else:
        tmp = Operand(n, N, nat.PREC_FP32, dev)
        nat.ctf_pack(fy, df, N, ctf, nat.PREC_FP32, tmp.Khp, 0, tmp.hi, tmp.hi, ctf_half)
    return dict(imgAll=y if want_images else None, fy=fy, ctf=ctf_half)


def relion_stack(images, shift0, shift1):
    """RELION particle images (mrcfile's layout, float32 (m, N, N), numpy or device tensor)
    shifted as getDistanceCTF_local_Conj9combinedS2.py:259-264 does (ndimage.shift by
    (shift0, shift1) = (sh[1] - 0.5, sh[0] - 0.5), order 3, mode 'wrap') on the GPU, returned
    transposed: the SPIDER stack layout the loaders read (they transpose it back)."""
    dev = device()
    img = torch.as_tensor(np.ascontiguousarray(images, dtype=np.float32) if
                          isinstance(images, np.ndarray) else images, device=dev).contiguous()
    out = torch.empty_like(img)
    s0 = torch.as_tensor(np.asarray(shift0, dtype=np.float64), device=dev)
    s1 = torch.as_tensor(np.asarray(shift1, dtype=np.float64), device=dev)
    nat.shift_images(img, s0, s1, out, transpose_out=True)
    return out


def averages(imgAll, fy, ctf_half, N, msk2=None, snr=5.0):
    """a6-a7 outputs of one PD: imgAllFlip, imgAvgFlip, imgAllIntensity, imgAvg (Wiener)."""
    n = imgAll.shape[0]
    dev = imgAll.device
    flip = torch.empty_like(fy)
    nat.phase_flip(fy, ctf_half, N, flip)
    imgAllFlip = torch.fft.irfft2(flip, s=(N, N)).contiguous()
    s = torch.empty(N * N, dtype=torch.float64, device=dev)
    s2 = torch.empty_like(s)
    nat.image_moments(imgAllFlip, s, s2)
    m2 = 1.0 if msk2 is None else msk2.double().reshape(-1)
    imgAvgFlip = (s / n * m2).view(N, N)
    imgAllIntensity = (s2 / n).view(N, N)
    spec = fy if msk2 is None else torch.fft.rfft2(imgAll)
    wnum = torch.empty((N, N // 2 + 1), dtype=torch.complex64, device=dev)
    nat.wiener_average(spec.contiguous(), ctf_half, N, snr, wnum)
    imgAvg = torch.fft.irfft2(wnum, s=(N, N)).double().reshape(-1) * m2 / n
    return dict(imgAllFlip=imgAllFlip, imgAvgFlip=imgAvgFlip, imgAllIntensity=imgAllIntensity,
                imgAvg=imgAvg.view(N, N))


def full_ctf(ctf_half, N):
    """Half-spectrum CTF (n, N*Nh) -> full FFT layout (n, N*N); C depends on |k| only."""
    n = ctf_half.shape[0]
    Nh = N // 2 + 1
    h = ctf_half.view(n, N, Nh)
    cols = torch.arange(N, device=h.device)
    src = torch.where(cols < Nh, cols, N - cols)
    return h[:, :, src].reshape(n, N * N)


class Distances:
    """Distance matrices of a batch of PDs (flat float32 buffer, PD blocks per the plan)."""

    def __init__(self, plan, D, flag_count):
        self.plan = plan
        self.D = D
        self.flag_count = flag_count

    def block(self, p):
        return self.plan.block(self.D, p)


GRAM_VARIANT = {nat.PREC_FP32: 0, nat.PREC_SPLIT_BF16: 0}   # 0 = library default


def ctf_distances(plan, operand, fy, ctf_half, N, kappa=None, ws=None, events=None,
                  variant=None, flags=None, D=None):
    """a8 for every PD of the plan: returns Distances (float32 D on the device).  `events`: an
    optional (start, stop) hipEvent_t pair recorded around the MFMA Gram kernel.  ws / flags /
    D: optional preallocated buffers (partial tiles, flagged-entry queue, output)."""
    dev = operand.hi.device
    info = plan.info
    kappa = KAPPA[operand.precision] if kappa is None else kappa
    if D is None:
        D = torch.empty(info.d_elems, dtype=torch.float32, device=dev)
    if ws is None or ws.numel() < info.ws_floats:
        ws = torch.empty(max(info.ws_floats, 1), dtype=torch.float32, device=dev)
    if flags is None or flags.numel() < 3 * info.flag_cap:
        flags = torch.empty(3 * info.flag_cap, dtype=torch.int32, device=dev)
    cnt = torch.empty(1, dtype=torch.int32, device=dev)     # zeroed by the library
    nat.ctf_distance(plan, operand.hi, operand.lo if operand.lo is not None else operand.hi,
                     operand.Khp, operand.precision, kappa, fy, ctf_half, N, D, ws, flags, cnt,
                     events, GRAM_VARIANT[operand.precision] if variant is None else variant)
    return Distances(plan, D, cnt)


# ----------------------------------------------------------------------- Stage 2 ----------
def tanh_model(xx, aa0, aa1, aa2, aa3):
    """``fergusonE.py:21-23``."""
    return aa3 + aa2 * np.tanh(aa0 * xx + aa1)


def ferguson_fit(logEps, logSumWij, a0, rand=None):
    """``fergusonE.py:45-57``: curve_fit restarted from rand(4,1) - 0.5 while the summed
    parameter standard errors exceed 100.  ``rand`` defaults to numpy's global generator, as
    in the reference."""
    rand = np.random.rand if rand is None else rand
    resnorm = np.inf
    while resnorm > 100:
        popt, pcov = curve_fit(tanh_model, logEps, logSumWij, p0=np.ravel(a0))
        resnorm = np.sum(np.sqrt(np.fabs(np.diag(pcov))))
        a0 = 1 * (rand(4, 1) - 0.5)
        residuals = logSumWij - tanh_model(logEps, *popt)
        ss_res = np.sum(residuals ** 2)
        ss_tot = np.sum((logSumWij - np.mean(logSumWij)) ** 2)
        R_squared = 1 - ss_res / ss_tot
    return popt, resnorm, R_squared


_const_cache = {}
_const_last = {}


def _device_const(a, dev):
    """A device copy of a small constant host array (the eps grid), made once per content (and
    found without hashing the content again when the same array object comes back)."""
    hit = _const_last.get(str(dev))
    if hit is not None and hit[0] is a:
        return hit[1]
    src = a
    a = np.ascontiguousarray(a, dtype=np.float64)
    key = (str(dev), a.shape, a.tobytes())
    t = _const_cache.get(key)
    if t is None:
        if len(_const_cache) > 16:
            _const_cache.clear()
        # torch.tensor copies (as_tensor would wrap a read-only host array and warn)
        t = _const_cache[key] = torch.tensor(a, dtype=torch.float64, device=dev)
    if isinstance(src, np.ndarray) and not src.flags.writeable:
        _const_last[str(dev)] = (src, t)
    return t


def batched_inv(A):
    """np.linalg.inv of each matrix of a stack, NaN for the exactly singular ones (those inv
    rejects: an exact zero pivot, so det == 0), in one batched call instead of one call per
    matrix once any is singular."""
    with np.errstate(all="ignore"):
        ok = np.linalg.det(A) != 0
        out = np.full_like(A, np.nan)
        try:
            if ok.any():
                out[ok] = np.linalg.inv(A[ok])
        except np.linalg.LinAlgError:        # a zero pivot beside an infinite one: one at a time
            out = np.full_like(A, np.nan)
            for q in range(len(A)):
                try:
                    out[q] = np.linalg.inv(A[q])
                except np.linalg.LinAlgError:
                    pass
    return out


def ferguson_fit_batch(logEps, logS, a0s, rand=None, max_restarts=200, starts_per_round=8,
                       logS_dev=None, first=None):
    """The fit of ``fergusonE.py:45-57`` for many PDs at once on the GPU (memi_tanh_fit):
    fit from a0, pcov = inv(J^T J) ss_res / (m - 4) as scipy's curve_fit forms it, and refit
    from rand(4,1) - 0.5 while sum(sqrt|diag pcov|) > 100.  Returns (popt (npd, 4), R^2).

    The first round fits every PD's own start; each later round fits up to `starts_per_round`
    starts of every unfinished PD in parallel: its pending start followed by the next draws in
    the order the reference's loop makes them (one after every fit), keeping the first
    acceptable one in that order -- the same result as sequential restarts, in fewer round
    trips.  logS_dev: the same curves already on the device (npd, m) float64, if at hand.
    first: memi_tanh_fit's output (npd * 16 float64, device) for every PD's own start, if
    already computed: the first round then reuses it instead of fitting again."""
    rand = np.random.rand if rand is None else rand
    dev = device()
    npd, m = logS.shape
    x = _device_const(logEps, dev)
    y_all = (logS_dev if logS_dev is not None else
             torch.as_tensor(np.ascontiguousarray(logS, dtype=np.float64), device=dev))
    pending = np.stack([np.ravel(a) for a in a0s]).astype(np.float64)
    popt = np.zeros((npd, 4))
    R2 = np.zeros(npd)
    todo = np.ones(npd, dtype=bool)
    ss_tot = np.sum((logS - logS.mean(axis=1, keepdims=True)) ** 2, axis=1)
    best = np.full(npd, np.inf)
    iu = np.triu_indices(4)
    fits = 0
    while todo.any() and fits < max_restarts:
        idx = np.nonzero(todo)[0]
        # most PDs accept their first start: one fit each first, then several per round
        R = 1 if fits == 0 else int(max(1, min(starts_per_round, max_restarts - fits)))
        starts = np.empty((len(idx), R, 4))
        starts[:, 0] = pending[idx]
        if R > 1:       # PD after PD, in the order per-PD rand(4, 1) calls would consume them
            starts[:, 1:] = np.reshape(rand(len(idx) * (R - 1) * 4, 1),
                                       (len(idx), R - 1, 4)) - 0.5
        B = len(idx) * R
        if R == 1 and len(idx) == npd:           # every PD once, in order: no gather
            yy = y_all
        else:
            yy = y_all[torch.as_tensor(np.repeat(idx, R), device=dev)].contiguous()
        if fits == 0 and first is not None:
            out = first
        else:
            out = torch.empty(B * 16, dtype=torch.float64, device=dev)
            nat.tanh_fit(x, yy, _to_device(starts.reshape(B, 4), dev), out)
        o = _to_host(out[:B * 16]).reshape(len(idx), R, 16)
        J = np.zeros((len(idx), R, 4, 4))
        J[:, :, iu[0], iu[1]] = o[:, :, 4:14]
        J = J + np.triu(J, 1).transpose(0, 1, 3, 2)
        ss_res = o[:, :, 14]
        with np.errstate(all="ignore"):
            Jinv = batched_inv(J.reshape(-1, 4, 4))
            diag = np.diagonal(Jinv, axis1=1, axis2=2).reshape(len(idx), R, 4)
            resnorm = np.sum(np.sqrt(np.fabs(diag * (ss_res / (m - 4))[:, :, None])), axis=2)
        resnorm[~np.isfinite(resnorm)] = np.inf
        # per PD, in the order the sequential loop tries them: the first acceptable start;
        # if none, the best of this round competes with the best so far
        ok = resnorm <= 100
        hit = ok.any(axis=1)
        first = np.argmax(ok, axis=1)
        pick = np.where(hit, first, np.argmin(resnorm, axis=1))
        rq = np.arange(len(idx))
        rn = resnorm[rq, pick]
        upd = hit | (rn < best[idx])
        pu = idx[upd]
        best[pu] = rn[upd]
        popt[pu] = o[rq[upd], pick[upd], 0:4]
        R2[pu] = 1 - ss_res[rq[upd], pick[upd]] / ss_tot[pu]
        todo[idx[hit]] = False
        rest = idx[~hit]
        if len(rest):   # the draw after each unfinished PD's last fit
            pending[rest] = np.reshape(rand(len(rest) * 4, 1), (len(rest), 4)) - 0.5
        fits += R
    return popt, R2


def fit_bandwidths(logEps, logS_dev, a0s, tune, rand=None):
    """a13 for every PD: the tanh fit from each PD's start (memi_tanh_fit) and the restart
    test with the bandwidth on the device (memi_tanh_select).  Only the acceptance flags come
    back before the bandwidths are used; when a PD's first start fails the test, the host
    restart loop (ferguson_fit_batch, from these first fits) takes over for the batch.
    Returns (sigma (npd,) float64 device tensor, finish) with finish(logS host) -> (popt, R2).
    (Reading the flags after the eigensolve instead, so that nothing waits here, measured
    slower on one PD of 2000 images: 4.14 against 4.05 ms per step.)"""
    dev = logS_dev.device
    npd, m = logS_dev.shape
    x = _device_const(logEps, dev)
    starts = np.stack([np.ravel(a) for a in a0s]).astype(np.float64)
    p0 = _to_device(starts, dev)
    out = torch.empty(npd * 16, dtype=torch.float64, device=dev)
    nat.tanh_fit(x, logS_dev, p0, out)
    ok = torch.empty(npd, dtype=torch.int32, device=dev)
    sigma = torch.empty(npd, dtype=torch.float64, device=dev)
    nat.tanh_select(out, npd, 1, m, tune, 100.0, ok, sigma)
    fit_h = PendingCopy(out, "fit")
    if _to_host(ok).all():
        def finish(logS):
            o = fit_h.get().reshape(npd, 16)
            ss_tot = np.sum((logS - logS.mean(axis=1, keepdims=True)) ** 2, axis=1)
            return o[:, 0:4].copy(), 1 - o[:, 14] / ss_tot
        return sigma, finish
    logS = _to_host(logS_dev).reshape(npd, m)
    popt, R2 = ferguson_fit_batch(np.asarray(logEps), logS, a0s, rand, logS_dev=logS_dev,
                                  first=out)
    sig = tune * np.sqrt(2 * np.exp(-popt[:, 1] / popt[:, 0]))
    return _to_device(sig, dev), lambda _logS: (popt, R2)


class DMResult:
    __slots__ = ("lamb", "psi", "sigma", "mu", "logEps", "logSumWij", "popt", "R_squared",
                 "slope", "V", "lanczos_steps", "lanczos_residual")

    def as_tuple(self):
        return (self.lamb, self.psi, self.sigma, self.mu, self.logEps, self.logSumWij, self.popt,
                self.R_squared)


_POOL = None


def _pool():
    """Host thread pool for per-PD small dense work (LAPACK calls release the GIL)."""
    global _POOL
    if _POOL is None:
        from concurrent.futures import ThreadPoolExecutor
        import os
        _POOL = ThreadPoolExecutor(max_workers=max(1, min(16, os.cpu_count() or 1)))
    return _POOL


def lm_ritz(alpha, beta, k):
    """The k largest-magnitude eigenpairs of the symmetric tridiagonal T(alpha, beta): both
    ends of the spectrum (k each) by index range, then the k of largest |theta|.  Returns
    (theta, S) with S the eigenvectors of T (columns)."""
    from scipy.linalg import eigh_tridiagonal
    J = alpha.shape[0]
    if J <= 2 * k + 8:
        theta, S = eigh_tridiagonal(alpha, beta)
    else:
        t1, S1 = eigh_tridiagonal(alpha, beta, select="i", select_range=(0, k - 1))
        t2, S2 = eigh_tridiagonal(alpha, beta, select="i", select_range=(J - k, J - 1))
        theta = np.concatenate([t1, t2])
        S = np.concatenate([S1, S2], axis=1)
    sel = np.argsort(-np.abs(theta), kind="stable")[:k]
    return theta[sel], S[:, sel]


LANCZOS = dict(graph=True, split=None, check_every=None)   # dm_embed's Lanczos settings
LANCZOS_TOL = 1e-10         # residual target of lanczos_eigs, relative to max |theta|


def _next_check_gap(j, res, thmax, active, tol, step, hist, slack=0.6, cap=4):
    """Steps to the next convergence check, and the step by which every live PD is predicted
    to have converged.  Each check reports every live PD's largest residual r against its
    target tol * max|theta|; with two checks, log(r / target) is extrapolated linearly in the
    step count.  The next check goes to the earliest predicted convergence over the live PDs
    (Lanczos residuals usually fall faster than linearly, hence `slack`), in whole blocks of
    `step`, between 1 and `cap` blocks away.  The last predicted convergence (no slack) is None
    while any live PD has no falling trend yet.  Returns (gap, history, last)."""
    live = active.astype(bool)
    with np.errstate(divide="ignore", invalid="ignore"):
        lr = np.log(np.maximum(res, 1e-300) / (tol * np.maximum(thmax, 1e-300)))
    gap = step
    last = None
    if hist is not None and live.any():
        j0, lr0 = hist
        rate = (lr - lr0) / max(j - j0, 1)                # per step; negative while converging
        ok = live & np.isfinite(rate) & (rate < 0) & (lr > 0)
        if ok.any() and ok.sum() == live.sum():
            need = np.min(lr[ok] / -rate[ok]) * slack
            gap = int(min(cap, max(1, need // step))) * step
            last = j + float(np.max(lr[ok] / -rate[ok]))
    return gap, (j, lr), last


def lanczos_eigs(plan, L, ks, m=None, tol=LANCZOS_TOL, check_every=None, seed=20261202,
                 split=None, check="gpu", graph=True, full=False):
    """Largest-magnitude eigenpairs of every PD's kernel block of L (float64 or float32) with
    sembeddingonFly.py semantics (the k eigenpairs of largest |lambda|, sorted by descending
    lambda) by batched Lanczos with full reorthogonalisation on the GPU.

    Converged when every selected Ritz pair has residual |beta_J s_J| <= tol * max|theta| or
    the Krylov space is the whole space.  split: orthogonalise with the whole grid (default
    for batches of fewer than 128 PDs or with a PD above 1024 rows) instead of one workgroup
    per PD.  check: "gpu" (memi_lanczos_ritz: Ritz pairs and the decision on the device, a
    summary per PD read back per check; the first check after 2k + 8 steps, later ones every 8
    steps or, once two checks show how fast every live PD's residual falls, at the extrapolated
    convergence step, up to 32 steps on; the next block is queued behind every check that is
    not predicted to end the run) or "host" (scipy tridiagonal eigensolver per PD, every 32
    steps by default).  graph: the launches of one block of
    check_every steps are captured once in a HIP graph and replayed (the step index lives on
    the device).  Returns a list of (lam, V (n x k) numpy, steps, max residual); with full,
    also the (n_rows x kmax) array the V are row blocks of."""
    sparse = isinstance(L, tuple)               # (col, val, cnt) of nat.kernel_ell
    dev = (L[1] if sparse else L).device
    npd = plan.npd
    ns = plan.ns.astype(int)
    rows0 = plan.rows0
    n_rows = int(ns.sum())
    if m is None:
        m = int(min(ns.max(), 512))
    m = int(max(2, min(m, ns.max())))
    if split is None:
        split = npd < 128 or int(ns.max()) > 1024
    if check_every is None:
        check_every = 8 if check == "gpu" else 32
    # no Ritz pair can have converged before the Krylov space holds both ends of the spectrum
    # (a first check at 2k + 16 instead saved nothing measurable on one large PD and made the
    # NLSA's second embedding run 5 steps longer on average)
    first_check = min(m, max(check_every, 2 * max(ks) + 8)) if check == "gpu" else check_every
    if max(ks) > nat.RITZ_KMAX:
        raise ValueError(f"lanczos_eigs: at most {nat.RITZ_KMAX} eigenpairs per PD")
    K = nat.RITZ_KMAX
    # buffers (and the captured block of steps) persist on the plan between calls with the
    # same L buffer and settings, so repeated embeddings only replay the graph
    key = (tuple(t.data_ptr() for t in L) if sparse else (L.data_ptr(), L.numel(), str(L.dtype)),
           m, bool(split), check_every, bool(graph), check)
    st = getattr(plan, "_lanczos", None)
    if st is None or st["key"] != key:
        st = dict(key=key)
        st["V"] = torch.zeros(m * n_rows, dtype=torch.float64, device=dev)
        st["W"] = torch.empty(2 * n_rows, dtype=torch.float64, device=dev)
        st["H"] = torch.empty(npd * (m + 2), dtype=torch.float64, device=dev) if split else None
        st["ab"] = torch.zeros(npd * m * 2, dtype=torch.float64, device=dev)
        st["active"] = torch.zeros(npd, dtype=torch.int32, device=dev)
        st["j"] = torch.zeros(2, dtype=torch.int32, device=dev)
        if check == "gpu":
            st["theta"] = torch.zeros(npd * K, dtype=torch.float64, device=dev)
            st["S"] = torch.zeros(npd * m * K, dtype=torch.float64, device=dev)
            st["steps"] = torch.zeros(npd, dtype=torch.int32, device=dev)
            st["res"] = torch.zeros(npd, dtype=torch.float64, device=dev)
            st["piv"] = torch.empty(npd * K * 2 * m, dtype=torch.float64, device=dev)
            st["summary"] = torch.zeros(3 * npd, dtype=torch.float64, device=dev)
            st["summary_h"] = torch.zeros(3 * npd, dtype=torch.float64, pin_memory=True)
            st["summary_ev"] = torch.cuda.Event()
        st["graph"] = None
        if graph:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                nat.lanczos_steps(plan, L, st["V"], m, st["j"], check_every, st["active"],
                                  st["W"], st["ab"], H=st["H"])
            st["graph"] = g
        plan._lanczos = st
    V, W, H, ab, active, j_dev, g = (st["V"], st["W"], st["H"], st["ab"], st["active"], st["j"],
                                     st["graph"])
    if check == "gpu":
        theta, S_dev, steps_t, res_t, piv = st["theta"], st["S"], st["steps"], st["res"], st["piv"]
    # start vectors (column 0 of every PD's basis): fixed by the seed, so built and uploaded once
    # per plan state.  Columns past 0 need no clearing: every column a kernel reads was written
    # earlier in the same run.
    if st.get("v0_seed") != seed:
        rng = np.random.default_rng(seed)
        v0 = np.empty(n_rows)
        idx = np.empty(n_rows, dtype=np.int64)
        for p in range(npd):
            n = ns[p]
            v = 1.0 + 0.1 * rng.standard_normal(n)
            v0[rows0[p]:rows0[p] + n] = v / np.linalg.norm(v)
            idx[rows0[p]:rows0[p] + n] = m * rows0[p] + np.arange(n)
        if idx.size and (idx.min() < 0 or idx.max() >= V.numel()):
            raise RuntimeError("lanczos_eigs: start vector index outside the basis")
        st["v0"] = torch.as_tensor(v0, device=dev)
        st["v0_idx"] = torch.as_tensor(idx, device=dev)
        st["v0_seed"] = seed
    nat.lanczos_init(V, st["v0"], st["v0_idx"], ab, j_dev, active)
    active_h = np.ones(npd, dtype=np.int32)
    ks_key = tuple(int(k) for k in ks)
    if st.get("ks_key") != ks_key:
        st["ks_t"] = torch.as_tensor(np.asarray(ks, dtype=np.int32), device=dev)
        st["ks_key"] = ks_key
    ks_t = st["ks_t"]
    steps = np.minimum(ns, m)
    out = [None] * npd
    def block():
        if g is not None:
            g.replay()
        else:
            nat.lanczos_steps(plan, L, V, m, j_dev, check_every, active, W, ab, H=H)

    j = 0                            # steps queued on the stream
    jc = 0                           # step count of the last check
    gap = check_every
    hist = None                      # (j, log of residual / target) at the previous check
    last = None                      # step by which every live PD is predicted to converge
    while active_h.any():
        target = first_check if jc == 0 else min(m, jc + gap)
        while j < target:
            block()
            j = min(m, j + check_every)
        jc = j
        if check == "gpu":
            # the check writes {active, largest residual, max |theta|} per PD; one async copy
            # into pinned memory and a wait on its event bring them back.  Unless this check
            # is predicted to end the run, the next block of steps is queued behind it before
            # the wait, so the GPU works through the read-back (PDs the check finishes are
            # inactive by then and skipped by those steps)
            nat.lanczos_ritz(plan, ab, m, jc, ks_t, tol, active, theta, S_dev, steps_t, res_t,
                             piv, summary=st["summary"])
            st["summary_h"].copy_(st["summary"], non_blocking=True)
            st["summary_ev"].record()
            if jc < m and (last is None or last > jc):
                block()
                j = min(m, j + check_every)
            st["summary_ev"].synchronize()
            h = st["summary_h"].numpy().reshape(3, npd)
            active_h = h[0].astype(np.int32)
            gap, hist, last = _next_check_gap(jc, h[1], h[2], active_h, tol, check_every, hist)
            continue
        abh = ab.view(npd, m, 2).cpu().numpy()

        def check_host(p):
            J = int(min(j, steps[p]))
            alpha, beta = abh[p, :J, 0], abh[p, :J, 1]
            theta_p, S = lm_ritz(alpha, beta[:J - 1], ks[p])
            res = np.abs(beta[J - 1] * S[J - 1])
            done = (J >= ns[p]) or np.all(res <= tol * np.max(np.abs(theta_p))) or (J >= m)
            if done:
                order = np.argsort(-theta_p, kind="stable")
                return p, (theta_p[order], S[:, order], J,
                           float(np.max(res[order])) if J < ns[p] else 0.0)
            return p, None

        todo = np.nonzero(active_h)[0]
        if len(todo) > 8:
            results_c = list(_pool().map(check_host, todo))
        else:
            results_c = [check_host(p) for p in todo]
        for p, r in results_c:
            if r is not None:
                out[p] = r
                active_h[p] = 0
        active.copy_(torch.as_tensor(active_h, device=dev))
    # Ritz vectors of all PDs in one launch and one copy to the host
    if check == "gpu":
        # steps, residuals and Ritz values in one copy
        fin = _to_host(torch.cat([steps_t.double(), res_t, theta[:npd * K]]))
        J_h = fin[:npd].astype(np.int64)
        res_h = fin[npd:2 * npd]
        lam_h = fin[2 * npd:].reshape(npd, K)
        k_h = np.minimum(np.asarray(ks, dtype=np.int64), J_h)
        meta_h = np.stack([J_h, k_h, np.arange(npd, dtype=np.int64) * m * K], axis=1)
        S_all = S_dev
        for p in range(npd):
            out[p] = (lam_h[p, :k_h[p]].copy(), None, int(J_h[p]), float(res_h[p]))
    else:
        meta_h = np.zeros((npd, 3), dtype=np.int64)
        blocks, off = [], 0
        for p in range(npd):
            S = np.ascontiguousarray(out[p][1], dtype=np.float64)
            meta_h[p] = (out[p][2], S.shape[1], off)
            blocks.append(S.ravel())
            off += S.size
        S_all = torch.as_tensor(np.concatenate(blocks), device=dev)
    kmax = int(meta_h[:, 1].max())
    meta = torch.as_tensor(meta_h, device=dev)
    if st.get("Y") is None or st["Y"].numel() < n_rows * kmax:
        st["Y"] = torch.empty(n_rows * kmax, dtype=torch.float64, device=dev)
        st["Y_h"] = torch.empty(n_rows * kmax, dtype=torch.float64, pin_memory=True)
    Y = st["Y"][:n_rows * kmax]
    nat.ritz_vectors(plan, V, m, S_all, meta_h, meta, kmax, Y)
    # through pinned memory (a pageable copy of C3's 17 MB of vectors ran at half the rate); a
    # fresh array, since the pinned buffer is reused by the next call
    Y_h = st["Y_h"][:n_rows * kmax]
    Y_h.copy_(Y, non_blocking=True)
    torch.cuda.current_stream().synchronize()
    Yh = Y_h.numpy().reshape(n_rows, kmax).copy()
    results = []
    for p in range(npd):
        lam, _, J, res = out[p]
        k = int(meta_h[p, 1])
        results.append((lam, Yh[rows0[p]:rows0[p] + ns[p], :k], J, res))
    return (results, Yh) if full else results


def top_eigs_lm(L, k):
    """``sembeddingonFly.py:27-35`` semantics on a dense float64 GPU matrix: the k
    eigenpairs of largest magnitude, sorted by descending eigenvalue."""
    lam, V = torch.linalg.eigh(L)
    sel = torch.argsort(-lam.abs(), stable=True)[:k]
    lam, V = lam[sel], V[:, sel]
    order = torch.argsort(lam, descending=True, stable=True)
    return lam[order], V[:, order]


def knn_mask(plan, D):
    """a10: kNN membership bitmask of every PD (int32 words on the device)."""
    mask = torch.empty(max(plan.info.mask_words, 1), dtype=torch.int32, device=D.device)
    nat.knn_select(plan, D, mask)
    return mask


FERGUSON = "auto"   # "moments" (log-uniform eps grids), "direct", or "auto"


_grid_last = [None, None]


def uniform_grid(logEps):
    """(le0, dle) when logEps[e] = le0 + e*dle with dle > 0 (as np.arange builds it), else
    None.  The answer for a read-only array is kept for the next call with the same object."""
    if _grid_last[0] is logEps:
        return _grid_last[1]
    g = _uniform_grid(logEps)
    if isinstance(logEps, np.ndarray) and not logEps.flags.writeable:
        _grid_last[0], _grid_last[1] = logEps, g
    return g


def _uniform_grid(logEps):
    le = np.asarray(logEps, dtype=np.float64)
    if le.size < 2:
        return None
    dle = (le[-1] - le[0]) / (le.size - 1)
    if not dle > 0:
        return None
    if np.max(np.abs(le - (le[0] + np.arange(le.size) * dle))) > 1e-9 * max(1.0, abs(dle)):
        return None
    return float(le[0]), float(dle)


# Pinned staging buffers, one per dtype and direction, grown to the largest size seen (a
# pinned allocation costs far more than a copy, so sizes that change from call to call, like
# the fit's restart rounds, reuse one buffer).
_PINNED = {}
_PINNED_UP = {}


def _pinned(cache, dtype, n):
    buf = cache.get(dtype)
    if buf is None or buf.numel() < n:
        cap = max(n, 2 * (buf.numel() if buf is not None else 0), 1024)
        buf = cache[dtype] = torch.empty(cap, dtype=dtype, pin_memory=True)
    return buf[:n]


def _to_host(t):
    """A fresh numpy copy of a device tensor, through a pinned buffer (about twice the rate of a
    copy into pageable memory for the multi-MB arrays of large batches)."""
    buf = _pinned(_PINNED, t.dtype, t.numel())
    buf.copy_(t.reshape(-1), non_blocking=True)
    torch.cuda.current_stream().synchronize()
    return buf.numpy().copy()


def _to_device(a, dev):
    """A device copy of a small host array, staged through a pinned buffer and copied without
    waiting (a copy from pageable memory blocks the host).  The staging buffer is reused by the
    next call, which in every caller here comes after a synchronising read-back."""
    a = np.ascontiguousarray(a)
    dtype = torch.from_numpy(a[:0].reshape(0)).dtype
    buf = _pinned(_PINNED_UP, dtype, a.size)
    buf.numpy()[...] = a.reshape(-1)
    return buf.to(dev, non_blocking=True).view(a.shape)


_PENDING = {}


class PendingCopy:
    """An asynchronous device-to-host copy into the pinned buffer of `slot`; ``get()`` waits for
    it and returns a numpy copy (the copy is ordered on the stream after the producers).  A
    slot's buffer and event persist (a pinned allocation costs far more than the copy); a new
    copy into a slot must come after ``get()`` of the previous one."""

    def __init__(self, t, slot):
        st = _PENDING.get(slot)
        if st is None:
            st = _PENDING[slot] = [{}, torch.cuda.Event()]
        self.buf = _pinned(st[0], t.dtype, t.numel())
        self.buf.copy_(t.reshape(-1), non_blocking=True)
        self.ev = st[1]
        self.ev.record()

    def get(self):
        self.ev.synchronize()
        return self.buf.numpy().copy()


def ferguson_sums(plan, D, mask, logEps, method=None, want_log_dev=False, defer=False):
    """a12 sums: (npd, n_eps) arrays of sum W, d log(sum W)/d log eps and #terms (and, with
    want_log_dev, log sum W as a device tensor).  method: "moments" (bucket moments, needs a
    log-uniform grid), "direct" (every entry at every eps) or "auto" (moments when the grid
    allows).  defer: return a PendingCopy of the (3, npd, n_eps) host arrays instead of
    waiting for them (with want_log_dev: (pending, logS_dev))."""
    npd = plan.npd
    le = _device_const(logEps, D.device)
    n_eps = le.shape[0]
    fer = torch.empty(npd * n_eps * 3 + 2 * npd, dtype=torch.float64, device=D.device)
    method = method or FERGUSON
    grid = uniform_grid(logEps) if method in ("auto", "moments") else None
    if method == "moments" and grid is None:
        raise ValueError("ferguson_sums: the moments method needs a log-uniform eps grid")
    if grid is not None:
        ws = torch.empty(npd * nat.FM_WS_PER_PD, dtype=torch.float64, device=D.device)
        nat.ferguson_moments(plan, D, mask, le, grid[0], grid[1], fer, ws)
    else:
        nat.ferguson_sweep(plan, D, mask, le, fer)
    S_dev = fer[:npd * n_eps * 3].view(npd, n_eps, 3)[:, :, 0]
    logS_dev = torch.log(S_dev).contiguous() if want_log_dev else None
    # the slope (d sum W / d log eps over sum W) divided on the device (IEEE float64 division
    # there and in numpy give the same values), then S, slope and the counts in one copy
    f3 = fer[:npd * n_eps * 3].view(npd, n_eps, 3)
    st = torch.stack([f3[:, :, 0], f3[:, :, 1] / f3[:, :, 0], f3[:, :, 2]])
    if defer:
        pend = PendingCopy(st, "fer")
        return (pend, logS_dev) if want_log_dev else pend
    h = _to_host(st)
    h = h.reshape(3, npd, n_eps)
    if want_log_dev:
        return h[0], h[1], h[2], logS_dev
    return h[0], h[1], h[2]


class StageTimer:
    """Opt-in wall-clock timer of pipeline stages (synchronises the device at each mark)."""

    def __init__(self):
        import time
        self._now = time.perf_counter
        self.t = {}
        self._last = None

    def start(self):
        torch.cuda.synchronize()
        self._last = self._now()

    def mark(self, name):
        torch.cuda.synchronize()
        now = self._now()
        self.t[name] = self.t.get(name, 0.0) + (now - self._last)
        self._last = now


# The Lanczos matvec reads the kernel's nonzeros only (nat.kernel_ell) for the PDs that keep
# k <= n / SPARSE_RATIO neighbours: a kNN kernel then has about 2k nonzeros per row, and 12 B
# per nonzero beat 8 B per entry of the dense row.  The other PDs of the batch keep the dense
# matvec (both run in every step, each skipping the other's PDs).
SPARSE_RATIO = 4

# Storage type of the normalised kernel L that the Lanczos matvecs read.  float64 is the
# reference's kernel (slaplacianonFly.py builds it in float64); float32 rounds each entry once
# (~6e-8 relative), which moves the eigenvectors by about that over the eigengap, for half the
# matvec bytes.
KERNEL_DTYPE = torch.float64


def dm_embed(plan, D, tune, num_eigs=15, logEps=LOG_EPS, rand=None, a0s=None, mask=None,
             eig="lanczos", timer=None, fit="gpu"):
    """a10-a16 for every PD of the plan.  D: flat float32 device buffer.  Returns a list of
    DMResult (numpy outputs).  eig: "lanczos" (batched GPU Lanczos) or "dense" (rocSOLVER
    syevd on each PD, for checking)."""
    dev = D.device
    info = plan.info
    npd = plan.npd
    if timer:
        timer.start()
    if mask is None:
        mask = knn_mask(plan, D)
    if timer:
        timer.mark("knn")
    # the sums and log sums come back to the host while the device goes on (they are outputs;
    # the fit reads the device copy of log sum W)
    fer_h, logS_dev = ferguson_sums(plan, D, mask, logEps, want_log_dev=True, defer=True)
    logS_h = PendingCopy(logS_dev, "logS")
    n_eps = logS_dev.shape[1]
    if timer:
        timer.mark("ferguson")
    rand = np.random.rand if rand is None else rand
    results = [DMResult() for _ in range(npd)]
    a0l = [((rand(4, 1) - 0.5) if a0s is None else a0s[p]) for p in range(npd)]
    logEps_a = np.asarray(logEps)
    if fit == "gpu":
        sigma, fit_finish = fit_bandwidths(logEps_a, logS_dev, a0l, tune, rand)
        sig_h = PendingCopy(sigma, "sigma")
    else:
        logS = logS_h.get().reshape(npd, n_eps)
        sig = np.zeros(npd)
        for p, r in enumerate(results):
            r.popt, _, r.R_squared = ferguson_fit(logEps_a, logS[p], a0l[p], rand)
            sig[p] = tune * np.sqrt(2 * np.exp(-r.popt[1] / r.popt[0]))
        sigma = _to_device(sig, dev)
    if timer:
        timer.mark("tanh_fit_host")
    # the kernel matrix buffer is kept on the plan so that the Lanczos graph stays valid
    L = getattr(plan, "_L", None)
    if (L is None or L.numel() < max(info.d_elems, 1) or L.device != dev or
            L.dtype != KERNEL_DTYPE):
        L = torch.empty(max(info.d_elems, 1), dtype=KERNEL_DTYPE, device=dev)
        plan._L = L
    deg = torch.empty(2 * info.n_rows, dtype=torch.float64, device=dev)
    ks = [min(num_eigs, int(n) - 3) + 1 for n in plan.ns]

    sel = tuple(int(k) * SPARSE_RATIO <= int(n) for k, n in zip(plan.ks, plan.ns))
    sparse = eig == "lanczos" and KERNEL_DTYPE == torch.float64 and any(sel)
    sel_t = None
    if sparse:
        ell = getattr(plan, "_ell", None)
        if ell is None or ell[0].numel() < max(info.d_elems, 1) or ell[0].device != dev:
            ell = (torch.empty(max(info.d_elems, 1), dtype=torch.int32, device=dev),
                   torch.empty(max(info.d_elems, 1), dtype=torch.float64, device=dev),
                   torch.empty(max(info.n_rows, 1), dtype=torch.int32, device=dev))
            plan._ell = ell
        if not all(sel):                 # mixed batch: the selection rides with the buffers
            cached = getattr(plan, "_ell_sel", None)
            if cached is None or cached[0] != sel or cached[1].device != dev:
                cached = (sel, torch.as_tensor(np.asarray(sel, dtype=np.int32), device=dev))
                plan._ell_sel = cached
            sel_t = cached[1]
            ell = ell + (sel_t, L)

    def spectral(sigma):
        """a14-a15 from the bandwidths: (eigs, Yall) of the Lanczos path, else None."""
        nat.markov_normalize(plan, D, mask, sigma, L, deg)
        if sparse:
            nat.kernel_ell(plan, L, *ell[:3], sel=sel_t)
        if timer:
            timer.mark("normalize")
        if eig != "lanczos":
            return None
        out = lanczos_eigs(plan, ell if sparse else L, ks, full=True, **LANCZOS)
        if timer:
            timer.mark("lanczos")
        return out

    eig_out = spectral(sigma)
    if fit == "gpu":
        sig = sig_h.get()
    if eig == "lanczos":
        eigs, Yall = eig_out
        # psi = V[:, 1:] / V[:, 0] and mu = V[:, 0]^2 over all rows at once; PDs take row blocks
        with np.errstate(divide="ignore", invalid="ignore"):
            psi_all = Yall[:, 1:] / Yall[:, 0:1]
        mu_all = Yall[:, 0] ** 2
    # the deferred outputs of a12 / a13 (their copies finished long before the eigensolve)
    fh = fer_h.get().reshape(3, npd, n_eps)
    logS = logS_h.get().reshape(npd, n_eps)
    if fit == "gpu":
        popt_b, R2_b = fit_finish(logS)
    for p, r in enumerate(results):
        r.logSumWij = logS[p]
        r.slope = fh[1, p]
        r.logEps = logEps_a
        r.sigma = sig[p]
        if fit == "gpu":
            r.popt, r.R_squared = popt_b[p], R2_b[p]
    for p in range(npd):
        n = int(plan.ns[p])
        nEigs = ks[p] - 1
        if eig == "lanczos":
            lam, V, r_steps, r_res = eigs[p]
            results[p].lanczos_steps, results[p].lanczos_residual = r_steps, r_res
            if r_steps < n and r_res > LANCZOS_TOL * max(np.max(np.abs(lam)), 1e-300):
                # the reference reports ARPACK's non-convergence and keeps the partial
                # results (sembeddingonFly.py:28-32); so does this path
                warnings.warn(f"Lanczos did not converge for PD {p} (n={n}) in {r_steps} "
                              f"steps: largest Ritz residual {r_res:.3e}; keeping the partial "
                              "results", RuntimeWarning, stacklevel=2)
        else:
            lam, V = top_eigs_lm(plan.block(L, p).double(), nEigs + 1)
            lam, V = lam.cpu().numpy(), V.cpu().numpy()
        r = results[p]
        r.lamb = lam
        r0 = int(plan.rows0[p])
        if eig == "lanczos" and V.shape[1] - 1 == nEigs:
            r.psi = psi_all[r0:r0 + n, :nEigs]
            r.mu = mu_all[r0:r0 + n]
        else:                            # fewer vectors than asked (tiny PDs): zero columns
            psi = np.zeros((n, nEigs))
            psi[:, :V.shape[1] - 1] = V[:, 1:] / V[:, 0:1]
            r.psi = psi
            r.mu = V[:, 0] ** 2
        r.V = V
    if timer:
        timer.mark("eig_post")
    return results

"""Synthetic projection-direction (PD) stacks for tests and benchmarks.

The reference ships no particle images (its demo holds alignments only), so every
workload here is generated.  The model:

* volume: ``n_blobs`` isotropic 3-D Gaussians in an ``N``-voxel box, sigma ~ U(2, 4) px,
  centres uniform in a ball of radius ``0.35 * N``, amplitude ~ U(0.5, 1);
  ``n_moving`` of them are displaced by ``shift_px * t * u`` with a fixed unit vector ``u``
  and a per-particle conformational coordinate ``t ~ U(0, 1)``;
* orientations: directions inside a ``cap_deg`` cap around the PD axis, in-plane angle
  psi ~ U(0, 360) deg; Euler (rot, tilt, psi) -> quaternion follows the reference's
  RELION convention (``util.py:76-92``, ``read_alignfile.py:93-94``):
  ``q = qzs * (qy * qz)`` with the flipped psi half-angle;
* images: analytic projections (a 3-D Gaussian projects to a 2-D Gaussian of amplitude
  ``a * sigma * sqrt(2 pi)``) of the rotated volume, multiplied in Fourier space by the CTF
  of ``ctemh_cryoFrank.py:24-44`` at the particle's defocus, plus white Gaussian noise of
  variance ``var(signal) / snr``;
* storage: float32 SPIDER raw layout, i.e. ``raw[i] = image.T`` so that the reference's
  ``np.memmap(...).T`` (``getDistanceCTF_local_Conj9combinedS2.py:256-258``) returns the image.

Everything is computed in float64 torch on the requested device and returned as numpy
(CPU) or torch (``as_torch=True``) arrays.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

SEED = 20261202

# microscope constants shared by all workloads (kV, mm, ratio, A/px)
EKV = 300.0
CS = 2.7
AMPC = 0.1
PIX = 1.5


@dataclass
class PDStack:
    raw: object            # (n_img, N, N) float32, SPIDER layout (raw[i] = image.T)
    q: np.ndarray          # (4, n) float64 quaternions
    df: np.ndarray         # (n,) defocus in Angstrom
    ind: np.ndarray        # (n,) global particle indices (>= nStot/2 marks a conjugate)
    nStot: int             # total particles after conjugate augmentation
    N: int
    t: np.ndarray          # (n,) conformational coordinate used for each particle
    axis: np.ndarray       # (3,) PD axis
    extra: dict = field(default_factory=dict)


def seeds(n):
    """Independent child seed sequences of the fixed root seed."""
    return np.random.SeedSequence(SEED).spawn(n)


def qmult(q, s):
    """Quaternion product of 4xM arrays (same algebra as ``qMult_bsx.py:15-40``)."""
    q0, qv = q[0], q[1:4]
    s0, sv = s[0], s[1:4]
    c = np.stack((qv[1] * sv[2] - qv[2] * sv[1],
                  qv[2] * sv[0] - qv[0] * sv[2],
                  qv[0] * sv[1] - qv[1] * sv[0]))
    return np.vstack((q0 * s0 - np.sum(qv * sv, axis=0), q0 * sv + s0 * qv + c))


def euler_to_q(phi, theta, psi):
    """RELION (rot, tilt, psi) in radians -> quaternions 4xM (``util.py:76-92`` with flip)."""
    z = np.zeros_like(phi)
    qz = np.vstack((np.cos(phi / 2), z, z, -np.sin(phi / 2)))
    qy = np.vstack((np.cos(theta / 2), z, -np.sin(theta / 2), z))
    qzs = np.vstack((np.cos(psi / 2), z, z, -np.sin(psi / 2)))
    return qmult(qzs, qmult(qy, qz))


def q_to_rot(q):
    """Rotation matrices (M, 3, 3) of unit quaternions 4xM; row 3 is the PD of
    ``calc_avg_pd`` (``getDistanceCTF_local_Conj9combinedS2.py:182-184``)."""
    q0, q1, q2, q3 = q
    R = np.empty((q.shape[1], 3, 3))
    R[:, 0, 0] = 1 - 2 * (q2 * q2 + q3 * q3)
    R[:, 0, 1] = 2 * (q1 * q2 - q0 * q3)
    R[:, 0, 2] = 2 * (q1 * q3 + q0 * q2)
    R[:, 1, 0] = 2 * (q1 * q2 + q0 * q3)
    R[:, 1, 1] = 1 - 2 * (q1 * q1 + q3 * q3)
    R[:, 1, 2] = 2 * (q2 * q3 - q0 * q1)
    R[:, 2, 0] = 2 * (q1 * q3 - q0 * q2)
    R[:, 2, 1] = 2 * (q2 * q3 + q0 * q1)
    R[:, 2, 2] = 1 - 2 * (q1 * q1 + q2 * q2)
    return R


def make_volume(rng, N, n_blobs=64, n_moving=16):
    """Random blob volume (positions relative to the box centre, in pixels)."""
    r = 0.35 * N * rng.uniform(0, 1, n_blobs) ** (1.0 / 3.0)
    v = rng.normal(size=(n_blobs, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    centres = v * r[:, None]
    sig = rng.uniform(2.0, 4.0, n_blobs)
    amp = rng.uniform(0.5, 1.0, n_blobs)
    u = rng.normal(size=3)
    u /= np.linalg.norm(u)
    moving = np.zeros(n_blobs, bool)
    moving[:n_moving] = True
    return dict(centres=centres, sig=sig, amp=amp, u=u, moving=moving)


def random_axis(rng):
    v = rng.normal(size=3)
    return v / np.linalg.norm(v)


def cap_orientations(rng, n, axis, cap_deg=3.0):
    """Euler angles (phi, theta, psi) of ``n`` directions inside a cap around ``axis``."""
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    cos_cap = math.cos(math.radians(cap_deg))
    cz = rng.uniform(cos_cap, 1.0, n)
    az = rng.uniform(0, 2 * np.pi, n)
    sz = np.sqrt(1 - cz * cz)
    local = np.stack((sz * np.cos(az), sz * np.sin(az), cz), axis=1)
    # frame whose z axis is `axis`
    a = np.array([1.0, 0, 0]) if abs(axis[0]) < 0.9 else np.array([0, 1.0, 0])
    e1 = np.cross(axis, a)
    e1 /= np.linalg.norm(e1)
    e2 = np.cross(axis, e1)
    d = local[:, 0:1] * e1 + local[:, 1:2] * e2 + local[:, 2:3] * axis
    theta = np.arccos(np.clip(d[:, 2], -1, 1))
    phi = np.arctan2(d[:, 1], d[:, 0])
    psi = rng.uniform(0, 2 * np.pi, n)
    return phi, theta, psi


def ctf_grid_k(N, pix):
    """|k| in 1/A on the unshifted FFT grid, built like ``create_grid`` then ``ifftshift``."""
    if N % 2 == 1:
        a = np.arange(-(N - 1) / 2, (N - 1) / 2 + 1)
    else:
        a = np.arange(-N / 2, N / 2)
    X, Y = np.meshgrid(a, a)
    Q = (1.0 / (N / 2.0)) * np.sqrt(X ** 2 + Y ** 2)
    return np.fft.ifftshift(Q / (2 * pix))


def ctf_values(k, df, ekv=EKV, cs=CS, ampc=AMPC, gauss_env=np.inf):
    """CTF of ``ctemh_cryoFrank.py:24-44``; df in Angstrom (broadcasts over k and df)."""
    csa = cs * 1.0e7
    wav = 12.3986 / np.sqrt((2 * 511.0 + ekv) * ekv)
    w1 = np.pi * csa * wav ** 3
    w2 = np.pi * wav * df
    k2 = k * k
    sigm = gauss_env / math.sqrt(2 * math.log(2))
    wi = np.exp(-k2 / (2 * sigm ** 2))
    wr = (0.5 * w1 * k2 - w2) * k2
    return (np.sin(wr) - ampc * np.cos(wr)) * wi


def project(vol, q, t, N, shift_px, device="cpu"):
    """Analytic projections (n, N, N) float64 torch of the blob volume."""
    R = torch.as_tensor(q_to_rot(q), dtype=torch.float64, device=device)        # (n,3,3)
    cen = torch.as_tensor(vol["centres"], dtype=torch.float64, device=device)   # (B,3)
    u = torch.as_tensor(vol["u"], dtype=torch.float64, device=device)
    mv = torch.as_tensor(vol["moving"], dtype=torch.float64, device=device)
    tt = torch.as_tensor(t, dtype=torch.float64, device=device)
    sig = torch.as_tensor(vol["sig"], dtype=torch.float64, device=device)
    amp = torch.as_tensor(vol["amp"], dtype=torch.float64, device=device)
    c = cen[None] + (shift_px * tt)[:, None, None] * mv[None, :, None] * u[None, None, :]
    cr = torch.einsum("nij,nbj->nbi", R, c)                                     # (n,B,3)
    h = N / 2.0
    g = torch.arange(N, dtype=torch.float64, device=device) - h
    # separable Gaussians: rows follow the second rotated coordinate, columns the first
    ex = torch.exp(-(g[None, None, :] - cr[:, :, 0:1]) ** 2 / (2 * sig[None, :, None] ** 2))
    ey = torch.exp(-(g[None, None, :] - cr[:, :, 1:2]) ** 2 / (2 * sig[None, :, None] ** 2))
    w = (amp * sig * math.sqrt(2 * math.pi))[None, :, None]
    return torch.bmm((ey * w).transpose(1, 2).contiguous(), ex.contiguous())   # (n,N,N)


def make_pd(seed, n, N, df_range_um=(0.5, 3.0), snr=0.1, shift_px=6.0, axis=None,
            cap_deg=3.0, device="cpu", as_torch=False, conj_frac=0.0, vol=None):
    """One synthetic PD: ``n`` particles of side ``N``.

    ``conj_frac`` marks that fraction of particles as conjugates (index >= nStot/2); their
    stored image is the vertically flipped projection, so the reference's ``np.flipud``
    on load (``getDistanceCTF_local_Conj9combinedS2.py:271-275``) restores it.
    """
    rng = np.random.default_rng(seed)
    if vol is None:
        vol = make_volume(rng, N)
    if axis is None:
        axis = np.array([0.0, 0.0, 1.0])
    phi, theta, psi = cap_orientations(rng, n, axis, cap_deg)
    q = euler_to_q(phi, theta, psi)
    t = rng.uniform(0, 1, n)
    df = rng.uniform(df_range_um[0], df_range_um[1], n) * 1.0e4
    img = project(vol, q, t, N, shift_px, device=device)
    C = torch.as_tensor(ctf_values(ctf_grid_k(N, PIX)[None], df[:, None, None]), device=device)
    img = torch.fft.ifft2(torch.fft.fft2(img) * C).real
    var = img.reshape(n, -1).var(dim=1)
    gen = torch.Generator(device=device).manual_seed(int(rng.integers(1 << 62)))
    noise = torch.randn(img.shape, generator=gen, dtype=torch.float64, device=device)
    img = img + noise * torch.sqrt(var / snr)[:, None, None]
    nconj = int(round(conj_frac * n))
    conj = np.zeros(n, bool)
    if nconj:
        conj[rng.choice(n, nconj, replace=False)] = True
    conj_t = torch.as_tensor(conj, device=device)
    img = torch.where(conj_t[:, None, None], torch.flip(img, dims=[1]), img)
    raw = img.transpose(1, 2).contiguous().to(torch.float32)
    nStot = 2 * n
    ind = np.arange(n, dtype=np.int64) + np.where(conj, nStot // 2, 0)
    if not as_torch:
        raw = raw.cpu().numpy()
    return PDStack(raw=raw, q=q, df=df, ind=ind, nStot=nStot, N=N, t=t, axis=np.asarray(axis),
                   extra=dict(conj=conj))


def lognormal_sizes(seed, n_pd, median=400, sigma=0.6, lo=100, hi=3000):
    """Particle counts of a batched workload: lognormal, clipped."""
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.lognormal(math.log(median), sigma, n_pd)), lo, hi).astype(int)


# the five workloads (image side, particles, defocus range, shift, PD count)
CONFIGS = {
    "c1_plain_one_pd": dict(n=200, N=70, df=(0.5, 3.0), shift=6.0, n_pd=1, k=None,
                            n_eps=101, n_eigs=8),
    "c2_single_large_pd": dict(n=2000, N=128, df=(0.5, 3.5), shift=6.0, n_pd=1, k=200,
                               n_eps=1501, n_eigs=10),
    "c3_batched_pds": dict(n="lognormal", N=96, df=(0.5, 3.0), shift=6.0, n_pd=500, k=150,
                           n_eps=1501, n_eigs=8),
    "c4_autograd": dict(n=1000, N=128, df=(0.5, 3.0), shift=6.0, n_pd=1, k=None,
                        n_eps=41, n_eigs=8),
    "c5_nlsa": dict(n=1500, N=160, df=(0.5, 3.0), shift=8.0, n_pd=200, k=None,
                    n_eps=1501, n_eigs=8, con_order=50),
}

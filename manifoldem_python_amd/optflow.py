"""Optical flow of NLSA movies on the GPU, batched over movies (the CC stage's per-movie work).

``movie_flows`` runs ``OpticalFlowMovie.op`` (CC/OpticalFlowMovie.py:385-555) for a whole batch
of movies in grouped launches:

1. 3-D Perona-Malik diffusion of every movie (``anisodiff3``, niter 5, kappa 50, gamma 0.1,
   step (5, 3, 3)): 4 launches;
2. float64 means of blocks of ``blockSize_avg`` frames: 1 launch;
3. in the 'GF-HS' mode, a float64 Gaussian pre-filter (sigma 1.5) of every averaged frame;
4. the Horn-Schunck op (CC/hornschunck_simple.py:43-84) for every frame pair of every movie:
   the float32 Gaussian (sigma 2.0 in 'HS', 1.5 in 'GF-HS') of every averaged frame once (the
   reference filters each frame twice, as im2 and then as im1, with the same result), the
   derivative images, then ``Niter`` iterations temporally blocked in LDS;
5. VxM, VyM = the sum over pairs, with the orientation and magnitude of FWD and REV.

Frame pair a of a movie is (frame max(a - 1, 0), frame a): the reference's first pair compares
the first averaged frame with itself.  The 'GF-HS' mode starts Horn-Schunck from the Farneback
flows (OpenCV), which this package does not compute: pass them as ``init_flows``.
"""
from __future__ import annotations

import numpy as np
import torch

from . import _native as nat

ANISO = dict(niter=5, kappa=50, gamma=0.1, step=(5.0, 3.0, 3.0), option=1)


def _dev():
    return torch.device("cuda")


def anisodiff3(stack, niter=1, kappa=50, gamma=0.1, step=(1.0, 1.0, 1.0), option=1):
    """Perona-Malik diffusion of float32 movies (nmov, F, H, W) on the GPU -> new tensor."""
    stack = stack.to(dtype=torch.float32).contiguous()
    out = torch.empty_like(stack)
    tmp = torch.empty_like(stack) if niter > 2 else None
    nat.flow_anisodiff(stack, out, tmp, niter, kappa, gamma, step, option)
    return out


def gaussian(images, sigma, out_dtype):
    """scipy gaussian_filter of (count, H, W) images (float64 -> float64 / float32, or
    float32 -> float32)."""
    images = images.contiguous()
    out = torch.empty(images.shape, dtype=out_dtype, device=images.device)
    tmp = torch.empty_like(out)
    nat.gaussian_filter(images, out, tmp, sigma)
    return out


def derivatives(frames, pairs):
    """fx, fy, ft float32 (nP, H, W) of frame pairs (indices into frames (nF, H, W))."""
    nP = pairs.shape[0]
    H, W = frames.shape[1:]
    fx, fy, ft = (torch.empty((nP, H, W), dtype=torch.float32, device=frames.device)
                  for _ in range(3))
    nat.hs_derivatives(frames.contiguous(), pairs, fx, fy, ft)
    return fx, fy, ft


def hs_iterate(fx, fy, ft, alpha, niter, U, V):
    """Horn-Schunck iterations in place on U, V (nP, H, W)."""
    nat.hs_iterate(fx, fy, ft, alpha, niter, U, V, torch.empty_like(U), torch.empty_like(V))
    return U, V


def frame_pairs(nmov, A, device):
    """int32 (nmov * A, 2): pair a of movie m = (m A + max(a - 1, 0), m A + a)."""
    a = torch.arange(A, dtype=torch.int32)
    prev = torch.clamp(a - 1, min=0)
    base = (torch.arange(nmov, dtype=torch.int32) * A)[:, None]
    return torch.stack([(base + prev).reshape(-1), (base + a).reshape(-1)], 1).to(device)


def movie_flows(movies, blockSize_avg=5, of_type="HS", init_flows=None, Niter=200, alpha=2.0):
    """FWD flows of a batch of movies.

    movies: (nmov, F, dim*dim) or (nmov, F, dim, dim), any real dtype, host or device.
    init_flows ('GF-HS' only): (Vx, Vy), each (nmov, A, dim, dim), the initial flow of every
    frame pair (A = ceil(F / blockSize_avg)); float32 keeps the iteration in float32 as the
    reference does for OpenCV's float32 flows.
    Returns a dict of float64 device tensors (nmov, dim, dim): Vx, Vy, Orient, Mag and
    Orient_rev (the REV orientation; REV's Vx, Vy are -Vx, -Vy and its Mag is Mag).
    """
    if of_type not in ("HS", "GF-HS"):
        raise NotImplementedError(f"optical flow type {of_type!r}: only 'HS' and 'GF-HS' "
                                  "(with init_flows) are built")
    dev = _dev()
    mv = torch.as_tensor(np.asarray(movies) if not isinstance(movies, torch.Tensor) else movies)
    mv = mv.to(device=dev, dtype=torch.float32)
    nmov, F = mv.shape[:2]
    if mv.dim() == 3:
        dim = int(np.sqrt(mv.shape[2]))
        if dim * dim != mv.shape[2]:
            raise ValueError("movie frames must be square images")
        mv = mv.reshape(nmov, F, dim, dim)
    H, W = mv.shape[2:]
    mv = mv.contiguous()
    dif = anisodiff3(mv, **ANISO)
    A = -(-F // blockSize_avg)
    avg = torch.empty((nmov, A, H, W), dtype=torch.float64, device=dev)
    nat.flow_block_average(dif, blockSize_avg, avg)
    del dif
    prefilter = of_type == "GF-HS"
    sig = 1.5 if prefilter else 2.0
    frames = avg.reshape(nmov * A, H, W)
    if prefilter:
        frames = gaussian(frames, sig, torch.float64)
    g = gaussian(frames, sig, torch.float32)
    pairs = frame_pairs(nmov, A, dev)
    fx, fy, ft = derivatives(g, pairs)
    del g
    if prefilter:
        if init_flows is None:
            raise NotImplementedError(
                "'GF-HS' starts from Farneback flows (OpenCV), which are not computed here: "
                "pass init_flows=(Vx, Vy) or use of_type='HS'")
        U = torch.as_tensor(np.asarray(init_flows[0]) if not isinstance(init_flows[0],
                            torch.Tensor) else init_flows[0]).to(dev)
        V = torch.as_tensor(np.asarray(init_flows[1]) if not isinstance(init_flows[1],
                            torch.Tensor) else init_flows[1]).to(dev)
        dt = torch.float32 if (U.dtype == torch.float32 and V.dtype == torch.float32) \
            else torch.float64
        U = U.to(dt).reshape(nmov * A, H, W).contiguous().clone()
        V = V.to(dt).reshape(nmov * A, H, W).contiguous().clone()
    else:
        U = torch.zeros((nmov * A, H, W), dtype=torch.float64, device=dev)
        V = torch.zeros_like(U)
    hs_iterate(fx, fy, ft, alpha, Niter, U, V)
    out = {k: torch.empty((nmov, H, W), dtype=torch.float64, device=dev)
           for k in ("Vx", "Vy", "Orient", "Mag", "Orient_rev")}
    nat.flow_sum(U.reshape(nmov, A, H, W), V.reshape(nmov, A, H, W), True, out["Vx"], out["Vy"],
                 out["Orient"], out["Mag"], out["Orient_rev"])
    return out


def flow_dicts(res, m):
    """The reference's (FWD, REV) FlowVec dicts of movie m from ``movie_flows`` output."""
    vx = res["Vx"][m].cpu().numpy()
    vy = res["Vy"][m].cpu().numpy()
    mag = res["Mag"][m].cpu().numpy()
    fwd = dict(Vx=vx, Vy=vy, Orient=res["Orient"][m].cpu().numpy(), Mag=mag)
    rev = dict(Vx=-1.0 * vx, Vy=-1.0 * vy, Orient=res["Orient_rev"][m].cpu().numpy(),
               Mag=mag.copy())
    return fwd, rev


def psi_movie_flows(movies, blockSize_avg=5, of_type="HS", init_flows=None):
    """``ComputeOpticalFlowPrDAll.ComputePsiMovieOpticalFlow`` (CC/ComputeOpticalFlowPrDAll.py:
    89-184, splitmovie = 0) for a list of equally shaped movies at once: a list of
    dict(FWD=..., REV=...)."""
    res = movie_flows(np.stack([np.asarray(m) for m in movies]), blockSize_avg, of_type,
                      init_flows)
    out = []
    for m in range(len(movies)):
        fwd, rev = flow_dicts(res, m)
        out.append(dict(FWD=fwd, REV=rev))
    return out

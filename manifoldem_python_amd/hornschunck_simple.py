"""Drop-in for ``CC/hornschunck_simple.py``: Horn-Schunck optical flow of one frame pair on the
GPU (the Gaussian pre-filter, the derivative images and the iterations are HIP kernels).

Same names, arguments and results as the reference (numpy arrays in and out).  The flow is
float64 unless both initial flows are float32 (then float32, as numpy keeps it); a mixed
float32 / float64 start runs in float64 from the first iteration.
"""
from __future__ import annotations

import numpy as np
import torch

from . import optflow

HSKERN = np.array([[1.0 / 12, 1.0 / 6, 1.0 / 12],
                   [1.0 / 6, 0., 1.0 / 6],
                   [1.0 / 12, 1.0 / 6, 1.0 / 12]], float)
kernelX = np.array([[-1.0, 1.0], [-1.0, 1.0]]) * .25
kernelY = np.array([[-1.0, -1.0], [1.0, 1.0]]) * .25
kernelT = np.ones((2, 2)) * .25


def _to_dev(a, dtype):
    return torch.as_tensor(np.array(a, dtype=dtype, order="C"), device="cuda")


def lowpassfilt(im, sig):
    """``hornschunck_simple.py:35-38`` (scipy gaussian_filter, mode 'reflect') for a 2-D float32
    or float64 image; the result has the input's dtype."""
    im = np.asarray(im)
    if im.ndim != 2 or im.dtype not in (np.float32, np.float64):
        raise TypeError("lowpassfilt: a 2-D float32 or float64 image")
    dt = torch.float64 if im.dtype == np.float64 else torch.float32
    x = _to_dev(im, im.dtype)[None]
    return optflow.gaussian(x, sig, dt)[0].cpu().numpy()


def computeDerivatives(im1, im2):
    """``hornschunck_simple.py:24-31`` for float32 frames (the dtype the reference's op passes)."""
    im1, im2 = np.asarray(im1), np.asarray(im2)
    if im1.dtype != np.float32 or im2.dtype != np.float32 or im1.shape != im2.shape \
            or im1.ndim != 2:
        raise TypeError("computeDerivatives: two 2-D float32 frames of one shape")
    frames = _to_dev(np.stack([im1, im2]), np.float32)
    pairs = torch.tensor([[0, 1]], dtype=torch.int32, device="cuda")
    fx, fy, ft = optflow.derivatives(frames, pairs)
    return fx[0].cpu().numpy(), fy[0].cpu().numpy(), ft[0].cpu().numpy()


def op(im1, im2, uInitial, vInitial, sigma=1.0, alpha=0.001, Niter=8, verbose=False):
    """``hornschunck_simple.py:43-84``: returns (U, V)."""
    im1 = np.asarray(im1)
    im2 = np.asarray(im2)
    if im1.ndim != 2 or im1.shape != im2.shape:
        raise ValueError("op: two 2-D frames of one shape")
    H, W = im1.shape
    uInitial = np.asarray(uInitial)
    vInitial = np.asarray(vInitial)
    if uInitial.shape[0] < 2:
        uInitial = np.zeros([H, W])
    if vInitial.shape[0] < 2:
        vInitial = np.zeros([H, W])
    f32 = uInitial.dtype == np.float32 and vInitial.dtype == np.float32
    dt = np.float32 if f32 else np.float64
    frames = _to_dev(np.stack([im1.astype(np.float32), im2.astype(np.float32)]), np.float32)
    g = optflow.gaussian(frames, sigma, torch.float32)
    pairs = torch.tensor([[0, 1]], dtype=torch.int32, device="cuda")
    fx, fy, ft = optflow.derivatives(g, pairs)
    U = _to_dev(np.broadcast_to(uInitial, (H, W)), dt)[None].clone()
    V = _to_dev(np.broadcast_to(vInitial, (H, W)), dt)[None].clone()
    optflow.hs_iterate(fx, fy, ft, alpha, int(Niter), U, V)
    return U[0].cpu().numpy(), V[0].cpu().numpy()

"""Drop-in for ``CC/OpticalFlowMovie.py``: the optical flow of one NLSA movie on the GPU.

``op`` follows the reference (CC/OpticalFlowMovie.py:385-555): diffusion, block averaging,
Horn-Schunck per frame pair, the summed flow, REV as the negated FWD flow.  ``OF_Type`` selects
the method as the reference's hard-coded variable does.  The reference's setting, 'GF-HS',
starts Horn-Schunck from OpenCV's Farneback flow, which this package does not compute: with
'GF-HS' set, ``op`` raises unless ``farneback`` is given a callable (prev, curr) -> (Vx, Vy)
float32 (dim, dim) for each frame pair.  'HS' (zero initial flow) is complete.  Figures
(``OFvisualPrint``) are not produced.
"""
from __future__ import annotations

import copy
import warnings

import numpy as np
import torch

from . import optflow

OF_Type = "GF-HS"
farneback = None


def getOrientMag(X, Y):
    """``OpticalFlowMovie.py:21-24``: ((atan2(Y, X) % 2) * pi, sqrt(X^2 + Y^2)) on the GPU."""
    X = np.asarray(X, dtype=np.float64)
    Y = np.asarray(Y, dtype=np.float64)
    shape = np.broadcast_shapes(X.shape, Y.shape)
    n = int(np.prod(shape)) if shape else 1
    # np.array copies the read-only broadcast view into a writable host array
    Xd = torch.as_tensor(np.array(np.broadcast_to(X, shape), order="C").reshape(1, 1, 1, n),
                         device="cuda")
    Yd = torch.as_tensor(np.array(np.broadcast_to(Y, shape), order="C").reshape(1, 1, 1, n),
                         device="cuda")
    vx, vy, o, m = (torch.empty(n, dtype=torch.float64, device="cuda") for _ in range(4))
    from . import _native as nat
    nat.flow_sum(Xd, Yd, False, vx, vy, o, m)
    return o.cpu().numpy().reshape(shape), m.cpu().numpy().reshape(shape)


def anisodiff3(stack, niter=1, kappa=50, gamma=0.1, step=(1., 1., 1.), option=1, ploton=False):
    """``OpticalFlowMovie.py:202-341`` on the GPU (float32 result, niter - 1 updates)."""
    if ploton:
        raise NotImplementedError("anisodiff3: ploton figures are not produced")
    stack = np.asarray(stack)
    if stack.ndim == 4:
        warnings.warn("Only grayscale stacks allowed, converting to 3D matrix")
        stack = stack.mean(3)
    x = torch.as_tensor(np.ascontiguousarray(stack, dtype=np.float32), device="cuda")[None]
    return optflow.anisodiff3(x, niter, kappa, gamma, step, option)[0].cpu().numpy()


def _farneback_flows(Mov, blockSize_avg):
    """Initial flows of every frame pair from the user's Farneback callable (pre-filtered,
    block-averaged frames, as the reference passes them)."""
    F = Mov.shape[0]
    dim = int(np.sqrt(Mov.shape[1]))
    x = torch.as_tensor(np.ascontiguousarray(np.resize(Mov, (F, dim, dim)), dtype=np.float32),
                        device="cuda")[None]
    dif = optflow.anisodiff3(x, **optflow.ANISO)
    A = -(-F // blockSize_avg)
    avg = torch.empty((1, A, dim, dim), dtype=torch.float64, device="cuda")
    optflow.nat.flow_block_average(dif, blockSize_avg, avg)
    fr = optflow.gaussian(avg.reshape(A, dim, dim), 1.5, torch.float64).cpu().numpy()
    us, vs = [], []
    for a in range(A):
        u, v = farneback(fr[max(a - 1, 0)], fr[a])
        us.append(np.asarray(u))
        vs.append(np.asarray(v))
    return np.stack(us)[None], np.stack(vs)[None]


def op(Mov, prd_psinum, blockSize_avg, label, OFvisualPrint, *argv):
    """``OpticalFlowMovie.py:385-555``: the FlowVec dict (Vx, Vy, Orient, Mag) of a movie
    (F, dim*dim); a 'REV' label negates the FWD flow (given as argv[0], or computed)."""
    if OFvisualPrint[0] or OFvisualPrint[1]:
        raise NotImplementedError("optical-flow figures (OFvisualPrint) are not produced")
    Mov = np.asarray(Mov)
    if argv:
        FlowVec = copy.deepcopy(argv[0])
        VxM, VyM = FlowVec["Vx"], FlowVec["Vy"]
    else:
        if OF_Type == "GF-HS":
            if farneback is None:
                raise NotImplementedError(
                    "OF_Type 'GF-HS' starts from OpenCV's Farneback flow, which is not computed "
                    "here: set OpticalFlowMovie.farneback to a callable (prev, curr) -> (Vx, Vy) "
                    "or OpticalFlowMovie.OF_Type = 'HS'")
            init = _farneback_flows(Mov, blockSize_avg)
            res = optflow.movie_flows(Mov[None], blockSize_avg, "GF-HS", init)
        elif OF_Type == "HS":
            res = optflow.movie_flows(Mov[None], blockSize_avg, "HS")
        else:
            raise NotImplementedError(f"OF_Type {OF_Type!r} is not built")
        VxM = res["Vx"][0].cpu().numpy()
        VyM = res["Vy"][0].cpu().numpy()
        FlowVec = dict(Vx=VxM, Vy=VyM, Orient=[], Mag=[])
    if label[0:3] == "REV":
        VxM = copy.deepcopy(-1.0 * VxM)
        VyM = copy.deepcopy(-1.0 * VyM)
        FlowVec["Vx"] = VxM
        FlowVec["Vy"] = VyM
    FlowVec["Orient"], FlowVec["Mag"] = getOrientMag(VxM, VyM)
    return FlowVec

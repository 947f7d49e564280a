"""Oracle for the CC-stage optical flow of NLSA movies (Horn-Schunck, float32/float64 as the
reference computes it).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``CC/hornschunck_simple.py:9-84`` (kernels, computeDerivatives, lowpassfilt, op) and
``CC/OpticalFlowMovie.py:21-24`` (getOrientMag), ``:202-341`` (anisodiff3) and ``:385-555``
(op: diffusion, block averaging, optional pre-filter, Horn-Schunck per frame pair, sum over
pairs, REV negation).  The filters are the reference's own library calls
(``scipy.ndimage.convolve`` and ``gaussian_filter``, mode 'reflect').  The Farneback initial
flow of the 'GF-HS' mode (OpenCV, absent here) is not restated: ``movie_flow`` takes the
initial flows as an argument in that mode.

Pinned by (``tests/test_oracle_optflow.py``):
  computeDerivatives - brute-force loops over the convolution definition with half-sample
                       reflection (catches correlate-for-convolve, a flipped kernel and a
                       wrong edge rule);
  lowpassfilt        - brute-force separable sum of the normalised 4-sigma Gaussian;
  op                 - identical frames give exactly zero flow; a ramp moved by d pixels gives
                       U = d (V = 0) in the interior; transposing both frames swaps U and V;
  anisodiff3         - niter = 1 leaves the stack unchanged (the loop runs niter - 1 times),
                       the stack sum is conserved, the variance falls, and one step equals the
                       flux form of the Perona-Malik update written per pixel;
  movie_flow         - the block average is the exact float64 mean, and a static movie has zero
                       flow;
  getOrientMag       - closed-form values, including the (atan2 % 2) * pi precedence.
"""
from __future__ import annotations

import numpy as np
from scipy.ndimage import convolve, gaussian_filter

HSKERN = np.array([[1.0 / 12, 1.0 / 6, 1.0 / 12],
                   [1.0 / 6, 0.0, 1.0 / 6],
                   [1.0 / 12, 1.0 / 6, 1.0 / 12]], float)
KX = np.array([[-1.0, 1.0], [-1.0, 1.0]]) * 0.25
KY = np.array([[-1.0, -1.0], [1.0, 1.0]]) * 0.25
KT = np.ones((2, 2)) * 0.25


def compute_derivatives(im1, im2):
    """``hornschunck_simple.py:24-31``."""
    fx = convolve(im1, KX) + convolve(im2, KX)
    fy = convolve(im1, KY) + convolve(im2, KY)
    ft = convolve(im1, KT) + convolve(im2, -KT)
    return fx, fy, ft


def lowpassfilt(im, sig):
    """``hornschunck_simple.py:35-38``."""
    return gaussian_filter(im, sigma=sig)


def hornschunck(im1, im2, uInitial, vInitial, sigma=1.0, alpha=0.001, Niter=8):
    """``hornschunck_simple.py:43-84``: the flow's dtype follows the initial flow's (float64
    for the zero start)."""
    im1 = lowpassfilt(np.asarray(im1).astype(np.float32), sigma)
    im2 = lowpassfilt(np.asarray(im2).astype(np.float32), sigma)
    if uInitial.shape[0] < 2:
        uInitial = np.zeros([im1.shape[0], im1.shape[1]])
    if vInitial.shape[0] < 2:
        vInitial = np.zeros([im1.shape[0], im1.shape[1]])
    U, V = uInitial, vInitial
    fx, fy, ft = compute_derivatives(im1, im2)
    for _ in range(Niter):
        uAvg = convolve(U, HSKERN)
        vAvg = convolve(V, HSKERN)
        der = (fx * uAvg + fy * vAvg + ft) / (alpha ** 2 + fx ** 2 + fy ** 2)
        U = uAvg - fx * der
        V = vAvg - fy * der
    return U, V


def anisodiff3(stack, niter=1, kappa=50, gamma=0.1, step=(1., 1., 1.), option=1):
    """``OpticalFlowMovie.py:202-341`` (float32; the update loop runs niter - 1 times)."""
    stack = stack.astype("float32")
    out = stack.copy()
    dD = np.zeros_like(out)
    dS = dD.copy()
    dE = dD.copy()
    for _ in np.arange(1, niter):
        dD[:-1] = np.diff(out, axis=0)
        dS[:, :-1] = np.diff(out, axis=1)
        dE[:, :, :-1] = np.diff(out, axis=2)
        if option == 1:
            gD = np.exp(-(dD / kappa) ** 2.) / step[0]
            gS = np.exp(-(dS / kappa) ** 2.) / step[1]
            gE = np.exp(-(dE / kappa) ** 2.) / step[2]
        else:
            gD = 1. / (1. + (dD / kappa) ** 2.) / step[0]
            gS = 1. / (1. + (dS / kappa) ** 2.) / step[1]
            gE = 1. / (1. + (dE / kappa) ** 2.) / step[2]
        D = gD * dD
        E = gE * dE
        S = gS * dS
        UD = D.copy()
        NS = S.copy()
        EW = E.copy()
        UD[1:] -= D[:-1]
        NS[:, 1:] -= S[:, :-1]
        EW[:, :, 1:] -= E[:, :, :-1]
        out += gamma * (UD + NS + EW)
    return out


def block_average(Mov, blockSize_avg):
    """``OpticalFlowMovie.py:447-457``: float64 means of consecutive blocks of frames."""
    F = Mov.shape[0]
    nA = int(np.ceil(float(F) / blockSize_avg))
    out = np.zeros((nA,) + Mov.shape[1:])
    for b in range(nA):
        out[b] = np.mean(Mov[b * blockSize_avg:min((b + 1) * blockSize_avg, F)], axis=0,
                         dtype=np.float64)
    return out


def get_orient_mag(X, Y):
    """``OpticalFlowMovie.py:21-24`` (the operator precedence is the reference's)."""
    return np.arctan2(Y, X) % 2 * np.pi, np.sqrt(X ** 2 + Y ** 2)


def movie_flow(Mov, blockSize_avg, of_type="HS", init_flows=None, Niter=200, diffuse=True):
    """``OpticalFlowMovie.py:385-523`` for a FWD movie: returns (VxM, VyM).

    of_type 'HS': zero initial flow, no pre-filter, sigma 2.0.  'GF-HS': every block-averaged
    frame is pre-filtered (float64 Gaussian, sigma 1.5) and Horn-Schunck (sigma 1.5) starts
    from init_flows[frame] = (Vx, Vy), the Farneback flows the reference computes with OpenCV.
    diffuse=False takes Mov as already diffused (a checker of the later steps).
    """
    F = Mov.shape[0]
    dim = int(np.sqrt(Mov.shape[1]))
    Mov = np.resize(Mov, (F, dim, dim))
    if diffuse:
        Mov = anisodiff3(Mov, niter=5, kappa=50, gamma=0.1, step=(5., 3., 3.), option=1)
    Avg = block_average(Mov, blockSize_avg)
    prefilter = of_type == "GF-HS"
    sig = 1.5 if prefilter else 2.0
    prev = lowpassfilt(Avg[0], sig) if prefilter else Avg[0]
    VxM = np.zeros((dim, dim))
    VyM = np.zeros((dim, dim))
    for f in range(Avg.shape[0]):
        curr = lowpassfilt(Avg[f], sig) if prefilter else Avg[f]
        if prefilter:
            u0, v0 = init_flows[f]
        else:
            u0 = np.zeros(prev.shape)
            v0 = np.zeros(prev.shape)
        Vx, Vy = hornschunck(prev, curr, u0, v0, sig, 2.0, Niter)
        VxM = VxM + Vx
        VyM = VyM + Vy
        prev = curr.copy()
    return VxM, VyM

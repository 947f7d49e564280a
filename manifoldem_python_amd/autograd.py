"""Differentiable fused op: CTF-weighted distances -> Gaussian kernel -> log-sum.

For Fourier images X (the rfft2 half spectra of the aligned particles, complex64 (n, N, Nh)),
defocus df and bandwidths eps (float64, n_eps):

    D_ij   = sum_k w_k |C_jk X_ik - C_ik X_jk|^2          (getDistanceCTF...:391-397)
    K^e_ij = exp(-D_ij / (2 eps_e))                      (the kernel fergusonE.py:40 sums)
    L_e    = sum_i log sum_j K^e_ij

Forward runs the MFMA Gram and a HIP row-sum kernel.  Backward is hand written and recomputes
K from the stored D instead of storing K:

    dL_e/d eps_e   = sum_i T_ei / (2 eps_e^2 S_ei),  S = sum_j K, T = sum_j K D
    Wt_ij          = sum_e g_e (-1/(2 eps_e)) K^e_ij (1/S_ei + 1/S_ej)       (HIP kernel)
    dL/d conj(X_ik) = w_k [X_ik (Wt B)_ik - C_ik (Wt Y)_ik],  B = C^2, Y = C X  (two GEMMs)

returned in torch's convention for complex inputs (2 dL/d conj X).  The Ferguson log-log slope
d log sum_ij K / d log eps is ``slope(D, eps)`` below.
"""
from __future__ import annotations

import torch

from . import _native as nat
from . import pipeline as pl


def _half_weights(N, device):
    Nh = N // 2 + 1
    w = torch.full((Nh,), 2.0, dtype=torch.float32, device=device)
    w[0] = 1.0
    if N % 2 == 0:
        w[-1] = 1.0
    return w


class KernelLogSum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, fy, df, eps, ctf, precision=nat.PREC_FP32):
        dev = fy.device
        n, N, Nh = fy.shape
        fyd = fy.detach().contiguous()
        op = pl.Operand(n, N, precision, dev)
        ctf_half = torch.empty((n, N * Nh), dtype=torch.float32, device=dev)
        nat.ctf_pack(fyd, df.detach().contiguous(), N, ctf, precision, op.Khp, 0, op.hi,
                     op.lo if op.lo is not None else op.hi, ctf_half)
        plan = nat.Plan([0], [n], [n], chunks=pl.CHUNKS[precision], device=dev)
        D = pl.ctf_distances(plan, op, fyd, ctf_half, N).block(0)
        epsd = eps.detach().to(torch.float64).contiguous()
        S = torch.empty((epsd.numel(), n), dtype=torch.float64, device=dev)
        T = torch.empty_like(S)
        nat.kernel_rowsums(D, epsd, S, T)
        ctx.save_for_backward(fyd, ctf_half, D, epsd, S, T)
        ctx.N = N
        return torch.log(S).sum(dim=1)

    @staticmethod
    def backward(ctx, gL):
        fyd, ctf_half, D, epsd, S, T = ctx.saved_tensors
        N = ctx.N
        n = fyd.shape[0]
        g = gL.detach().to(torch.float64).contiguous()
        g_eps = None
        if ctx.needs_input_grad[2]:
            g_eps = g * (T / (2.0 * epsd[:, None] ** 2 * S)).sum(dim=1)
        g_fy = None
        if ctx.needs_input_grad[0]:
            Wt = torch.empty((n, n), dtype=torch.float32, device=D.device)
            nat.kernel_grad_weights(D, epsd, S, g, Wt)
            C = ctf_half.view(n, N, N // 2 + 1)
            X = fyd
            B = (C * C).reshape(n, -1)
            Y = (C * X).reshape(n, -1)
            WB = (Wt @ B).view_as(C)
            WYr = (Wt @ Y.real.contiguous()).view_as(C)
            WYi = (Wt @ Y.imag.contiguous()).view_as(C)
            w = _half_weights(N, D.device)
            gr = w * (X.real * WB - C * WYr)
            gi = w * (X.imag * WB - C * WYi)
            g_fy = 2.0 * torch.complex(gr, gi)
        return g_fy, None, g_eps, None, None


def kernel_logsum(fy, df, eps, ctf, precision=nat.PREC_FP32):
    """L_e = sum_i log sum_j exp(-D_ij / (2 eps_e)) with a hand-written backward."""
    return KernelLogSum.apply(fy, df, eps, ctf, precision)


def slope(D, eps):
    """Exact d log(sum_ij K) / d log eps for each eps from the row sums (no threshold)."""
    n = D.shape[0] if D.dim() == 2 else int(round(D.numel() ** 0.5))
    epsd = torch.as_tensor(eps, dtype=torch.float64, device=D.device).contiguous()
    S = torch.empty((epsd.numel(), n), dtype=torch.float64, device=D.device)
    T = torch.empty_like(S)
    nat.kernel_rowsums(D.reshape(n, n).contiguous(), epsd, S, T)
    return (T.sum(1) / (2.0 * epsd)) / S.sum(1)

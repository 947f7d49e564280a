"""1-D open-manifold fit x_j = a_j cos(j pi tau) + b_j (a25).

Reference: ``fit_1D_open_manifold_3D.py:64-160``, ``get_fit_1D_open_manifold_3D_param.py:12-95``,
``solve_d_R_d_tau_p_3D.py:35-60``, ``R_p.py:41-49``.  The reference solves one quintic per point
with ``np.roots`` inside a Python loop, every iteration.  Here each tau update of all points and
the sums the next (a, b) needs run in one HIP launch (memi_fit1d_tau: real roots by monotone
segments, arg min of R_p), and only the 2x2 least-squares solves and the convergence test stay
on the host.  np.roots keeps its real roots where they are exactly real in LAPACK's output; the
kernel flags the rare points where that decision (or the arg min) is within rounding of a tie,
and ``solve_tau`` re-solves exactly those with np.roots' own LAPACK routine, so the
classification is the reference's.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from . import _native as nat

EPS = 1e-4


def _residual(x, a, b, tau):
    """R_p for points x (n, 3) at candidate taus (n, c) -> (n, c)."""
    j = np.arange(1, 4)
    err = x[:, None, :] - b - a * np.cos(tau[:, :, None] * j * np.pi)
    return np.sum(err ** 2, axis=2)


def _coefficients(x, a, b):
    n = x.shape[0]
    c = np.empty((n, 6))
    c[:, 0] = 48 * a[2] ** 2
    c[:, 1] = 0.0
    c[:, 2] = 8 * a[1] ** 2 - 48 * a[2] ** 2
    c[:, 3] = -12 * a[2] * (x[:, 2] - b[2])
    c[:, 4] = a[0] ** 2 - 4 * a[1] ** 2 + 9 * a[2] ** 2 - 4 * a[1] * (x[:, 1] - b[1])
    c[:, 5] = -a[0] * (x[:, 0] - b[0]) + 3 * a[2] * (x[:, 2] - b[2])
    return c


def solve_tau(x, a, b):
    """tau per point: argmin of R_p over arccos(real roots in [-1, 1]) / pi, then 0 and 1."""
    n = x.shape[0]
    coef = _coefficients(x, a, b)
    roots = np.full((n, 5), np.nan + 0j)
    regular = (coef[:, 0] != 0) & (coef[:, 5] != 0)
    if regular.any():
        p = coef[regular]
        comp = np.zeros((p.shape[0], 5, 5))
        comp[:, 0, :] = -p[:, 1:] / p[:, 0:1]
        comp[:, np.arange(1, 5), np.arange(0, 4)] = 1.0
        roots[regular] = np.linalg.eigvals(comp)
    for i in np.nonzero(~regular)[0]:
        r = np.roots(coef[i])
        roots[i, :len(r)] = r
    ok = ~np.isnan(roots.real) & ~(np.abs(roots.imag) > 0) & ~(np.abs(roots) > 1)
    beta = np.where(ok, roots.real, 0.0)
    cand = np.concatenate([np.arccos(np.clip(beta, -1, 1)) / np.pi, np.zeros((n, 1)),
                           np.ones((n, 1))], axis=1)
    R = _residual(x, a, b, cand)
    R[:, :5][~ok] = np.inf
    return cand[np.arange(n), np.argmin(R, axis=1)]


def initial_guess(psi):
    """``get_fit_1D_open_manifold_3D_param.py:29-86``."""
    nS = psi.shape[0]
    X, Z, Y = psi[:, 0], psi[:, 2], psi[:, 1]
    s = [np.sum(X ** k) for k in range(7)]
    A = np.array([[s[6], s[5], s[4], s[3]], [s[5], s[4], s[3], s[2]],
                  [s[4], s[3], s[2], s[1]], [s[3], s[2], s[1], nS]])
    rhs = np.array([np.dot(X ** 3, Z), np.dot(X ** 2, Z), np.dot(X, Z), np.sum(Z)])
    Dc, E, F, G = np.linalg.lstsq(A, rhs, rcond=None)[0]
    disc = max(E * E - 3 * Dc * F, 0.0)
    if np.absolute(Dc) < 1e-8:
        Dc = 1e-8
    a1 = (2.0 * np.sqrt(disc)) / (3.0 * Dc)
    a3 = (2.0 * disc ** 1.5) / (27.0 * Dc * Dc)
    b1 = -E / (3 * Dc)
    b3 = (2.0 * E ** 3) / (27.0 * Dc * Dc) - (E * F) / (3 * Dc) + G
    XXb = X * (X - 2 * b1)
    A2 = np.array([[np.sum(XXb * XXb), np.sum(XXb)], [np.sum(XXb), nS]])
    Ac, C = np.linalg.lstsq(A2, np.array([np.dot(XXb, Y), np.sum(Y)]), rcond=None)[0]
    a2 = 2.0 * Ac * disc / (9.0 * Dc * Dc)
    b2 = C + (Ac * E * E) / (9.0 * Dc * Dc) - (2.0 * Ac * F) / (3.0 * Dc)
    return np.array([a1, a2, a3]), np.array([b1, b2, b3])


def _lstsq_curve(s, nS):
    """(a, b) from the tau sums, ``fit_1D_open_manifold_3D.py:85-100``: for each coordinate the
    2x2 normal equations [[A11, A12], [A12, nS]] (a, b) = (b1, b2).  Where a system is well
    conditioned its least-squares solution is its solution, taken here by Cramer's rule in
    plain floats (it agrees with lstsq to ~1e-10 relative; numpy's per-call overhead on 2x2
    arrays cost more than the iteration's other host work); an ill-conditioned or singular one
    keeps the reference's lstsq (its cutoff picks the minimum-norm solution)."""
    a, b = np.zeros(3), np.zeros(3)
    n = float(nS)
    for q in range(3):
        p11, p12, r1, r2 = float(s[q]), float(s[3 + q]), float(s[6 + q]), float(s[9 + q])
        # singular values of the symmetric system: |mid -+ rad|
        mid, rad = 0.5 * (p11 + n), math.hypot(0.5 * (p11 - n), p12)
        lo, hi = sorted((abs(mid - rad), abs(mid + rad)))
        if lo > 1e-6 * hi:
            det = p11 * n - p12 * p12
            a[q] = (r1 * n - p12 * r2) / det
            b[q] = (p11 * r2 - p12 * r1) / det
        else:
            a[q], b[q] = np.linalg.lstsq(np.array([[p11, p12], [p12, n]]), np.array([r1, r2]),
                                         rcond=None)[0]
    return a, b

def fit(psi, max_iter=100, delta_a_max=1, delta_b_max=1, delta_tau_max=0.01, device=None):
    """(a, b, tau (nS, 1)).  Like the reference, the tau-change test is identically zero
    (``fit_1D_open_manifold_3D.py:218`` binds tau_old to the array it then updates)."""
    x = np.ascontiguousarray(np.asarray(psi)[:, 0:3], dtype=np.float64)
    a, b = initial_guess(psi)
    nS = x.shape[0]
    dev = device if device is not None else torch.device("cuda")
    xd = torch.as_tensor(x, device=dev)
    tau_d = torch.empty(nS, dtype=torch.float64, device=dev)
    flag_d = torch.empty(nS, dtype=torch.int32, device=dev)
    sums_d = torch.empty(13, dtype=torch.float64, device=dev)
    ws = torch.empty(nat.fit1d_ws_len(nS), dtype=torch.float64, device=dev)
    jpi = np.pi * np.array([[1, 2, 3]])

    def update(a, b):
        """tau for (a, b) and its sums; tau on the host only when points were re-solved."""
        nat.fit1d_tau(xd, a, b, tau_d, flag_d, sums_d, ws)
        s = sums_d.cpu().numpy()
        if s[12] == 0:
            return s, None
        tau = tau_d.cpu().numpy()
        fl = flag_d.cpu().numpy().astype(bool)
        tau[fl] = solve_tau(x[fl], a, b)
        cj = np.cos(np.dot(tau[:, None], jpi))
        s = np.concatenate([np.sum(cj ** 2, axis=0), np.sum(cj, axis=0),
                            np.sum(x * cj, axis=0), np.sum(x, axis=0), [fl.sum()]])
        return s, tau

    s, tau_h = update(a, b)
    for _ in range(max_iter):
        a_old, b_old = a, b
        a, b = _lstsq_curve(s, nS)
        s, tau_h = update(a, b)
        da = np.max(np.fabs(a - a_old) / (np.fabs(a) + EPS)) * 100
        db = np.max(np.fabs(b - b_old) / (np.fabs(b) + EPS)) * 100
        if da < delta_a_max and db < delta_b_max and 0.0 < delta_tau_max:
            break
    tau = tau_h if tau_h is not None else tau_d.cpu().numpy()
    return a, b, tau.reshape(nS, 1)

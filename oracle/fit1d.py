"""Oracle for the 1-D open-manifold fit x_j = a_j cos(j pi tau) + b_j (float64).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Restates ``fit_1D_open_manifold_3D.py:64-160``, ``get_fit_1D_open_manifold_3D_param.py:12-95``,
``solve_d_R_d_tau_p_3D.py:35-60`` and ``R_p.py:41-49``, with the module-global state of
``a.py`` replaced by local variables.  Per point, tau is the arccos/pi of a real root in
[-1, 1] of the quintic dR/dbeta (``np.roots``; roots with any imaginary part are dropped),
or 0 or 1, whichever minimises the residual.

Reference quirk kept: the tau-change stopping test is always 0 (``tau_old`` aliases ``tau``).

Pinned by (``tests/test_oracle_nlsa.py``): points lying exactly on a_j cos(j pi tau) + b_j
recover tau; the residual at the returned tau is <= the residual at 0, 1 and a dense grid.
"""
from __future__ import annotations

import numpy as np

EPS = 1e-4


def residual(x_p, a, b, tau):
    """``R_p.py:41-49``: squared residual of point x_p at each candidate tau (column)."""
    jj = np.arange(1, 4)
    err = x_p - b - a * np.cos(np.asarray(tau).reshape(-1, 1) * jj * np.pi)
    return np.sum(err ** 2, axis=1)


def solve_tau(x_p, a, b):
    """``solve_d_R_d_tau_p_3D.py:35-60``."""
    coef = np.array([48 * a[2] ** 2,
                     0,
                     8 * a[1] ** 2 - 48 * a[2] ** 2,
                     -12 * a[2] * (x_p[2] - b[2]),
                     a[0] ** 2 - 4 * a[1] ** 2 + 9 * a[2] ** 2 - 4 * a[1] * (x_p[1] - b[1]),
                     -a[0] * (x_p[0] - b[0]) + 3 * a[2] * (x_p[2] - b[2])])
    beta = np.roots(coef)
    beta = beta[~(np.abs(np.imag(beta)) > 0)]
    beta = np.real(beta[~(np.abs(beta) > 1)])
    cand = np.vstack((np.arccos(beta.reshape(-1, 1)) / np.pi, 0, 1))
    return cand[np.argmin(residual(x_p, a, b, cand))], beta


def initial_guess(psi):
    """``get_fit_1D_open_manifold_3D_param.py:12-95``: a, b from cubic/quadratic fits."""
    nS = psi.shape[0]
    X, Z = psi[:, 0], psi[:, 2]
    X2 = X * X
    X3 = X2 * X
    s = [np.sum(X ** p) for p in range(7)]
    A = np.array([[s[6], s[5], s[4], s[3]],
                  [s[5], s[4], s[3], s[2]],
                  [s[4], s[3], s[2], s[1]],
                  [s[3], s[2], s[1], nS]])
    rhs = np.array([np.dot(X3, Z), np.dot(X2, Z), np.dot(X, Z), np.sum(Z)])
    Dc, E, F, G = np.linalg.lstsq(A, rhs, rcond=None)[0]
    disc = E * E - 3 * Dc * F
    if disc < 0:
        disc = 0.0
    if np.absolute(Dc) < 1e-8:
        Dc = 1e-8
    a1 = (2.0 * np.sqrt(disc)) / (3.0 * Dc)
    a3 = (2.0 * disc ** (3 / 2.0)) / (27.0 * Dc * Dc)
    b1 = -E / (3 * Dc)
    b3 = (2.0 * E * E * E) / (27.0 * Dc * Dc) - (E * F) / (3 * Dc) + G
    Xb = X - 2 * b1
    Y = psi[:, 1]
    XXb = X * Xb
    A = np.array([[np.sum(XXb * XXb), np.sum(XXb)], [np.sum(XXb), nS]])
    rhs = np.array([np.dot(XXb, Y), np.sum(Y)])
    Ac, C = np.linalg.lstsq(A, rhs, rcond=None)[0]
    a2 = 2.0 * Ac * disc / (9.0 * Dc * Dc)
    b2 = C + (Ac * E * E) / (9.0 * Dc * Dc) - (2.0 * Ac * F) / (3.0 * Dc)
    return np.array([a1, a2, a3]), np.array([b1, b2, b3])


def op(psi, max_iter=100, delta_a_max=1, delta_b_max=1, delta_tau_max=0.01):
    """``fit_1D_open_manifold_3D.op`` -> (a, b, tau (nS, 1))."""
    x = psi[:, 0:3]
    a, b = initial_guess(psi)
    nS = x.shape[0]
    tau = np.zeros((nS, 1))
    for p in range(nS):
        tau[p], _ = solve_tau(x[p], a, b)
    for _ in range(1, max_iter + 1):
        a_old, b_old = a, b
        cj = np.cos(np.dot(tau, np.pi * np.array([[1, 2, 3]])))
        A11 = np.sum(cj ** 2, axis=0)
        A12 = np.sum(cj, axis=0)
        b1 = np.sum(x * cj, axis=0)
        b2 = np.sum(x, axis=0)
        coeff = np.zeros((2, 3))
        for qq in range(3):
            A = np.array([[A11[qq], A12[qq]], [A12[qq], nS]])
            coeff[:, qq] = np.linalg.lstsq(A, np.array([b1[qq], b2[qq]]), rcond=None)[0]
        a, b = coeff[0, :], coeff[1, :]
        # fit_1D_open_manifold_3D.py:218-222 binds tau_old to the same array it then fills
        # in place, so the reference's tau change is identically zero.
        tau_old = tau
        for p in range(nS):
            tau[p], _ = solve_tau(x[p], a, b)
        da = np.max(np.fabs(a - a_old) / (np.fabs(a) + EPS)) * 100
        db = np.max(np.fabs(b - b_old) / (np.fabs(b) + EPS)) * 100
        dt = np.max(np.fabs(tau - tau_old))
        if da < delta_a_max and db < delta_b_max and dt < delta_tau_max:
            break
    return a, b, tau

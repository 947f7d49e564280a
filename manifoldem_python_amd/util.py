"""The alignment helpers of ``util.py`` the tessellation needs: RELION Euler angles to
quaternions (``util.py:76-92`` with ``read_alignfile.py:93-94``) and the conjugate
augmentation (``util.py:94-106``)."""
from __future__ import annotations

import numpy as np

from .synth import qmult


def eul_to_quat(phi, theta, psi, flip=True):
    """(qz, qy, qzs) of RELION's (rot, tilt, psi) in radians, 4 x n each."""
    phi, theta, psi = (np.asarray(a, dtype=np.float64) for a in (phi, theta, psi))
    if phi.size == 0 or theta.size == 0 or psi.size == 0:
        raise ValueError("eul_to_quat: some Euler angles are missing")
    z = np.zeros(phi.shape[0])
    qz = np.vstack((np.cos(phi / 2), z, z, -np.sin(phi / 2)))
    qy = np.vstack((np.cos(theta / 2), z, -np.sin(theta / 2), z))
    sp = np.sin(psi / 2)
    if flip:
        sp = -sp
    qzs = np.vstack((np.cos(psi / 2), z, z, sp))
    return qz, qy, qzs


def euler_to_q(phi, theta, psi, flip=True):
    """q = qzs (qy qz) (``read_alignfile.py:93-94``), 4 x n."""
    qz, qy, qzs = eul_to_quat(phi, theta, psi, flip)
    return qmult(qzs, qmult(qy, qz))


def augment(q):
    """Append the conjugate orientation of every particle, [-q1, q0, -q3, q2]: its
    projection direction is the antipode, and the tessellation bins both copies."""
    q = np.asarray(q)
    if q.shape[0] <= 3:
        raise ValueError("augment: q must be 4 x n")
    qc = np.vstack((-q[1, :], q[0, :], -q[3, :], q[2, :]))
    return np.hstack((q, qc))

"""Drop-in for ``distribute3Sphere`` (``distribute3Sphere.py:19-49``): numPts points spread
over the unit sphere on rings of polar angle w1 = (k + 1/2) delta with azimuths
(j + 1/2) delta / sin w1, delta rescaled by sqrt(count / numPts) until the count is numPts
(at most 100 iterations).  Host code (a few thousand points, once per project); each ring is
one ``np.arange`` of the same arguments as the reference's loop, so the points and the count
per iteration are the reference's."""
from __future__ import annotations

import math

import numpy as np


def _rings(delta):
    out = []
    for w1 in np.arange(0.5 * delta, np.pi, delta):
        cw1, sw1 = math.cos(w1), math.sin(w1)
        dw2 = delta / sw1
        w2 = np.arange(0.5 * dw2, 2 * np.pi, dw2)
        out.append(np.stack((np.full(w2.shape, cw1), sw1 * np.cos(w2), sw1 * np.sin(w2)), 1))
    return out


def op(numPts):
    K = int(numPts)
    delta = math.exp(math.log(4 * np.pi / K) / 2.0)
    # one buffer of 2K rows for all iterations, as in the reference: if the last iteration
    # places fewer than K points, rows beyond them keep an earlier iteration's values
    results = np.zeros((2 * K, 3))
    it, count = 0, 0
    while count != K and it < 100:
        it += 1
        rings = _rings(delta)
        count = sum(r.shape[0] for r in rings)
        if count > 2 * K:          # the reference's buffer overflows here
            raise IndexError(f"distribute3Sphere: {count} points for a buffer of {2 * K}")
        if count:
            results[:count] = np.concatenate(rings)
        delta = delta * math.exp(math.log(float(count) / K) / 2.0)
    return results[0:K, :], it

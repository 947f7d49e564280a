"""Drop-in for the Stage-3 driver ``psiAnalysis`` (``psiAnalysis.py:29-117``): NLSA of every
(PD, psi) pair without a progress marker, on the GPU."""
from __future__ import annotations

import os

import numpy as np

from . import p, psiAnalysisParS2


def fileCheck(N):
    fin = np.zeros((N, p.num_psis), dtype=int)
    for _, _, files in os.walk(p.psi2_prog):
        for f in sorted(files):
            if not f.startswith("."):
                a, b = f.split("_")
                fin[int(a), int(b)] = 1
    return fin


def divid(N, rc, fin):
    ll = []
    for prD in range(N):
        psinums = rc["psiNumsAll"][prD, :]
        senses = rc["sensesAll"][prD, :]
        psi_list = [k for k in range(len(psinums)) if fin[prD, k] != 1]
        ll.append(["{}prD_{}".format(p.dist_file, prD), "{}prD_{}".format(p.psi_file, prD),
                   "{}prD_{}".format(p.psi2_file, prD), "{}prD_{}".format(p.EL_file, prD),
                   psinums, senses, prD, psi_list])
    return ll


def op(*argv):
    rc = {"psiNumsAll": np.tile(np.arange(p.num_psis), (p.numberofJobs, 1)),
          "sensesAll": np.tile(np.ones(p.num_psis), (p.numberofJobs, 1))}
    jobs = divid(p.numberofJobs, rc, fileCheck(p.numberofJobs))
    for job in jobs:
        if job[7]:
            psiAnalysisParS2.op(job, p.conOrderRange, p.trajName, 0, p.num_psiTrunc)

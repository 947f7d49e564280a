"""Drop-in for the second NLSA pass ``EL1D`` (``EL1D.py:26-101``): for every PD not removed by
the user and without an ``ELConc`` progress marker, ``psiAnalysisParS2.op`` with isFull = 1 on
the PD's chosen reaction coordinate and sense (the ``CC_file`` pickle's psinums / senses), on
the GPU.  The reference then computes the 1-D energy landscape (``ComputeEnergy1D``), which
is outside this package."""
from __future__ import annotations

import os

import numpy as np

from . import myio, p, psiAnalysisParS2


def fileCheck():
    fin = []
    for _, _, files in os.walk(p.EL_prog):
        for f in sorted(files):
            if not f.startswith("."):
                fin.append(int(f))
    return fin


def divide1(R, psiNumsAll, sensesAll):
    """``:34-47``."""
    fin = fileCheck()
    ll = []
    for prD in R:
        if prD not in fin:
            ll.append(["{}prD_{}".format(p.dist_file, prD), "{}prD_{}".format(p.psi_file, prD),
                       "{}prD_{}".format(p.psi2_file, prD), "{}prD_{}".format(p.EL_file, prD),
                       [psiNumsAll[0, prD]], [sensesAll[0, prD]], prD])
    return ll


def count1(R):
    return len(R) - len(fileCheck())


def op(*argv):
    R = np.arange(p.numberofJobs)
    trash = np.asarray(p.trash_list)
    if trash.size:
        R = np.delete(R, np.nonzero(trash == 1)[0])
    data = myio.fin1(p.CC_file)
    jobs = divide1(R, np.atleast_2d(data["psinums"]), np.atleast_2d(data["senses"]))
    done = len(R) - len(jobs)
    for job in jobs:
        psiAnalysisParS2.op(job, p.conOrderRange, p.trajName, 1, p.num_psiTrunc)
        done += 1
        if argv:
            argv[0].emit(int(done / float(len(R)) * 100))

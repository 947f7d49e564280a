"""Drop-in for the Stage-2 driver ``manifoldAnalysis`` (``manifoldAnalysis.py:34-116``): runs
manifoldTrimmingAuto on every PD without a progress marker, on the GPU."""
from __future__ import annotations

import os

from . import manifoldTrimmingAuto, p


def fileCheck():
    fin = []
    for _, _, files in os.walk(p.psi_prog):
        for f in sorted(files):
            if not f.startswith("."):
                fin.append(int(f))
    return fin


def op(*argv):
    fin = fileCheck()
    jobs = []
    for prD in range(p.numberofJobs):
        if prD in fin:
            continue
        jobs.append(["{}prD_{}".format(p.dist_file, prD), "{}prD_{}".format(p.psi_file, prD),
                     "{}prD_{}_eig_spec.txt".format(p.psi_file, prD), prD])
    for i, job in enumerate(jobs):
        manifoldTrimmingAuto.op(job, 0, p.tune, p.rad, False, {"Is": True})
        if argv:
            argv[0].emit(int(((p.numberofJobs - len(jobs) + i + 1) / float(p.numberofJobs)) * 100))

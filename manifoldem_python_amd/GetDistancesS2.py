"""Drop-in for the Stage-1 driver ``GetDistancesS2`` (``GetDistancesS2.py:29-124``): builds the
per-PD jobs from the tessellation pickle, skips PDs that already have a progress marker, and
runs every remaining PD on the GPU in this process (the reference's process pool and MPI
fan-out are replaced by the GPU; a multi-GPU run would give each GPU its own PD list)."""
from __future__ import annotations

import os

import numpy as np

from . import getDistanceCTF_local_Conj9combinedS2 as gd
from . import myio, p


def fileCheck():
    fin = []
    for _, _, files in os.walk(p.dist_prog):
        for f in sorted(files):
            if not f.startswith("."):
                fin.append(int(f))
    return fin


def divide(CG, q, df, N):
    ll = []
    fin = fileCheck()
    for prD in range(N):
        ind = CG[prD]
        if prD not in fin:
            ll.append([ind, q[:, ind], df[ind], "{}prD_{}".format(p.dist_file, prD), prD])
    return ll


def op(*argv):
    data = myio.fin1(p.tess_file)
    CG, df, q = data["CG"], data["df"], data["q"]
    sh = data.get("sh")
    filterPar = dict(type="Butter", Qc=0.5, N=8)
    options = dict(verbose=False, avgOnly=False, visual=False, parallel=False,
                   relion_data=p.relion_data, thres=p.PDsizeThH)
    if p.relion_data is False:
        p.nPix = int(np.sqrt(os.path.getsize(p.img_stack_file) / (4 * p.num_part)))
    jobs = divide(CG, q, df, p.numberofJobs)
    print("Processing {} projection directions.".format(len(jobs)))
    for i, job in enumerate(jobs):
        gd.op(job, filterPar, p.img_stack_file, sh, len(df), options)
        if argv:
            argv[0].emit(int(((p.numberofJobs - len(jobs) + i + 1) / float(p.numberofJobs)) * 100))

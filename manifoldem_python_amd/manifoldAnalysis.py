"""Drop-in for the Stage-2 driver ``manifoldAnalysis`` (``manifoldAnalysis.py:34-116``): the
trimmed diffusion map of every PD without a progress marker, on the GPU.

Where the reference runs ``manifoldTrimmingAuto.op`` once per PD (a loop, or a process pool),
this driver embeds the pending PDs in batches: every trimming round of a batch is one grouped
set of launches over all its PDs (``trimming.trim_batch``).  A batch closes once it holds
``p.pd_batch_rows`` particles or more.  Each PD's outputs are the reference's: the ``gC_trimmed_psi``
pickle, the progress marker and ``topos/PrD_<n>/eig_spec.txt``."""
from __future__ import annotations

import os

import numpy as np

from . import manifoldTrimmingAuto, myio, p


def fileCheck():
    fin = []
    for _, _, files in os.walk(p.psi_prog):
        for f in sorted(files):
            if not f.startswith("."):
                fin.append(int(f))
    return fin


def count(N):
    return N - len(fileCheck())


def divide(N):
    """``:46-55``: [dist_file, psi_file, eig_file, prD] of every PD without a marker."""
    fin = fileCheck()
    ll = []
    for prD in range(N):
        if prD not in fin:
            ll.append(["{}prD_{}".format(p.dist_file, prD), "{}prD_{}".format(p.psi_file, prD),
                       "{}/topos/PrD_{}/eig_spec.txt".format(p.out_dir, prD + 1), prD])
    return ll


def _flush(batch, doSave):
    Ds = [b[1] for b in batch]
    out = manifoldTrimmingAuto.embed_trimmed_batch(Ds, p.tune, p.rad)
    for (job, _, ind), (posPath, res) in zip(batch, out):
        manifoldTrimmingAuto.write_outputs(job[1], job[2], job[3], posPath, ind, res, doSave)


def op(*argv):
    jobs = divide(p.numberofJobs)
    for i in range(p.numberofJobs):
        os.makedirs(os.path.join(p.out_dir, "topos", "PrD_{}".format(i + 1)), exist_ok=True)
    doSave = dict(outputFile="", Is=True)
    cap = max(int(p.pd_batch_rows), 1)
    done = p.numberofJobs - len(jobs)
    batch, rows = [], 0
    for j, job in enumerate(jobs):
        d = myio.fin1(job[0])          # only D and ind are kept from the stage-1 pickle
        D = np.asarray(d["D"], dtype=np.float64)
        batch.append((job, D, d["ind"]))
        rows += D.shape[0]
        del d
        if j == len(jobs) - 1 or rows >= cap:
            _flush(batch, doSave)
            done += len(batch)
            batch, rows = [], 0
            if argv:
                argv[0].emit(int(done / float(p.numberofJobs) * 100))


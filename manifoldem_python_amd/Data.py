"""Drop-in for the alignment-import stage ``Data.op`` (``Data.py:72-125``) on RELION input:
read the STAR file, double the particles by conjugate augmentation, tessellate S^2 on the GPU
and write the ``selecGCs`` pickle (keys CG1, CG, nG, q, df, S2, S20, sh, NC) that the
distance stage reads.  SPIDER alignment documents, the angle files for the GUI and the
CC-graph step after it are not part of this package."""
from __future__ import annotations

import os

import numpy as np

from . import S2tessellation, myio, p, star
from .util import augment

TESS_KEYS = ["CG1", "CG", "nG", "q", "df", "S2", "S20", "sh", "NC"]


def tessellate(q, df, sh, ang_width, PDsizeThL, PDsizeThH, tess_file):
    """Augment, tessellate and write tess_file; returns the tessellation tuple.  q (4 x n),
    df (n,) and sh are the particles as read (before augmentation)."""
    q = augment(q)
    df = np.concatenate((df, df))
    CG1, CG, nG, S2, S20_th, S20, NC = S2tessellation.op(q, ang_width, PDsizeThL, False,
                                                         PDsizeThH)
    myio.fout1(tess_file, TESS_KEYS, [CG1, CG, nG, q, df, S2, S20_th, sh, NC])
    return CG1, CG, nG, S2, S20_th, S20, NC


def op(align_param_file):
    if not p.relion_data:
        raise NotImplementedError("Data.op reads RELION STAR alignments (p.relion_data); "
                                  "SPIDER alignment documents are not read by this package")
    sh, q, U, V, micro = star.get_from_relion(align_param_file, flip=True)
    p.EkV, p.Cs, p.AmpContrast = micro["EkV"], micro["Cs"], micro["AmpContrast"]
    df = (U + V) / 2
    if not p.tess_file:
        p.tess_file = os.path.join(p.user_dir, "outputs_{}/selecGCs".format(p.proj_name))
    res = tessellate(q, df, sh, p.ang_width, p.PDsizeThL, p.PDsizeThH, p.tess_file)
    p.numberofJobs = len(res[1])
    return None

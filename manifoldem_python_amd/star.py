"""RELION STAR tables (the alignment input of ``read_alignfile.get_from_relion``,
``read_alignfile.py:20-96``), read with plain Python.

A STAR file holds data blocks (``data_<name>``); a block's table starts with ``loop_``, then
one ``_rlnName #k`` header line per column, then whitespace-separated rows up to a blank line,
the next block or the end.  RELION 3.1 files begin with a ``data_optics`` block (microscope
parameters per optics group) before the particle table; older files hold the microscope
columns in the particle table itself.
"""
from __future__ import annotations

import numpy as np


def read_tables(path):
    """Every loop table of the file in order: list of (block name, {column: list of str})."""
    tables = []
    block = None
    cols, rows, in_loop = [], [], False

    def flush():
        nonlocal cols, rows, in_loop
        if cols:
            tables.append((block, {c: [r[i] if i < len(r) else "" for r in rows]
                                   for i, c in enumerate(cols)}))
        cols, rows, in_loop = [], [], False

    with open(path) as f:
        for raw in f:
            line = raw.strip()
            if line.startswith("data_"):
                flush()
                block = line.split()[0]
                continue
            if line.startswith("loop_"):
                flush()
                in_loop = True
                continue
            if not in_loop:
                continue
            if line.startswith("_rln"):
                if rows:                       # a header after rows: a new table
                    flush()
                    in_loop = True
                cols.append(line.split()[0].lstrip("_"))
                continue
            if not line or line.startswith("#"):
                if rows:
                    flush()
                continue
            if cols:
                rows.append(line.split())
    flush()
    return tables


def column(table, name, dtype=np.float64):
    return np.asarray(table[name], dtype=dtype)


def get_from_relion(path, flip=True):
    """``read_alignfile.get_from_relion``: (sh, q, U, V, micro) where sh = (rlnOriginX,
    rlnOriginY) (zeros when absent), q the particles' quaternions 4 x n (q = qzs (qy qz) of
    rot, tilt, psi), U / V the defoci and micro = {EkV, Cs, AmpContrast} from the optics table
    (RELION 3.1) or the particle table."""
    from .util import euler_to_q
    tables = read_tables(path)
    if not tables:
        raise ValueError(f"{path}: no STAR loop table")
    names = [b for b, _ in tables]
    if "data_optics" in names:
        optics = tables[names.index("data_optics")][1]
        parts = [t for b, t in tables if b != "data_optics" and "rlnAngleRot" in t]
        if not parts:
            raise ValueError(f"{path}: no particle table after data_optics")
        part, micro_src = parts[0], optics
    else:
        part = next((t for _, t in tables if "rlnAngleRot" in t), tables[0][1])
        micro_src = part
    for c in ("rlnDefocusU", "rlnDefocusV"):
        if c not in part:
            raise ValueError(f"{path}: missing defocus ({c})")
    for c in ("rlnAngleRot", "rlnAngleTilt", "rlnAnglePsi"):
        if c not in part:
            raise ValueError(f"{path}: missing Euler angles ({c})")
    try:
        micro = dict(EkV=float(micro_src["rlnVoltage"][0]),
                     Cs=float(micro_src["rlnSphericalAberration"][0]),
                     AmpContrast=float(micro_src["rlnAmplitudeContrast"][0]))
    except KeyError as e:
        raise ValueError(f"{path}: missing microscope parameters ({e})") from None
    U, V = column(part, "rlnDefocusU"), column(part, "rlnDefocusV")
    if "rlnOriginX" in part and "rlnOriginY" in part:
        sh = (column(part, "rlnOriginX"), column(part, "rlnOriginY"))
    else:
        z = U * 0.0
        sh = (z, z)
    q = euler_to_q(np.deg2rad(column(part, "rlnAngleRot")),
                   np.deg2rad(column(part, "rlnAngleTilt")),
                   np.deg2rad(column(part, "rlnAnglePsi")), flip)
    return sh, q, U, V, micro

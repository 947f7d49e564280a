"""Module-global parameters read by the per-PD path.

Same names and defaults as the reference's ``p.py:13-114`` (with the GUI defaults of
``ManifoldEM_GUI.py:112-123``), limited to what the path reads, plus the GPU options at the end
(no counterpart in the reference).
"""
import numpy as np

# microscope / data (set by the data-import stage in the reference)
relion_data = False
nPix = None
pix_size = None
Cs = None
EkV = None
AmpContrast = None
gaussEnv = np.inf
img_stack_file = ""
mask_vol_file = ""
num_part = 0

# computational resources
ncpu = 1
machinefile = False
eps = 1e-10

# tessellation
ang_width = 0.0               # bin width on S^2 in radians (set by the caller, as the GUI does)
PDsizeThL = 100
PDsizeThH = 2000
numberofJobs = 0

# eigenfunction parameters
num_eigs = 15
num_psiTrunc = 8
num_psis = 8
tune = 3
rad = 5
conOrderRange = 50
nClass = 50
trajName = "1"

# file layout (filled by create_dir in the reference)
user_dir = ""
proj_name = ""
tess_file = ""
dist_file = ""
dist_prog = ""
psi_file = ""
psi_prog = ""
psi2_file = ""
psi2_prog = ""
EL_file = ""
EL_prog = ""
out_dir = ""
CC_file = ""
trash_list = np.zeros(0, dtype=int)     # 1 = PD excluded by the user (EL1D.py:61)

# --- GPU options -------------------------------------------------------------------------
precision = "split_bf16"      # "split_bf16" or "fp32" distance Gram
kappa = None                  # cancellation ratio above which entries are recomputed directly
pd_batch_rows = 262144        # particles per grouped launch in the batch driver


def init():
    """Reset the defaults above (the reference's p.init sets the same values)."""
    global num_eigs, num_psiTrunc, num_psis, tune, rad, conOrderRange, nClass, gaussEnv, eps
    num_eigs, num_psiTrunc, num_psis, tune, rad = 15, 8, 8, 3, 5
    conOrderRange, nClass, gaussEnv, eps = 50, 50, np.inf, 1e-10
    return None


def create_dir():
    """Output directories and file stems of the reference's ``p.create_dir`` (``p.py:116-225``)
    for the stages this package writes."""
    import os
    global dist_dir, dist_prog, psi_dir, psi_prog, psi2_dir, psi2_prog, EL_dir, EL_prog
    global dist_file, psi_file, psi2_file, EL_file, tess_file, out_dir, CC_file
    out = os.path.join(user_dir, "outputs_{}".format(proj_name))
    out_dir = out + "/"
    os.makedirs(os.path.join(out, "topos"), exist_ok=True)
    os.makedirs(os.path.join(out, "CC"), exist_ok=True)
    dist_dir = os.path.join(out, "distances/")
    psi_dir = os.path.join(out, "diff_maps/")
    psi2_dir = os.path.join(out, "psi_analysis/")
    EL_dir = os.path.join(out, "ELConc{}/".format(conOrderRange))
    dist_prog = os.path.join(dist_dir, "progress/")
    psi_prog = os.path.join(psi_dir, "progress/")
    psi2_prog = os.path.join(psi2_dir, "progress/")
    EL_prog = os.path.join(EL_dir, "progress/")
    for d in (dist_prog, psi_prog, psi2_prog, EL_prog):
        os.makedirs(d, exist_ok=True)
    tess_file = os.path.join(out, "selecGCs")
    dist_file = "{}/IMGs_".format(dist_dir)
    psi_file = "{}/gC_trimmed_psi_".format(psi_dir)
    psi2_file = "{}/S2_".format(psi2_dir)
    EL_file = "{}/S2_".format(EL_dir)
    CC_file = os.path.join(out, "CC", "CC_file")
    return None

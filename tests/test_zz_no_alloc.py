"""memi_prep_images_noalloc (tables in the library's static device scratch) against
memi_prep_images (tables in a caller workspace): identical aligned images, spectra, CTFs and
operand rows on two calls in a row, the second with fewer, smaller images (the scratch keeps the
first call's larger tables, which the second call must rewrite before use)."""
import math

import numpy as np
import pytest

from conftest import FILTER, PARAMS

pytestmark = pytest.mark.gpu


def _inputs(cuda, n, N, seed):
    import torch
    from manifoldem_python_amd import pipeline as pl, synth
    s = synth.make_pd(synth.seeds(seed + 1)[seed], n, N, conj_frac=0.25)
    conj = s.ind >= s.nStot // 2
    src = np.where(conj, s.ind - s.nStot // 2, s.ind)
    geo = pl.pd_geometry(s.q)
    return (torch.as_tensor(s.raw, device=cuda),
            torch.as_tensor(src.astype(np.int64), device=cuda),
            torch.as_tensor(conj.astype(np.int32), device=cuda),
            torch.as_tensor(-(180 / math.pi) * geo["Psi"], device=cuda),
            torch.full((n,), -geo["psi_p"], dtype=torch.float64, device=cuda),
            torch.as_tensor(s.df, device=cuda))


def _outputs(cuda, n, N, precision):
    import torch
    from manifoldem_python_amd import pipeline as pl
    op = pl.Operand(n, N, precision, cuda)
    op.hi.zero_()
    if op.lo is not None:
        op.lo.zero_()
    fy = torch.zeros((n, N, N // 2 + 1), dtype=torch.complex64, device=cuda)
    ch = torch.zeros((n, pl.half_len(N)), dtype=torch.float32, device=cuda)
    img = torch.zeros((n, N, N), dtype=torch.float32, device=cuda)
    return op, fy, ch, img


@pytest.mark.parametrize("precision", [0, 1])
def test_noalloc_prep_matches_workspace_entry(cuda, precision):
    import torch
    from manifoldem_python_amd import _native as nat, pipeline as pl
    ctf = nat.ctf_params(PARAMS["pix_size"], PARAMS["Cs"], PARAMS["EkV"], PARAMS["AmpContrast"],
                         PARAMS["gaussEnv"])
    kind = pl.filter_kind(FILTER)
    for n, N, seed in ((48, 160, 3), (20, 64, 4)):          # the second call is smaller
        args = _inputs(cuda, n, N, seed)
        ref = _outputs(cuda, n, N, precision)
        got = _outputs(cuda, n, N, precision)
        ws = torch.empty(6 * pl.half_len(N), dtype=torch.float64, device=cuda)
        Khp = ref[0].Khp
        nat.prep_images(*args, N, kind, FILTER["Qc"], FILTER["N"], ctf, precision, Khp, 0,
                        ref[0].hi, ref[0].lo if ref[0].lo is not None else ref[0].hi, ref[1],
                        ref[2], ws, img_out=ref[3])
        nat.prep_images_noalloc(*args, N, kind, FILTER["Qc"], FILTER["N"], ctf, precision, Khp, 0,
                                got[0].hi, got[0].lo if got[0].lo is not None else got[0].hi,
                                got[1], got[2], img_out=got[3])
        torch.cuda.synchronize()
        assert torch.equal(got[0].hi, ref[0].hi), (n, N)
        if precision == nat.PREC_SPLIT_BF16:
            assert torch.equal(got[0].lo, ref[0].lo), (n, N)
        assert torch.equal(torch.view_as_real(got[1]), torch.view_as_real(ref[1])), (n, N)
        assert torch.equal(got[2], ref[2]) and torch.equal(got[3], ref[3]), (n, N)
        assert float(ref[3].abs().max()) > 0                 # the comparison saw real output

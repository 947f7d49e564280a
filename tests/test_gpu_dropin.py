"""The reference-named drop-in modules on the GPU against the oracle, and the three stage
drivers end to end on a small synthetic project (pickles + progress markers)."""
import os
from collections import namedtuple

import numpy as np
import pytest

from conftest import FILTER, PARAMS, rel_offdiag

pytestmark = pytest.mark.gpu


def _set_params(N):
    from manifoldem_python_amd import p
    p.init()
    p.nPix, p.pix_size, p.Cs, p.EkV = N, PARAMS["pix_size"], PARAMS["Cs"], PARAMS["EkV"]
    p.AmpContrast, p.gaussEnv = PARAMS["AmpContrast"], np.inf
    p.mask_vol_file = ""
    return p


def test_ferguson_slaplacian_sembedding(cuda):
    from manifoldem_python_amd import fergusonE, slaplacianonFly, sembeddingonFly
    from oracle import dm
    rng = np.random.default_rng(1)
    n = 90
    X = rng.normal(size=(n, 3))
    D = ((X[:, None] - X[None]) ** 2).sum(-1) * 50
    yRow, yCol, yVal = dm.assemble(D.copy(), 40)
    le = np.linspace(-150, 150, 301)
    a0 = np.array([[0.1], [-0.2], [0.3], [0.05]])
    np.random.seed(3)
    popt, lsw, resnorm, R2 = fergusonE.op(np.sqrt(yVal), le, a0)
    ref_lsw, _ = dm.ferguson_sums(np.sqrt(yVal), le)
    np.testing.assert_allclose(lsw, ref_lsw, atol=1e-12)
    Opt = namedtuple("O", "sigma alpha nEigs autotune")
    o = Opt(sigma=np.sqrt(np.median(D)), alpha=1, nEigs=8, autotune=0)
    l, st = slaplacianonFly.op(yVal, yCol, yRow, n, o)
    lref = dm.slaplacian(yVal, yCol, yRow, n, o.sigma)
    np.testing.assert_allclose(l.toarray(), lref.toarray(), rtol=1e-12, atol=1e-15)
    vals, vecs = sembeddingonFly.op(yVal, yCol, yRow, n, o)
    rv, rV = dm.top_eigs(lref, 9, dense=True)
    np.testing.assert_allclose(vals, rv, atol=1e-10)
    for c in range(9):
        assert abs(abs(vecs[:, c] @ rV[:, c]) - 1) < 1e-8


def test_small_helpers(cuda):
    from manifoldem_python_amd import get_wiener, svdRF, L2_distance
    from oracle import nlsa
    rng = np.random.default_rng(2)
    CTF = rng.normal(size=(20, 6, 6))
    pp, ps = np.arange(20), rng.permutation(20)
    W, C1 = get_wiener.op(CTF, pp, ps, 4, 20)
    Wr, C1r = nlsa.get_wiener(CTF, pp, ps, 4, 20)
    np.testing.assert_allclose(W, Wr, rtol=1e-12)
    A = rng.normal(size=(500, 8))
    U, S, V = svdRF.op(A)
    Ur, Sr, Vr = nlsa.svdRF(A)
    np.testing.assert_allclose(S, Sr, rtol=1e-12)
    np.testing.assert_allclose(np.abs(U), np.abs(Ur), atol=1e-10)
    a = rng.normal(size=(30, 12))
    np.testing.assert_allclose(L2_distance.op(a, a), nlsa.l2_distance(a, a), atol=1e-7)


def test_getdistance_op_writes_reference_pickle(cuda, tmp_path):
    from manifoldem_python_amd import getDistanceCTF_local_Conj9combinedS2 as gd, myio, synth
    from oracle import distance as od
    n, N = 60, 40
    p = _set_params(N)
    p.dist_prog = str(tmp_path) + "/"
    s = synth.make_pd(synth.seeds(5)[3], n, N, conj_frac=0.3)
    stack_file = os.path.join(tmp_path, "stack.bin")
    s.raw.astype(np.float32).tofile(stack_file)
    out_file = os.path.join(tmp_path, "IMGs_prD_7")
    options = dict(verbose=False, avgOnly=False, visual=False, parallel=False, relion_data=False,
                   thres=2000)
    gd.op([s.ind, s.q, s.df, out_file, 7], FILTER, stack_file, None, s.nStot, options)
    assert os.path.exists(os.path.join(tmp_path, "7"))
    got = myio.fin1(out_file)
    assert sorted(got) == sorted(gd.KEYS)
    ref = od.op([s.ind, s.q, s.df, None, 7], FILTER, s.raw, s.nStot, options,
                dict(PARAMS, nPix=N))
    assert rel_offdiag(got["D"], ref["D"]).max() < 1e-5
    for k in ("imgAll", "imgAllFlip"):
        assert np.max(np.abs(got[k] - ref[k])) < 1e-5 * np.max(np.abs(ref[k]))
    for k in ("imgAvg", "imgAvgFlip", "imgAllIntensity"):
        assert np.max(np.abs(got[k] - ref[k])) < 1e-5 * np.max(np.abs(ref[k])), k
    np.testing.assert_allclose(got["CTF"], ref["CTF"], atol=1e-6)
    np.testing.assert_array_equal(got["imgLabels"], ref["imgLabels"])
    np.testing.assert_allclose(got["Psis"], ref["Psis"], atol=1e-12)
    np.testing.assert_allclose(got["PD"], ref["PD"], atol=1e-12)


def test_stage_drivers_end_to_end(cuda, tmp_path):
    """Tessellation pickle -> GetDistancesS2 -> manifoldAnalysis -> psiAnalysis, with the
    resume-by-marker rule: a PD with a marker is skipped."""
    from manifoldem_python_amd import (GetDistancesS2, manifoldAnalysis, psiAnalysis, myio,
                                       synth)
    N = 32
    p = _set_params(N)
    p.user_dir, p.proj_name = str(tmp_path), "t"
    p.create_dir()
    p.num_psis, p.conOrderRange, p.num_eigs, p.rad = 2, 20, 8, 50
    sizes = [80, 70, 90]
    raws, CG, qs, dfs, off = [], [], [], [], 0
    for i, n in enumerate(sizes):
        s = synth.make_pd(synth.seeds(9)[i + 1], n, N)
        raws.append(s.raw)
        CG.append(np.arange(off, off + n))
        qs.append(s.q)
        dfs.append(s.df)
        off += n
    stack = np.concatenate(raws).astype(np.float32)
    p.img_stack_file = os.path.join(tmp_path, "stack.bin")
    stack.tofile(p.img_stack_file)
    p.num_part = off
    q = np.concatenate(qs, axis=1)
    df = np.concatenate(dfs)
    # conjugate augmentation doubles the particle count (util.augment): nStot = 2 * num_part
    q = np.hstack([q, q])
    df = np.concatenate([df, df])
    myio.fout1(p.tess_file, ["CG", "q", "df", "sh"], [CG, q, df, None])
    p.numberofJobs = len(sizes)
    open(os.path.join(p.dist_prog, "1"), "a").close()      # PD 1 already done
    GetDistancesS2.op()
    assert not os.path.exists("{}prD_1".format(p.dist_file))
    os.remove(os.path.join(p.dist_prog, "1"))
    GetDistancesS2.op()
    for i in range(3):
        d = myio.fin1("{}prD_{}".format(p.dist_file, i))
        assert d["D"].shape == (sizes[i], sizes[i])
    manifoldAnalysis.op()
    for i in range(3):
        d = myio.fin1("{}prD_{}".format(p.psi_file, i))
        assert d["psi"].shape[1] == 8 and len(d["posPath"]) <= sizes[i]
    psiAnalysis.op()
    for i in range(3):
        for k in range(2):
            d = myio.fin1("{}prD_{}_psi_{}".format(p.psi2_file, i, k))
            assert d["IMG1"].shape[0] == N * N
            assert np.all(np.isfinite(d["tau"]))
            assert os.path.exists(os.path.join(p.psi2_prog, "%d_%d" % (i, k)))

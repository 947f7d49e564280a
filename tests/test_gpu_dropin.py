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
    # wide, square and a short side above the tall-Gram kernels' 8 columns
    for shape in ((6, 300), (12, 12), (400, 20), (30, 900)):
        A = rng.normal(size=shape)
        U, S, V = svdRF.op(A)
        Ur, Sr, Vr = nlsa.svdRF(A)
        assert U.shape == Ur.shape and V.shape == Vr.shape
        np.testing.assert_allclose(S, Sr, rtol=1e-10)
        np.testing.assert_allclose(U @ S @ V.T, A, atol=1e-9)
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


def _norm01(t):
    t = np.ravel(t)
    return (t - t.min()) / (t.max() - t.min())


def _tau_gap(t, r):
    """Largest difference of min-max normalised tau up to direction."""
    t, r = _norm01(t), _norm01(r)
    return min(np.max(np.abs(t - r)), np.max(np.abs(t - (1 - r))))


def test_stage_drivers_end_to_end(cuda, tmp_path):
    """Tessellation pickle -> GetDistancesS2 -> manifoldAnalysis (rad = 5) -> psiAnalysis ->
    EL1D, with the resume-by-marker rule (a PD with a marker is skipped); the trimmed-psi and
    psi-analysis pickles against the oracle, and the isFull pass's tau sense rule."""
    from manifoldem_python_amd import (GetDistancesS2, manifoldAnalysis, psiAnalysis, myio,
                                       synth, EL1D)
    N = 32
    p = _set_params(N)
    p.user_dir, p.proj_name = str(tmp_path), "t"
    p.create_dir()
    p.num_psis, p.conOrderRange, p.num_eigs, p.rad = 2, 20, 8, 5
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
    from oracle import dm, nlsa
    from oracle import fit1d as ofit
    for i in range(3):
        d = myio.fin1("{}prD_{}".format(p.psi_file, i))
        assert d["psi"].shape[1] == 8 and len(d["posPath"]) <= sizes[i]
        # the gC_trimmed_psi pickle against the oracle's trimming of the stage-1 D
        D = myio.fin1("{}prD_{}".format(p.dist_file, i))["D"]
        pos_r, res_r = dm.trim(D, p.tune, p.rad, num_eigs=8, rng=np.random.default_rng(i))
        np.testing.assert_array_equal(d["posPath"], pos_r)
        np.testing.assert_allclose(d["lamb"], res_r[0], atol=1e-6)
        assert d["sigma"] == pytest.approx(res_r[2], rel=1e-5)
        assert os.path.exists(os.path.join(p.out_dir, "topos", "PrD_%d" % (i + 1),
                                           "eig_spec.txt"))
    psiAnalysis.op()
    first = {}
    unstable = 0
    for i in range(3):
        img = myio.fin1("{}prD_{}".format(p.dist_file, i))
        psd = myio.fin1("{}prD_{}".format(p.psi_file, i))
        pos = psd["posPath"]
        nS = len(pos)
        C = nS // p.conOrderRange
        Dp = img["D"][pos][:, pos]
        for k in range(2):
            d = myio.fin1("{}prD_{}_psi_{}".format(p.psi2_file, i, k))
            first[i, k] = d
            assert d["IMG1"].shape[0] == N * N
            assert os.path.exists(os.path.join(p.psi2_prog, "%d_%d" % (i, k)))
            # the S2_prD_psi pickle against the oracle's NLSA of the same psi ordering
            order = np.argsort(psd["psi"][:, k])
            par = dict(num=nS, ConOrder=C, k=nS - C, tune=p.tune, nS=nS, save=False,
                       psiTrunc=p.num_psiTrunc)
            ref = nlsa.op(par, Dp[order][:, order], pos, order, img["imgAll"], img["msk2"],
                          img["CTF"].reshape(-1, N, N), num_eigs=8,
                          rng=np.random.default_rng(k))
            np.testing.assert_allclose(d["sdiag"], ref[4], rtol=1e-4)
            for c in range(ref[3].shape[1]):
                a, b = d["psiC1"][:, c], ref[3][:, c]
                assert np.max(np.abs(np.sign(a @ b) * a - b)) < 1e-3 * np.max(np.abs(b))
            # tau up to direction.  Where the reference's 1-D fit converges, tau is a stable
            # function of psirec: the GPU's tau equals the oracle's fit of the same psirec and
            # agrees with the end-to-end oracle to its 1 % stopping rule.  On some psi of PDs
            # this small the fit does not converge within its 100 iterations (the curve
            # coefficients drift) and its tau moves by ~0.2 under 1e-15 relative changes of
            # psirec -- in the oracle as on the GPU; there tau is not compared.
            tau = np.ravel(d["tau"])
            psirec = np.asarray(d["psirec"], dtype=np.float64)
            pert = np.random.default_rng(5).standard_normal(psirec.shape)
            fits = [_norm01(ofit.op(psirec * (1 + e * pert))[2]) for e in (0.0, 1e-13, -1e-13)]
            if max(_tau_gap(fits[0], f) for f in fits[1:]) > 1e-3:
                unstable += 1
                continue
            assert _tau_gap(tau, fits[0]) < 1e-6
            assert _tau_gap(tau, _norm01(ref[7])) < 2e-2
    assert unstable <= 2          # at least 4 of the 6 psi compared
    # second pass (EL1D, psiAnalysisParS2 isFull = 1): PD 0 keeps its sense, PD 1 flips, PD 2
    # is removed by the user
    p.trash_list = np.array([0, 0, 1])
    myio.fout1(p.CC_file, ["psinums", "senses"], [np.array([[1, 0, 1]]), np.array([[1, -1, 1]])])
    open(os.path.join(p.EL_prog, "0"), "a").close()       # PD 0 already done: skipped
    EL1D.op()
    assert not os.path.exists("{}prD_0_{}_1".format(p.EL_file, p.trajName))
    assert not os.path.exists("{}prD_2_{}_1".format(p.EL_file, p.trajName))
    os.remove(os.path.join(p.EL_prog, "0"))
    EL1D.op()
    for i, k, sense in ((0, 1, 1), (1, 0, -1)):
        e = myio.fin1("{}prD_{}_{}_1".format(p.EL_file, i, p.trajName))
        assert sorted(e) == sorted(["IMG1", "IMGT", "posPath", "PosPsi1", "psirec", "tau",
                                    "psiC1", "mu", "VX", "sdiag", "Topo_mean", "tauinds"])
        a = first[i, k]
        # the re-run reproduces the first pass's frames, so diff_corr > 0: sense -1 flips tau
        expect = np.ravel(a["tau"]) if sense == 1 else 1 - np.ravel(a["tau"])
        assert np.max(np.abs(np.ravel(e["tau"]) - expect)) < 1e-4
        np.testing.assert_allclose(e["sdiag"], a["sdiag"], rtol=1e-5)
        assert e["IMGT"].shape[0] == N * N
        assert os.path.exists(os.path.join(p.EL_prog, "%d" % i))

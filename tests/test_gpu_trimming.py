"""GPU trimming (manifoldTrimmingAuto.py:48-77) with the distances resident on the device,
the embedd re-embedding (embedd.py:16-79) and the isFull pass of psiAnalysisParS2
(psiAnalysisParS2.py:126-142, as EL1D.py:80-101 runs it), against the oracle."""
import os

import numpy as np
import pytest

from conftest import curve_with_outlier

pytestmark = pytest.mark.gpu


def _subspace_close(a, b, tol):
    """Columns of a equal those of b up to sign, to tol relative to b's largest entry."""
    for c in range(b.shape[1]):
        s = np.sign(a[:, c] @ b[:, c]) or 1.0
        assert np.max(np.abs(s * a[:, c] - b[:, c])) < tol * np.max(np.abs(b[:, c])), c


def test_gather_blocks_bitwise(cuda):
    import torch
    from manifoldem_python_amd import _native as nat, trimming
    rng = np.random.default_rng(3)
    src_ns = [37, 300, 129]
    src = trimming.packed_plan(src_ns, cuda)
    flat = torch.as_tensor(rng.normal(size=src.info.d_elems).astype(np.float32), device=cuda)
    host = flat.cpu().numpy()
    pick = [2, 0, 1]                                   # destination PD j <- source PD pick[j]
    sels = [np.sort(rng.choice(src_ns[q], m, replace=False)) for q, m in zip(pick, (100, 5, 250))]
    dst = trimming.packed_plan([len(s) for s in sels], cuda)
    out = torch.empty(dst.info.d_elems, dtype=torch.float32, device=cuda)
    nat.gather_blocks(dst, flat, src.pds["d_off"][pick], np.asarray(src_ns)[pick],
                      np.concatenate(sels), out)
    got = out.cpu().numpy()
    for j, (q, s) in enumerate(zip(pick, sels)):
        n = src_ns[q]
        blk = host[src.pds["d_off"][q]:src.pds["d_off"][q] + n * n].reshape(n, n)
        d0 = dst.pds["d_off"][j]
        np.testing.assert_array_equal(got[d0:d0 + len(s) ** 2].reshape(len(s), len(s)),
                                      blk[s][:, s])
    with pytest.raises(RuntimeError):
        nat.gather_blocks(dst, flat, src.pds["d_off"][pick], np.asarray(src_ns)[pick],
                          np.concatenate(sels) + 1000, out)


def test_trim_rad5_matches_oracle(cuda):
    """One trimming round at rad = 5: the outlier is removed, the curve re-embedded."""
    from manifoldem_python_amd import manifoldTrimmingAuto as mt, p
    from oracle import dm
    p.init()
    D = curve_with_outlier()
    n = D.shape[0]
    pos_r, res_r = dm.trim(D, 3, 5, num_eigs=15, rng=np.random.default_rng(1))
    pos, res = mt.embed_trimmed(D, 3, 5)
    np.testing.assert_array_equal(pos, pos_r)
    assert len(pos) == n - 1
    lamb, psi, sigma, mu = res[:4]
    assert sigma == pytest.approx(res_r[2], rel=1e-5)
    np.testing.assert_allclose(lamb, res_r[0], atol=1e-6)
    _subspace_close(psi[:, :8], res_r[1][:, :8], 1e-4)
    np.testing.assert_allclose(mu, res_r[3], rtol=1e-5, atol=1e-12)


def test_trim_batch_equals_single_pd_runs(cuda):
    """Three PDs in one batch (two trimmed, one not) give each PD's own result."""
    from manifoldem_python_amd import manifoldTrimmingAuto as mt, p
    p.init()
    Ds = [curve_with_outlier(400, 0.13, 0), curve_with_outlier(250, 0.5, 1)[:250, :250],
          curve_with_outlier(350, 0.13, 2)]          # outlier radii 6.6 and 5.75 (oracle)
    batch = mt.embed_trimmed_batch(Ds, 3, 5)
    for D, (keep, res) in zip(Ds, batch):
        pos1, res1 = mt.embed_trimmed(D, 3, 5)
        np.testing.assert_array_equal(keep, pos1)
        np.testing.assert_allclose(res[0], res1[0], atol=1e-6)
        assert res[2] == pytest.approx(res1[2], rel=1e-5)
    assert len(batch[1][0]) == 250                    # the curve alone: nothing trimmed
    assert len(batch[0][0]) == 400 and len(batch[2][0]) == 350


def _pd_pickles(tmp_path, D, n):
    from manifoldem_python_amd import myio, p
    p.init()
    p.user_dir, p.proj_name = str(tmp_path), "t"
    p.create_dir()
    myio.fout1("{}prD_0".format(p.dist_file), ["D", "ind"], [D, np.arange(n) + 1000])
    return p


def test_manifold_analysis_writes_trimmed_pickle(cuda, tmp_path):
    """manifoldAnalysis -> manifoldTrimmingAuto at rad = 5: the gC_trimmed_psi pickle against
    the oracle's trimming of the same D, the marker and topos/PrD_1/eig_spec.txt."""
    from manifoldem_python_amd import manifoldAnalysis, myio
    from oracle import dm
    D = curve_with_outlier()
    n = D.shape[0]
    p = _pd_pickles(tmp_path, D, n)
    p.numberofJobs = 1
    manifoldAnalysis.op()
    got = myio.fin1("{}prD_0".format(p.psi_file))
    assert sorted(got) == sorted(["lamb", "psi", "sigma", "mu", "posPath", "ind", "logEps",
                                  "logSumWij", "popt", "R_squared"])
    pos_r, res_r = dm.trim(D, 3, 5, num_eigs=15, rng=np.random.default_rng(1))
    np.testing.assert_array_equal(got["posPath"], pos_r)
    np.testing.assert_array_equal(got["ind"], np.arange(n) + 1000)
    np.testing.assert_allclose(got["lamb"], res_r[0], atol=1e-6)
    assert got["sigma"] == pytest.approx(res_r[2], rel=1e-5)
    assert np.max(np.abs(got["logSumWij"] - res_r[5])) < 1e-5   # float32 D entries
    _subspace_close(got["psi"][:, :8], res_r[1][:, :8], 1e-4)
    assert os.path.exists(os.path.join(p.psi_prog, "0"))
    lines = open(os.path.join(p.out_dir, "topos", "PrD_1", "eig_spec.txt")).read().split("\n")
    assert lines[0] == "1\t%.5f" % got["lamb"][1]


def test_embedd_reembeds_the_kept_points(cuda, tmp_path):
    """embedd.op: posPath restricted to the points the user kept, their DM against the
    oracle's, the 6-key pickle, and the PD's psi-analysis files removed."""
    from manifoldem_python_amd import embedd, myio
    from oracle import dm
    D = curve_with_outlier(200, 0.5, 4)[:200, :200]
    n = 200
    p = _pd_pickles(tmp_path, D, n)
    posPath = np.arange(0, n, 2)[:90]                 # a trimmed PD: 90 of the 200 points
    psi = np.random.default_rng(5).normal(size=(90, 15))
    myio.fout1("{}prD_0".format(p.psi_file), ["psi", "posPath", "ind"],
               [psi, posPath, np.arange(n)])
    stale = "{}prD_0_psi_1".format(p.psi2_file)
    open(stale, "w").close()
    open(os.path.join(p.psi2_prog, "0_1"), "w").close()
    keep = np.setdiff1d(np.arange(90), np.arange(30, 45))   # user removed 15 points
    x, y = psi[:, 0], psi[:, 1]
    embedd.op(list(zip(x, y)), list(zip(x[keep], y[keep])), 0)
    got = myio.fin1("{}prD_0".format(p.psi_file))
    assert sorted(got) == sorted(["lamb", "psi", "sigma", "mu", "posPath", "ind"])
    np.testing.assert_array_equal(got["posPath"], posPath[keep])
    D1 = D[posPath[keep]][:, posPath[keep]]
    ref = dm.op(D1, len(keep), 3, 60000, 15, np.random.default_rng(2))
    np.testing.assert_allclose(got["lamb"], ref[0][ref[0] > 0], atol=1e-6)
    _subspace_close(got["psi"][:, :8], ref[1][:, :8], 1e-4)
    assert not os.path.exists(stale) and not os.path.exists(os.path.join(p.psi2_prog, "0_1"))
    assert os.path.exists(os.path.join(p.out_dir, "topos", "PrD_1", "eig_spec.txt"))

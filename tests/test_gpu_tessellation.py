"""GPU S^2 tessellation (memi_nearest_bins behind S2tessellation.op) against the oracle, whose
classS2 is the reference's scikit-learn ball-tree 1-NN (S2tessellation.py:59-147)."""
import os

import numpy as np
import pytest

from oracle import tessellation as ot

pytestmark = pytest.mark.gpu


def _random_q(n, seed):
    q = np.random.default_rng(seed).normal(size=(4, n))
    return q / np.linalg.norm(q, axis=0)


@pytest.mark.parametrize("n,m", [(5, 1), (1000, 7), (200000, 5000)])
def test_nearest_bins_match_ball_tree(cuda, n, m):
    import torch
    from manifoldem_python_amd import _native as nat
    X = ot.distribute3Sphere(m)[0] if m > 1 else np.array([[0.0, 0.0, 1.0]])
    Q = np.ascontiguousarray(ot.get_S2(_random_q(n, n)).T)
    X = np.ascontiguousarray(X)
    ref, _ = ot.classS2(X, Q)
    idx = torch.empty(n, dtype=torch.int32, device=cuda)
    nat.nearest_bins(torch.as_tensor(Q, device=cuda), torch.as_tensor(X, device=cuda), idx)
    np.testing.assert_array_equal(idx.cpu().numpy(), ref[:, 0])


def test_tessellation_matches_oracle(cuda):
    from manifoldem_python_amd import S2tessellation, util
    q = util.augment(_random_q(60000, 9))
    got = S2tessellation.op(q, 0.1, 40, False, 60)
    ref = ot.op(q, 0.1, 40, False, 60)
    CG1, CG, nG, S2, S20_th, S20, NC = got
    assert nG == ref[2] == int(np.floor(4 * np.pi / 0.01))
    assert len(CG1) == len(ref[0])
    for a, b in zip(CG1, ref[0]):
        np.testing.assert_array_equal(a, b)
    assert len(CG) == len(ref[1]) > 0
    for a, b in zip(CG, ref[1]):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(NC, ref[6])
    np.testing.assert_allclose(S20_th, ref[4], rtol=0, atol=1e-15)
    np.testing.assert_array_equal(S2, ref[3])


def test_data_tessellate_writes_the_selection_pickle(cuda, tmp_path):
    """Data.tessellate (Data.py:72-118): augmented q and df, the 9 keys of selecGCs, and bins
    that GetDistancesS2 reads (CG of indices into the augmented particle list)."""
    from manifoldem_python_amd import Data, myio
    q = _random_q(20000, 4)
    df = np.random.default_rng(4).uniform(5000, 30000, 20000)
    sh = (np.zeros(20000), np.zeros(20000))
    f = os.path.join(tmp_path, "selecGCs")
    CG1, CG, nG, S2, S20_th, S20, NC = Data.tessellate(q, df, sh, 0.15, 30, 80, f)
    d = myio.fin1(f)
    assert sorted(d) == sorted(Data.TESS_KEYS)
    assert d["q"].shape == (4, 40000) and d["df"].shape == (40000,)
    np.testing.assert_array_equal(d["df"][20000:], df)
    np.testing.assert_array_equal(d["S20"], S20_th)
    assert len(d["CG"]) == S20_th.shape[1] and all(c.max() < 40000 for c in d["CG"])

"""Batched path (many PDs of uneven size in grouped launches) and its graph-captured form.

Each PD block of the batched distances is checked against the oracle's per-PD distance
matrix (getDistanceCTF_local_Conj9combinedS2.py:391-397 on the same images), and replays of the
captured HIP graph against the eager run of the same work, with eager work in between."""
import math

import numpy as np
import pytest

from conftest import FILTER, PARAMS, rel_offdiag

pytestmark = pytest.mark.gpu

SIZES = (150, 261, 90, 129)      # full tiles, edge tiles, a PD smaller than one tile
N = 40


def _pds(conj_frac=0.25):
    from manifoldem_python_amd import synth
    ss = []
    for p, n in enumerate(SIZES):
        ss.append(synth.make_pd(synth.seeds(20 + p)[p], n, N, conj_frac=conj_frac,
                                axis=synth.random_axis(np.random.default_rng(p))))
    return ss


def _batch(ss, cuda):
    import torch
    from manifoldem_python_amd import _native as nat, batch as B
    stacks, pds, off = [], [], 0
    for s in ss:
        conj = s.ind >= s.nStot // 2
        src = np.where(conj, s.ind - s.nStot // 2, s.ind) + off
        pds.append(dict(src=src, conj=conj.astype(np.int32), q=s.q, df=s.df,
                        k=min(len(s.ind), 50)))
        stacks.append(s.raw)
        off += s.raw.shape[0]
    stack = torch.as_tensor(np.concatenate(stacks), device=cuda)
    ctf = nat.ctf_params(PARAMS["pix_size"], PARAMS["Cs"], PARAMS["EkV"],
                         PARAMS["AmpContrast"], PARAMS["gaussEnv"])
    return B.make_batch(stack, pds, N, ctf, FILTER, cuda)


@pytest.fixture(scope="module")
def oracle_D():
    from oracle import distance as od
    out = []
    for s in _pds():
        p = dict(PARAMS, nPix=s.N)
        out.append(od.op([s.ind, s.q, s.df, None, 0], FILTER, s.raw, s.nStot,
                         dict(avgOnly=False), p)["D"])
    return out


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("tile", [128, 256])
def test_batched_distances_match_oracle_per_pd(cuda, oracle_D, precision, tile):
    from manifoldem_python_amd import batch as B
    batch = _batch(_pds(), cuda)
    wsp = B.Workspace(batch, precision, chunk_rows=200, gram_tile=tile)   # several row chunks
    dist = B.distances(batch, wsp)
    for p, ref in enumerate(oracle_D):
        D = dist.block(p).cpu().numpy()
        assert np.array_equal(D, D.T) and np.all(np.diag(D) == 0)
        err = rel_offdiag(D, ref)
        assert err.max() < 1e-5, (p, err.max())


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("tile", [128, 256])
def test_gram_schedules_agree_bitwise(cuda, precision, tile):
    """Every schedule of the Gram kernel (one- or two-step register prefetch, direct-to-LDS
    loads) runs the same MFMAs in the same order on the same staged bytes: D is bitwise the
    same, over several K chunks and PDs with edge tiles."""
    from manifoldem_python_amd import batch as B, pipeline as pl
    batch = _batch(_pds(), cuda)
    wsp = B.Workspace(batch, precision, chunk_rows=200, gram_tile=tile)
    B.distances(batch, wsp)
    variants = (1, 3) if tile == 256 else (1, 2, 3)
    Ds = [pl.ctf_distances(wsp.plan, wsp.operand, wsp.fy, wsp.ctf, N, ws=wsp.ws,
                           variant=v).D.cpu().numpy() for v in variants]
    for v, D in zip(variants[1:], Ds[1:]):
        assert np.array_equal(D, Ds[0]), v


@pytest.mark.parametrize("precision", [0, 1])
def test_graph_replays_match_eager(cuda, precision):
    """Three replays of the captured prep + Gram, each after eager work on the same stream,
    give the eager run's D and flag count."""
    import torch
    from manifoldem_python_amd import batch as B
    batch = _batch(_pds(), cuda)
    wsp = B.Workspace(batch, precision, chunk_rows=200)
    eager = B.distances(batch, wsp)
    D0 = eager.D.clone()
    c0 = int(eager.flag_count.item())
    g = B.GraphedDistances(batch, wsp)
    for r in range(3):
        x = torch.randn(1 << 20, device=cuda)
        (x * 2).sum().item()                 # eager work and a synchronisation in between
        dist = g.run()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(dist.D.cpu().numpy(), D0.cpu().numpy())
        assert int(dist.flag_count.item()) == c0, r


@pytest.mark.parametrize("split", [0, 16])
def test_overlapped_steps_match_serial(cuda, split):
    """Steps overlapped on two streams (step k+1's distances during step k's embedding, two
    buffer sets) give every step the eigenvalues and bandwidths of the same steps run one after
    another from the same random fit starts; also with the two streams on disjoint CU sets
    (16 of every 32 CUs for the embeddings, CU-masked streams)."""
    import torch
    from manifoldem_python_amd import batch as B, _native as nat
    batch = _batch(_pds(), cuda)

    def seeded():
        r = np.random.default_rng(4)
        return lambda *sh: r.random(sh)

    wa = B.Workspace(batch, nat.PREC_SPLIT_BF16)
    wb = B.Workspace(batch, nat.PREC_SPLIT_BF16)
    B.distances(batch, wa)
    B.distances(batch, wb)
    ov = B.OverlappedSteps(batch, wa, wb, split=split)
    got = []
    ov.run(5, num_eigs=8, rand=seeded(),
           on_result=lambda k, res: got.append([(r.lamb.copy(), r.sigma) for r in res]))
    torch.cuda.synchronize()
    rs = seeded()
    for k in range(5):
        res = B.embed(wa, B.distances(batch, wa), num_eigs=8, rand=rs)
        for (lam, sig), r in zip(got[k], res):
            np.testing.assert_allclose(lam, r.lamb, rtol=0, atol=1e-10)
            assert abs(sig - r.sigma) <= 1e-9 * abs(r.sigma)

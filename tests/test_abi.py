"""The C-ABI library: loads without a GPU, exports every entry point include/memi.h declares,
and its host-only plan builder and argument checks behave.  CPU only (no kernel launches)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_symbols():
    text = open(os.path.join(ROOT, "include", "memi.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(memi_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    from manifoldem_python_amd import _native as nat
    lib = nat.load()
    syms = _header_symbols()
    assert len(syms) >= 24
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) == set(nat.EXPORTS)
    assert lib.memi_version() >= 10000


@pytest.mark.parametrize("tile", [128, 256])
def test_plan_tables(tile):
    from manifoldem_python_amd import _native as nat
    ns, ks = [300, 129, 1], [150, 200, 1]
    rows0 = np.concatenate(([0], np.cumsum(ns)[:-1]))
    pl = nat.Plan(rows0, ns, ks, chunks=2, device="cpu", gram_tile=tile)
    I = pl.info
    T = [(n + tile - 1) // tile for n in ns]
    TF = [(n + 127) // 128 for n in ns]
    assert I.gram_tile == tile
    assert I.n_utiles == sum(t * (t + 1) // 2 for t in T)
    assert I.n_ftiles == sum(t * t for t in TF)
    assert I.n_items == I.n_utiles * 2 * 2
    assert I.ws_floats == I.n_items * tile * tile
    assert I.n_rows == sum(ns)
    assert I.d_elems == sum(n * n for n in ns)
    assert I.mask_words == sum(n * ((n + 31) // 32) for n in ns)
    assert I.list_elems == sum(min(k, n) * n for n, k in zip(ns, ks))
    assert I.flag_cap == sum(n * (n - 1) // 2 for n in ns)
    np.testing.assert_array_equal(pl.pds["d_off"], [0, 90000, 90000 + 129 * 129])
    items = pl.host[I.off_items:I.off_items + I.n_items * 24].view(np.int32).reshape(-1, 6)
    # every (tile, half, chunk) slot appears exactly once, slot = utile * 2 chunks + half *
    # chunks + chunk
    assert len(set(items[:, 5].tolist())) == I.n_items
    assert items[:, 5].max() == I.n_utiles * 2 * 2 - 1
    assert set(items[:, 3].tolist()) == {0, 1}
    assert set(items[:, 4].tolist()) == {0, 1}
    ut = pl.host[I.off_utiles:I.off_utiles + I.n_utiles * 16].view(np.int32).reshape(-1, 4)
    assert np.all(ut[:, 1] <= ut[:, 2])
    for x in items:
        u = ut[np.nonzero((ut[:, 0] == x[0]) & (ut[:, 1] == x[1]) & (ut[:, 2] == x[2]))[0][0]]
        assert x[5] == u[3] + x[3] * 2 + x[4]


def test_plan_rejects_bad_input():
    from manifoldem_python_amd import _native as nat
    with pytest.raises(RuntimeError, match="packed contiguously"):
        nat.Plan([0, 5], [10, 10], [5, 5], device="cpu")
    with pytest.raises(RuntimeError, match="n=0"):
        nat.Plan([0], [0], [5], device="cpu")
    with pytest.raises(RuntimeError, match="gram_tile"):
        nat.Plan([0], [10], [5], device="cpu", gram_tile=64)


def test_wrappers_refuse_host_tensors():
    import torch
    from manifoldem_python_amd import _native as nat
    x = torch.zeros(4, 8, 8)
    with pytest.raises(RuntimeError, match="GPU tensors"):
        nat.mask_multiply(x, 8)


def test_no_cpu_fallback_in_product_modules():
    """The product package never imports the oracle."""
    pkg = os.path.join(ROOT, "manifoldem_python_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in src and "from oracle" not in src, f


def test_wrappers_check_sizes_before_launch():
    """Shape mistakes raise on the host (regression: a flat n*n distance buffer passed where an
    (n, n) matrix is expected must not reach the row-sum kernel)."""
    import torch
    from manifoldem_python_amd import _native as nat
    D = torch.zeros(48 * 48)
    eps = torch.ones(1, dtype=torch.float64)
    S = torch.zeros(1, 48, dtype=torch.float64)
    with pytest.raises(RuntimeError, match=r"\(n, n\) matrix"):
        nat.kernel_rowsums(D, eps, S, S.clone())
    with pytest.raises(RuntimeError, match=r"\(n, n\) matrix"):
        nat.kernel_grad_weights(D, eps, S, eps, torch.zeros(48, 48))
    with pytest.raises(RuntimeError, match="spline_rotate"):
        nat.spline_rotate(torch.zeros(4, 8, 8), torch.zeros(3, dtype=torch.float64), 8,
                          torch.zeros(4, 8, 8))
    with pytest.raises(RuntimeError, match="ctf_pack"):
        nat.ctf_pack(torch.zeros(2, 8, 5, dtype=torch.complex64),
                     torch.zeros(2, dtype=torch.float64), 8, nat.ctf_params(1, 2, 300, .1, 1e9),
                     nat.PREC_FP32, 64, 0, torch.zeros(2 * 4 * 63), None, torch.zeros(2 * 40))
    pl = nat.Plan([0], [100], [50], device="cpu")
    with pytest.raises(RuntimeError, match="knn_select: D too small"):
        nat.knn_select(pl, torch.zeros(99 * 100), torch.zeros(pl.info.mask_words,
                                                            dtype=torch.int32))


def test_distance_rejects_a_split_with_an_empty_chunk():
    """Uneven K splits are allowed (last chunk shorter) but one that leaves a chunk empty is
    refused before anything touches the device: Khp = 64 split-BF16 is 2 K steps."""
    from manifoldem_python_amd import _native as nat
    fake = ctypes.c_void_p(256)       # never dereferenced: the argument checks fail first
    for chunks, ok in ((2, True), (3, False)):
        plan = nat.Plan([0], [10], [5], chunks=chunks, device="cpu")
        if ok:
            continue
        with pytest.raises(RuntimeError, match="empty chunk"):
            nat.call("memi_ctf_distance", fake, ctypes.byref(plan.info), fake, fake, 64,
                     nat.PREC_SPLIT_BF16, 8.0, fake, fake, 8, fake, fake, fake, fake, 0, None,
                     None, None)
    from manifoldem_python_amd import autotune
    assert autotune.valid_chunks(nat.PREC_SPLIT_BF16, 64, 2)
    assert not autotune.valid_chunks(nat.PREC_SPLIT_BF16, 64, 3)
    # C2-sized operand (N = 128): 260 split-BF16 steps, 3 chunks of 87/87/86
    assert autotune.valid_chunks(nat.PREC_SPLIT_BF16, 8320, 3)


def test_ritz_vectors_checks_meta_on_the_host():
    """A Ritz assembly whose basis length J exceeds the PD size (or k exceeds kmax) is refused
    before launch: the kernel indexes the basis and S by these numbers."""
    import torch
    from manifoldem_python_amd import _native as nat
    plan = nat.Plan([0, 10], [10, 20], [5, 5], device="cpu")
    m, kmax = 8, 4
    V = torch.zeros(m * 30, dtype=torch.float64)
    S = torch.zeros(100, dtype=torch.float64)
    Y = torch.zeros(30 * kmax, dtype=torch.float64)
    meta = torch.zeros(6, dtype=torch.int64)
    bad = [np.array([[11, 2, 0], [8, 2, 22]]),     # J > n of PD 0
           np.array([[8, 5, 0], [8, 2, 40]]),      # k > kmax
           np.array([[8, 2, 0], [8, 4, 90]])]      # s_off + J*k beyond S
    for mh in bad:
        with pytest.raises(RuntimeError, match="ritz_vectors"):
            nat.ritz_vectors(plan, V, m, S, mh, meta, kmax, Y)


def test_noalloc_prep_rejects_images_beyond_its_scratch():
    """memi_prep_images_noalloc checks N against its static scratch before any device call."""
    from manifoldem_python_amd import _native as nat
    lib = nat.load()
    ctf = nat.ctf_params(1.5, 2.7, 300.0, 0.1, np.inf)
    rc = lib.memi_prep_images_noalloc(None, 0, None, None, None, None, None, 0, 170, 1, 0.5, 8,
                                      ctypes.byref(ctf), 0, 64 * 1024, 0, None, None, None, None,
                                      None, None)
    assert rc == -1
    assert b"static scratch" in lib.memi_last_error()

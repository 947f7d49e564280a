"""Hand-written backward of the fused distance -> kernel -> log-sum op against a float64
torch.autograd reference of the same formula and against central finite differences."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ref_loss(img, C, w, eps):
    """float64 torch reference: images -> rfft2 -> weighted half-spectrum distances -> L_e."""
    import torch
    X = torch.fft.rfft2(img)
    n = X.shape[0]
    A = (X.abs() ** 2 * w).reshape(n, -1)
    B = (C ** 2).reshape(n, -1)
    Y = (C * X * w.sqrt()).reshape(n, -1)
    P = A @ B.T
    D = P + P.T - 2 * (Y @ Y.conj().T.contiguous()).real
    # D_ii = sum_k |C_i X_ik - C_i X_ik|^2 is exactly 0 (and so is its gradient); the expanded
    # form leaves ~1e-15 relative rounding noise of either sign there, which alone would make
    # dL/d eps at eps ~ 2^-20 median (where every off-diagonal K underflows) -- so the
    # reference uses the op's definition, an exact zero diagonal.
    D = D.masked_fill(torch.eye(n, dtype=torch.bool), 0.0)
    K = torch.exp(-D[None] / (2 * eps[:, None, None]))
    return torch.log(K.sum(2)).sum(1), D


@pytest.mark.parametrize("n,N", [(48, 32), (160, 64)])
def test_backward_matches_float64_autograd_and_fd(cuda, n, N):
    import torch
    from manifoldem_python_amd import _native as nat, autograd, synth
    s = synth.make_pd(synth.seeds(7)[2], n, N)
    rng = np.random.default_rng(0)
    img64 = torch.as_tensor(rng.normal(size=(n, N, N)), dtype=torch.float64)
    df = torch.as_tensor(s.df, device=cuda)
    ctf = nat.ctf_params(synth.PIX, synth.CS, synth.EKV, synth.AMPC, np.inf)
    # median distance sets the eps scale: 2^-20 .. 2^20 x median (41 values)
    img = img64.to(torch.float32).to(cuda).requires_grad_(True)
    fy = torch.fft.rfft2(img)
    with torch.no_grad():
        L0 = autograd.kernel_logsum(fy, df, torch.ones(1, dtype=torch.float64, device=cuda), ctf)
    k = synth.ctf_grid_k(N, synth.PIX)[:, :N // 2 + 1]
    C64 = torch.as_tensor(synth.ctf_values(k[None], s.df[:, None, None]))
    w = torch.full((N // 2 + 1,), 2.0, dtype=torch.float64)
    w[0] = 1.0
    w[-1] = 1.0
    _, Dref = _ref_loss(img64, C64, w, torch.ones(1, dtype=torch.float64))
    med = float(torch.median(Dref[~torch.eye(n, dtype=torch.bool)]))
    eps = torch.as_tensor(med * 2.0 ** np.linspace(-20, 20, 41), device=cuda).requires_grad_(True)
    g = torch.as_tensor(rng.normal(size=41), device=cuda)
    L = autograd.kernel_logsum(fy, df, eps, ctf)
    (L * g).sum().backward()
    # float64 autograd reference
    x = img64.clone().requires_grad_(True)
    e = eps.detach().cpu().clone().requires_grad_(True)
    Lr, _ = _ref_loss(x, C64, w, e)
    (Lr * g.cpu()).sum().backward()
    np.testing.assert_allclose(L.detach().cpu().numpy(), Lr.detach().numpy(), rtol=1e-5,
                               atol=1e-5)
    ge = eps.grad.cpu().numpy()
    ger = e.grad.numpy()
    err_e = np.max(np.abs(ge - ger) / (np.abs(ger) + 1e-30 + 1e-6 * np.max(np.abs(ger))))
    gx = img.grad.cpu().numpy().astype(np.float64)
    gxr = x.grad.numpy()
    rel = np.linalg.norm(gx - gxr) / np.linalg.norm(gxr)
    err_x = np.max(np.abs(gx - gxr)) / np.max(np.abs(gxr))
    print(f"[autograd n={n} N={N}] dL/deps max rel {err_e:.3e}; dL/dX norm rel {rel:.3e}, "
          f"max/max {err_x:.3e}")
    # measured on the MI355X: 1.2e-6 / 1.4e-6 (eps), 2.1e-7 / 2.6e-7 (X, norm), 2.5e-7 /
    # 3.3e-7 (X, max) at n = 48 / 160: float32 rounding of the images and spectra
    assert err_e < 1e-5, err_e
    assert rel < 2e-6, rel
    assert err_x < 3e-6, err_x
    # the float64 reference itself against central differences (step 1e-3 relative)
    for e_idx in (10, 20, 30):
        h = 1e-3 * float(e.detach()[e_idx])
        ep, em = e.detach().clone(), e.detach().clone()
        ep[e_idx] += h
        em[e_idx] -= h
        fd = float(((_ref_loss(img64, C64, w, ep)[0] - _ref_loss(img64, C64, w, em)[0])
                    * g.cpu()).sum()) / (2 * h)
        assert abs(fd - ger[e_idx]) <= 1e-5 * max(1.0, abs(ger[e_idx]))
    for (i, a, b) in ((0, 3, 5), (7, 10, 20)):
        h = 1e-3 * float(img64.abs().mean())
        xp, xm = img64.clone(), img64.clone()
        xp[i, a, b] += h
        xm[i, a, b] -= h
        fd = float(((_ref_loss(xp, C64, w, e.detach())[0] - _ref_loss(xm, C64, w, e.detach())[0])
                    * g.cpu()).sum()) / (2 * h)
        assert abs(fd - gxr[i, a, b]) <= 1e-4 * max(1.0, np.max(np.abs(gxr)))

"""GPU adjoint (backward) checks: dot-product identity, dense Y^T against
scipy at Nside=32 / lmax=64, and central finite differences of the weighted
least-squares loss at Nside 512 / lmax 1024 / 3 maps through torch.autograd."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rand_alm(L, K, seed):
    import wavemoth_amd as wm
    return wm.power_law_alm(L, L, K, seed)


@pytest.mark.parametrize("N,L,K", [(2, 4, 1), (8, 16, 2), (32, 64, 3), (64, 128, 1), (16, 40, 9)])
def test_dot_product_identity(N, L, K):
    import wavemoth_amd as wm
    print("case N=%d L=%d K=%d" % (N, L, K), flush=True)
    a = _rand_alm(L, K, 3)
    rng = np.random.default_rng(4)
    g = rng.standard_normal((12 * N * N, K))
    out = np.zeros((12 * N * N, K))
    plan = wm.ShtPlan(N, L, L, a.copy(), out, 'mmajor')
    f = plan.execute().copy()
    plan.output[...] = g
    b = plan.execute_adjoint().copy()
    lhs = float(np.sum(f * g))
    rhs = float(np.sum((np.conj(a) * b).real))
    print("lhs %.17g rhs %.17g" % (lhs, rhs))
    assert abs(lhs - rhs) <= 1e-12 * (np.linalg.norm(f) * np.linalg.norm(g))


def test_adjoint_vs_dense_Yt():
    """b_lm = c_m sum_p g_p Y_lm(theta_p, phi_p)^* against scipy at the pixel
    centres (Nside=32, lmax=64)."""
    from scipy.special import sph_harm_y
    import wavemoth_amd as wm
    from wavemoth_amd import healpix as H
    from wavemoth_amd.alm import l_of_index
    N, L = 32, 64
    th = H.get_ring_thetas(N)
    counts = H.get_ring_pixel_counts(N)
    phi0 = H.get_grid_phi0s(N)
    thetas = np.concatenate([np.full(c, th[i]) for i, c in enumerate(counts)])
    phis = np.concatenate([phi0[i] + 2 * np.pi * np.arange(c) / c for i, c in enumerate(counts)])
    rng = np.random.default_rng(9)
    g = rng.standard_normal((12 * N * N, 1))
    ls, ms = l_of_index(L, L)
    ref = np.zeros(ls.shape[0], complex)
    for p0 in range(0, thetas.shape[0], 1024):
        Y = sph_harm_y(ls[None, :], ms[None, :], thetas[p0:p0 + 1024, None],
                       phis[p0:p0 + 1024, None])
        ref += (np.conj(Y) * g[p0:p0 + 1024]).sum(0)
    ref *= np.where(ms == 0, 1.0, 2.0)
    ref[ms == 0] = ref[ms == 0].real
    plan = wm.ShtPlan(N, L, L, np.zeros((ls.shape[0], 1), complex), g.copy(), 'mmajor')
    b = plan.execute_adjoint()[:, 0]
    err = np.linalg.norm(b - ref) / np.linalg.norm(ref)
    print("dense Y^T rel err", err)
    assert err < 1e-12


def test_autograd_finite_differences_nside512():
    """Loss = 1/2 sum_p w_p ((Y a)_p - d_p)^2, w_p = 0 for |z| < 0.2 else
    U(0.5, 2), d = Y a_true + N(0,1); gradient vs central differences (step
    1e-6) on 64 random coefficients (Nside=512, lmax=1024, nmaps=3)."""
    import torch
    import wavemoth_amd as wm
    from wavemoth_amd import healpix as H
    from wavemoth_amd.torch_sht import synthesis
    N, L, K = 512, 1024, 3
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(2024)
    th = H.get_ring_thetas(N)
    counts = H.get_ring_pixel_counts(N)
    z = np.concatenate([np.full(c, np.cos(th[i])) for i, c in enumerate(counts)])
    w = np.where(np.abs(z) < 0.2, 0.0, rng.uniform(0.5, 2.0, z.shape[0]))
    a_true = torch.from_numpy(wm.power_law_alm(L, L, K, 11)).to(dev)
    d = synthesis(a_true, N, L).detach() + torch.from_numpy(
        rng.standard_normal((12 * N * N, K))).to(dev)
    wt = torch.from_numpy(w).to(dev)[:, None]
    a0 = torch.from_numpy(wm.power_law_alm(L, L, K, 12)).to(dev)

    def loss(a):
        r = synthesis(a, N, L) - d
        return 0.5 * torch.sum(wt * r * r)

    a = a0.clone().requires_grad_(True)
    L0 = loss(a)
    L0.backward()
    grad = a.grad.detach().cpu().numpy()
    nal = a0.shape[0]
    idx = rng.integers(0, nal, 64)
    kk = rng.integers(0, K, 64)
    part = rng.integers(0, 2, 64)  # 0: real part, 1: imaginary part
    h = 1e-6
    worst = 0.0
    gscale = float(np.abs(grad).max())
    errs = []
    with torch.no_grad():
        for i in range(64):
            e = torch.zeros_like(a0)
            e[idx[i], kk[i]] = h if part[i] == 0 else 1j * h
            rp = synthesis(a0 + e, N, L) - d
            rm = synthesis(a0 - e, N, L) - d
            # central difference [L(a+e) - L(a-e)] / 2h, with the difference of
            # the two ~1e7 losses formed pixel by pixel as
            # 1/2 w (r+ - r-)(r+ + r-) so it does not cancel catastrophically
            fd = torch.sum(0.5 * wt * (rp - rm) * (rp + rm)).item() / (2 * h)
            an = grad[idx[i], kk[i]].real if part[i] == 0 else grad[idx[i], kk[i]].imag
            if part[i] == 1 and idx[i] <= L:
                # m = 0 rows: the synthesis ignores Im(a_l0) (c2r drops Im X_0)
                assert an == 0.0 and fd == 0.0
                continue
            # relative error, floored at 1e-6 of the gradient's scale: the
            # FD sum over 9.4M pixels carries ~1e-12 of rounding per step
            err = abs(fd - an) / (abs(an) + 1e-6 * gscale)
            errs.append(err)
            worst = max(worst, err)
    print("FD relative mismatch: median %.3e worst %.3e (grad scale %.3e, loss %.6e)"
          % (float(np.median(errs)), worst, gscale, L0.item()))
    assert worst < 1e-5

"""Parity at the headline size, Nside 2048 / lmax = mmax 4096.

- FP64 path at 1 and 3 maps against the oracle's fixture
  (golden_n2048_l4096.npz): whole rings (poles, cap rings of split
  Bluestein length 4096, the cap/belt boundaries, belt rings of both phi0
  parities, the equator), every 2003rd pixel, per-ring rms, per-map norm;
  bound rel rms <= 1e-10.
- 64 maps (the mixed-precision contraction workload), power-law and flat
  C_l: every scheme within 1e-6 rel rms of the FP64 path (BF16x3 within
  5e-5), and the first three maps of the power-law case, which carry the
  fixture's a_lm, against the fixture directly (FP64 1e-10, schemes 1e-6).
- The adjoint at Nside 2048 (ring FFTs of 8192-point Bluestein in the
  caps): dot-product identity at 1 and 3 maps.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
N, L = 2048, 4096
F32, BF3, BF6, F16 = 0x1000, 0x2000, 0x4000, 0x8000
NAME = {0: "FP64", F32: "FP32", BF3: "BF16x3", BF6: "BF16x6", F16: "FP16x3"}
BOUND = {F32: 1e-6, BF6: 1e-6, BF3: 5e-5, F16: 1e-6}


def _golden():
    return np.load(os.path.join(HERE, "golden", "golden_n2048_l4096.npz"))


def _checksum(alm):
    return np.array([np.sum(np.abs(alm) ** 2), np.sum(alm.real), np.sum(alm.imag)])


def _fixture_alm(g):
    import wavemoth_amd as wm
    alm = wm.power_law_alm(L, L, 3, int(g["seed"][0]))
    np.testing.assert_allclose(_checksum(alm), g["alm_checksum"], rtol=1e-13)
    return alm


def _offsets():
    from wavemoth_amd.healpix import get_ring_offsets
    return get_ring_offsets(N)


def _vs_fixture(mp, g, maps):
    """Errors of host map columns `maps` (fixture columns) against the
    fixture: (whole rings of map 0, strided pixels, ring rms, map norms)."""
    offs = _offsets()
    out = {}
    if 0 in maps:
        c = maps.index(0)
        ring_pix = np.concatenate([mp[offs[i]:offs[i + 1], c] for i in g["rings"]])
        out["rings"] = float(np.linalg.norm(ring_pix - g["ring_pix"]) /
                             np.linalg.norm(g["ring_pix"]))
    sp = g["stride_pix"][:, maps]
    out["stride"] = float(np.linalg.norm(mp[g["stride_idx"]] - sp) / np.linalg.norm(sp))
    nr = 4 * N - 1
    sq = np.add.reduceat(mp * mp, offs[:nr], axis=0)
    rms = np.sqrt(sq / np.diff(offs)[:, None])
    ref_rms = g["ring_rms"][:, maps]
    out["ring_rms"] = float(np.abs(rms - ref_rms).max() / ref_rms.max())
    norm = np.linalg.norm(mp, axis=0)
    out["norm"] = float(np.abs(norm - g["map_norm"][maps]).max() / g["map_norm"].max())
    return out


@pytest.mark.parametrize("K", [1, 3])
def test_fp64_vs_fixture(K):
    import torch
    from wavemoth_amd.torch_sht import DevicePlan
    g = _golden()
    alm = _fixture_alm(g)[:, :K].copy()
    if K == 1:
        np.testing.assert_allclose(_checksum(alm), g["alm1_checksum"], rtol=1e-13)
    print("case FP64 Nside=%d lmax=%d nmaps=%d vs fixture" % (N, L, K), flush=True)
    p = DevicePlan(N, L, L, K)
    d_alm = torch.from_numpy(alm).cuda()
    d_map = torch.empty((12 * N * N, K), dtype=torch.float64, device="cuda")
    p.synthesize_into(d_alm, d_map)
    mp = d_map.cpu().numpy()
    errs = _vs_fixture(mp, g, list(range(K)))
    print(errs, flush=True)
    assert all(v <= 1e-10 for v in errs.values()), errs


def _err(x, ref, chunk=1 << 22):
    """rel rms and max/rms of x - ref over all entries, in row chunks (no
    full-size temporary)."""
    import torch
    num = torch.zeros((), dtype=torch.float64, device=x.device)
    den = torch.zeros_like(num)
    mx = torch.zeros_like(num)
    for r0 in range(0, x.shape[0], chunk):
        d = x[r0:r0 + chunk] - ref[r0:r0 + chunk]
        num += (d * d).sum()
        den += (ref[r0:r0 + chunk] ** 2).sum()
        mx = torch.maximum(mx, d.abs().max())
    n = x.numel()
    return float(torch.sqrt(num / den)), float(mx / torch.sqrt(den / n))


@pytest.mark.parametrize("flat", [False, True])
def test_schemes_nside2048_lmax4096_64maps(flat):
    """Nside 2048, lmax 4096, 64 maps: every contraction scheme against the
    FP64 path; with the power law, maps 0-2 also against the fixture."""
    import torch
    from wavemoth_amd.torch_sht import DevicePlan, random_alm
    K = 64
    dev = torch.device("cuda:0")
    g = _golden()
    alm = random_alm(L, L, K, 2048 + int(flat), dev, flat=flat)
    if not flat:
        alm[:, :3] = torch.from_numpy(_fixture_alm(g)).to(dev)
    print("case FP64 Nside=%d lmax=%d nmaps=%d flat=%s" % (N, L, K, flat), flush=True)
    ref = torch.empty((12 * N * N, K), dtype=torch.float64, device=dev)
    p = DevicePlan(N, L, L, K)
    p.synthesize_into(alm, ref)
    torch.cuda.synchronize()
    del p
    if not flat:
        errs = _vs_fixture(ref[:, :3].cpu().numpy(), g, [0, 1, 2])
        print("FP64 64-map plan, maps 0-2 vs fixture", errs, flush=True)
        assert all(v <= 1e-10 for v in errs.values()), errs
    got = torch.empty_like(ref)
    failures = []
    for scheme in (F32, F16, BF6, BF3):
        print("case %s Nside=%d lmax=%d nmaps=%d flat=%s" % (NAME[scheme], N, L, K, flat),
              flush=True)
        got.fill_(float("nan"))
        p = DevicePlan(N, L, L, K, scheme)
        p.synthesize_into(alm, got)
        torch.cuda.synchronize()
        del p
        rms, mx = _err(got, ref)
        ok = bool(torch.isfinite(got).all()) and rms <= BOUND[scheme]
        msg = "%s vs FP64: rel rms %.3e max/rms %.3e" % (NAME[scheme], rms, mx)
        if not flat:
            errs = _vs_fixture(got[:, :3].cpu().numpy(), g, [0, 1, 2])
            msg += "; maps 0-2 vs fixture %s" % errs
            ok = ok and errs["stride"] <= BOUND[scheme] and errs["rings"] <= BOUND[scheme]
        print(msg, flush=True)
        if not ok:
            failures.append(msg)
    assert not failures, failures


@pytest.mark.parametrize("K", [1, 3])
def test_adjoint_dot_product_nside2048(K):
    """<Y a, g> = <a, Y^T g> at Nside 2048: the adjoint's ring kernel runs
    8192-point Bluestein FFTs for the caps whose synthesis uses the
    4096-point register kernels."""
    import torch
    from wavemoth_amd.torch_sht import DevicePlan, random_alm
    Lsmall = 256
    dev = torch.device("cuda:0")
    print("case adjoint Nside=%d lmax=%d nmaps=%d" % (N, Lsmall, K), flush=True)
    a = random_alm(Lsmall, Lsmall, K, 5 + K, dev)
    gen = torch.Generator(device=dev)
    gen.manual_seed(11 + K)
    gm = torch.randn((12 * N * N, K), generator=gen, device=dev, dtype=torch.float64)
    p = DevicePlan(N, Lsmall, Lsmall, K)
    f = torch.empty_like(gm)
    p.synthesize_into(a, f)
    b = torch.empty_like(a)
    p.adjoint_into(gm, b)
    torch.cuda.synchronize()
    lhs = float((f * gm).sum())
    rhs = float((a.real * b.real + a.imag * b.imag).sum())
    scale = float(torch.linalg.norm(f) * torch.linalg.norm(gm))
    print("lhs %.17g rhs %.17g rel %.3e" % (lhs, rhs, abs(lhs - rhs) / scale), flush=True)
    assert abs(lhs - rhs) <= 1e-12 * scale
    # per map as well: each column is an independent identity
    for k in range(K):
        lk = float((f[:, k] * gm[:, k]).sum())
        rk = float((a[:, k].real * b[:, k].real + a[:, k].imag * b[:, k].imag).sum())
        sk = float(torch.linalg.norm(f[:, k]) * torch.linalg.norm(gm[:, k]))
        assert abs(lk - rk) <= 1e-12 * sk, (k, lk, rk)

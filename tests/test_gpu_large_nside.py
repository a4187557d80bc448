"""Nside 4096 and 8192.

Nside 4096 (201,326,592 pixels, 16383 rings; rings up to 16384 pixels):
the ring FFTs at the LDS limit of 8192 points.

- Belt rings (2 Nside = 8192-point FFTs) in the register-FFT belt kernel
  with one job per workgroup; polar caps past ring 2048, whose split
  Bluestein halves need 8192 points, in the one-job-per-block cap kernel
  (cap_block_seq); at mmax > 2 Nside the wrapping rings of those lengths
  and the belt (n = 16384 <= mmax) in the same kernel's wrapping fold.
- FP64 at 1 and 2 maps against the oracle (bound 1e-10 rel rms); the
  FP32 ring path (MIXED_F32 plans) at 1 and 2 maps within 1e-6.
- a_lm is the seeded power law restricted to three disjoint arithmetic
  progressions of m (every 331st m from 0, m = 1..3, the three largest):
  the oracle's Legendre stage sums only those (its plan precomputes an
  m progression), the ring stage runs over every ring and bin.
- A batch holding an Nside 4096 request matches its single plan bitwise.
- The adjoint is refused at Nside 4096 (its ring kernel holds at most
  8192-point Bluestein FFTs, Nside <= 2048).
- Nside 8192 (805,306,368 pixels): rings past the LDS limit (belt
  transforms of 16384 points, polar-cap Bluestein lengths of 32768) in the
  four-step long path, FP64 and FP32 against the oracle, at lmax 4096 and
  at lmax = mmax = 20000 (the belt and the long caps wrap).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 4096
F32 = 0x1000


def _progressions(M):
    return [(0, M + 1, 331), (1, 4, 1), (M - 2, M + 1, 1)]


def _sparse_alm(L, M, K, seed):
    import wavemoth_amd as wm
    from wavemoth_amd.alm import l_of_index
    alm = wm.power_law_alm(L, M, K, seed)
    keep = np.zeros(M + 1, bool)
    for a, b, s in _progressions(M):
        keep[a:b:s] = True
    _, ms = l_of_index(L, M)
    alm[~keep[ms]] = 0
    return alm, np.flatnonzero(keep)


_REF = {}


def _oracle_map(L, M, K, seed, Ns=N):
    """Oracle map of the sparse a_lm: the Legendre stage of each m
    progression into one work array (disjoint m slots), then the full ring
    stage."""
    from oracle import oracle as O
    key = (Ns, L, M, K, seed)
    if key not in _REF:
        _REF.clear()
        alm, _ = _sparse_alm(L, M, K, seed)
        work = None
        for a, b, s in _progressions(M):
            op = O.OraclePlan(Ns, L, M, K, a, b, s)
            if work is None:
                work = np.zeros((M + 1, 2, 2 * Ns, K), np.complex128)
            op.legendre_stage_fast(alm, work=work)
        _REF[key] = op.c2r(op.ring_spectra(work))
        del work
    return _REF[key]


def _run(L, M, K, flags, seed, Ns=N):
    import torch
    from wavemoth_amd.torch_sht import DevicePlan
    alm, ms = _sparse_alm(L, M, K, seed)
    p = DevicePlan(Ns, L, M, K, flags)
    info = p.info()
    d_alm = torch.from_numpy(alm).cuda()
    d_map = torch.full((12 * Ns * Ns, K), float("nan"), dtype=torch.float64, device="cuda")
    p.synthesize_into(d_alm, d_map)
    torch.cuda.synchronize()
    out = d_map.cpu().numpy()
    del p, d_alm, d_map
    return out, info, ms


def _errors(got, ref, Ns=N):
    from wavemoth_amd.healpix import get_ring_offsets
    offs = get_ring_offsets(Ns)
    rms = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    # per ring class: polar caps past ring 2048, the rest of the caps, belt
    d2 = np.add.reduceat((got - ref) ** 2, offs[:-1], axis=0).sum(axis=1)
    r2 = np.add.reduceat(ref ** 2, offs[:-1], axis=0).sum(axis=1)
    ring = np.arange(4 * Ns - 1)
    north = np.minimum(ring, 4 * Ns - 2 - ring) + 1  # ring number from its pole
    cut = Ns // 2  # caps past it: split Bluestein halves of 2 Nside points or more
    cls = {"cap>Nside/2": (north > cut) & (north < Ns), "cap<=Nside/2": north <= cut,
           "belt": north >= Ns}
    per = {k: float(np.sqrt(d2[c].sum() / r2[c].sum())) for k, c in cls.items()}
    return rms, per


@pytest.mark.parametrize("L,M,K", [(8192, 8192, 2), (12288, 12288, 1)])
def test_fp64_vs_oracle(L, M, K):
    print("case FP64 Nside=%d lmax=%d mmax=%d nmaps=%d" % (N, L, M, K), flush=True)
    got, info, ms = _run(L, M, K, 0, 4096 + K)
    print(info, "m values:", len(ms), flush=True)
    assert info["max_register_fft"] == 8192
    # Ring pairs by kernel.  Belt pairs r = 0..N (n = 4N, h = 8192): the belt
    # kernel when mmax <= h, else the wrapping fold.  Cap ring c (n = 4c):
    # c > mmax/4 in the cap kernel (split Bluestein of 8192 points past
    # c = 2048); c <= 2048 in the general kernel (its 8192-point Bluestein);
    # 2048 < c <= mmax/4 in the wrapping kernel, the general one being
    # too short for them.
    if M == 8192:
        want = {"rings_belt": N + 1, "rings_cap": 2047, "rings_wrap": 0, "rings_general": 2048}
    else:
        want = {"rings_belt": 0, "rings_cap": 1023, "rings_wrap": N + 1 + 1024,
                "rings_general": 2048}
    assert {k: info[k] for k in want} == want
    assert info["adjoint_groups"] == 0 and info["adjoint_groups_noalloc"] == 0
    assert np.isfinite(got).all()
    ref = _oracle_map(L, M, K, 4096 + K)
    rms, per = _errors(got, ref)
    print("FP64 rel rms %.3e per class %s" % (rms, per), flush=True)
    assert rms <= 1e-10 and all(v <= 1e-10 for v in per.values()), (rms, per)


@pytest.mark.parametrize("L,M,K", [(8192, 8192, 2), (12288, 12288, 1)])
def test_fp32_ring_path(L, M, K):
    """MIXED_F32: q in FP32 and the FP32 ring kernels (the two-map belt job
    at K = 2, the single-buffer cap job in FP32)."""
    print("case MIXED_F32 Nside=%d lmax=%d mmax=%d nmaps=%d" % (N, L, M, K), flush=True)
    got, info, _ = _run(L, M, K, F32, 4096 + K)
    assert info["scheme"] == 0 and info["max_register_fft"] == 8192
    assert np.isfinite(got).all()
    ref = _oracle_map(L, M, K, 4096 + K)
    rms, per = _errors(got, ref)
    print("MIXED_F32 rel rms %.3e per class %s" % (rms, per), flush=True)
    assert rms <= 1e-6 and all(v <= 1e-6 for v in per.values()), (rms, per)


def test_batch_with_nside4096_request():
    """A batch of an Nside 4096 request (lmax 64) and an Nside 64 one:
    each map equals its single plan's bitwise."""
    import torch
    import wavemoth_amd as wm
    from wavemoth_amd import BatchPlan
    from wavemoth_amd.torch_sht import DevicePlan
    reqs = [(N, 64, 64, 1), (64, 128, 128, 3)]
    alms = [torch.from_numpy(wm.power_law_alm(L, M, K, 17 + i)).cuda()
            for i, (n, L, M, K) in enumerate(reqs)]
    print("case batch Nside 4096 + 64", flush=True)
    b = BatchPlan(reqs)
    maps = [torch.full((12 * n * n, K), float("nan"), dtype=torch.float64, device="cuda")
            for (n, L, M, K) in reqs]
    b.synthesize_into(alms, maps)
    torch.cuda.synchronize()
    for (n, L, M, K), a, mp in zip(reqs, alms, maps):
        ref = torch.empty_like(mp)
        DevicePlan(n, L, M, K).synthesize_into(a, ref)
        torch.cuda.synchronize()
        assert torch.equal(mp, ref), (n, L, K)


def test_adjoint_refused_past_nside2048():
    import torch
    from wavemoth_amd.torch_sht import DevicePlan
    print("case adjoint refusal Nside=%d" % N, flush=True)
    p = DevicePlan(N, 32, 32, 1)
    g = torch.zeros((12 * N * N, 1), dtype=torch.float64, device="cuda")
    out = torch.empty((33 * 34 // 2, 1), dtype=torch.complex128, device="cuda")
    with pytest.raises(RuntimeError, match="Nside > 2048"):
        p.adjoint_into(g, out)


N8 = 8192


@pytest.mark.parametrize("L,M,K,flags", [(4096, 4096, 1, 0), (20000, 20000, 1, 0),
                                         (4096, 4096, 2, F32)])
def test_nside8192_long_path(L, M, K, flags):
    """Nside 8192: the belt (16384-point transforms) and the caps past ring
    4096 (Bluestein length 32768) in the four-step long path; at mmax 20000
    both wrap (n <= mmax).  FP64 within 1e-10 of the oracle, FP32 ring
    stage (MIXED_F32) within 1e-6."""
    name = "MIXED_F32" if flags else "FP64"
    print("case %s Nside=%d lmax=%d mmax=%d nmaps=%d" % (name, N8, L, M, K), flush=True)
    got, info, ms = _run(L, M, K, flags, 8192 + K, Ns=N8)
    print({k: info[k] for k in ("rings_general", "rings_belt", "rings_cap", "rings_wrap",
                                "rings_long", "max_register_fft")}, flush=True)
    # long pairs: the belt (r = 0 .. N) and caps c = 4097 .. 8191
    assert info["rings_long"] == (N8 + 1) + (N8 - 1 - 4096)
    assert info["adjoint_groups"] == 0
    assert np.isfinite(got).all()
    ref = _oracle_map(L, M, K, 8192 + K, Ns=N8)
    rms, per = _errors(got, ref, Ns=N8)
    bound = 1e-6 if flags else 1e-10
    print("%s rel rms %.3e per class %s" % (name, rms, per), flush=True)
    assert rms <= bound and all(v <= bound for v in per.values()), (rms, per)

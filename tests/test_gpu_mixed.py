"""GPU tests of the mixed-precision contraction (WAVEMOTH_MIXED_*): the
map of each scheme against the FP64 path on the same a_lm, for the power
law and flat C_l, with the error bound of each scheme written here:
  FP32 MFMA  rel rms <= 1e-6
  BF16x6     rel rms <= 1e-6
  BF16x3     rel rms <= 5e-5 (two bf16 parts carry ~16 bits)
  FP16x3     rel rms <= 1e-6 (two fp16 parts carry ~22 bits; scaled)
plus the oracle at a small size, determinism, mmax < lmax and nmaps not a
multiple of the column block."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32, BF3, BF6, F16 = 0x1000, 0x2000, 0x4000, 0x8000
BOUND = {F32: 1e-6, BF6: 1e-6, BF3: 5e-5, F16: 1e-6}
NAME = {F32: "FP32", BF3: "BF16x3", BF6: "BF16x6", F16: "FP16x3"}


def _maps(N, L, M, K, flags, alm):
    import torch
    from wavemoth_amd.torch_sht import DevicePlan
    p = DevicePlan(N, L, M, K, flags)
    out = torch.empty((12 * N * N, K), dtype=torch.float64, device=alm.device)
    p.synthesize_into(alm, out)
    torch.cuda.synchronize()
    return out, p


def _err(x, ref):
    import torch
    d = x - ref
    rms = float(torch.linalg.norm(d) / torch.linalg.norm(ref))
    mx = float(d.abs().max() / torch.sqrt(torch.mean(ref * ref)))
    return rms, mx


@pytest.mark.parametrize("scheme", [F32, BF3, BF6, F16])
@pytest.mark.parametrize("N,L,M,K,flat", [(32, 64, 64, 16, False), (64, 128, 128, 32, False),
                                          (64, 128, 128, 32, True), (16, 40, 20, 3, False),
                                          (128, 256, 256, 17, False),
                                          (128, 512, 512, 8, False)])
def test_scheme_vs_fp64(scheme, N, L, M, K, flat):
    import torch
    from wavemoth_amd.torch_sht import random_alm
    dev = torch.device("cuda:0")
    print("case %s N=%d L=%d M=%d K=%d flat=%s" % (NAME[scheme], N, L, M, K, flat), flush=True)
    alm = random_alm(L, M, K, 123 + N + K, dev, flat=flat)
    ref, _ = _maps(N, L, M, K, 0, alm)
    got, p = _maps(N, L, M, K, scheme, alm)
    assert p.info()["scheme"] == {F32: 0, BF3: 1, BF6: 2, F16: 3}[scheme]
    rms, mx = _err(got, ref)
    print("rel rms %.3e  max/rms %.3e" % (rms, mx), flush=True)
    assert rms <= BOUND[scheme]


@pytest.mark.parametrize("scheme", [F32, F16])
@pytest.mark.parametrize("N,L,K", [(512, 1024, 8), (256, 512, 5), (1024, 2048, 2),
                                   (2048, 256, 4), (2048, 256, 3), (2048, 256, 2)])
def test_cap_block_stores(scheme, N, L, K):
    """FP32 cap and wrapping-ring blocks whose jobs are consecutive maps of
    one ring pair store their pixels together (Bluestein lengths 512 to
    2048, 2 to 8 jobs per block); K = 5 leaves padding jobs in such blocks.
    At Nside 2048 the caps of Bluestein length 4096 and the belt run two
    maps per job; such a cap job parks its southern Z entries in the
    southern ring's pixel rows of its two maps (K = 2: one job per pair;
    K = 3: the last job has one map and folds q again instead)."""
    import torch
    from wavemoth_amd.torch_sht import random_alm
    dev = torch.device("cuda:0")
    print("case %s N=%d L=%d K=%d" % (NAME[scheme], N, L, K), flush=True)
    alm = random_alm(L, L, K, 77 + K, dev)
    ref, _ = _maps(N, L, L, K, 0, alm)
    got, _ = _maps(N, L, L, K, scheme, alm)
    rms, mx = _err(got, ref)
    print("rel rms %.3e  max/rms %.3e" % (rms, mx), flush=True)
    assert rms <= BOUND[scheme]
    # every map is written (a skipped column would leave the empty() values)
    assert bool(torch.isfinite(got).all())
    per_map = torch.linalg.norm(got - ref, dim=0) / torch.linalg.norm(ref, dim=0)
    assert float(per_map.max()) <= 10 * BOUND[scheme]


@pytest.mark.parametrize("scheme", [F32, F16, BF6, BF3])
@pytest.mark.parametrize("N,L,K,flat", [(16, 32, 16, False), (64, 128, 32, False),
                                        (64, 128, 32, True)])
def test_scheme_vs_oracle(oracle, scheme, N, L, K, flat):
    """Each scheme against the CPU oracle directly (not through the FP64
    path), to the scheme's bound."""
    import torch
    from wavemoth_amd import power_law_alm
    print("case %s N=%d L=%d K=%d flat=%s vs oracle" % (NAME[scheme], N, L, K, flat), flush=True)
    alm_h = power_law_alm(L, L, K, 5 + N, flat=flat)
    got, _ = _maps(N, L, L, K, scheme, torch.from_numpy(alm_h).to("cuda:0"))
    ref = oracle.OraclePlan(N, L, L, K).synthesize(alm_h)
    err = np.linalg.norm(got.cpu().numpy() - ref) / np.linalg.norm(ref)
    print("rel rms vs oracle %.3e" % err)
    assert err <= BOUND[scheme]


@pytest.mark.parametrize("scheme", [F32, BF6, F16])
def test_scheme_deterministic(scheme):
    import torch
    from wavemoth_amd.torch_sht import DevicePlan, random_alm
    dev = torch.device("cuda:0")
    N, L, K = 64, 128, 32
    alm = random_alm(L, L, K, 9, dev)
    p = DevicePlan(N, L, L, K, scheme)
    a = torch.empty((12 * N * N, K), dtype=torch.float64, device=dev)
    b = torch.empty_like(a)
    p.synthesize_into(alm, a)
    p.synthesize_into(alm, b)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("amp", [1e-12, 1.0, 1e9])
def test_fp16_scheme_scale_invariant(amp):
    """The FP16 scheme's column scales follow the data: the relative error
    does not depend on the amplitude of a_lm (fp16 alone spans 6e-8..65504)."""
    import torch
    from wavemoth_amd.torch_sht import random_alm
    dev = torch.device("cuda:0")
    N, L, K = 32, 64, 16
    alm = random_alm(L, L, K, 31, dev) * amp
    ref, _ = _maps(N, L, L, K, 0, alm)
    got, _ = _maps(N, L, L, K, F16, alm)
    rms, mx = _err(got, ref)
    print("amp %g rel rms %.3e" % (amp, rms), flush=True)
    assert rms <= 1e-6


@pytest.mark.parametrize("scheme", [F32, F16])
def test_mixed_get_work(scheme):
    """A mixed plan keeps the Legendre-stage output in FP32; get_work widens
    it into the reference's complex128 layout, which matches the FP64 plan's
    to the scheme's accuracy."""
    import wavemoth_amd as wm
    N, L, K = 32, 64, 16
    alm = wm.power_law_alm(L, L, K, 17)
    works = []
    for flags in (0, scheme):
        p = wm.ShtPlan(N, L, L, alm, np.zeros((12 * N * N, K)), 'mmajor', flags=flags)
        p.perform_legendre_transform()
        works.append(p.get_work())
    ref, got = works
    assert got.dtype == np.complex128 and got.shape == ref.shape
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    print("%s work rel rms %.3e" % (NAME[scheme], err), flush=True)
    assert err <= 1e-6


@pytest.mark.parametrize("N,L,M,K,flags", [(96, 200, 150, 40, 0), (48, 100, 100, 100, 0),
                                           (64, 128, 128, 64, 0x100), (2048, 128, 128, 1, 0),
                                           (8, 300, 300, 16, 0)])
def test_fp16_register_kernel_vs_producer_kernel(monkeypatch, N, L, M, K, flags):
    """The FP16x3 contraction with Lambda generated in registers (four
    l-segments per ring from plan-time checkpoints; the default) and the
    producer/consumer kernel (WAVEMOTH_MIX_PRODUCERS=1) both meet the bound
    against the FP64 path, on shapes with a partial 128-ring block (Nside
    96), several column groups (100 maps), mmax < lmax, seeds generated per
    execute (checkpoints recomputed; flag 0x100), one map, and more l than
    rings (Nside 8, lmax 300)."""
    import torch
    from wavemoth_amd.torch_sht import random_alm
    dev = torch.device("cuda:0")
    print("case FP16x3 kernels N=%d L=%d M=%d K=%d flags=%#x" % (N, L, M, K, flags), flush=True)
    alm = random_alm(L, M, K, 500 + K, dev)
    ref, _ = _maps(N, L, M, K, 0, alm)
    got, _ = _maps(N, L, M, K, F16 | flags, alm)
    monkeypatch.setenv("WAVEMOTH_MIX_PRODUCERS", "1")
    old, _ = _maps(N, L, M, K, F16 | flags, alm)
    r_new, _ = _err(got, ref)
    r_old, _ = _err(old, ref)
    r_pair, _ = _err(got, old)
    print("rel rms: register kernel %.3e  producer kernel %.3e  between %.3e"
          % (r_new, r_old, r_pair), flush=True)
    assert bool(torch.isfinite(got).all())
    assert r_new <= 1e-6 and r_old <= 1e-6 and r_pair <= 1e-6
    per_map = torch.linalg.norm(got - ref, dim=0) / torch.linalg.norm(ref, dim=0)
    assert float(per_map.max()) <= 1e-5

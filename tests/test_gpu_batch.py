"""GPU tests of the batched engine (wavemoth_plan_batch): every request's
map must equal, bit for bit, the map a single plan of the same request
produces (same arithmetic per ring and l), and match the oracle at small
sizes; graph replays, pointer changes between executes, the autotuner
flag, the host-buffer variant and stage timing are covered too."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = [(4, 8, 8, 1), (8, 16, 16, 3), (16, 32, 32, 8), (32, 64, 64, 1), (16, 40, 20, 2),
         (1, 2, 2, 1), (64, 128, 128, 3), (32, 64, 64, 8), (2, 4, 4, 5), (8, 16, 16, 3),
         (64, 100, 100, 1), (128, 256, 256, 4)]


def _dev_alm(L, M, K, seed, dev):
    """Seeded power-law a_lm generated on the device (C_l = (l+1)^-2,
    a_l0 real); only used as test input, so any seeded law would do."""
    import torch
    from wavemoth_amd.alm import l_of_index
    ls, ms = l_of_index(L, M)
    g = torch.Generator(device=dev)
    g.manual_seed(seed)
    re = torch.randn((len(ls), K), generator=g, device=dev, dtype=torch.float64)
    im = torch.randn((len(ls), K), generator=g, device=dev, dtype=torch.float64)
    l_t = torch.from_numpy(ls).to(dev)
    m0 = torch.from_numpy(ms == 0).to(dev)
    sd = ((l_t + 1.0) ** -1 / np.sqrt(2.0))[:, None]
    im = torch.where(m0[:, None], torch.zeros_like(im), im)
    sd = torch.where(m0[:, None], sd * np.sqrt(2.0), sd)
    return torch.complex(re * sd, im * sd).contiguous()


def _single(req, alm):
    import torch
    from wavemoth_amd.torch_sht import DevicePlan
    N, L, M, K = req
    p = DevicePlan(N, L, M, K)
    out = torch.empty((12 * N * N, K), dtype=torch.float64, device=alm.device)
    p.synthesize_into(alm, out)
    torch.cuda.synchronize()
    return out


def _run_batch(reqs, flags, seed0, dev):
    import torch
    from wavemoth_amd import BatchPlan
    b = BatchPlan(reqs, flags)
    alms = [_dev_alm(L, M, K, seed0 + i, dev) for i, (N, L, M, K) in enumerate(reqs)]
    outs = [torch.full((12 * N * N, K), float("nan"), dtype=torch.float64, device=dev)
            for (N, L, M, K) in reqs]
    b.synthesize_into(alms, outs)
    torch.cuda.synchronize()
    return b, alms, outs


@pytest.mark.parametrize("flags", [0, 0x200])
def test_batch_equals_single_plans(flags):
    import torch
    dev = torch.device("cuda:0")
    print("batch of %d small requests, flags %#x" % (len(SMALL), flags), flush=True)
    b, alms, outs = _run_batch(SMALL, flags, 500, dev)
    assert b.info()["graph"] == (0 if flags else 1)
    for r, req in enumerate(SMALL):
        ref = _single(req, alms[r])
        assert torch.equal(outs[r], ref), (r, req, (outs[r] - ref).abs().max().item())


def test_batch_matches_oracle(oracle):
    import torch
    from wavemoth_amd import BatchPlan, power_law_alm
    reqs = [(4, 8, 8, 2), (16, 32, 32, 3), (32, 64, 64, 1), (16, 40, 20, 8)]
    dev = torch.device("cuda:0")
    alms_h = [power_law_alm(L, M, K, 77 + i) for i, (N, L, M, K) in enumerate(reqs)]
    b = BatchPlan(reqs)
    outs = [torch.empty((12 * N * N, K), dtype=torch.float64, device=dev)
            for (N, L, M, K) in reqs]
    b.synthesize_into([torch.from_numpy(a).to(dev) for a in alms_h], outs)
    torch.cuda.synchronize()
    for r, (N, L, M, K) in enumerate(reqs):
        ref = oracle.OraclePlan(N, L, M, K).synthesize(alms_h[r])
        got = outs[r].cpu().numpy()
        err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print("request", r, "rel rms", err)
        assert err <= 1e-12


def test_graph_replays_and_pointer_changes():
    """Replays of the captured graph with eager torch work in between, new
    input values, and new output buffers (the pointer array is re-uploaded):
    every replay must equal the direct-launch result."""
    import torch
    from wavemoth_amd import BatchPlan
    dev = torch.device("cuda:0")
    reqs = SMALL[:6]
    g = BatchPlan(reqs)
    e = BatchPlan(reqs, 0x200)
    assert g.info()["graph"] == 1 and e.info()["graph"] == 0
    for rep in range(4):
        print("replay", rep, flush=True)
        alms = [_dev_alm(L, M, K, 900 + 10 * rep + i, dev)
                for i, (N, L, M, K) in enumerate(reqs)]
        outs_g = [torch.empty((12 * N * N, K), dtype=torch.float64, device=dev)
                  for (N, L, M, K) in reqs]
        outs_e = [torch.empty_like(o) for o in outs_g]
        junk = torch.randn(1 << 20, device=dev).sum()  # eager work between replays
        g.synthesize_into(alms, outs_g)
        e.synthesize_into(alms, outs_e)
        torch.cuda.synchronize()
        assert torch.isfinite(junk)
        for r in range(len(reqs)):
            assert torch.equal(outs_g[r], outs_e[r]), (rep, r)
    # same buffers again (no re-upload), new values
    for r, (N, L, M, K) in enumerate(reqs):
        alms[r].copy_(_dev_alm(L, M, K, 7 + r, dev))
    g.synthesize_into(alms, outs_g)
    e.synthesize_into(alms, outs_e)
    torch.cuda.synchronize()
    for r in range(len(reqs)):
        assert torch.equal(outs_g[r], outs_e[r])


def test_measure_flag_autotunes():
    """WAVEMOTH_MEASURE: every request measured.  The sweep shape keeps the
    map bit-identical to the single plan's; the ring kernel of the wrapping
    rings (both variants timed for (128, 512, 512, K), whose caps of 65 ..
    128 pixels per quarter wrap at Bluestein length >= 256) does too when
    it is the single plan's choice, else agrees to rounding."""
    import torch
    from wavemoth_amd import WAVEMOTH_MEASURE
    dev = torch.device("cuda:0")
    reqs = [(64, 128, 128, 1), (32, 64, 64, 8), (64, 128, 128, 3), (128, 512, 512, 3),
            (128, 512, 512, 8)]
    b, alms, outs = _run_batch(reqs, WAVEMOTH_MEASURE, 40, dev)
    for r, req in enumerate(reqs):
        info = b.request_info(r)
        print(req, "-> RL", info["rings_per_lane"], "KB", info["maps_per_pass"],
              "wrap_register", info["wrap_register"], flush=True)
        assert info["measured"] == 1
        ref = _single(req, alms[r])
        if info["wrap_register"] == (1 if req[3] >= 8 else 0):
            assert torch.equal(outs[r], ref)
        else:
            err = float(torch.linalg.norm(outs[r] - ref) / torch.linalg.norm(ref))
            assert err <= 1e-13, (req, err)


def test_host_execute_and_profiling():
    import torch
    from wavemoth_amd import BatchPlan, power_law_alm
    reqs = [(8, 16, 16, 2), (16, 32, 32, 1)]
    ins = [power_law_alm(L, M, K, 3 + i) for i, (N, L, M, K) in enumerate(reqs)]
    outs = [np.zeros((12 * N * N, K)) for (N, L, M, K) in reqs]
    b = BatchPlan(reqs)
    b.execute(ins, outs)
    dev = torch.device("cuda:0")
    for r, req in enumerate(reqs):
        ref = _single(req, torch.from_numpy(ins[r]).to(dev)).cpu().numpy()
        assert np.array_equal(outs[r], ref)
    b.set_profiling(True)
    for _ in range(3):
        b.execute(ins, outs)
    ms, n = b.stage_times()
    b.set_profiling(False)
    assert n == 3 and all(v >= 0 for v in ms) and ms[1] > 0


def test_batch_rejects_bad_buffers():
    import torch
    from wavemoth_amd import BatchPlan
    b = BatchPlan([(4, 8, 8, 1)])
    dev = torch.device("cuda:0")
    with pytest.raises(ValueError):
        b.synthesize_into([torch.zeros((3, 1), dtype=torch.complex128, device=dev)],
                          [torch.zeros((192, 1), dtype=torch.float64, device=dev)])


def test_uneven_batch_full_size():
    """96 requests, Nside uniform over {128 .. 2048}, lmax = mmax = 2 Nside,
    nmaps over {1, 3, 8}: one request of every (Nside, nmaps) class present
    is checked bit for bit against its single plan; all maps finite."""
    import torch
    from wavemoth_amd import BatchPlan, random_requests
    dev = torch.device("cuda:0")
    reqs = random_requests(96, seed=3)
    print("requests:", sorted(set(reqs)), flush=True)
    b = BatchPlan(reqs)
    info = b.info()
    print("batch info", info, flush=True)
    alms = [_dev_alm(L, M, K, 10_000 + i, dev) for i, (N, L, M, K) in enumerate(reqs)]
    outs = [torch.empty((12 * N * N, K), dtype=torch.float64, device=dev)
            for (N, L, M, K) in reqs]
    b.synthesize_into(alms, outs)
    torch.cuda.synchronize()
    print("batch executed", flush=True)
    seen = set()
    for r, req in enumerate(reqs):
        assert bool(torch.isfinite(outs[r]).all())
        if req in seen:
            continue
        seen.add(req)
        ref = _single(req, alms[r])
        assert torch.equal(outs[r], ref), (r, req)
        print("request", r, req, "equal", flush=True)
        del ref

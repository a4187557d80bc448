"""The reference boundary's remaining calls through the C ABI:
wavemoth_get_legendre_flops (src/wavemoth.c:876-885), the stage-wise hooks
wavemoth_perform_legendre_transforms + wavemoth_perform_backward_ffts
(src/wavemoth.c:933-941, 1193-1195) and wavemoth_disable_phase_shifting
(src/wavemoth.c:1240-1248), plus odd Nside (phi0 from wavemoth.c's grid)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _plan(N, L, M, K, seed, **kw):
    import wavemoth_amd as wm
    alm = wm.power_law_alm(L, M, K, seed)
    out = np.zeros((12 * N * N, K))
    return wm.ShtPlan(N, L, M, alm, out, 'mmajor', **kw), alm


def rel_rms(x, ref):
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("N,L,M", [(16, 32, 32), (64, 200, 150), (256, 512, 512)])
def test_legendre_flops(oracle, N, L, M):
    """4 x the (l, ring-pair) elements evaluated for (m, odd), as the
    reference's 4 * element_count: the sum over (m, odd) is 4 T_eps; per m it
    lies between 4 x the oracle's trimmed count (threshold 1e-30) and 4 x the
    dense count nk(m, odd) * 2 Nside."""
    print("case N=%d L=%d M=%d" % (N, L, M), flush=True)
    plan, _ = _plan(N, L, M, 1, 3)
    info = plan.info()
    R = 2 * N
    fl = np.array([[plan.legendre_flops(m, odd) for odd in (0, 1)] for m in range(M + 1)])
    assert fl.sum() == 4 * info["triples_eps"]
    nk = np.array([[(L - m - odd) // 2 + 1 if L - m - odd >= 0 else 0 for odd in (0, 1)]
                   for m in range(M + 1)])
    assert np.all(fl >= 0) and np.all(fl <= 4 * nk * R)
    op = oracle.OraclePlan(N, L, M, 1)
    lib = oracle.lib()
    for m in range(0, M + 1, max(1, M // 17)):
        t_or = lib.or_plan_trimmed_triples(op._p, m, m + 1, 1)
        assert 4 * t_or <= fl[m].sum() <= 4 * (nk[m].sum() * R), m
    # out-of-range arguments
    assert plan.legendre_flops(M + 1, 0) == -1 and plan.legendre_flops(0, 2) == -1


@pytest.mark.parametrize("N,L,K", [(16, 32, 1), (64, 128, 3), (512, 1024, 2)])
def test_stage_hooks_equal_execute(N, L, K):
    """perform_legendre_transform followed by perform_backward_ffts writes
    the same map as execute, bit for bit."""
    print("case N=%d L=%d K=%d" % (N, L, K), flush=True)
    plan, _ = _plan(N, L, L, K, 41)
    ref = plan.execute().copy()
    plan.output[...] = 0
    plan.perform_legendre_transform()
    plan.perform_backward_ffts()
    assert np.array_equal(plan.output, ref)


@pytest.mark.parametrize("N,L,K", [(8, 16, 1), (32, 64, 2), (128, 256, 1)])
def test_phase_shifts_off_matches_oracle(oracle, N, L, K):
    """ShtPlan(phase_shifts=False) evaluates every ring from phi = 0; the
    oracle's ring stage with phase_shifts=False is the reference's
    wavemoth_disable_phase_shifting."""
    print("case N=%d L=%d K=%d" % (N, L, K), flush=True)
    plan, alm = _plan(N, L, L, K, 5, phase_shifts=False)
    got = plan.execute().copy()
    op = oracle.OraclePlan(N, L, L, K)
    ref = op.synthesize(alm, phase_shifts=False)
    err = rel_rms(got, ref)
    print("rel rms %.3e" % err)
    assert err <= 1e-12
    with_phase = op.synthesize(alm)
    assert rel_rms(with_phase, ref) > 1e-3  # the switch changes the map


def test_disable_phase_shifting_after_executes(oracle):
    """Switching phases off on a plan that has already executed (and has a
    device execute queued) affects only the executes issued afterwards."""
    import ctypes
    import torch
    import wavemoth_amd as wm
    from wavemoth_amd import _native
    N, L, K = 32, 64, 2
    plan, alm = _plan(N, L, L, K, 9)
    on = plan.execute().copy()
    d_alm = torch.from_numpy(alm).cuda()
    d_map = torch.empty((12 * N * N, K), dtype=torch.float64, device="cuda")
    plan.execute_device(d_alm.data_ptr(), d_map.data_ptr(),
                        torch.cuda.current_stream().cuda_stream)
    _native.lib().wavemoth_disable_phase_shifting(ctypes.c_void_p(plan._plan))
    assert wm.last_error() == ""
    torch.cuda.synchronize()
    assert np.array_equal(d_map.cpu().numpy(), on)
    off = plan.execute().copy()
    ref = oracle.OraclePlan(N, L, L, K).synthesize(alm, phase_shifts=False)
    assert rel_rms(off, ref) <= 1e-12


@pytest.mark.parametrize("N,L,K", [(3, 6, 1), (5, 10, 2), (7, 20, 1), (33, 66, 3)])
def test_odd_nside_matches_oracle(oracle, N, L, K):
    """Odd Nside: belt phi0 follows src/wavemoth.c:1223 (iring % 2)."""
    print("case N=%d L=%d K=%d" % (N, L, K), flush=True)
    plan, alm = _plan(N, L, L, K, 50 + N)
    got = plan.execute().copy()
    ref = oracle.OraclePlan(N, L, L, K).synthesize(alm)
    err = rel_rms(got, ref)
    print("rel rms %.3e" % err)
    assert err <= 1e-12

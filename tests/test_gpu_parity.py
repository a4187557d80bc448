"""GPU parity: the HIP path against the CPU oracle on the same seeded
a_lm (small sizes) and against the committed golden fixtures."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def rel_rms(x, ref):
    return float(np.linalg.norm(x - ref) / np.linalg.norm(ref))


def _plan(N, L, M, K, seed, **kw):
    import wavemoth_amd as wm
    alm = wm.power_law_alm(L, M, K, seed)
    out = np.zeros((12 * N * N, K))
    return wm.ShtPlan(N, L, M, alm, out, 'mmajor', **kw), alm


@pytest.mark.parametrize("N,L,K", [(1, 2, 1), (2, 4, 1), (4, 8, 2), (8, 16, 1), (16, 32, 3),
                                   (32, 64, 1), (64, 128, 1), (64, 128, 4)])
def test_map_matches_oracle(oracle, N, L, K):
    print("case N=%d L=%d K=%d" % (N, L, K), flush=True)
    plan, alm = _plan(N, L, L, K, seed=1000 + N)
    gpu = plan.execute().copy()
    ref = oracle.OraclePlan(N, L, L, K).synthesize(alm)
    err = rel_rms(gpu, ref)
    print("rel rms", err)
    assert err <= 1e-12


@pytest.mark.parametrize("N,L,M,K", [(4, 8, 8, 2), (16, 40, 20, 1), (32, 64, 64, 2)])
def test_work_matches_oracle(oracle, N, L, M, K):
    print("case N=%d L=%d M=%d K=%d" % (N, L, M, K), flush=True)
    plan, alm = _plan(N, L, M, K, seed=7)
    plan.perform_legendre_transform()
    work = plan.get_work()
    ref = oracle.OraclePlan(N, L, M, K).legendre_stage(alm)
    err = np.abs(work - ref).max() / np.abs(ref).max()
    print("max rel", err)
    assert err <= 1e-12


def _golden(name):
    return np.load(os.path.join(HERE, "golden", name))


def _checksum(alm):
    return np.array([np.sum(np.abs(alm) ** 2), np.sum(alm.real), np.sum(alm.imag)])


def test_golden_small():
    import wavemoth_amd as wm
    g = _golden("golden_small.npz")
    for N in [1, 2, 4, 8, 16, 32]:
        for K in [1, 3]:
            key = "n%d_k%d" % (N, K)
            print("case", key, flush=True)
            L = 2 * N
            alm = wm.power_law_alm(L, L, K, int(g[key + "_seed"][0]))
            np.testing.assert_allclose(_checksum(alm), g[key + "_alm_checksum"], rtol=1e-13)
            out = np.zeros((12 * N * N, K))
            gpu = wm.ShtPlan(N, L, L, alm, out, 'mmajor').execute()
            assert rel_rms(gpu, g[key + "_map"]) <= 1e-12


def test_golden_nside64_lmax128():
    import wavemoth_amd as wm
    g = _golden("golden_n64_l128.npz")
    N, L = 64, 128
    alm = wm.power_law_alm(L, L, 1, int(g["seed"][0]))
    np.testing.assert_allclose(_checksum(alm), g["alm_checksum"], rtol=1e-13)
    out = np.zeros((12 * N * N, 1))
    plan = wm.ShtPlan(N, L, L, alm, out, 'mmajor')
    gpu = plan.execute()
    err = rel_rms(gpu, g["map"])
    print("map rel rms", err)
    assert err <= 1e-12
    work = plan.get_work()[g["work_m"]]
    werr = np.abs(work - g["work"]).max() / np.abs(g["work"]).max()
    print("work max rel", werr)
    assert werr <= 1e-12


def test_golden_nside1024_lmax2048():
    """Nside 1024 / lmax 2048 parity: rel rms <= 1e-10 against the CPU
    reference algorithm."""
    import wavemoth_amd as wm
    g = _golden("golden_n1024_l2048.npz")
    N, L = 1024, 2048
    alm = wm.power_law_alm(L, L, 1, int(g["seed"][0]))
    np.testing.assert_allclose(_checksum(alm), g["alm_checksum"], rtol=1e-13)
    out = np.zeros((12 * N * N, 1))
    gpu = wm.ShtPlan(N, L, L, alm, out, 'mmajor').execute()[:, 0]
    from wavemoth_amd.healpix import get_ring_offsets
    offs = get_ring_offsets(N)
    ring_pix = np.concatenate([gpu[offs[i]:offs[i + 1]] for i in g["rings"]])
    e_rings = rel_rms(ring_pix, g["ring_pix"])
    e_stride = rel_rms(gpu[g["stride_idx"]], g["stride_pix"])
    rms = np.array([np.sqrt(np.mean(gpu[offs[i]:offs[i + 1]] ** 2)) for i in range(4 * N - 1)])
    e_rms = np.abs(rms - g["ring_rms"]).max() / g["ring_rms"].max()
    e_norm = abs(np.linalg.norm(gpu) - g["map_norm"][0]) / g["map_norm"][0]
    print("rings %.3e stride %.3e ring-rms %.3e norm %.3e" % (e_rings, e_stride, e_rms, e_norm))
    assert e_rings <= 1e-10 and e_stride <= 1e-10 and e_rms <= 1e-10 and e_norm <= 1e-10


def test_seeds_on_the_fly_identical():
    import wavemoth_amd as wm
    N, L, K = 32, 64, 2
    alm = wm.power_law_alm(L, L, K, 3)
    a = wm.ShtPlan(N, L, L, alm, np.zeros((12 * N * N, K)), 'mmajor').execute().copy()
    b = wm.ShtPlan(N, L, L, alm, np.zeros((12 * N * N, K)), 'mmajor',
                   flags=wm.WAVEMOTH_SEEDS_ON_THE_FLY).execute().copy()
    assert np.array_equal(a, b)


def test_repeat_executes_bitwise_identical():
    """The reference's determinism check (test_wavemoth.py:92-117)."""
    import hashlib
    import wavemoth_amd as wm
    N, L = 16, 32
    alm = np.zeros((wm.nalm(L, L), 1), np.complex128)
    alm[wm.lm_to_idx_mmajor(10, 4, L)] = 1 + 1j
    plan = wm.ShtPlan(N, L, L, alm, np.zeros((12 * N * N, 1)), 'mmajor')
    h0 = hashlib.md5(plan.execute().tobytes()).hexdigest()
    for _ in range(50):
        assert hashlib.md5(plan.execute().tobytes()).hexdigest() == h0


@pytest.mark.parametrize("N,L,K", [(128, 512, 8), (96, 384, 8)])
def test_wrapping_rings_many_maps(oracle, N, L, K):
    """Rings with n <= mmax on an FFT length of 256 or more, at 8 maps or
    more: the register kernel that folds the wrapped coefficients (belt
    rings of Nside 128 with mmax 512; Bluestein belts of Nside 96)."""
    print("case N=%d L=%d K=%d" % (N, L, K), flush=True)
    plan, alm = _plan(N, L, L, K, seed=77 + N)
    gpu = plan.execute().copy()
    ref = oracle.OraclePlan(N, L, L, K).synthesize(alm)
    err = rel_rms(gpu, ref)
    print("rel rms", err)
    assert err <= 1e-12


@pytest.mark.parametrize("N,L,K", [(512, 512, 2), (512, 512, 10)])
def test_belt_paired_maps(oracle, N, L, K):
    """Belt rings with FFT length >= 1024 at an even map count: the register
    kernel's paired jobs exchange lanes so that each store covers two maps
    (K = 10 also leaves a half-padded group of maps)."""
    print("case N=%d L=%d K=%d" % (N, L, K), flush=True)
    plan, alm = _plan(N, L, L, K, seed=31 + K)
    gpu = plan.execute().copy()
    ref = oracle.OraclePlan(N, L, L, K).synthesize(alm)
    err = rel_rms(gpu, ref)
    print("rel rms", err)
    assert err <= 1e-12


@pytest.mark.parametrize("N,L,M,K", [(16, 32, 32, 10), (8, 24, 12, 5), (32, 60, 60, 8)])
def test_many_maps_and_mmax(oracle, N, L, M, K):
    print("case N=%d L=%d M=%d K=%d" % (N, L, M, K), flush=True)
    plan, alm = _plan(N, L, M, K, seed=11)
    gpu = plan.execute().copy()
    ref = oracle.OraclePlan(N, L, M, K).synthesize(alm)
    assert rel_rms(gpu, ref) <= 1e-12


def test_torch_device_path_matches_host_path():
    import torch
    import wavemoth_amd as wm
    from wavemoth_amd.torch_sht import synthesis
    N, L, K = 32, 64, 3
    alm = wm.power_law_alm(L, L, K, 21)
    host = wm.ShtPlan(N, L, L, alm, np.zeros((12 * N * N, K)), 'mmajor').execute().copy()
    dev = synthesis(torch.from_numpy(alm).cuda(), N, L).cpu().numpy()
    assert np.array_equal(host, dev)


def test_large_mmax_small_nside(oracle):
    """mmax above one LDS staging round of the ring stage (4608 values of m)."""
    N, L = 4, 4700
    print("case N=%d L=%d" % (N, L), flush=True)
    plan, alm = _plan(N, L, L, 1, seed=77)
    gpu = plan.execute().copy()
    ref = oracle.OraclePlan(N, L, L, 1).synthesize(alm)
    err = rel_rms(gpu, ref)
    print("rel rms", err)
    assert err <= 1e-12

"""The CPU baseline's timed legs against the oracle (CPU only):
- oracle/sht_fast.c's vectorised Legendre stage (the reference's SIMD kernel
  restated, src/legendre_transform.c.in:63-616) against or_legendre_stage,
  the scalar restatement the parity tests pin, at each vector width the CPU
  has: one map (rings in the vector lanes) and several (columns in the
  lanes, with the single-vector and scalar column tails), partial last
  strips, m samples, rings whose first row leaves <= 4 rows (dense), and
  parity 1 at m = lmax (no rows);
- the sampled ring stage (fold + pocketfft c2r of all maps per ring pair)
  against the oracle's full ring stage on the same rings."""
import numpy as np
import pytest


def _widths(oracle):
    w = oracle.fast_width()
    return [x for x in (4, 8) if x <= w]


@pytest.mark.parametrize("N,L,M,K,step", [
    (4, 8, 8, 1, 1),      # R = 8: one partial strip at either width
    (8, 16, 12, 3, 1),    # 6 columns: single-vector and scalar tails
    (16, 40, 40, 1, 1),
    (16, 32, 32, 8, 1),   # 16 columns
    (32, 64, 64, 12, 3),  # 24 columns, sampled m
    (64, 160, 160, 2, 5),
])
def test_fast_legendre_stage_matches_oracle(oracle, N, L, M, K, step):
    widths = _widths(oracle)
    if not widths:
        pytest.skip("CPU has neither AVX2+FMA nor AVX-512F")
    from wavemoth_amd.alm import power_law_alm
    alm = power_law_alm(L, M, K, 17 + N)
    op = oracle.OraclePlan(N, L, M, K, m_step=step)
    ref = op.legendre_stage(alm, compact=True)
    scale = np.abs(ref).max()
    for w in widths:
        got, ran = op.legendre_stage_fast(alm, compact=True, width=w)
        assert ran == w
        err = np.abs(got - ref).max() / scale
        assert err <= 1e-12, (w, err)
        # rings with no rows (parity 1 at m = lmax) come out zero
        if step == 1 and M == L:
            assert not got[L, 1].any()


def test_fast_legendre_stage_full_layout(oracle):
    """compact=False writes work[m, odd, ring, map] at every sampled m and
    leaves the others alone."""
    if not _widths(oracle):
        pytest.skip("CPU has neither AVX2+FMA nor AVX-512F")
    from wavemoth_amd.alm import power_law_alm
    N, L, K = 8, 20, 2
    alm = power_law_alm(L, L, K, 3)
    op = oracle.OraclePlan(N, L, L, K, m_step=2)
    ref = op.legendre_stage(alm)
    work = np.full((L + 1, 2, 2 * N, K), 7.0 + 0j)
    got, _ = op.legendre_stage_fast(alm, work=work)
    assert np.abs(got[::2] - ref[::2]).max() <= 1e-12 * np.abs(ref).max()
    assert np.all(got[1::2] == 7.0)


@pytest.mark.parametrize("N,K,step", [(4, 3, 1), (8, 2, 3), (16, 5, 4), (16, 1, 16)])
def test_ring_stage_sample_matches_oracle(oracle, N, K, step):
    if not _widths(oracle):
        pytest.skip("CPU has neither AVX2+FMA nor AVX-512F")
    L = 2 * N
    rng = np.random.default_rng(N * 10 + K)
    op = oracle.OraclePlan(N, L, L, K, m_start=0, m_stop=1)
    work = (rng.standard_normal((L + 1, 2, 2 * N, K)) +
            1j * rng.standard_normal((L + 1, 2, 2 * N, K)))
    full = op.c2r(op.ring_spectra(work))
    for workers in (1, 3):
        rings, px = op.ring_stage_sample(work[:, :, ::step], step, workers=workers)
        mid = 2 * N - 1
        assert rings[0] == mid and len(rings) == 2 * len(range(0, 2 * N, step)) - 1
        ref = np.concatenate([full[op.ring_offsets[i]:op.ring_offsets[i + 1]] for i in rings])
        assert np.abs(px - ref).max() <= 1e-12 * np.abs(ref).max()

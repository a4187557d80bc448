"""Pins for the CPU oracle (no GPU): every function of oracle/ is checked
against the reference's own compiled code (oracle/_ref), the reference's
test cases and known-answer tests, independent libraries (scipy, mpmath)
and closed forms, chosen so that a dropped term, a wrong sign or index or
a transposed operand fails at least one of them."""
import math

import numpy as np
import pytest
from scipy.special import sph_harm_y


def _ref_or_skip(O):
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref not built (no reference tree)")


# ------------------------------------------------------------- Lambda

@pytest.mark.parametrize("N", [4, 64, 512])
def test_lambda_matches_reference_ylmgen_on_rings(oracle, N):
    """Restated Ylmgen == compiled libpshtlight/ylmgen_c.c, bit for bit, on
    the HEALPix ring thetas (including the m_crit skip across columns)."""
    O = oracle
    _ref_or_skip(O)
    thetas = O.ring_thetas_positive(N)
    L = 2 * N
    for m in sorted(set([0, 1, 2, 3, N // 2, N - 1, N, L - 1, L])):
        ref = O.ref_lambda_matrix(m, thetas, L, eps=1e-300)
        mine = O.lambda_matrix_shared(m, thetas, L, eps=1e-300)
        assert np.array_equal(mine, ref), (N, m)


def test_lambda_matches_reference_ylmgen_extremes(oracle):
    O = oracle
    _ref_or_skip(O)
    for (m, th, L, eps) in [(4096, 0.3, 4096, 1e-300), (3000, 1.2, 4096, 1e-30),
                            (2000, 0.01, 4096, 1e-300), (0, 1e-9, 100, 1e-300),
                            (5, np.pi - 1e-16, 40, 1e-300), (700, np.pi / 2, 900, 1e-30)]:
        a = O.lambda_matrix(m, [th], L, eps)
        b = O.ref_lambda_matrix(m, [th], L, eps)
        assert np.array_equal(a, b), (m, th)


@pytest.mark.parametrize("theta,m,lmax", [(np.pi / 2, 0, 10), (np.pi / 4, 0, 10),
                                          (3 * np.pi / 4, 0, 10), (np.pi / 4, 1, 4),
                                          (np.pi / 4, 2, 4), (np.pi / 4, 50, 50),
                                          (np.pi / 2, 49, 50), (0.001, 250, 300)])
def test_lambda_vs_scipy(oracle, theta, m, lmax):
    """wavemoth/test/test_legendre.py:9-26 restated with scipy's sph_harm_y."""
    lam = oracle.lambda_matrix(m, [theta], lmax)[0]
    ref = sph_harm_y(np.arange(m, lmax + 1), m, theta, 0.0).real
    np.testing.assert_allclose(lam, ref, rtol=1e-12, atol=1e-15)


def test_lambda_closed_forms(oracle):
    O = oracle
    # Lambda_00 = 1/sqrt(4 pi) at any theta
    for th in [0.1, 1.0, 2.5]:
        assert abs(O.lambda_column(0, 3, th)[0][0] - 1 / math.sqrt(4 * math.pi)) < 1e-16
    # Lambda_mm = (-1)^m sqrt((2m+1)!!/(2m)!! / (4 pi)) sin^m ; Lambda_{m+1,m} = sqrt(2m+3) cos Lambda_mm
    for m in [1, 2, 7, 30, 200]:
        for th in [0.4, 1.3, 2.0]:
            col = O.lambda_column(m, m + 1, th)[0]
            logv = 0.5 * sum(math.log((2 * i + 1) / (2 * i)) for i in range(1, m + 1))
            lmm = (-1) ** m * math.exp(logv + m * math.log(math.sin(th))) / math.sqrt(4 * math.pi)
            assert abs(col[m] - lmm) <= 1e-13 * abs(lmm)
            assert abs(col[m + 1] - math.sqrt(2 * m + 3) * math.cos(th) * col[m]) <= 1e-13 * abs(col[m + 1]) + 1e-300


def test_lambda_vs_mpmath_high_l(oracle):
    import mpmath
    mpmath.mp.dps = 50
    for (l, m, th) in [(4096, 100, 0.7), (4000, 3000, 1.3), (2500, 1, 0.05), (300, 250, 1.0)]:
        x = mpmath.cos(mpmath.mpf(th))
        exact = mpmath.sqrt((2 * l + 1) / (4 * mpmath.pi) * mpmath.factorial(l - m)
                            / mpmath.factorial(l + m)) * mpmath.legenp(l, m, x)
        v = oracle.lambda_column(m, l, th)[0][l]
        scale = max(abs(float(exact)), 1e-3 * math.sqrt((2 * l + 1) / (4 * math.pi)) / math.sqrt(l))
        assert abs(v - float(exact)) <= 1e-11 * scale, (l, m, th, v, float(exact))


# --------------------------------------------------- x^2 recurrence

def test_auxdata_recurrence_reproduces_lambda_rows(oracle):
    """The reference's payload self-check (wavemoth/lib.pyx:470-485): the
    x^2 recurrence from two seed rows reproduces the later rows."""
    O = oracle
    N = 2048
    thetas = O.ring_thetas_positive(N)[::97]
    x2 = np.cos(thetas) ** 2
    for m, odd in [(0, 0), (0, 1), (10, 0), (300, 1), (1500, 0)]:
        L = 4096
        lam = O.lambda_matrix(m, thetas, L).T  # rows l - m
        rows = lam[odd::2]
        keep = np.all(np.abs(rows[:2]) > 1e-200, axis=0)
        nk = rows.shape[0]
        a = np.zeros((nk, 2))
        a[-1, 0] = 1.0
        y = O.legendre_transform(m, m + odd, a, x2[keep], rows[0, keep], rows[1, keep])
        err = np.linalg.norm(y[:, 0] - rows[-1, keep]) / np.linalg.norm(rows[-1, keep])
        # the reference raises above 1e-9 (wavemoth/lib.pyx:479-485)
        assert err < 1e-10, (m, odd, err)


def test_auxdata_planted_sign_error_is_caught(oracle):
    """A flipped sign in alpha (a plausible restatement mistake) must break
    the check above by many orders of magnitude."""
    O = oracle
    thetas = O.ring_thetas_positive(64)[5:20]
    x2 = np.cos(thetas) ** 2
    m, L = 3, 128
    rows = O.lambda_matrix(m, thetas, L).T[0::2]
    aux = O.auxdata(m, m, rows.shape[0]).copy()
    aux[0::3] *= -1.0
    a = np.zeros((rows.shape[0], 2))
    a[-1, 0] = 1
    y = O.legendre_transform(m, m, a, x2, rows[0], rows[1], aux=aux)
    assert np.linalg.norm(y[:, 0] - rows[-1]) / np.linalg.norm(rows[-1]) > 1e-2


@pytest.mark.parametrize("nvecs", [2, 4, 6, 10, 12, 22])
def test_legendre_transform_vs_dense(oracle, nvecs):
    """wavemoth/test/test_legendre_transforms.py:22-72 restated: Nside=2048,
    m=10, l from m+200, rings from 340, against np.dot(a.T, P).T."""
    O = oracle
    Nside, ixmin, m = 2048, 340, 10
    for nx in [1, 2, 3, 7, 11]:
        for nk in [2, 3, 4, 7, 11, 40, 44]:
            lmin = m + 200
            lstop = lmin + 2 * nk
            nodes = O.ring_thetas_positive(Nside)[ixmin:ixmin + nx]
            P = O.lambda_matrix(m, nodes, lstop, eps=1e-30).T[(lmin - m):(lstop - m):2, :]
            x2 = np.cos(nodes) ** 2
            a = np.arange(nk * nvecs, dtype=np.double).reshape(nk, nvecs)
            y0 = np.dot(a.T, P).T
            y = O.legendre_transform(m, lmin, a, x2, P[0].copy(), P[1].copy())
            np.testing.assert_allclose(y, y0, rtol=1e-11, atol=1e-9 * np.abs(y0).max())


# ---------------------------------------------------------------- grid

@pytest.mark.parametrize("N", [1, 2, 4, 16, 2048])
def test_grid_invariants(oracle, N):
    """wavemoth/test/test_healpix.py:26-33: 4N-1 rings, 4 pixels at the
    poles, 4N at the equator, 12 N^2 in total, symmetric."""
    phi0s, offs = oracle.grid(N)
    counts = np.diff(offs)
    assert counts.shape == (4 * N - 1,)
    assert counts[0] == 4 and counts[-1] == 4 and counts[2 * N - 1] == 4 * N
    assert counts.sum() == 12 * N * N
    assert np.array_equal(counts, counts[::-1])


@pytest.mark.parametrize("N", [2, 4, 8, 64])
def test_grid_phi0_is_healpix(oracle, N):
    """phi0 of every ring equals HEALPix's half-pixel offset convention
    (even Nside): cap ring i -> pi/(4i), belt ring i -> pi/(4N) when
    (i - N + 1) is odd, else 0."""
    phi0s, _ = oracle.grid(N)
    want = np.zeros(4 * N - 1)
    for idx in range(4 * N - 1):
        i = idx + 1
        if i < N:
            want[idx] = math.pi / (4 * i)
        elif i > 3 * N:
            want[idx] = math.pi / (4 * (4 * N - i))
        else:
            want[idx] = (math.pi / (4 * N)) * ((i - N + 1) % 2)
    np.testing.assert_allclose(phi0s, want, rtol=0, atol=1e-15)


def test_ring_z_matches_healpix_formula(oracle):
    N = 8
    z = oracle.ring_z(N)
    assert z[0] == 0.0
    for r in range(2 * N):
        i = 2 * N - r
        want = 1 - i * i / (3 * N * N) if i <= N else 4 / 3 - 2 * i / (3 * N)
        assert abs(z[r] - want) < 1e-16


def _all_ring_z(z_eq_to_pole):
    """z of all 4N-1 rings, north to south, from the 2N northern values
    ordered equator -> pole."""
    north = np.asarray(z_eq_to_pole)[::-1]
    return np.concatenate([north, -north[:-1][::-1]])


def _equal_area_residual(N, z, counts):
    """HEALPix pixels all have area 4 pi / (12 N^2); a ring's centre splits
    its own pixels in half, so the spherical cap above ring i has area
    2 pi (1 - z_i) = (pi / (3 N^2)) (pixels of rings 1..i-1 + n_i / 2).
    Returns the largest deviation from that identity, in units of one
    pixel's area."""
    north_pix = np.concatenate([[0], np.cumsum(counts)[:-1]]) + counts / 2.0
    cap_area = 2 * math.pi * (1 - np.asarray(z))
    return float(np.abs(cap_area / (math.pi / (3.0 * N * N)) - north_pix).max())


@pytest.mark.parametrize("N", [1, 2, 3, 8, 64, 2048])
def test_ring_z_equal_area(oracle, N):
    """The ring z of the oracle (and the product's helper) satisfy the
    equal-area identity at every ring: this pins the cap and belt formulas
    and the equator -> pole ordering without restating them."""
    from wavemoth_amd import healpix as H
    _, offs = oracle.grid(N)
    counts = np.diff(offs).astype(np.float64)
    z = _all_ring_z(oracle.ring_z(N))
    assert _equal_area_residual(N, z, counts) < 1e-9 * N * N
    zp = np.cos(H.get_ring_thetas(N))
    assert _equal_area_residual(N, zp, counts) < 1e-9 * N * N


def test_ring_z_equal_area_catches_planted_mistakes(oracle):
    """Plausible mistakes in the ring formulas break the identity by at
    least half a pixel: cap index off by one, belt slope 1/(3N) instead of
    2/(3N), rings in pole -> equator order."""
    N = 16
    _, offs = oracle.grid(N)
    counts = np.diff(offs).astype(np.float64)
    i = np.arange(1, 2 * N + 1, dtype=np.float64)
    good = np.where(i <= N, 1 - i ** 2 / (3 * N * N), 4 / 3 - 2 * i / (3 * N))
    assert _equal_area_residual(N, _all_ring_z(good[::-1]), counts) < 1e-9
    plants = [np.where(i <= N, 1 - (i + 1) ** 2 / (3 * N * N), 4 / 3 - 2 * i / (3 * N)),
              np.where(i <= N, 1 - i ** 2 / (3 * N * N), 1 - (i - N) / (3 * N) - 1 / 3),
              good[::-1]]
    for bad in plants:
        assert _equal_area_residual(N, _all_ring_z(bad[::-1]), counts) > 0.4


@pytest.mark.parametrize("N", [3, 5, 7])
def test_grid_phi0_odd_nside_follows_wavemoth_c(oracle, N):
    """For odd Nside the reference's grid (src/wavemoth.c:1223, belt phi0
    from iring % 2) and HEALPix ((i - N + 1) mod 2, the convention of
    wavemoth/healpix.py) differ on every belt ring: the grid here follows
    wavemoth.c, whose maps it reproduces.  Cap rings agree."""
    phi0s, _ = oracle.grid(N)
    idx = np.arange(4 * N - 1)
    i = idx + 1
    belt = (i > N) & (i <= 3 * N)  # rings N and 3N close the caps (4N pixels)
    healpix = (math.pi / (4 * N)) * ((i - N + 1) % 2)
    wavemoth_c = (math.pi / (4 * N)) * (idx % 2)
    np.testing.assert_allclose(phi0s[belt], wavemoth_c[belt], atol=1e-16)
    assert np.all(np.abs(phi0s[belt] - healpix[belt]) > 1e-3)
    from wavemoth_amd import healpix as H
    np.testing.assert_allclose(H.get_grid_phi0s(N), phi0s, atol=1e-16)


def test_product_geometry_helpers_agree_with_oracle(oracle):
    from wavemoth_amd import healpix as H
    for N in [1, 2, 5, 16]:
        phi0s, offs = oracle.grid(N)
        assert np.array_equal(H.get_ring_offsets(N), offs)
        np.testing.assert_allclose(H.get_grid_phi0s(N), phi0s, atol=1e-16)
        np.testing.assert_allclose(np.cos(H.get_ring_thetas(N, positive_only=True)),
                                   oracle.ring_z(N), atol=2e-16)


# -------------------------------------------------------- ring stage

def test_c2r_is_fftw_semantics(oracle):
    """numpy irfft * n == the direct unnormalised c2r sum with Hermitian
    extension and the imaginary parts of DC/Nyquist ignored."""
    rng = np.random.default_rng(0)
    for n in [4, 8, 12, 20]:
        X = rng.standard_normal(n // 2 + 1) + 1j * rng.standard_normal(n // 2 + 1)
        full = np.zeros(n, complex)
        full[:n // 2 + 1] = X
        full[0] = X[0].real
        full[n // 2] = X[n // 2].real
        for k in range(1, n // 2):
            full[n - k] = np.conj(X[k])
        p = np.arange(n)
        direct = np.array([np.sum(full * np.exp(2j * np.pi * np.arange(n) * q / n)) for q in p]).real
        np.testing.assert_allclose(np.fft.irfft(X, n) * n, direct, atol=1e-12)


# -------------------------------------------------------- whole path

def _pixel_angles(N):
    from wavemoth_amd import healpix as H
    th = H.get_ring_thetas(N)
    counts = H.get_ring_pixel_counts(N)
    phi0 = H.get_grid_phi0s(N)
    thetas, phis = [], []
    for i in range(4 * N - 1):
        n = counts[i]
        thetas.append(np.full(n, th[i]))
        phis.append(phi0[i] + 2 * np.pi * np.arange(n) / n)
    return np.concatenate(thetas), np.concatenate(phis)


@pytest.mark.parametrize("N", [2, 4, 8])
def test_synthesis_vs_direct_sum(oracle, N):
    """Whole path against an independent dense sum over scipy's Y_lm at the
    pixel centres: f = sum a_l0 Y_l0 + 2 Re sum_{m>0} a_lm Y_lm."""
    from wavemoth_amd.alm import l_of_index, power_law_alm
    L = 2 * N
    K = 2
    alm = power_law_alm(L, L, K, 99)
    ls, ms = l_of_index(L, L)
    th, ph = _pixel_angles(N)
    Y = sph_harm_y(ls[None, :], ms[None, :], th[:, None], ph[:, None])
    w = np.where(ms == 0, 1.0, 2.0)
    direct = np.stack([(Y * (w * alm[:, k].real)[None]).real.sum(1)
                       - (Y.imag * (w * alm[:, k].imag)[None]).sum(1) for k in range(K)], 1)
    # m = 0 imaginary parts are zero in the generator, so the two forms agree
    mp = oracle.OraclePlan(N, L, L, K).synthesize(alm)
    assert np.linalg.norm(mp - direct) / np.linalg.norm(direct) < 1e-13


def test_single_modes_and_linearity(oracle):
    from wavemoth_amd.alm import lm_to_idx_mmajor, nalm
    N, L = 8, 16
    op = oracle.OraclePlan(N, L, L, 1)
    th, ph = _pixel_angles(N)
    na = nalm(L, L)
    a = np.zeros((na, 1), complex)
    a[0] = 1.0
    np.testing.assert_allclose(op.synthesize(a)[:, 0], 1 / math.sqrt(4 * math.pi), atol=1e-15)
    maps = []
    for (l, m, c) in [(5, 3, 0.3 + 0.7j), (16, 16, -1.1 + 0.2j), (9, 0, 2.0)]:
        a = np.zeros((na, 1), complex)
        a[lm_to_idx_mmajor(l, m, L)] = c
        mp = op.synthesize(a)[:, 0]
        Y = sph_harm_y(l, m, th, ph)
        want = (c * Y).real * (1 if m == 0 else 2)
        np.testing.assert_allclose(mp, want, atol=1e-13)
        maps.append((a, mp))
    asum = sum(x[0] for x in maps)
    np.testing.assert_allclose(op.synthesize(asum)[:, 0], sum(x[1] for x in maps), atol=1e-13)


def test_get_work_vs_dense(oracle):
    """wavemoth/test/test_wavemoth.py:204-230 restated: the Legendre-stage
    work for all (m, odd, ring) vs dense Lambda products (Nside=4, lmax=8,
    nmaps=2, the reference test's input pattern)."""
    from wavemoth_amd.alm import lm_to_idx_mmajor, nalm
    N, L, K = 4, 8, 2
    inp = np.zeros((nalm(L, L), K), complex)
    k = np.arange(1, 3)
    for l in range(L + 1):
        for m in range(l, L + 1):
            inp[lm_to_idx_mmajor(l, m, L), :] = 10 * k * np.cos(l * 0.1) * np.sqrt(m)
    inp[1, :] = (5 + 5j) * k
    inp[2, :] = 34 + 3j
    work = oracle.OraclePlan(N, L, L, K).legendre_stage(inp)
    nodes = oracle.ring_thetas_positive(N)
    work0 = np.zeros_like(work)
    idx = 0
    for m in range(L + 1):
        lam = sph_harm_y(np.arange(m, L + 1)[:, None], m, nodes[None, :], 0.0).real
        ncoef = L - m + 1
        a_l = inp[idx:idx + ncoef]
        idx += ncoef
        for odd in range(2):
            work0[m, odd] = np.dot(a_l[odd::2].T, lam[odd::2]).T
    np.testing.assert_allclose(work, work0, atol=1e-13)


def test_golden_small_fixtures_regenerate(oracle):
    """The committed fixtures are what the oracle computes today."""
    import os
    from wavemoth_amd.alm import power_law_alm
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "golden_small.npz"))
    for N in [1, 2, 4, 8]:
        for K in [1, 3]:
            key = "n%d_k%d" % (N, K)
            alm = power_law_alm(2 * N, 2 * N, K, int(g[key + "_seed"][0]))
            mp = oracle.OraclePlan(N, 2 * N, 2 * N, K).synthesize(alm)
            assert np.array_equal(mp, g[key + "_map"])

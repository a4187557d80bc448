"""The per-strip Legendre transform (wavemoth_legendre_transform, the call
wavemoth/lib.pyx:220-280 wraps), with the cases of the reference's
wavemoth/test/test_legendre_transforms.py:22-72 restated: Nside 2048 ring
nodes from index 340, m = 10, lmin = m + 200, a = arange, against the dense
product np.dot(a.T, P).T, for nvecs = 2 and many vectors, nx and nk over
the reference's lists, auxdata taken at an offset of 0 or 1 (the
reference's alignment cases).  The GPU sums also equal the oracle's scalar
kernel (src/legendre_transform.c.in:680-716, built without FMA
contraction) bit for bit."""
import numpy as np
import pytest

Nside, IXMIN, M = 2048, 340, 10
NK_BLOCK = (1024 * 4) // (4 * 8 * 2)  # LEGENDRE_TRANSFORM_WORK_SIZE // (4 * 8 * 2)


def _case(oracle, nvecs, nx, nk):
    from wavemoth_amd.healpix import get_ring_thetas
    lmin = M + 200
    lstop = lmin + 2 * nk
    nodes = get_ring_thetas(Nside, positive_only=True)[IXMIN:IXMIN + nx]
    lam = oracle.lambda_matrix(M, nodes, lstop, eps=1e-30)  # (nx, lstop - m + 1)
    P = lam.T[(lmin - M):(lstop - M):2, :].copy('C')       # (nk, nx)
    x_squared = np.cos(nodes) ** 2
    a = np.arange(nk * nvecs, dtype=np.double).reshape(nk, nvecs)
    return lmin, P, x_squared, a


def test_auxdata_matches_oracle(oracle):
    """Host arithmetic: needs no GPU."""
    from wavemoth_amd import legendre_transform_auxdata
    for m, lmin, nk in [(0, 0, 3), (0, 1, 10), (10, 210, 64), (300, 301, 500), (4095, 4095, 3),
                        (7, 7, 2000)]:
        got = legendre_transform_auxdata(m, lmin, nk)
        assert np.array_equal(got, oracle.auxdata(m, lmin, nk)), (m, lmin, nk)
    assert legendre_transform_auxdata(3, 3, 2).shape == (0,)


def test_argument_checks():
    """The reference's ValueErrors (wavemoth/lib.pyx:230-261); raised before
    any device work."""
    from wavemoth_amd import legendre_transform
    a = np.zeros((5, 2))
    x = np.zeros(3)
    with pytest.raises(ValueError, match="nonconforming"):
        legendre_transform(0, 0, a, np.zeros((3, 2)), x, np.zeros(2), np.zeros(3))
    with pytest.raises(ValueError, match="nonconforming"):
        legendre_transform(0, 0, a, np.zeros((3, 4)), x, np.zeros(3), np.zeros(3))
    with pytest.raises(ValueError, match="auxdata"):
        legendre_transform(0, 0, a, np.zeros((3, 2)), x, np.zeros(3), np.zeros(3),
                           auxdata=np.zeros(4))
    with pytest.raises(ValueError, match="divisible by 2"):
        legendre_transform(0, 0, np.zeros((5, 3)), np.zeros((3, 3)), x, np.zeros(3),
                           np.zeros(3))


def _run(oracle, nvecs, nx, nk, auxalign):
    from wavemoth_amd import legendre_transform, legendre_transform_auxdata
    lmin, P, x_squared, a = _case(oracle, nvecs, nx, nk)
    a_copy = a.copy()
    y0 = np.dot(a.T, P).T
    auxdata = np.zeros(max(3 * (nk - 2), 0) + auxalign)
    auxdata[auxalign:] = legendre_transform_auxdata(M, lmin, nk)
    ys = []
    for use_sse in (False, True):
        y = np.zeros((nx, nvecs))
        legendre_transform(M, lmin, a, y, x_squared, P[0, :].copy('C'), P[1, :].copy('C'),
                           use_sse=use_sse, auxdata=auxdata[auxalign:])
        for j in range(nvecs):
            np.testing.assert_almost_equal(y0[:, j], y[:, j])
        ys.append(y)
    assert np.all(a_copy == a)
    ref = oracle.legendre_transform(M, lmin, a, x_squared, P[0], P[1],
                                    aux=auxdata[auxalign:].copy())
    assert np.array_equal(ys[0], ref) and np.array_equal(ys[1], ref)


@pytest.mark.gpu
def test_nvec2(oracle):
    for nx in [1, 2, 3, 4, 6, 7, 10, 11]:
        for nk in [2, 3, 4, 6, 7, 10, 11]:
            for auxalign in [0, 1]:
                _run(oracle, 2, nx, nk, auxalign)


@pytest.mark.gpu
def test_multivec(oracle):
    for nx in [1, 2, 3, 4, 6, 7, 10, 11]:
        for nk in [2, 3, 4, 6, 7, 10, 11, NK_BLOCK - 2, NK_BLOCK, NK_BLOCK + 2]:
            for nvecs in [2, 4, 6, 10, 12, 22]:
                for auxalign in [0, 1]:
                    _run(oracle, nvecs, nx, nk, auxalign)


@pytest.mark.gpu
def test_device_entry_long_strip(oracle):
    """The device-pointer entry on a long strip (all 2N ring nodes of
    Nside 512 from the equator, 700 rows): equal to the oracle bit for bit."""
    import ctypes
    import torch
    from wavemoth_amd import _native, legendre_transform_auxdata
    from wavemoth_amd.healpix import get_ring_thetas
    N, m, nk, nvecs = 512, 40, 700, 6
    nodes = get_ring_thetas(N, positive_only=True)
    lmin = m + 1
    lam = oracle.lambda_matrix(m, nodes, lmin + 2 * nk, eps=1e-300)
    P = lam.T[(lmin - m)::2, :][:2].copy('C')
    x2 = np.cos(nodes) ** 2
    rng = np.random.default_rng(2)
    a = rng.standard_normal((nk, nvecs))
    abig = np.zeros((2 * nk, nvecs))
    abig[0::2] = a
    aux = legendre_transform_auxdata(m, lmin, nk)
    d = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda()
         for k, v in dict(a=abig, x2=x2, aux=aux, P=P[0], P1=P[1]).items()}
    y = torch.empty((nodes.shape[0], nvecs), dtype=torch.float64, device="cuda")
    rc = _native.lib().wavemoth_legendre_transform_device(
        nodes.shape[0], nk, nvecs, *[ctypes.c_void_p(d[k].data_ptr())
                                      for k in ("a",)], ctypes.c_void_p(y.data_ptr()),
        *[ctypes.c_void_p(d[k].data_ptr()) for k in ("x2", "aux", "P", "P1")],
        ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == 0, _native.last_error()
    torch.cuda.synchronize()
    ref = oracle.legendre_transform(m, lmin, a, x2, P[0], P[1], aux=aux)
    assert np.array_equal(y.cpu().numpy(), ref)

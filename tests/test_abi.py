"""C-ABI checks that need no GPU: the library loads, exports every symbol
include/wavemoth.h declares, and its host-side argument checks behave like
the reference's boundary (src/wavemoth.c:411-414, wavemoth/lib.pyx:125-148)."""
import ctypes
import os
import re
import struct

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "wavemoth.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(wavemoth_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from wavemoth_amd import _native
    L = _native.lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_native.SIGNATURES), set(names) ^ set(_native.SIGNATURES)


def test_nthreads_must_be_positive():
    from wavemoth_amd import _native
    L = _native.lib()
    p = L.wavemoth_plan_to_healpix(4, 8, 8, 1, 0, None, None, 0, 0, None)
    assert not p
    assert "nthreads" in _native.last_error()


@pytest.mark.parametrize("args,msg", [((4, 8, 8, 1, 1, None, None, 1, 0, None), "ordering"),
                                      ((0, 8, 8, 1, 1, None, None, 0, 0, None), "Nside"),
                                      ((8193, 8, 8, 1, 1, None, None, 0, 0, None),
                                       "Nside must be in [1, 8192]"),
                                      ((4, 8, 9, 1, 1, None, None, 0, 0, None), "mmax"),
                                      ((4, 8, 8, 0, 1, None, None, 0, 0, None), "nmaps")])
def test_plan_argument_errors_return_null(args, msg):
    from wavemoth_amd import _native
    L = _native.lib()
    assert not L.wavemoth_plan_to_healpix(*args)
    assert msg in _native.last_error()


def _write_header(path, lmax, mmax, nside):
    with open(path, "wb") as f:
        f.write(struct.pack("<3q", lmax, mmax, nside))
        f.write(b"\0" * 64)


def test_query_resourcefile(tmp_path):
    from wavemoth_amd import query_resourcefile
    fn = tmp_path / "64.dat"
    _write_header(str(fn), 128, 128, 64)
    assert query_resourcefile(str(fn)) == (64, 128)
    assert query_resourcefile(str(tmp_path / "missing.dat")) is None


def test_resource_file_mismatch_rejected_before_device_use(tmp_path):
    from wavemoth_amd import _native
    from wavemoth_amd.resource import write_resource
    L = _native.lib()
    fn = tmp_path / "r.dat"
    write_resource(str(fn), 8, 16, 16)
    assert not L.wavemoth_plan_to_healpix(4, 8, 8, 1, 1, None, None, 0, 0, str(fn).encode())
    assert "Incompatible Nside" in _native.last_error()
    assert not L.wavemoth_plan_to_healpix(8, 8, 8, 1, 1, None, None, 0, 0, str(fn).encode())
    assert "Incompatible" in _native.last_error()
    assert not L.wavemoth_plan_to_healpix(8, 16, 16, 1, 1, None, None, 0, 0,
                                          str(tmp_path / "nope.dat").encode())
    assert "Error in loading resource" in _native.last_error()
    # a header without its offset table is malformed
    _write_header(str(tmp_path / "h.dat"), 16, 16, 8)
    assert not L.wavemoth_plan_to_healpix(8, 16, 16, 1, 1, None, None, 0, 0,
                                          str(tmp_path / "h.dat").encode())
    assert "offset table past the end" in _native.last_error()


def test_shtplan_validation_mirrors_reference():
    from wavemoth_amd import ShtPlan, nalm
    N, L = 4, 8
    inp = np.zeros((nalm(L, L), 2), np.complex128)
    out = np.zeros((12 * N * N, 2))
    with pytest.raises(ValueError, match="Invalid ordering"):
        ShtPlan(N, L, L, inp, out, 'ring')
    with pytest.raises(ValueError, match="Nonconforming"):
        ShtPlan(N, L, L, inp, np.zeros((12 * N * N, 3)), 'mmajor')
    with pytest.raises(ValueError, match="Npix"):
        ShtPlan(N, L, L, inp, np.zeros((10, 2)), 'mmajor')
    with pytest.raises(ValueError):
        ShtPlan(N, L, L, inp[:-1], out, 'mmajor')


def test_alm_generator_layout():
    from wavemoth_amd.alm import l_of_index, lm_to_idx_mmajor, nalm, power_law_alm
    L, M = 40, 30
    a = power_law_alm(L, M, 3, 5)
    assert a.shape == (nalm(L, M), 3) and a.dtype == np.complex128
    ls, ms = l_of_index(L, M)
    assert lm_to_idx_mmajor(L, M, L) == a.shape[0] - 1
    assert np.all(a[ms == 0].imag == 0)
    assert np.array_equal(a, power_law_alm(L, M, 3, 5))
    flat = power_law_alm(600, 600, 2, 1, flat=True)
    ls, ms = l_of_index(600, 600)
    v = np.mean(np.abs(flat[ms > 0]) ** 2)
    assert 0.97 < v < 1.03


@pytest.mark.parametrize("reqs,msg", [([], "nreq"),
                                      ([(4, 8, 8, 1), (0, 8, 8, 1)], "Nside"),
                                      ([(4097, 8, 8, 1)], "Nside must be in [1, 4096]"),
                                      ([(4, 8, 9, 1)], "mmax"),
                                      ([(4, 8, 8, 1), (4, 8, 8, 0)], "nmaps")])
def test_batch_argument_errors_return_null(reqs, msg):
    """wavemoth_plan_batch validates every request before touching a device
    and names the offending request."""
    from wavemoth_amd import _native
    L = _native.lib()
    n = len(reqs)
    cols = [(ctypes.c_int * max(n, 1))(*[r[i] for r in reqs]) for i in range(4)]
    assert not L.wavemoth_plan_batch(n, *cols, 0)
    err = _native.last_error()
    assert msg in err
    if n > 1:
        assert "request %d" % (n - 1) in err


def test_batch_python_checks_requests():
    from wavemoth_amd import random_requests
    reqs = random_requests(96, seed=3)
    assert len(reqs) == 96
    assert {r[0] for r in reqs} <= {128, 256, 512, 1024, 2048}
    assert all(r[1] == r[2] == 2 * r[0] for r in reqs)
    assert {r[3] for r in reqs} <= {1, 3, 8}
    assert reqs == random_requests(96, seed=3)


def test_two_mixed_flags_rejected():
    """At most one WAVEMOTH_MIXED_* scheme; checked before any device work."""
    from wavemoth_amd import WAVEMOTH_MIXED_BF16X3, WAVEMOTH_MIXED_F32, _native
    L = _native.lib()
    flags = WAVEMOTH_MIXED_F32 | WAVEMOTH_MIXED_BF16X3
    assert not L.wavemoth_plan_to_healpix(4, 8, 8, 1, 1, None, None, 0, flags, None)
    assert "MIXED" in _native.last_error()


def test_noalloc_adjoint_rejects_null_arguments():
    """The allocation-free adjoint checks its arguments before any device
    call."""
    from wavemoth_amd import _native
    L = _native.lib()
    assert L.wavemoth_execute_adjoint_device_noalloc(None, None, None, None) == -1
    assert "NULL" in _native.last_error()

"""rev1 resource files (src/wavemoth.c:144-285, wavemoth/lib.pyx:397-592,
src/butterfly.h.in:28-87): the writer (wavemoth_amd.resource) against the
oracle through an independent decoder of the format written here, the
library's validating reader (wavemoth_resource_info) on good and on
corrupted files, the walk of a hand-built compressed tree, and plan
creation's use of the file (the checks that run before any device work).
CPU only."""
import ctypes
import os
import struct

import numpy as np
import pytest


# ------------------------------------------------ independent decoder

def _decode(path):
    """Parse a rev1 file with identity-root matrices the way the reference
    reads it (src/wavemoth.c:203-264 for the table, src/butterfly.c.in:
    284-445 for the matrix header and root, src/wavemoth.c:958-1009 for the
    strips).  Returns (lmax, mmax, Nside, {(m, odd): (row_start, nk,
    [strips])}) with strips (rstart, cstart, cstop, kind, arrays)."""
    buf = open(path, "rb").read()
    i64 = lambda o: struct.unpack_from("<q", buf, o)[0]
    i32 = lambda o: struct.unpack_from("<i", buf, o)[0]
    al = lambda o: (o + 15) & ~15
    arr = lambda o, n: np.frombuffer(buf, np.float64, n, o)
    lmax, mmax, Nside = i64(0), i64(8), i64(16)
    mats = {}
    for m in range(mmax + 1):
        for odd in (0, 1):
            start = i64(24 + 8 * (4 * m + 2 * odd))
            if start == 0:
                continue
            nrows, ncols = i32(start), i32(start + 4)
            first_size, heap_size = i32(start + 24), i32(start + 28)
            assert first_size == 1 and heap_size == 1
            resid = start + i64(start + 40)
            node = start + i64(start + 48)
            assert i32(node) == 0 and i32(node + 4) == ncols
            assert i64(resid) == 1
            p = al(start + i64(resid + 8))
            row_start, row_stop = i64(p), i64(p + 8)
            p += 16
            nk = row_stop - row_start
            strips = []
            if nk <= 4:
                strips.append((0, 0, ncols, "dense",
                               arr(al(p), ncols * nk).reshape((nk, ncols), order="F")))
            else:
                nstrips = i64(p)
                p = al(p + 8)
                aux = arr(p, 3 * (nk - 2))
                p += 8 * 3 * (nk - 2)
                cstart = 0
                for _ in range(nstrips):
                    rstart, cstop = i64(p), i64(p + 8)
                    p += 16
                    nx, nks = cstop - cstart, nk - rstart
                    if nks <= 4:
                        p = al(p)
                        strips.append((rstart, cstart, cstop, "dense",
                                       arr(p, nks * nx).reshape((nks, nx), order="F")))
                        p += 8 * nks * nx
                    else:
                        out = []
                        for _ in range(3):
                            p = al(p)
                            out.append(arr(p, nx))
                            p += 8 * nx
                        strips.append((rstart, cstart, cstop, "rec", (aux[3 * rstart:],) +
                                       tuple(out)))
                    cstart = cstop
                assert cstart == ncols
            mats[(m, odd)] = (row_start, nk, strips)
    return lmax, mmax, Nside, mats


def _work_from_file(path, alm):
    """work[m, odd, ring, map] from the file's strips: the reference's
    pull_a_through_legendre_block (src/wavemoth.c:958-1028) on the scalar
    transform (src/legendre_transform.c.in:680-716)."""
    lmax, mmax, Nside, mats = _decode(path)
    K = alm.shape[1]
    work = np.zeros((mmax + 1, 2, 2 * Nside, K), np.complex128)
    for (m, odd), (row_start, nk, strips) in mats.items():
        base = m * (2 * lmax - m + 3) // 2  # a_lm row of (l = m, m)
        ls = m + odd + 2 * (row_start + np.arange(nk))
        a = alm[base + ls - m]
        for rstart, c0, c1, kind, data in strips:
            if kind == "dense":
                work[m, odd, c0:c1] = data.T @ a[rstart:]
            else:
                aux, x2, P0, P1 = data
                pp, pc = P0.copy(), P1.copy()
                acc = np.outer(pp, a[rstart]) + np.outer(pc, a[rstart + 1])
                for k in range(2, nk - rstart):
                    al_, be, ga = aux[3 * (k - 2):3 * (k - 2) + 3]
                    pp, pc = pc, (x2 + al_) * be * pc + ga * pp
                    acc += np.outer(pc, a[rstart + k])
                work[m, odd, c0:c1] = acc
    return work


@pytest.mark.parametrize("N,L,M,K", [(4, 8, 8, 1), (8, 20, 12, 3), (16, 40, 40, 2),
                                     (32, 64, 50, 1)])
def test_writer_legendre_stage_matches_oracle(tmp_path, oracle, N, L, M, K):
    """The file's strips reproduce the oracle's Legendre stage (work_q)."""
    from wavemoth_amd import power_law_alm
    from wavemoth_amd.resource import resource_info, write_resource
    fn = str(tmp_path / ("%d.dat" % N))
    write_resource(fn, N, L, M)
    assert resource_info(fn)["present"] == 2 * (M + 1)
    alm = power_law_alm(L, M, K, 100 + N)
    got = _work_from_file(fn, alm)
    ref = oracle.OraclePlan(N, L, M, K).legendre_stage(alm)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 1e-13, err


def test_writer_lambda_matches_oracle(oracle):
    """The writer's Lambda generator against the oracle's Ylmgen
    restatement (itself bit-exact against the compiled Ylmgen): equal to
    1e-12 of each column's largest value, zero where the oracle's is zero
    except at entries of size ~eps."""
    from wavemoth_amd.healpix import get_ring_thetas
    from wavemoth_amd.resource import lambda_matrix
    for N, L in [(8, 40), (32, 128), (64, 300)]:
        th = get_ring_thetas(N, positive_only=True)
        for m in [0, 1, 5, L // 2, L - 1, L]:
            a = lambda_matrix(m, th, L, 1e-300)
            b = oracle.lambda_matrix(m, th, L, eps=1e-300)
            assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), (N, L, m)
            differ = (a == 0) != (b == 0)
            assert np.all(np.maximum(np.abs(a), np.abs(b))[differ] < 1e-290), (N, L, m)


def test_reader_accepts_written_file(tmp_path):
    from wavemoth_amd import query_resourcefile
    from wavemoth_amd.resource import resource_info, write_resource
    fn = str(tmp_path / "8.dat")
    write_resource(fn, 8, 20, 16)
    info = resource_info(fn)
    assert (info["lmax"], info["mmax"], info["Nside"]) == (20, 16, 8)
    assert info["present"] == 2 * 17 and info["missing"] == 0 and info["trees"] == 0
    assert info["recurrence_strips"] > 0 and info["max_depth"] == 0
    # the strips' x^2 are the rings' cos^2 theta (numpy's and libm's
    # cos(arccos z) may differ in the last bit)
    assert 0 <= info["x2_deviation_e18"] <= 1000
    assert query_resourcefile(fn) == (8, 20)


def _patched(src, dst, offset, fmt, value):
    b = bytearray(open(src, "rb").read())
    struct.pack_into(fmt, b, offset, value)
    open(dst, "wb").write(bytes(b))


def _table_start(path, m, odd):
    return struct.unpack_from("<q", open(path, "rb").read(), 24 + 8 * (4 * m + 2 * odd))[0]


@pytest.fixture(scope="module")
def good_file(tmp_path_factory):
    from wavemoth_amd.resource import write_resource
    fn = str(tmp_path_factory.mktemp("res") / "4.dat")
    write_resource(fn, 4, 10, 10)
    return fn


def test_reader_rejects_corrupted_files(good_file, tmp_path):
    from wavemoth_amd.resource import resource_info
    bad = str(tmp_path / "bad.dat")
    data = open(good_file, "rb").read()
    # truncated in the middle of the last matrix
    open(bad, "wb").write(data[:len(data) - 40])
    with pytest.raises(ValueError, match="bad \\(start, len\\)"):
        resource_info(bad)
    s0 = _table_start(good_file, 3, 1)
    for what, off, fmt, val, msg in [
            ("misaligned start", 24 + 8 * (4 * 3 + 2), "<q", s0 + 8, "bad \\(start, len\\)"),
            ("wrong row count", s0, "<i", 99, "rows, expected"),
            ("wrong column count", s0 + 4, "<i", 7, "columns, expected"),
            ("bad tree header", s0 + 28, "<i", 0, "bad tree header"),
            ("node outside", s0 + 48, "<q", 1 << 40, "tree node outside the matrix"),
            ("header mmax", 8, "<q", -1, "bad header")]:
        _patched(good_file, bad, off, fmt, val)
        with pytest.raises(ValueError, match=msg):
            resource_info(bad)
    with pytest.raises(ValueError, match="Error in loading resource"):
        resource_info(str(tmp_path / "missing.dat"))


def test_reader_rejects_bad_strip(good_file, tmp_path):
    """A strip whose cstop runs past the block's columns."""
    from wavemoth_amd.resource import resource_info
    lmax, mmax, Nside, mats = _decode(good_file)
    (m, odd), (row_start, nk, strips) = next(((k, v) for k, v in sorted(mats.items())
                                              if v[1] > 4))
    start = _table_start(good_file, m, odd)
    buf = open(good_file, "rb").read()
    p = (start + struct.unpack_from("<q", buf, start + struct.unpack_from(
        "<q", buf, start + 40)[0] + 8)[0] + 15) & ~15
    p += 16 + 8                                     # rows, nstrips
    p = (p + 15) & ~15
    p += 8 * 3 * (nk - 2)                           # aux
    bad = str(tmp_path / "bad.dat")
    _patched(good_file, bad, p + 8, "<q", 2 * Nside + 1)  # first strip's cstop
    with pytest.raises(ValueError, match="strip columns out of order"):
        resource_info(bad)


def _tree_file(path, filt_zeros):
    """A one-matrix file (lmax = mmax = 0, Nside 1: 1 row, 2 columns) whose
    matrix is a compressed tree of depth 1, serialized as
    wavemoth/butterfly.pyx:743-825 does: an inner root with two blocks
    (T, B) of rank 1 over two identity leaves of one column each, both
    residual blocks dense (1 row).  filt_zeros: zeros in the T filter
    (1 for a well-formed file)."""
    out = bytearray()
    w64 = lambda v: out.extend(struct.pack("<q", v))
    w32 = lambda v: out.extend(struct.pack("<i", v))
    pad = lambda: out.extend(b"\0" * ((16 - len(out) % 16) % 16))
    w64(0), w64(0), w64(1)                      # lmax, mmax, Nside
    table = len(out)
    for _ in range(4):
        w64(0)
    pad()
    start = len(out)
    for v in (1, 2, 1, 2):                      # nrows, ncols, k_max, nblocks_max
        w32(v)
    w64(0)                                      # element_count
    for v in (1, 3, 1, 0):                      # first_level_size, heap_size, first index
        w32(v)
    resid_slot = len(out)
    w64(0)
    heap_slots = len(out)
    w64(0), w64(0), w64(0)
    resid_pos = len(out)
    w64(2)                                      # two remainder blocks (T, B)
    offs = len(out)
    w64(0), w64(0), w64(0)
    pays = []
    for _ in range(2):
        pad()
        pays.append(len(out))
        w64(0), w64(1)                          # rows 0..1 (dense)
        pad()
        out.extend(np.array([0.5]).tobytes())   # 1 column x 1 row
    pays.append(len(out))
    nodes = []
    pad()
    nodes.append(len(out))                      # root: 2 blocks of rank 1
    w32(2), w32(1), w32(1)
    for zeros in (filt_zeros, 1):               # T, B filters over n = 2 columns
        out.extend(bytes([0] * zeros + [1] * (2 - zeros)))
        pad()
        out.extend(np.array([0.25] * ((2 - 1) * 1)).tobytes())
    for _ in range(2):                          # identity leaves of 1 column
        pad()
        nodes.append(len(out))
        w32(0), w32(1)
    end = len(out)
    struct.pack_into("<q", out, resid_slot, resid_pos - start)
    for i, n in enumerate(nodes):
        struct.pack_into("<q", out, heap_slots + 8 * i, n - start)
    for i, p in enumerate(pays):
        struct.pack_into("<q", out, offs + 8 * i, p - start)
    struct.pack_into("<q", out, table, start)
    struct.pack_into("<q", out, table + 8, end - start)
    open(path, "wb").write(bytes(out))


def test_reader_walks_compressed_tree(tmp_path):
    from wavemoth_amd.resource import resource_info
    fn = str(tmp_path / "tree.dat")
    _tree_file(fn, 1)
    info = resource_info(fn)
    assert info["trees"] == 1 and info["max_depth"] == 1 and info["missing"] == 1
    assert info["elements"] == 2 * 1 + 2 * 1  # two 1x1 residual blocks, two 1x1 interpolants
    _tree_file(fn, 2)
    with pytest.raises(ValueError, match="filter does not select k columns"):
        resource_info(fn)


def test_plan_checks_resource_before_device_use(good_file, tmp_path):
    """Plan creation reads the resource first (src/wavemoth.c:556-569):
    mismatches and broken files are refused before any device work."""
    from wavemoth_amd import _native
    L = _native.lib()
    f = good_file.encode()
    assert not L.wavemoth_plan_to_healpix(8, 10, 10, 1, 1, None, None, 0, 0, f)
    assert "Incompatible Nside" in _native.last_error()
    assert not L.wavemoth_plan_to_healpix(-1, 10, 9, 1, 1, None, None, 0, 0, f)
    assert "Incompatible mmax" in _native.last_error()
    assert not L.wavemoth_plan_to_healpix(-1, 11, 10, 1, 1, None, None, 0, 0, f)
    assert "Incompatible lmax" in _native.last_error()
    bad = str(tmp_path / "bad.dat")
    _patched(good_file, bad, _table_start(good_file, 2, 0), "<i", 99)
    assert not L.wavemoth_plan_to_healpix(4, 10, 10, 1, 1, None, None, 0, 0, bad.encode())
    assert "rows, expected" in _native.last_error()
    assert not L.wavemoth_plan_to_healpix(-1, 10, 10, 1, 1, None, None, 0, 0, None)
    assert "Invalid Nside" in _native.last_error()


def test_configured_directory_is_searched(good_file, tmp_path):
    """With wavemoth_configure(dir) and no file name, <dir>/rev1/<N>.dat is
    read (src/wavemoth.c:137-142, 266-285): a broken one there is
    refused; a missing one is not needed."""
    from wavemoth_amd import _native
    from wavemoth_amd.resource import resource_default_path
    L = _native.lib()
    d = tmp_path / "res"
    os.makedirs(d / "rev1")
    bad = resource_default_path(str(d), 4)
    _patched(good_file, bad, _table_start(good_file, 1, 1), "<i", 77)
    try:
        L.wavemoth_configure(str(d).encode())
        assert not L.wavemoth_plan_to_healpix(4, 10, 10, 1, 1, None, None, 0, 0, None)
        assert "rows, expected" in _native.last_error()
        # a mismatching file in the directory is refused like a given one
        from wavemoth_amd.resource import write_resource
        write_resource(resource_default_path(str(d), 2), 2, 10, 10)
        assert not L.wavemoth_plan_to_healpix(2, 10, 9, 1, 1, None, None, 0, 0, None)
        assert "Incompatible mmax" in _native.last_error()
    finally:
        L.wavemoth_configure(b"")

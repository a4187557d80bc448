"""rev1 resource files: the reference's precomputed Legendre matrices.

write_resource() writes the file the reference's precompute step writes
(wavemoth/lib.pyx:501-592, ResourceComputer.compute), with every (m, odd)
matrix stored uncompressed, as the butterfly serializer does for
num_levels = 0 (wavemoth/butterfly.pyx:743-825): an identity root whose one
remainder block holds the Legendre strips of every row and ring pair
(wavemoth/lib.pyx:397-487, LegendreMatrixProvider.serialize_block_payload).
The GPU path does not need the file (it generates Lambda on the device);
the file is for callers of the reference's API, which accepts one in
wavemoth_plan_to_healpix(..., resource_filename) or finds it under the
directory given to wavemoth_configure.

resource_info() validates a file through the library's reader
(wavemoth_resource_info) and returns its summary.

Layout (int64 little-endian unless noted; offsets relative to the start of
the file for the table, of the matrix inside a matrix):
  lmax, mmax, Nside, then table[4 (mmax + 1)] of (start, len) per (m, odd);
  per matrix, 16-byte aligned: int32 nrows, ncols, k_max, nblocks_max,
  int64 element_count, int32 first_level_size, heap_size, heap_first_index,
  0; residual payload header offsets; heap node offsets; the residual
  payload header (n, offsets[n + 1]); the block payload; the root node
  (int32 0, int32 ncols).
"""
import ctypes
import io
import os
import struct

import numpy as np

from ._native import last_error, lib
from .healpix import get_ring_thetas

# ------------------------------------------------------------------ Lambda

_LARGE_EXP2 = 90
_MINSCALE = -4
_MAXSCALE = 11


def lambda_matrix(m, thetas, lmax, eps=1e-300):
    """Normalized associated Legendre functions Lambda_l^m(theta) for
    l = m..lmax, shape (len(thetas), lmax - m + 1): the quantity of
    wavemoth/legendre.pyx:24-57 (Ylmgen, libpshtlight/ylmgen_c.c:643-732).
    Lambda_mm = (-1)^m sqrt(prod_{i<=m} (2i+1)/(2i)) / sqrt(4 pi) sin^m theta
    is formed in log2 space and the three-term recurrence
    Lambda_{l+1} = cos(theta) r0(l) Lambda_l - r1(l) Lambda_{l-1} runs on
    values scaled by 2^(90 s) until they reach double range; entries before
    the first l with |Lambda| > eps are 0.  Vectorized over theta."""
    thetas = np.asarray(thetas, dtype=np.float64)
    if lmax < m:
        raise ValueError("lmax < m")
    th = np.where(thetas == 0.0, 1e-16, thetas)
    th = np.where(np.abs(th - np.pi) < 1e-15, np.pi - 1e-15, th)
    cth, sth = np.cos(th), np.sin(th)
    # log2 of sqrt(prod (2i+1)/(2i)) / sqrt(4 pi)
    mf = 1.0
    for i in range(1, m + 1):
        mf *= np.sqrt((2 * i + 1.0) / (2 * i))
    logval = np.full(th.shape, np.log2(mf / np.sqrt(4 * np.pi)))
    if m > 0:
        with np.errstate(divide="ignore"):
            logval = logval + m * np.log2(sth)
    ls = np.arange(m, lmax + 1, dtype=np.float64)
    rf0 = np.sqrt(4.0 * (ls + 1) ** 2 - 1.0) / np.sqrt((ls + m + 1.0) * (ls - m + 1.0))
    rf1 = np.empty_like(rf0)
    rf1[0] = rf0[0]
    rf1[1:] = rf0[1:] / rf0[:-1]
    fbig, fsmall = 2.0 ** _LARGE_EXP2, 2.0 ** -_LARGE_EXP2
    finite = np.isfinite(logval)
    lv = np.where(finite, logval, 0.0)
    scale = np.trunc(lv / _LARGE_EXP2).astype(np.int64) - _MINSCALE
    cur = np.exp2(lv - (scale + _MINSCALE) * _LARGE_EXP2)
    if m & 1:
        cur = -cur
    cur = np.where(finite, cur, 0.0)
    prev = np.zeros_like(cur)
    scaled = scale < 0
    # thetas already in range: undo the scale once
    cf = np.exp2((np.clip(scale, 0, _MAXSCALE - _MINSCALE) + _MINSCALE) * _LARGE_EXP2)
    cur = np.where(scaled, cur, cur * cf)
    started = np.zeros(th.shape, dtype=bool)
    out = np.zeros((th.shape[0], lmax - m + 1))
    for j in range(lmax - m + 1):
        if j > 0:
            nxt = cth * cur * rf0[j - 1] - prev * rf1[j - 1]
            prev, cur = cur, nxt
            # the scaled range is left only at even offsets from m, where
            # Ylmgen checks its second recurrence value
            while j % 2 == 0:
                big = scaled & (np.abs(cur) > fbig)
                if not big.any():
                    break
                cur = np.where(big, cur * fsmall, cur)
                prev = np.where(big, prev * fsmall, prev)
                scale = scale + big
                done = big & (scale >= 0)
                if done.any():
                    f = np.exp2((np.clip(scale, 0, _MAXSCALE - _MINSCALE) + _MINSCALE)
                                * _LARGE_EXP2)
                    cur = np.where(done, cur * f, cur)
                    prev = np.where(done, prev * f, prev)
                    scaled = scaled & ~done
        started |= ~scaled & (np.abs(cur) > eps)
        out[:, j] = np.where(started, cur, 0.0)
    return out


# --------------------------------------------------------------- stripify

def _first_true(x):
    nz = np.nonzero(x)[0]
    return x.shape[0] if nz.shape[0] == 0 else nz[0]


def stripify(A, include_above=1e-30, exclude_below=1e-80, jump_treshold=10, col_divisor=6):
    """Partition the elements of a matrix whose columns grow in magnitude
    down the rows into strips (row_start, row_stop, col_start, col_stop):
    elements below exclude_below are left out, those above include_above
    kept, and a new strip starts when more than jump_treshold rows could be
    dropped, at a column divisible by col_divisor where possible.  The
    greedy algorithm of wavemoth/lib.pyx:301-368."""
    inc = np.log2(include_above)
    exc = np.log2(exclude_below)
    M = A.copy('F')
    mask = M == 0
    M[mask] = 1
    M = np.log2(np.abs(M))
    M[mask] = exc - 1
    col_starts = []
    for col in range(M.shape[1]):
        a = _first_true(M[:, col] >= exc)
        b = _first_true(M[:, col] >= inc)
        if np.any(M[a:, col] < exc):
            raise ValueError("Magnitude of column %d not increasing regularly enough in "
                             "magnitude" % col)
        col_starts.append((a, b))
    inf = M.shape[0] + 1
    max_a, min_b, start_col = -1, inf, 0
    strips = []
    col_starts.append((inf, inf))  # sentinel
    for col, (a, b) in enumerate(col_starts):
        needed = a > min_b or b < max_a
        wanted = col > 0 and (b - min_b) > jump_treshold
        if needed or (wanted and col % col_divisor == 0):
            strips.append((min_b, M.shape[0], start_col, col))
            start_col = col
            max_a, min_b = a, b
        max_a = max(a, max_a)
        min_b = min(b, min_b)
    return strips


# ------------------------------------------------------------- serializer

def _pad128(s):
    r = s.tell() % 16
    if r:
        s.write(b"\0" * (16 - r))


def _i64(s, v):
    s.write(struct.pack("<q", int(v)))


def _i32(s, v):
    s.write(struct.pack("<i", int(v)))


def _aligned(s, arr):
    _pad128(s)
    s.write(np.ascontiguousarray(arr, dtype=np.float64).tobytes(order="A"))


def auxdata(m, lmin, nk):
    """Recurrence coefficients of a parity strip (wavemoth/lib.pyx:282-290),
    from the library's host routine."""
    if nk < 3:
        return np.zeros(0)
    out = np.empty(3 * (nk - 2))
    lib().wavemoth_legendre_transform_auxdata(int(m), int(lmin), int(nk), out.ctypes.data)
    return out


def _strip_last_row(aux, x2, P0, P1, nk):
    """Last row of a strip from its start rows by the x^2 recurrence
    P_k = (x^2 + alpha_k) beta_k P_{k-1} + gamma_k P_{k-2}
    (src/legendre_transform.c.in:680-716), for the stability check of
    wavemoth/lib.pyx:470-485."""
    pp, p = P0.copy(), P1.copy()
    for k in range(2, nk):
        a, b, c = aux[3 * (k - 2):3 * (k - 2) + 3]
        pp, p = p, (x2 + a) * b * p + c * pp
    return p if nk > 1 else pp


def _block_payload(s, m, odd, thetas, nrows):
    """One remainder block over all rows and ring pairs of (m, odd):
    wavemoth/lib.pyx:397-487 (serialize_block_payload with row_start = 0,
    row_stop = nrows and every column)."""
    row_start, row_stop = 0, nrows
    ncols = thetas.shape[0]
    if ncols == 0 or row_start == row_stop:
        _pad128(s)
        _i64(s, 0)
        _i64(s, 0)
        return
    lmin = m + odd + 2 * row_start
    lmax = m + odd + 2 * (row_stop - 1)
    x_squared = np.cos(thetas) ** 2
    Lam = lambda_matrix(m, thetas, lmax, eps=1e-300).T[lmin - m::2, :]
    strips = stripify(Lam, include_above=1e-30, exclude_below=1e-250, jump_treshold=10,
                      col_divisor=6)
    min_rstart = min(r0 for r0, _, _, _ in strips)
    row_start += min_rstart
    lmin += 2 * min_rstart
    Lam = Lam[min_rstart:, :]
    strips = [(r0 - min_rstart, r1 - min_rstart, c0, c1) for r0, r1, c0, c1 in strips]
    _pad128(s)
    _i64(s, row_start)
    _i64(s, row_stop)
    nk = row_stop - row_start
    if nk <= 4:
        _aligned(s, Lam.copy('F').ravel(order='F'))
        return
    _i64(s, len(strips))
    aux = auxdata(m, lmin, nk)
    _aligned(s, aux)
    prev_stop = 0
    for rstart, rstop, cstart, cstop in strips:
        assert rstop == Lam.shape[0] and cstart == prev_stop
        prev_stop = cstop
        _i64(s, rstart)
        _i64(s, cstop)
        if rstop - rstart <= 4:
            _aligned(s, Lam[rstart:rstop, cstart:cstop].ravel(order='F'))
        else:
            L0 = Lam[rstart, cstart:cstop].copy()
            L2 = Lam[rstart + 1, cstart:cstop].copy()
            _aligned(s, x_squared[cstart:cstop])
            _aligned(s, L0)
            _aligned(s, L2)
            y = _strip_last_row(aux[3 * rstart:], x_squared[cstart:cstop], L0, L2,
                                rstop - rstart)
            ref = Lam[-1, cstart:cstop]
            err = np.linalg.norm(y - ref) / max(np.linalg.norm(y), 1e-300)
            if err > 1e-9:
                raise ArithmeticError("unstable Legendre strip (m=%d odd=%d rel err %g)"
                                      % (m, odd, err))
    assert prev_stop == Lam.shape[1]


def _matrix(s, m, odd, Nside, lmax):
    """Butterfly matrix of (m, odd) with num_levels = 0: an identity root
    and one remainder block (wavemoth/butterfly.pyx:743-825; the root's
    element_count is 0, as IdentityNode.size() returns)."""
    start = s.tell()
    assert start % 16 == 0
    thetas = get_ring_thetas(Nside, positive_only=True)
    nrows = max((lmax - m - odd) // 2 + 1, 0)
    ncols = thetas.shape[0]
    for v in (nrows, ncols, 0, 1):       # nrows, ncols, k_max, nblocks_max
        _i32(s, v)
    _i64(s, 0)                           # element_count
    for v in (1, 1, 1, 0):               # first_level_size, heap_size, heap_first_index
        _i32(s, v)
    resid_slot = s.tell()
    _i64(s, 0)
    heap_slot = s.tell()
    _i64(s, 0)
    resid_pos = s.tell()
    _i64(s, 1)                           # one remainder block
    offs_pos = s.tell()
    _i64(s, 0)
    _i64(s, 0)
    p0 = s.tell()
    _block_payload(s, m, odd, thetas, nrows)
    p1 = s.tell()
    _pad128(s)
    node_pos = s.tell()
    _i32(s, 0)                           # identity node
    _i32(s, ncols)
    end = s.tell()
    for slot, v in ((resid_slot, resid_pos - start), (heap_slot, node_pos - start),
                    (offs_pos, p0 - start), (offs_pos + 8, p1 - start)):
        s.seek(slot)
        _i64(s, v)
    s.seek(end)


def write_resource(target, Nside, lmax, mmax=None):
    """Write the rev1 resource of (Nside, lmax, mmax) to a path or a binary
    stream (wavemoth/lib.pyx:553-592).  Every (m, odd) matrix is present
    and uncompressed."""
    mmax = lmax if mmax is None else mmax
    if not (0 <= mmax <= lmax) or Nside < 1:
        raise ValueError("require Nside >= 1 and 0 <= mmax <= lmax")
    if isinstance(target, (str, bytes, os.PathLike)):
        with open(target, "wb") as f:
            return write_resource(f, Nside, lmax, mmax)
    s = target if target.seekable() else None
    if s is None:
        buf = io.BytesIO()
        write_resource(buf, Nside, lmax, mmax)
        target.write(buf.getvalue())
        return
    base = s.tell()
    if base % 16:
        raise ValueError("the stream must be 16-byte aligned")
    _i64(s, lmax)
    _i64(s, mmax)
    _i64(s, Nside)
    table = s.tell()
    for _ in range(4 * (mmax + 1)):
        _i64(s, 0)
    for m in range(mmax + 1):
        for odd in (0, 1):
            _pad128(s)
            start = s.tell()
            _matrix(s, m, odd, Nside, lmax)
            end = s.tell()
            s.seek(table + 8 * (4 * m + 2 * odd))
            _i64(s, start - base)
            _i64(s, end - start)
            s.seek(end)


def resource_default_path(directory, Nside):
    """<dir>/rev1/<Nside>.dat, where wavemoth_configure(dir) looks
    (src/wavemoth.c:137-142)."""
    return os.path.join(directory, "rev1", "%d.dat" % Nside)


INFO_KEYS = ["lmax", "mmax", "Nside", "present", "missing", "trees", "recurrence_strips",
             "dense_blocks", "elements", "max_depth", "x2_deviation_e18"]


def resource_info(filename):
    """Validate a resource file with the library's reader; returns its
    summary (INFO_KEYS) or raises ValueError with the reader's message."""
    info = (ctypes.c_int64 * len(INFO_KEYS))()
    f = filename if isinstance(filename, bytes) else os.fsencode(filename)
    if lib().wavemoth_resource_info(f, info, len(INFO_KEYS)) != 0:
        raise ValueError(last_error())
    return dict(zip(INFO_KEYS, list(info)))

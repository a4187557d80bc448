"""rev1 resource files: the reference's precomputed Legendre matrices.

write_resource() writes the file the reference's precompute step writes
(wavemoth/lib.pyx:501-592, ResourceComputer.compute).  Each (m, odd) matrix
is stored either uncompressed, as the butterfly serializer does for
num_levels = 0 (wavemoth/butterfly.pyx:743-825: an identity root whose one
remainder block holds the Legendre strips of every row and ring pair,
wavemoth/lib.pyx:397-487), or butterfly-compressed
(wavemoth/butterfly.pyx:591-720: interpolative decompositions of column
blocks, merged pairwise up a binary tree) and cut to a number of levels.
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


def _block_payload(s, m, odd, thetas, row_start, row_stop):
    """One remainder block: rows [row_start, row_stop) of the (m, odd)
    matrix over the columns (rings) whose colatitudes are `thetas`:
    wavemoth/lib.pyx:397-487 (LegendreMatrixProvider.serialize_block_payload)."""
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


# ------------------------------------------------------ butterfly compression

class Interpolation:
    """X ~= X[:, filter == 0] [I | A_ip] with the interpolated columns
    X[:, filter == 1] = X[:, filter == 0] A_ip (wavemoth/butterfly.pyx:
    InterpolationBlock).  transpose_apply maps k values onto the block's n
    columns: identity columns take them, the others A_ip^T of them."""

    def __init__(self, filter, A_ip):
        self.filter = np.ascontiguousarray(filter, dtype=np.int8)
        self.A_ip = np.asfortranarray(A_ip, dtype=np.float64)
        self.n = self.filter.shape[0]
        self.k = self.n - int(self.filter.sum())
        assert self.A_ip.shape == (self.k, self.n - self.k)

    def transpose_apply(self, x):
        y = np.empty((self.n,) + x.shape[1:], x.dtype)
        y[self.filter == 0] = x
        y[self.filter == 1] = self.A_ip.T @ x if self.k else 0
        return y

    def size(self):
        return self.A_ip.size + self.n


def interpolative_decomposition(X, eps):
    """(sorted skeleton columns, Interpolation) of X to precision eps: the
    reference's sparse_interpolative_decomposition
    (wavemoth/interpolative_decomposition.pyx:41-98) over libidlight's
    iddp_id; here scipy.linalg.interpolative's deterministic ID (rand=False,
    the same pivoted-QR interpolative decomposition) stands in for it.  An
    all-zero block has rank 0."""
    import scipy.linalg.interpolative as sli
    n = X.shape[1]
    if n == 0:
        return np.zeros(0, np.int64), Interpolation(np.zeros(0), np.zeros((0, 0)))
    if X.shape[0] == 0 or not np.any(X):
        return np.zeros(0, np.int64), Interpolation(np.ones(n), np.zeros((0, n)))
    k, idx, proj = sli.interp_decomp(np.asfortranarray(X, dtype=np.float64), eps, rand=False)
    iden, ipol = np.asarray(idx[:k]), np.asarray(idx[k:])
    proj = np.asarray(proj).reshape(k, n - k)
    A_ip = proj[:, np.argsort(ipol)] if n > k else np.zeros((k, 0))
    A_ip = A_ip[np.argsort(iden), :]
    filt = np.zeros(n, np.int8)
    filt[ipol] = 1
    return np.sort(iden), Interpolation(filt, A_ip)


class _Leaf:
    """Identity node over n columns (wavemoth/butterfly.pyx: IdentityNode)."""

    def __init__(self, n, remainders):
        self.ncols = n
        self.heights = [n]
        self.remainders = remainders   # [(row_start, row_stop, columns)]
        self.children = []

    def depth(self):
        return 0

    def size(self):
        return 0


class _Inner:
    """Butterfly node: per block pair (T, B) of its output blocks, one
    interpolation of the pair's left and right child blocks
    (wavemoth/butterfly.pyx: InnerNode)."""

    def __init__(self, pairs, children, remainders):
        self.pairs = pairs
        self.children = children
        self.heights = [x for T, B in pairs for x in (T.k, B.k)]
        self.ncols = sum(c.ncols for c in children)
        self.remainders = remainders

    def depth(self):
        return 1 + max(c.depth() for c in self.children)

    def size(self):
        return (sum(c.size() for c in self.children) +
                sum(T.size() + B.size() for T, B in self.pairs))

    def transpose_apply(self, x):
        """Outputs of the interpolation tree below this node for its input
        blocks x (concatenated in block order): the reference's
        transpose_apply_interpolations."""
        L, R = self.children
        zl, zr = [], []
        i = 0
        for (T, B), hl, hr in zip(self.pairs, L.heights, R.heights):
            t = T.transpose_apply(x[i:i + T.k])
            i += T.k
            t = t + B.transpose_apply(x[i:i + B.k])
            i += B.k
            zl.append(t[:hl])
            zr.append(t[hl:])
        out = []
        for c, z in ((L, zl), (R, zr)):
            z = np.concatenate(z) if z else np.zeros((0,) + x.shape[1:])
            out.append(z if isinstance(c, _Leaf) else c.transpose_apply(z))
        return np.concatenate(out)


def _butterfly(col, nrows, partition, A, eps):
    """Compress columns [col, partition[-1]) of A (rows 0..nrows): the
    reference's butterfly_core (wavemoth/butterfly.pyx:591-645).  Returns
    the skeleton column blocks and the node."""
    if len(partition) == 1:
        cols = np.arange(col, partition[0])
        return [A[:, cols]], _Leaf(len(cols), [(0, nrows, cols)])
    mid = len(partition) // 2
    lk, ln = _butterfly(col, nrows, partition[:mid], A, eps)
    rk, rn = _butterfly(col + ln.ncols, nrows, partition[mid:], A, eps)
    remainders, pairs, blocks = [], [], []
    for (r0, r1, lc), (_, _, rc), LX, RX in zip(ln.remainders, rn.remainders, lk, rk):
        cols = np.concatenate([lc, rc]).astype(np.int64)
        X = np.hstack([LX, RX])
        vmid = X.shape[0] // 2
        pair = []
        i = r0
        for part in (X[:vmid], X[vmid:]):
            rel, ip = interpolative_decomposition(part, eps)
            pair.append(ip)
            remainders.append((i, i + part.shape[0], cols[rel]))
            blocks.append(part[:, rel].copy())
            i += part.shape[0]
        pairs.append(tuple(pair))
    return blocks, _Inner(pairs, (ln, rn), remainders)


def _partition(ncols, chunk_size):
    """End columns of chunk_size-column blocks, padded to a power-of-two
    count by repeating ends (empty blocks): make_partition and
    pad_with_empty_columns of wavemoth/butterfly.pyx:678-706."""
    ends = list(range(chunk_size, ncols, chunk_size))
    if not ends or ends[-1] != ncols:
        ends.append(ncols)
    target = 1
    while target < len(ends):
        target *= 2
    extra = target - len(ends)
    out = []
    for e in ends:
        out.append(e)
        if extra > 0:
            out.append(e)
            extra -= 1
    return out


def compress_matrix(A, chunk_size, eps):
    """Butterfly compression of a matrix whose columns are rings
    (butterfly_compress, wavemoth/butterfly.pyx:708-720)."""
    return _butterfly(0, A.shape[0], _partition(A.shape[1], chunk_size), A, eps)[1]


def _nodes_at(node, level):
    if level == 0:
        return [node]
    return [x for c in node.children for x in _nodes_at(c, level - 1)]


def choose_levels(tree, memop_cost=20.0):
    """Compression levels to keep: the argmin over levels of interpolation
    elements plus residual cost rows x columns x (5/2 + 1) / memop_cost
    (ResourceComputer.compute_matrix, wavemoth/lib.pyx:519-534, with
    get_stats of wavemoth/butterfly.pyx:455-480)."""
    depth = tree.depth()
    nrows = sum(r1 - r0 for r0, r1, _ in tree.remainders)

    def rcost(rows, cols):
        return rows * cols * (5.0 / 2.0 + 1.0) / memop_cost

    costs = [rcost(nrows, tree.ncols)]
    for level in range(1, depth + 1):
        nodes = _nodes_at(tree, depth - level)
        res = sum(rcost(r1 - r0, len(c)) for nd in nodes for r0, r1, c in nd.remainders)
        ip = sum(nd.size() for nd in nodes)
        costs.append(res + ip)
    return int(np.argmin(costs))


def _interp_bytes(s, ip):
    s.write(ip.filter.tobytes())
    _pad128(s)
    s.write(ip.A_ip.tobytes(order="F"))


def _matrix(s, m, odd, Nside, lmax, chunk_size=None, eps=1e-10, num_levels=None,
            memop_cost=20.0):
    """Butterfly matrix of (m, odd) (wavemoth/butterfly.pyx:743-825).
    Without chunk_size, or with num_levels 0: an identity root and one
    remainder block.  Otherwise the compressed tree cut to num_levels
    levels (default: choose_levels): 2^(depth - num_levels) roots, each
    with its remainder blocks, then the heap of nodes.  The residual
    payload headers list their block offsets once (the reference writes
    its payloads several times over and points at the last copy; readers
    only follow the offsets)."""
    start = s.tell()
    assert start % 16 == 0
    thetas = get_ring_thetas(Nside, positive_only=True)
    nrows = max((lmax - m - odd) // 2 + 1, 0)
    ncols = thetas.shape[0]
    tree = None
    if chunk_size and nrows > 0:
        lam = lambda_matrix(m, thetas, m + odd + 2 * (nrows - 1), eps=1e-30).T[odd::2, :]
        tree = compress_matrix(lam, chunk_size, eps)
        levels = choose_levels(tree, memop_cost) if num_levels is None else num_levels
        levels = min(levels, tree.depth())
        if levels == 0:
            tree = None
    if tree is None:
        roots = [_Leaf(ncols, [(0, nrows, np.arange(ncols))])]
        heap = list(roots)
        k_max, nblocks_max, size = 0, 1, 0
        count = 1
    else:
        skip = tree.depth() - levels
        count = 2 ** skip
        roots = _nodes_at(tree, skip)
        heap = [None] * (2 ** (tree.depth() + 1) - count)

        def fill(node, idx):
            heap[idx - count] = node
            for j, c in enumerate(node.children):
                fill(c, 2 * idx + j)

        for i, r in enumerate(roots):
            fill(r, count + i)

        def kmax(node):
            if isinstance(node, _Leaf):
                return node.ncols
            return max([max(node.heights)] + [kmax(c) for c in node.children])

        k_max = max(kmax(r) for r in roots)
        nblocks_max = max(len(r.heights) for r in roots)
        size = tree.size()
    for v in (nrows, ncols, k_max, nblocks_max):
        _i32(s, v)
    _i64(s, size)                        # element_count
    for v in (count, len(heap), count, 0):   # first_level_size, heap_size, heap_first_index
        _i32(s, v)
    resid_slot = s.tell()
    for _ in roots:
        _i64(s, 0)
    heap_slot = s.tell()
    for _ in heap:
        _i64(s, 0)
    resid_pos, node_pos = [], []
    for r in roots:
        resid_pos.append(s.tell() - start)
        _i64(s, len(r.remainders))
        offs = s.tell()
        for _ in range(len(r.remainders) + 1):
            _i64(s, 0)
        pos = []
        for r0, r1, cols in r.remainders:
            pos.append(s.tell() - start)
            _block_payload(s, m, odd, thetas[cols], r0, r1)
        pos.append(s.tell() - start)
        end = s.tell()
        s.seek(offs)
        for v in pos:
            _i64(s, v)
        s.seek(end)
    for node in heap:
        _pad128(s)
        node_pos.append(s.tell() - start)
        if node is None or isinstance(node, _Leaf):
            _i32(s, 0)
            _i32(s, 0 if node is None else node.ncols)
        else:
            _i32(s, len(node.heights))
            s.write(np.asarray(node.heights, np.int32).tobytes())
            for T, B in node.pairs:
                _interp_bytes(s, T)
                _interp_bytes(s, B)
    end = s.tell()
    s.seek(resid_slot)
    for v in resid_pos:
        _i64(s, v)
    s.seek(heap_slot)
    for v in node_pos:
        _i64(s, v)
    s.seek(end)


def write_resource(target, Nside, lmax, mmax=None, chunk_size=None, eps=1e-10,
                   num_levels=None, memop_cost=20.0):
    """Write the rev1 resource of (Nside, lmax, mmax) to a path or a binary
    stream (wavemoth/lib.pyx:553-592).  Every (m, odd) matrix is present.
    Without chunk_size the matrices are uncompressed (num_levels = 0);
    with it each is butterfly-compressed over column blocks of chunk_size
    rings to precision eps and cut to num_levels levels (default: the
    reference's cost rule with memop_cost, as scripts/precompute.py
    defaults: chunk size 64, tolerance 1e-10, memop cost 20)."""
    mmax = lmax if mmax is None else mmax
    if not (0 <= mmax <= lmax) or Nside < 1:
        raise ValueError("require Nside >= 1 and 0 <= mmax <= lmax")
    if chunk_size is not None and chunk_size < 1:
        raise ValueError("chunk_size must be >= 1")
    opts = dict(chunk_size=chunk_size, eps=eps, num_levels=num_levels, memop_cost=memop_cost)
    if isinstance(target, (str, bytes, os.PathLike)):
        with open(target, "wb") as f:
            return write_resource(f, Nside, lmax, mmax, **opts)
    s = target if target.seekable() else None
    if s is None:
        buf = io.BytesIO()
        write_resource(buf, Nside, lmax, mmax, **opts)
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
            _matrix(s, m, odd, Nside, lmax, **opts)
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

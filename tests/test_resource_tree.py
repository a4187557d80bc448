"""Butterfly-compressed rev1 matrices (wavemoth/butterfly.pyx:537-825,
src/butterfly.c.in:200-432, src/butterfly.h.in:28-87), CPU only:
- the compressor's tree against the dense Lambda^T x it stands for;
- written files decoded by an independent walk of the format (the
  reference's transpose_apply_node in numpy) against the oracle's Legendre
  stage, at every cut of the tree (num_levels 1 .. depth, forests of
  several roots) and at the reference's cost-chosen levels;
- the library's reader on compressed files: statistics and a corrupted
  filter refused."""
import struct

import numpy as np
import pytest


# ------------------------------------------ independent tree decoder

class _File:
    def __init__(self, path):
        self.buf = open(path, "rb").read()

    def i64(self, o):
        return struct.unpack_from("<q", self.buf, o)[0]

    def i32(self, o):
        return struct.unpack_from("<i", self.buf, o)[0]

    def arr(self, o, n):
        return np.frombuffer(self.buf, np.float64, n, o)


def _al(o):
    return (o + 15) & ~15


def _pull(f, p, ncols, a_of_row):
    """A block payload (src/wavemoth.c:958-1009): Lambda_block^T a for its
    ncols columns, from the strips' start rows by the x^2 recurrence or from
    dense blocks.  a_of_row(rows) gives the a_lm rows (K maps) of matrix
    rows `rows`."""
    p = _al(p)
    row_start, row_stop = f.i64(p), f.i64(p + 8)
    p += 16
    nk = row_stop - row_start
    a = a_of_row(row_start + np.arange(nk))
    K = a.shape[1]
    out = np.zeros((ncols, K), np.complex128)
    if nk <= 4 or ncols == 0:
        D = f.arr(_al(p), ncols * nk).reshape((nk, ncols), order="F")
        return D.T @ a if nk else out
    nstrips = f.i64(p)
    p = _al(p + 8)
    aux = f.arr(p, 3 * (nk - 2))
    p += 8 * 3 * (nk - 2)
    c0 = 0
    for _ in range(nstrips):
        rstart, c1 = f.i64(p), f.i64(p + 8)
        p += 16
        nx, nks = c1 - c0, nk - rstart
        if nks <= 4:
            p = _al(p)
            D = f.arr(p, nks * nx).reshape((nks, nx), order="F")
            p += 8 * nks * nx
            out[c0:c1] = D.T @ a[rstart:] if nks else 0
        else:
            x2, P0, P1 = [], None, None
            vals = []
            for _ in range(3):
                p = _al(p)
                vals.append(f.arr(p, nx))
                p += 8 * nx
            x2, P0, P1 = vals
            au = aux[3 * rstart:]
            pp, pc = P0.copy(), P1.copy()
            acc = np.outer(pp, a[rstart]) + np.outer(pc, a[rstart + 1])
            for k in range(2, nks):
                al_, be, ga = au[3 * (k - 2):3 * (k - 2) + 3]
                pp, pc = pc, (x2 + al_) * be * pc + ga * pp
                acc += np.outer(pc, a[rstart + k])
            out[c0:c1] = acc
        c0 = c1
    assert c0 == ncols
    return out


def _apply_matrix(f, start, a_of_row, target):
    """target[ring] = (Lambda^T a)[ring] by the reference's traversal
    (src/butterfly.c.in:284-432): roots pull their blocks, inner nodes
    scatter each (T, B) input through its filter (identity columns) and
    A_ip^T (interpolated columns), summing T and B, leaves take their rows."""
    ncols = f.i32(start + 4)
    first_size, heap_size, first_idx = (f.i32(start + 24), f.i32(start + 28),
                                        f.i32(start + 32))
    resid = [start + f.i64(start + 40 + 8 * i) for i in range(first_size)]
    hbase = start + 40 + 8 * first_size
    heap = [start + f.i64(hbase + 8 * i) for i in range(heap_size)]
    node = lambda i: heap[i - first_idx]

    def visit(inode, target_start, inputs, root_idx):
        p = node(inode)
        nblocks = f.i32(p)
        p += 4
        payloads = None
        if inputs is None:
            ph = resid[root_idx]
            n = f.i64(ph)
            payloads = [start + f.i64(ph + 8 + 8 * i) for i in range(n + 1)]
        if nblocks == 0:
            n = f.i32(p)
            if inputs is None:
                target[target_start:target_start + n] = _pull(f, payloads[0], n, a_of_row)
            else:
                target[target_start:target_start + n] = inputs[0]
            return target_start + n
        heights = [f.i32(p + 4 * i) for i in range(nblocks)]
        p += 4 * nblocks
        if inputs is None:
            inputs = [_pull(f, payloads[b], heights[b], a_of_row) for b in range(nblocks)]
        lp, rp = node(2 * inode), node(2 * inode + 1)
        lh = [f.i32(lp + 4 + 4 * i) for i in range(max(f.i32(lp), 1))]
        rh = [f.i32(rp + 4 + 4 * i) for i in range(max(f.i32(rp), 1))]
        outl, outr = [], []
        for i in range(nblocks // 2):
            n = lh[i] + rh[i]
            tot = 0
            for j in range(2):
                k = heights[2 * i + j]
                mask = np.frombuffer(f.buf, np.int8, n, p)
                p = _al(p + n)
                A = f.arr(p, (n - k) * k).reshape((k, n - k), order="F")
                p += 8 * (n - k) * k
                x = inputs[2 * i + j]
                y = np.zeros((n, x.shape[1]), np.complex128)
                y[mask == 0] = x
                y[mask == 1] = A.T @ x if k else 0
                tot = tot + y
            outl.append(tot[:lh[i]])
            outr.append(tot[lh[i]:])
        mid = visit(2 * inode, target_start, outl, None)
        return visit(2 * inode + 1, mid, outr, None)

    s = 0
    for i in range(first_size):
        s = visit(first_idx + i, s, None, i)
    assert s == ncols


def _work_from_tree_file(path, alm):
    f = _File(path)
    lmax, mmax, Nside = f.i64(0), f.i64(8), f.i64(16)
    K = alm.shape[1]
    work = np.zeros((mmax + 1, 2, 2 * Nside, K), np.complex128)
    depth = {}
    for m in range(mmax + 1):
        for odd in (0, 1):
            start = f.i64(24 + 8 * (4 * m + 2 * odd))
            if start == 0:
                continue
            base = m * (2 * lmax - m + 3) // 2  # a_lm row of (l = m, m)
            rows_to_a = lambda r, base=base, odd=odd: alm[base + odd + 2 * r]
            _apply_matrix(f, start, rows_to_a, work[m, odd])
            depth[(m, odd)] = f.i32(start + 28)
    return work, depth


# --------------------------------------------------------------- tests

def test_compressor_tree_equals_dense_product():
    """The tree's transposed apply (remainder blocks' Lambda^T x, then the
    interpolations) against Lambda^T x, to the ID precision."""
    from wavemoth_amd.healpix import get_ring_thetas
    from wavemoth_amd.resource import _Inner, compress_matrix, lambda_matrix
    N, L = 32, 64
    th = get_ring_thetas(N, positive_only=True)
    rng = np.random.default_rng(5)
    for m, odd in [(0, 0), (3, 1), (20, 0), (41, 1)]:
        nrows = (L - m - odd) // 2 + 1
        lam = lambda_matrix(m, th, m + odd + 2 * (nrows - 1), eps=1e-30).T[odd::2, :]
        x = rng.standard_normal((nrows, 3))
        ref = lam.T @ x
        for cs in (4, 16):
            tree = compress_matrix(lam, cs, 1e-10)
            assert isinstance(tree, _Inner) and tree.depth() == int(np.log2(64 // cs))
            ins = np.concatenate([lam[r0:r1][:, c].T @ x[r0:r1] for r0, r1, c in tree.remainders])
            err = np.abs(tree.transpose_apply(ins) - ref).max() / np.abs(ref).max()
            assert err <= 1e-9, (m, odd, cs, err)


@pytest.mark.parametrize("N,L,M,K,cs,levels", [
    (8, 16, 16, 2, 2, 1),
    (8, 16, 16, 1, 2, 2),
    (16, 32, 32, 2, 4, 1),
    (16, 32, 24, 3, 4, 2),
    (16, 32, 32, 1, 4, 3),     # full depth: one root per matrix
    (32, 64, 64, 2, 8, 2),     # forest: 2 roots per matrix
    (32, 64, 64, 1, 8, None),  # the reference's cost rule
])
def test_compressed_file_matches_oracle(tmp_path, oracle, N, L, M, K, cs, levels):
    from wavemoth_amd import power_law_alm
    from wavemoth_amd.resource import resource_info, write_resource
    fn = str(tmp_path / "c.dat")
    write_resource(fn, N, L, M, chunk_size=cs, eps=1e-12, num_levels=levels)
    info = resource_info(fn)
    if levels is not None:
        assert info["trees"] > 0 and info["max_depth"] == levels
    alm = power_law_alm(L, M, K, 7 + N)
    got, _ = _work_from_tree_file(fn, alm)
    ref = oracle.OraclePlan(N, L, M, K).legendre_stage(alm)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err <= 1e-9, err


def test_reader_refuses_corrupted_filter(tmp_path):
    """A filter byte flipped in a compressed node changes its rank count:
    the reader refuses the file."""
    from wavemoth_amd.resource import resource_info, write_resource
    fn = str(tmp_path / "c.dat")
    write_resource(fn, 8, 16, 16, chunk_size=2, eps=1e-12, num_levels=2)
    f = _File(fn)
    start = f.i64(24)                   # matrix (0, 0)
    first_size = f.i32(start + 24)
    node = start + f.i64(start + 40 + 8 * first_size)   # the first heap node (a root)
    nblocks = f.i32(node)
    assert nblocks > 0
    filt = node + 4 + 4 * nblocks
    data = bytearray(f.buf)
    data[filt] ^= 1
    open(fn, "wb").write(bytes(data))
    with pytest.raises(ValueError, match="filter does not select k columns"):
        resource_info(fn)


def test_hand_built_tree_known_answer(tmp_path):
    """The hand-built depth-1 tree (lmax = mmax = 0, Nside 1): both
    remainder blocks are 0.5 a_00, both filters keep column 0 and
    interpolate column 1 with 0.25, so ring 0 gets a_00 and ring 1
    0.25 a_00 (the odd matrix is missing: zero)."""
    from test_resource import _tree_file
    fn = str(tmp_path / "t.dat")
    _tree_file(fn, 1)
    a = np.array([[2.0 - 1.0j, 0.5 + 3.0j]])
    work, depth = _work_from_tree_file(fn, a)
    assert depth == {(0, 0): 3}
    assert np.allclose(work[0, 0], [a[0], 0.25 * a[0]], rtol=0, atol=1e-15)
    assert not work[0, 1].any()

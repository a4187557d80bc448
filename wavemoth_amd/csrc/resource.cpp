// rev1 resource files: mapping and validation (host code).
//
// The reference maps the file and trusts it (src/wavemoth.c:203-264); its
// butterfly apply then walks each matrix with pointer arithmetic
// (src/butterfly.c.in:284-432) and the strip decoder reads the payloads
// (src/wavemoth.c:958-1009).  Here the same walk runs once, at load time,
// with every offset and size checked against the matrix's extent, so a
// truncated or malformed file is refused with a message instead of being
// read out of bounds.
#include "wm_resource.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstring>

namespace wm {
void set_err(const std::string &s);
}

namespace wm {

Resource::~Resource() {
  if (base) munmap((void *)base, size);
}

namespace {

struct MatrixWalk {
  Resource *res;
  int m, odd;
  int64_t lo, hi;  // file extent of the matrix
  int64_t nrows = 0, ncols = 0;
  int64_t heap_first = 0;
  std::vector<int64_t> heap, resid;
  std::string err;

  bool fail(const std::string &msg) {
    err = "resource " + res->path + ": matrix m=" + std::to_string(m) +
          " odd=" + std::to_string(odd) + ": " + msg;
    return false;
  }
  bool rd32(int64_t &pos, int64_t *v) {
    if (pos < lo || pos + 4 > hi) return fail("read past the matrix end");
    int32_t x;
    std::memcpy(&x, res->base + pos, 4);
    *v = x;
    pos += 4;
    return true;
  }
  bool rd64(int64_t &pos, int64_t *v) {
    if (pos < lo || pos + 8 > hi) return fail("read past the matrix end");
    std::memcpy(v, res->base + pos, 8);
    pos += 8;
    return true;
  }
  // 16-byte aligned array of n doubles (the file is mapped page-aligned,
  // so file offsets and addresses share their alignment)
  bool arr(int64_t &pos, int64_t n, int64_t end, int64_t *off) {
    pos = (pos + 15) & ~(int64_t)15;
    if (n < 0 || pos + 8 * n > end) return fail("array past its payload");
    *off = pos;
    pos += 8 * n;
    return true;
  }
  bool node_at(int64_t inode, int64_t *pos) {
    const int64_t i = inode - heap_first;
    if (i < 0 || i >= (int64_t)heap.size()) return fail("tree node outside the heap table");
    *pos = heap[i];
    return true;
  }

  // A block payload (the Legendre rows of one remainder block) spanning
  // [pos, end), over `cols` columns: rings col0.. of q when chunk < 0 (an
  // identity root), else rows chunk.. of the chunk buffer.
  bool block_payload(int64_t pos, int64_t end, int64_t cols, int64_t col0, int64_t chunk) {
    if (pos < lo || end > hi || pos > end) return fail("block payload outside the matrix");
    int64_t r0, r1;
    pos = (pos + 15) & ~(int64_t)15;
    if (pos + 16 > end) return fail("block payload too short");
    if (!rd64(pos, &r0) || !rd64(pos, &r1)) return false;
    if (r0 < 0 || r1 < r0 || r1 > nrows) return fail("block rows outside the matrix");
    const int64_t nk = r1 - r0;
    auto strip = [&](int64_t c0, int64_t c1) {
      ResourceStrip s;
      s.m = m;
      s.odd = odd;
      s.row0 = r0;
      s.nk = nk;
      s.cstart = chunk < 0 ? col0 + c0 : c0;
      s.cstop = chunk < 0 ? col0 + c1 : c1;
      s.chunk = chunk < 0 ? -1 : chunk + c0;
      return s;
    };
    int64_t off;
    if (nk <= 4 || cols == 0) {
      if (!arr(pos, cols * nk, end, &off)) return false;
      res->dense_blocks += 1;
      res->elements += cols * nk;
      if (cols > 0) {
        // (a block of no rows writes zeros, as the reference's dgemm over
        // an empty inner dimension does)
        ResourceStrip s = strip(0, cols);
        s.dense_off = off;
        res->strips.push_back(s);
      }
      return true;
    }
    int64_t nstrips, aux;
    if (!rd64(pos, &nstrips)) return false;
    if (nstrips < 1 || nstrips > cols) return fail("bad strip count");
    if (!arr(pos, 3 * (nk - 2), end, &aux)) return false;
    int64_t cstart = 0;
    for (int64_t i = 0; i < nstrips; ++i) {
      int64_t rstart, cstop;
      if (!rd64(pos, &rstart) || !rd64(pos, &cstop)) return false;
      // rstart == nk: columns negligible on every row, a strip of no rows
      // (the reference's dgemm over no rows writes zeros)
      if (rstart < 0 || rstart > nk) return fail("strip start row outside the block");
      if (cstop < cstart || cstop > cols) return fail("strip columns out of order");
      const int64_t nx = cstop - cstart, nks = nk - rstart;
      ResourceStrip s = strip(cstart, cstop);
      s.rstart = rstart;
      s.aux_off = aux;
      if (nks <= 4) {
        if (!arr(pos, nks * nx, end, &s.dense_off)) return false;
      } else {
        if (!arr(pos, nx, end, &s.x2_off) || !arr(pos, nx, end, &s.p0_off) ||
            !arr(pos, nx, end, &s.p1_off))
          return false;
        res->recurrence_strips += 1;
      }
      res->elements += nks * nx;
      if (nx > 0) res->strips.push_back(s);
      cstart = cstop;
    }
    if (cstart != cols) return fail("strips do not cover every column");
    return true;
  }

  int64_t alloc_chunk(int64_t rows) {
    const int64_t r = res->chunk_rows;
    res->chunk_rows += rows;
    return r;
  }

  // The traversal of src/butterfly.c.in:284-391 (transpose_apply_node),
  // checking every offset and recording the work instead of applying it.
  // Roots pull their input blocks from the residual payloads (into chunks,
  // or straight into q for an identity root); inner nodes read one
  // interpolation block per input block and write their children's input
  // blocks (q rings for leaf children); leaves consume n columns of the
  // target.  in_base: the chunk rows of a non-root node's input blocks.
  bool walk(int64_t inode, bool is_root, int root_idx, int64_t target_start,
            int64_t *target_stop, int depth, const std::vector<int64_t> &in_base) {
    if (depth > 40) return fail("tree too deep");
    int64_t pos, nblocks;
    if (!node_at(inode, &pos) || !rd32(pos, &nblocks)) return false;
    if (nblocks < 0 || (nblocks > 0 && (nblocks & (nblocks - 1)) != 0) ||
        nblocks > (int64_t)1 << 24)
      return fail("bad block count");
    std::vector<int64_t> payloads;
    if (is_root) {
      int64_t ph = resid[root_idx], n;
      if (!rd64(ph, &n)) return false;
      if (!(n == nblocks || (nblocks == 0 && n == 1)))
        return fail("residual payload count does not match the root");
      payloads.resize(n + 1);
      for (int64_t i = 0; i <= n; ++i) {
        int64_t o;
        if (!rd64(ph, &o)) return false;
        payloads[i] = lo + o;
        if (payloads[i] < lo || payloads[i] > hi || (i > 0 && payloads[i] < payloads[i - 1]))
          return fail("residual payload offsets out of order");
      }
    } else if ((int64_t)in_base.size() != (nblocks == 0 ? 1 : nblocks)) {
      return fail("node inputs do not match its blocks");
    }
    if (nblocks == 0) {
      int64_t n;
      if (!rd32(pos, &n)) return false;
      if (n < 0 || target_start + n > ncols) return fail("leaf wider than the matrix");
      *target_stop = target_start + n;
      if (is_root && !block_payload(payloads[0], payloads[1], n, target_start, -1)) return false;
      res->max_depth = std::max<int64_t>(res->max_depth, depth);
      return true;
    }
    std::vector<int64_t> heights(nblocks);
    for (int64_t i = 0; i < nblocks; ++i)
      if (!rd32(pos, &heights[i])) return false;
    int64_t lpos, rpos, nl, nr;
    if (!node_at(2 * inode, &lpos) || !node_at(2 * inode + 1, &rpos)) return false;
    if (!rd32(lpos, &nl) || !rd32(rpos, &nr)) return false;
    if (!((nl == 0 && nr == 0) || (nl == nblocks / 2 && nr == nblocks / 2)))
      return fail("children do not match the node's blocks");
    if (nl == 0 && nblocks != 2) return fail("leaf children under more than two blocks");
    const bool leaves = nl == 0;
    std::vector<int64_t> base(nblocks);
    int64_t input_pos = 0;
    for (int64_t b = 0; b < nblocks; ++b) {
      if (heights[b] < 0) return fail("negative block height");
      if (is_root) {
        base[b] = alloc_chunk(heights[b]);
        if (!block_payload(payloads[b], payloads[b + 1], heights[b], input_pos, base[b]))
          return false;
        input_pos += heights[b];
      } else {
        base[b] = in_base[b];
      }
    }
    std::vector<int64_t> lbase, rbase;
    int64_t ring = target_start;
    for (int64_t i = 0; i < nblocks / 2; ++i) {
      int64_t hl, hr;
      int64_t lp = lpos + 4 * i, rp = rpos + 4 * i;
      if (!rd32(lp, &hl) || !rd32(rp, &hr)) return false;
      const int64_t n = hl + hr;
      if (hl < 0 || hr < 0) return fail("negative child block height");
      ResourceInterp ip;
      ip.m = m;
      ip.odd = odd;
      ip.depth = depth;
      ip.nl = hl;
      ip.nr = hr;
      ip.kT = heights[2 * i];
      ip.kB = heights[2 * i + 1];
      ip.inT = base[2 * i];
      ip.inB = base[2 * i + 1];
      ip.to_rings = leaves;
      if (leaves) {
        if (ring + n > ncols) return fail("leaf wider than the matrix");
        ip.outL = ring;
        ip.outR = ring + hl;
        ring += n;
      } else {
        ip.outL = alloc_chunk(hl);
        ip.outR = alloc_chunk(hr);
        lbase.push_back(ip.outL);
        rbase.push_back(ip.outR);
      }
      for (int j = 0; j < 2; ++j) {
        const int64_t k = heights[2 * i + j];
        if (k > n) return fail("interpolation block rank out of range");
        if (pos < lo || pos + n > hi) return fail("interpolation filter past the matrix end");
        int64_t zeros = 0;
        for (int64_t u = 0; u < n; ++u) zeros += res->base[pos + u] == 0;
        if (zeros != k) return fail("interpolation filter does not select k columns");
        (j ? ip.maskB : ip.maskT) = pos;
        pos += n;
        int64_t off;
        if (!arr(pos, (n - k) * k, hi, &off)) return false;
        (j ? ip.ipB : ip.ipT) = off;
        res->elements += (n - k) * k;
      }
      res->interps.push_back(ip);
      res->interp_depths = std::max(res->interp_depths, depth + 1);
    }
    int64_t mid;
    if (leaves) {
      std::vector<int64_t> one(1);
      if (!walk(2 * inode, false, 0, target_start, &mid, depth + 1, one)) return false;
      if (!walk(2 * inode + 1, false, 0, mid, target_stop, depth + 1, one)) return false;
      if (*target_stop != ring) return fail("leaves do not match their parent's blocks");
      return true;
    }
    if (!walk(2 * inode, false, 0, target_start, &mid, depth + 1, lbase)) return false;
    if (!walk(2 * inode + 1, false, 0, mid, target_stop, depth + 1, rbase)) return false;
    return true;
  }

  bool run(int64_t expect_rows, int64_t expect_cols) {
    int64_t pos = lo, k_max, nblocks_max, elems, first_size, heap_size, first_index, zero;
    if (!rd32(pos, &nrows) || !rd32(pos, &ncols) || !rd32(pos, &k_max) ||
        !rd32(pos, &nblocks_max) || !rd64(pos, &elems) || !rd32(pos, &first_size) ||
        !rd32(pos, &heap_size) || !rd32(pos, &first_index) || !rd32(pos, &zero))
      return false;
    if (nrows != expect_rows)
      return fail("has " + std::to_string(nrows) + " rows, expected " +
                  std::to_string(expect_rows));
    if (ncols != expect_cols)
      return fail("has " + std::to_string(ncols) + " columns, expected " +
                  std::to_string(expect_cols));
    if (first_size < 1 || heap_size < first_size || heap_size > (int64_t)1 << 24 ||
        first_index != first_size)
      return fail("bad tree header");
    resid.resize(first_size);
    heap.resize(heap_size);
    for (auto &o : resid) {
      if (!rd64(pos, &o)) return false;
      o += lo;
      if (o < lo || o >= hi) return fail("residual payload header outside the matrix");
    }
    for (auto &o : heap) {
      if (!rd64(pos, &o)) return false;
      o += lo;
      if (o < lo || o >= hi) return fail("tree node outside the matrix");
    }
    heap_first = first_index;
    int64_t start = 0;
    const std::vector<int64_t> none;
    for (int i = 0; i < first_size; ++i) {
      int64_t stop;
      if (!walk(first_index + i, true, i, start, &stop, 0, none)) return false;
      start = stop;
    }
    if (start != ncols) return fail("roots do not cover every column");
    if (heap_size > 1) res->trees += 1;
    return true;
  }
};

double healpix_ring_x(int N, int r) {
  // z of ring pair r (r = 0 the equator) and x = cos(arccos z), as the
  // reference's writer forms it (wavemoth/lib.pyx:382-383 from
  // wavemoth/healpix.py:5-18)
  const int i = 2 * N - r;
  const double z = (i <= N) ? 1.0 - (double)i * (double)i / (3.0 * (double)N * (double)N)
                            : 4.0 / 3.0 - 2.0 * (double)i / (3.0 * (double)N);
  return std::cos(std::acos(z));
}

}  // namespace

std::string resource_default_path(const std::string &dir, int Nside) {
  return dir + "/rev1/" + std::to_string(Nside) + ".dat";
}

Resource *resource_open(const char *path) {
  const int fd = ::open(path, O_RDONLY);
  if (fd < 0) {
    set_err(std::string("Error in loading resource ") + path);
    return nullptr;
  }
  struct stat st;
  if (fstat(fd, &st) != 0 || st.st_size < 24) {
    ::close(fd);
    set_err(std::string("Error in loading resource ") + path + ": shorter than its header");
    return nullptr;
  }
  void *mp = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
  ::close(fd);
  if (mp == MAP_FAILED) {
    set_err(std::string("Error in loading resource ") + path + ": mmap failed");
    return nullptr;
  }
  Resource *r = new Resource();
  r->path = path;
  r->base = (const char *)mp;
  r->size = (size_t)st.st_size;
  auto bad = [&](const std::string &msg) -> Resource * {
    set_err("resource " + r->path + ": " + msg);
    delete r;
    return nullptr;
  };
  int64_t hdr[3];
  std::memcpy(hdr, r->base, sizeof(hdr));
  r->lmax = hdr[0];
  r->mmax = hdr[1];
  r->Nside = hdr[2];
  if (r->mmax < 0 || r->lmax < r->mmax || r->lmax > (1 << 20) || r->Nside < 1 ||
      r->Nside > (1 << 15))
    return bad("bad header (lmax, mmax, Nside)");
  const int64_t table_end = 24 + 32 * (r->mmax + 1);
  if ((int64_t)r->size < table_end) return bad("offset table past the end of the file");
  for (int64_t m = 0; m <= r->mmax; ++m) {
    for (int odd = 0; odd < 2; ++odd) {
      int64_t se[2];
      std::memcpy(se, r->base + 24 + 8 * (4 * m + 2 * odd), sizeof(se));
      if (se[0] == 0) {
        r->missing += 1;
        continue;
      }
      if (se[0] < table_end || se[0] % 16 != 0 || se[1] < 40 ||
          se[0] + se[1] > (int64_t)r->size)
        return bad("matrix m=" + std::to_string(m) + " odd=" + std::to_string(odd) +
                   " has a bad (start, len)");
      MatrixWalk w;
      w.res = r;
      w.m = (int)m;
      w.odd = odd;
      w.lo = se[0];
      w.hi = se[0] + se[1];
      const int64_t span = r->lmax - m - odd;
      const int64_t rows = span >= 0 ? span / 2 + 1 : 0;
      if (!w.run(rows, 2 * r->Nside)) {
        set_err(w.err);
        delete r;
        return nullptr;
      }
      r->present += 1;
    }
  }
  return r;
}

bool resource_matches(const Resource &r, int *Nside, int lmax, int mmax) {
  // src/wavemoth.c:559-569
  if (!(*Nside < 0 || r.Nside == *Nside)) {
    set_err("Incompatible Nside");
    return false;
  }
  *Nside = (int)r.Nside;
  if (r.mmax != mmax) {
    set_err("Incompatible mmax");
    return false;
  }
  if (r.lmax != lmax) {
    set_err("Incompatible lmax");
    return false;
  }
  return true;
}

double resource_x2_deviation(const Resource &r) {
  bool rings = false;
  for (const ResourceStrip &s : r.strips) rings |= s.x2_off >= 0 && s.chunk < 0;
  if (!rings) return -1.0;
  const int N = (int)r.Nside;
  std::vector<double> x2(2 * N);
  for (int i = 0; i < 2 * N; ++i) {
    const double x = healpix_ring_x(N, i);
    x2[i] = x * x;
  }
  double worst = 0.0;
  for (const ResourceStrip &s : r.strips) {
    if (s.x2_off < 0 || s.chunk >= 0) continue;
    for (int64_t c = s.cstart; c < s.cstop; ++c) {
      double v;
      std::memcpy(&v, r.base + s.x2_off + 8 * (c - s.cstart), 8);
      worst = std::max(worst, std::fabs(v - x2[c]));
    }
  }
  return worst;
}

}  // namespace wm

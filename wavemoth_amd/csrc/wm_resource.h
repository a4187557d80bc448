// rev1 resource files (the reference's precomputed Legendre matrices):
// reader and validator.  Host code only; no device calls.
//
// Format (src/wavemoth.c:203-264, wavemoth/lib.pyx:553-592,
// src/butterfly.h.in:28-87, src/butterfly.c.in:284-445):
//   int64 lmax, mmax, Nside
//   int64 table[4 (mmax + 1)]: (start, len) of the matrix of (m, odd) at
//         [4 m + 2 odd], [4 m + 2 odd + 1]; start 0 = missing
//   per matrix (16-byte aligned; offsets below are relative to its start):
//     int32 nrows, ncols, k_max, nblocks_max; int64 element_count;
//     int32 first_level_size, heap_size, heap_first_index, 0;
//     int64 residual payload headers[first_level_size];
//     int64 heap node offsets[heap_size];
//     residual payload header: int64 n, int64 offsets[n + 1] of n block
//       payloads (the Legendre strips of one block of rows x columns);
//     nodes: int32 nblocks; identity node: int32 n; inner node: int32
//       block_heights[nblocks], then per block pair (T, B) an interpolation
//       block: char filter[n] (k zeros), padding to 16, double[(n - k) k].
//   block payload: pad to 16; int64 row_start, row_stop; nk = stop - start;
//     nk <= 4 or no columns: aligned double[ncols nk] (dense, column-major);
//     else int64 nstrips, aligned double aux[3 (nk - 2)], per strip int64
//     rstart, cstop and either an aligned dense double[(nk - rstart) nx]
//     (nk - rstart <= 4) or aligned x^2[nx], P0[nx], P1[nx].
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace wm {

// One Legendre strip of a remainder block: rows rstart.. of the block
// (l = m + odd + 2 (row_start + rstart + k)), its columns [cstart, cstop).
// The strips of an identity root's block (a num_levels = 0 matrix, or a
// root that is a leaf) produce rings cstart.. of q directly (chunk < 0);
// those of a compressed root's blocks produce rows chunk.. of the chunk
// buffer, the inputs of the root's interpolations.  Offsets are byte
// offsets into the file.
struct ResourceStrip {
  int m = 0, odd = 0;
  int64_t row0 = 0;        // first row of the block payload (row_start)
  int64_t nk = 0;          // rows of the block payload (row_stop - row_start)
  int64_t rstart = 0;      // first row of the strip within the block
  int64_t cstart = 0, cstop = 0;
  int64_t chunk = -1;      // chunk row of column cstart, or -1: columns are rings
  int64_t aux_off = -1;    // 3 (nk - 2) doubles (recurrence strips)
  int64_t x2_off = -1, p0_off = -1, p1_off = -1;  // recurrence strip data
  int64_t dense_off = -1;  // dense strip or dense block: (nk - rstart) x nx, column-major
};

// One block pair (T, B) of an inner node of a compressed matrix at tree
// depth `depth` (roots 0): the transposed interpolations of the T and B
// input blocks (kT and kB chunk rows), summed, give the n = nl + nr rows of
// the pair's left and right child blocks (src/butterfly.c.in:222-235 and
// 352-379).  A child that is a leaf takes its rows straight into q (rings
// outL.. / outR..); an inner child's block is a chunk of its own.
struct ResourceInterp {
  int m = 0, odd = 0, depth = 0;
  int64_t kT = 0, kB = 0, nl = 0, nr = 0;
  int64_t inT = 0, inB = 0;      // chunk rows of the inputs
  int64_t outL = 0, outR = 0;    // chunk rows, or rings when to_rings
  bool to_rings = false;
  int64_t maskT = 0, ipT = 0, maskB = 0, ipB = 0;  // filter and A_ip offsets
};

struct Resource {
  std::string path;
  int64_t lmax = 0, mmax = 0, Nside = 0;
  const char *base = nullptr;  // mapped file
  size_t size = 0;
  // statistics
  int64_t present = 0, missing = 0, trees = 0, recurrence_strips = 0, dense_blocks = 0,
          elements = 0, max_depth = 0;
  // every remainder block's strips, and the interpolations of the
  // compressed matrices in tree order (parents before children)
  std::vector<ResourceStrip> strips;
  std::vector<ResourceInterp> interps;
  int64_t chunk_rows = 0;  // rows of the chunk buffer the interpolations use
  int interp_depths = 0;   // 1 + the deepest interpolation
  ~Resource();
};

// Maps and validates a resource file.  Returns nullptr with the error set
// (set_err) on a missing, truncated or malformed file.
Resource *resource_open(const char *path);

// Checks the header against a plan's (Nside, lmax, mmax), in the
// reference's order and messages; Nside < 0 takes the file's.  Returns
// false with the error set.
bool resource_matches(const Resource &r, int *Nside, int lmax, int mmax);

// Largest |x^2 - cos^2 theta_r| over the recurrence strips' x^2 and the
// HEALPix ring pairs (equator -> pole) of r.Nside they stand for; -1
// without strips (x^2 <= 1, so this is also a relative measure).
double resource_x2_deviation(const Resource &r);

// <dir>/rev1/<Nside>.dat (src/wavemoth.c:137-142).
std::string resource_default_path(const std::string &dir, int Nside);

}  // namespace wm

// Internal declarations shared by the HIP translation units of
// libwavemoth_hip.so.  Not part of the public ABI (include/wavemoth.h).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

namespace wm {

// Length of an l-chunk of the Legendre recurrence.  Each m's l-range is
// padded to a multiple of it; seeds and the monic rescaling live on the
// chunk grid anchored at l = m.
constexpr int kChunk = 16;
// seed_l value of a (m, ring) column whose |Lambda| never reaches the
// start threshold for l <= lmax.
constexpr int kSeedNone = 0x3fffffff;
// Largest power-of-two FFT length held in LDS by the ring kernels
// (8192 complex doubles = 128 KiB).
constexpr int kMaxFFT = 8192;
// Largest Nside of the register-FFT ring kernels (and of batch plans): belt
// rings of 4 Nside pixels are folded to 2 Nside <= kMaxFFT points, polar
// caps' split Bluestein halves to P <= kMaxFFT.
constexpr int kMaxNside = kMaxFFT / 2;
// Largest Nside of a single plan: longer rings take the four-step path.
constexpr int kMaxNsidePlan = 8192;
// LDS of one workgroup (gfx950: 160 KiB) in complex double slots
constexpr int kLdsSlots = 160 * 1024 / 16;
constexpr int kRingThreads = 512;

// Device tables of the Legendre stage, all owned by one plan.
struct LegTables {
  int N = 0, L = 0, M = 0, K = 0, R = 0;
  int RL = 0;        // rings per lane
  int KB = 0;        // maps per sweep pass (atil holds ceil(K/KB)*KB maps)
  int W = 0;         // ring chunks (waves) per m
  int64_t nlpad = 0; // padded l entries over all m
  int64_t *off = nullptr;     // [M+1] start of m's padded l-range
  int *chunk_m = nullptr;     // [nlpad/kChunk] m owning each chunk
  double *ctil = nullptr;     // [nlpad] monic coefficient (l^2-m^2)/(4l^2-1)
  double *ascale = nullptr;   // [nlpad] A_l within its chunk, 0 past lmax
  double *cscale = nullptr;   // [nlpad/kChunk] A at the chunk end
  double2 *mfac = nullptr;    // [M+1] log2 Lambda_mm/sin^m as double-double
  double *ring_x = nullptr;   // [R] cos(theta_r)
  double2 *ring_ls = nullptr; // [R] log2 sin(theta_r) as double-double
  int *seed_l = nullptr;      // [(M+1)*R] chunk-aligned start l
  double2 *seed_v = nullptr;  // [(M+1)*R] (Lambda_{l-1} a_{l-1}, Lambda_l)
  int *wstart = nullptr;      // [(M+1)*W] min seed_l over a wave's rings
  int *wseedmax = nullptr;    // [(M+1)*W] max seed_l (<= lmax) over a wave
  double2 *atil = nullptr;    // [ceil(K/KB)][nlpad][KB] prescaled a_lm (per execute)
  double2 *qbuf = nullptr;    // [R][2][K][M+1] even/odd Legendre sums (q_index)
  int q32 = 0;                // q stored as float2 at the same indices (mixed plans)
  int64_t npix = 0;
};

// q(ring pair r, parity p, map k, m) lives at
//   [r][p][k / G][m][k % G],  G = q_group(K) = min(K, 8),
// with K padded to a multiple of G.  A group of G maps of one m is one
// 16G-byte row (a full 128-byte line for G = 8), which the Legendre kernels
// write whole; the ring stage reads map k of a group with stride G, and the
// G ring workgroups of a group run side by side on one XCD so they share
// the lines in its L2.  For K = 1 the layout is [R][2][M+1].
__host__ __device__ inline int q_group(int K) { return K < 8 ? K : 8; }
__host__ __device__ inline int q_groups(int K) { return (K + q_group(K) - 1) / q_group(K); }
__host__ __device__ inline int64_t q_index(int r, int p, int k, int m, int K, int Mp1) {
  const int G = q_group(K);
  return ((((int64_t)r * 2 + p) * q_groups(K) + k / G) * Mp1 + m) * G + k % G;
}
__host__ __device__ inline int64_t q_elems(int R, int K, int Mp1) {
  return (int64_t)R * 2 * q_groups(K) * q_group(K) * Mp1;
}
// Rows of q past the last (ring pair, parity, group) that the ring kernels
// may read (and discard): a row index up to kMaxFFT past mmax stays
// inside the allocation.
constexpr int64_t kQTail = (int64_t)kMaxFFT * 8;
__host__ __device__ inline int64_t q_alloc_elems(int R, int K, int Mp1) {
  return q_elems(R, K, Mp1) + kQTail;
}

// One ring pair of the output map (northern ring + its southern mirror),
// for the ring (phase + alias + c2r) stage.
struct RingPair {
  int64_t pix_n;     // first pixel of the northern ring
  int64_t pix_s;     // first pixel of the southern ring, -1 for the equator
  int64_t chirp_off; // Bluestein chirp table offset (h entries)
  int64_t vhat_off;  // Bluestein kernel spectrum offset (P entries)
  int64_t chirp2_off; // split Bluestein (length h/2) chirp offset, -1 if none
  int64_t vhat2_off;  // its kernel spectrum offset (P2 entries)
  int P2;            // split Bluestein FFT length next_pow2(h - 1), 0 if none
  int n;             // pixels per ring
  int s;             // phi0 = pi * s / n, s in {0, 1}
  int P;             // Bluestein FFT length, 0 when n/2 is a power of two
  int r;             // ring pair index, 0 = equator
};

constexpr int kMaxLaunches = 32;

// Ring kernel of a launch group: the general LDS kernel (k_ring_pairs) or a
// register-FFT kernel of ring16.hip.
enum RingKind {
  kRingGeneral = 0,
  kRingBelt16 = 1,
  kRingCap16 = 2,
  kRingWrap16 = 3,
  kRingLong = 4  // transform longer than the LDS holds: four-step FFT through HBM
};
// A ring of the long path: ring pair index (into RingTables::pairs) and
// hemisphere (0 north, g = E + O; 1 south, g = E - O).
struct LongVec {
  int pair, hemi;
};
// From this many maps the wrapping rings (n <= mmax) take the register-FFT
// kernel; with fewer, its few long-latency jobs leave the device idle and
// the general kernel is faster.
constexpr int kRingWideMaps = 8;

struct RingTables {
  int npairs = 0;
  RingPair *pairs = nullptr;  // [npairs]
  int *order = nullptr;       // [npairs] pairs grouped by launch, costliest first
  double2 *twiddle = nullptr; // [kMaxFFT] e^{-2 pi i t / kMaxFFT}
  double2 *chirp = nullptr;
  double2 *vhat = nullptr;
  double2 *chirp2 = nullptr;  // split Bluestein tables (k_ring16_cap)
  double2 *vhat2 = nullptr;
  // launch groups (host side): order[first[i] .. first[i]+count[i]) with
  // lds[i] complex slots of dynamic LDS
  int nlaunch = 0;
  int first[kMaxLaunches] = {0};
  int count[kMaxLaunches] = {0};
  int lds[kMaxLaunches] = {0};
  // complex slots k_ring_adjoint needs for the group: the largest
  // ring_lds_slots over its pairs (a register-FFT group's lds[i] is sized
  // for the synthesis kernel's layout, not for the adjoint's LDS FFTs)
  int adj_lds[kMaxLaunches] = {0};
  int kind[kMaxLaunches] = {0};  // RingKind
  int fftP[kMaxLaunches] = {0};  // FFT length of a register-FFT group
  int adj_ok = 0;                // every pair fits k_ring_adjoint's LDS
  int adj_maxp = 0;              // its size class: >= every Bluestein P and 2 h
};

// Batched execution of many independent transforms: prescale, Legendre
// sweep and ring stage each run as one persistent work-queue kernel over
// the union of the requests' work items, in the order the host put them.
struct BatchLegItem {
  int req, m, w, kb;  // request, m, ring chunk, map block
};
struct BatchRingItem {
  int req, pair, k;  // request, ring pair, map
};
// (request, launch group of its ring tables, block of that launch) of the
// batched register-FFT ring kernels
struct BatchRing16Item {
  int req, g, b;
};
// queue heads of the sweep shapes (one persistent kernel per shape)
constexpr int kBatchCounters = 64;
constexpr int kBatchSweepCtr = 48;  // queue head of sweep shape i: kBatchSweepCtr + i
struct BatchDev {
  int nreq = 0;
  const LegTables *tabs = nullptr;      // [nreq] per-request tables (K, KB, RL, buffers)
  const RingTables *rtabs = nullptr;    // [nreq] ring tables of the request's geometry
  const double2 *const *alm = nullptr;  // [nreq] device input of each request
  double *const *map = nullptr;         // [nreq] device output of each request
  const int64_t *prep_off = nullptr;    // [nreq + 1] prefix sums of nlpad * map slots
  int64_t prep_total = 0;
  const BatchLegItem *litems = nullptr;
  int nlitems = 0;
  const BatchRingItem *ritems = nullptr;
  const BatchRing16Item *r16items = nullptr;
  const double2 *twiddle = nullptr;     // [kMaxFFT] shared twiddle table
  int *counters = nullptr;              // [kBatchCounters] work-queue heads
};

// (RL, KB) sweep shapes of the batched sweep: one persistent kernel each,
// compiled like the single-plan sweep of that shape.
bool batch_shape_supported(int RL, int KB);
int batch_shape_index(int RL, int KB);  // position in the shape list, -1 if none

hipError_t launch_batch_reset(const BatchDev &b, hipStream_t s);
hipError_t launch_batch_prep(const BatchDev &b, hipStream_t s);
// Sweep items [first, first + count) of b.litems, all of shape (RL, KB),
// through queue head kBatchSweepCtr + batch_shape_index(RL, KB).
hipError_t launch_batch_legendre(const BatchDev &b, int RL, int KB, int first, int count,
                                 int nblocks, hipStream_t s);
// Ring items [first, first + count) of b.ritems (pairs of size class maxp
// or below), one block per item, with `slots` complex slots of dynamic LDS
// per block.
hipError_t launch_batch_rings(const BatchDev &b, int maxp, int first, int count, int slots,
                              hipStream_t s);
hipError_t batch_rings_set_lds_limit(int slots);
hipError_t batch_legendre_occupancy(int RL, int KB, int *blocks_per_cu);

// Mixed-precision contraction (matrix cores) of the Legendre stage.
// scheme 0 FP32 MFMA, 1 BF16x3, 2 BF16x6, 3 FP16x3; B operand `bmat` of
// mix_bmat_elems(t, NCp) elements per part (NCp = padded real columns, a
// multiple of mix_cols_per_block()); needs t.RL == mix_rings_per_lane_table().
int mix_parts(int scheme);
int mix_cols_per_block();
int mix_rings_per_lane_table();
int64_t mix_bmat_elems(const LegTables &t, int NCp);
hipError_t mix_set_lds_limits();
// cexp: [(M+1) NCp] power-of-two exponents of the FP16 scheme's a_lm
// scales per (m, column), written by the prescale (scheme 3), unused
// otherwise.
hipError_t launch_mix_prep(const LegTables &t, const double2 *alm, void *bmat, int NCp,
                           int scheme, int64_t part_stride, int *cexp, hipStream_t s);
// dtil / etil: d_l = c~_l + c~_{l+1}, e_l = c~_l c~_{l-1} on the padded
// l-grid of ctil (FP16: alpha, beta, and gamma in gtil of the normalised
// step), with at least 4 trailing entries past nlpad.
hipError_t launch_mix(const LegTables &t, const void *bmat, int NCp, int scheme,
                      int64_t part_stride, double lam_scale, const int *cexp,
                      const double *dtil, const double *etil, const double *gtil,
                      hipStream_t s);
double mix_lambda_scale(int lmax);
// FP16x3 contraction with the Lambda operand generated in registers
// (k_mixr): its recurrence runs as four l-segments per ring, segments 1..3
// starting from checkpoints written by launch_mix_ckpt (mix_ckpt_elems(t)
// entries; after the seeds and wave bounds, again whenever they change).  atab / btab / gtab: alpha, beta,
// gamma of the normalised step (dtil / etil / gtil of the FP16 scheme).
int64_t mix_ckpt_elems(const LegTables &t);
hipError_t launch_mix_ckpt(const LegTables &t, const double *atab, const double *btab,
                           const double *gtab, double lam_scale, double2 *ckpt, hipStream_t s);
hipError_t launch_mixr(const LegTables &t, const void *bmat, int NCp, int64_t part_stride,
                       double lam_scale, const int *cexp, const double *atab,
                       const double *btab, const double *gtab, const double2 *ckpt,
                       hipStream_t s);

// Shape of the Legendre sweep for nmaps = K.
int legendre_map_block(int K);
int legendre_default_rl(int K);
int legendre_padded_maps(int K);

// One strip of a resource matrix on the device (the Legendre stage read
// from a rev1 file): rows `row`.. of (m, odd) (l = m + odd + 2 row), nks of
// them, over nx columns; offsets in doubles into the file image on the
// device: a recurrence strip has x2 / p0 / p1 (nx each) and aux (3 per row
// past the second, of the strip's rows), a dense one a column-major
// nks x nx block at dense (x2 = -1).  Column c goes to ring pair cstart + c
// of q (chunk < 0) or to row chunk + c of the chunk buffer (an input block
// of a compressed matrix's root).
struct ResStrip {
  int m = 0, odd = 0, nks = 0, cstart = 0, nx = 0;
  int64_t row = 0;
  int64_t x2 = -1, p0 = -1, p1 = -1, aux = -1, dense = -1;
  int64_t chunk = -1;
};
// q (or chunk rows) <- the Legendre stage of the strips (FP64 q; items =
// prefix[nstrips] (column, strip) pairs); (m, odd) without a matrix are
// zeroed first when zero_q is set.
hipError_t launch_res_legendre(const LegTables &t, const double2 *alm, const double *img,
                               const ResStrip *strips, const int64_t *prefix, int nstrips,
                               int64_t items, double2 *chunk, bool zero_q, hipStream_t s);

// One block pair (T, B) of an inner node of a butterfly-compressed matrix
// (src/butterfly.c.in:222-235, 352-379): its nl + nr output rows are
//   out[p] = f_T(p) + f_B(p),  f_X(p) = x_X[c]                  (code c >= 0)
//                                      = sum_r A_X[r, c] x_X[r]  (code -1 - c),
// x_X the kX chunk rows at inX, A_X the column-major kX x (n - kX)
// interpolant at ipX (doubles into the file image), codes[codeX + p] the
// position of p among the filter's identity (0) or interpolated (1)
// columns.  Rows p < nl go to outL + p, the others to outR + p - nl: rings
// of q when to_rings (leaf children), else chunk rows.
struct ResInterp {
  int m = 0, odd = 0, kT = 0, kB = 0, nl = 0, nr = 0, to_rings = 0;
  int64_t inT = 0, inB = 0, outL = 0, outR = 0;
  int64_t codeT = 0, codeB = 0, ipT = 0, ipB = 0;
};
// The interpolations ops[op_lo, op_hi) (one tree depth); prefix[i] is the
// first output row of op i in a numbering over all ops.
hipError_t launch_res_interp(const LegTables &t, const double *img, const ResInterp *ops,
                             const int64_t *prefix, const int32_t *codes, int op_lo, int op_hi,
                             int64_t item_lo, int64_t item_hi, double2 *chunk, hipStream_t s);

// Launchers (return hipError_t of the launch).
hipError_t launch_seeds(const LegTables &t, hipStream_t s);
hipError_t launch_wave_bounds(const LegTables &t, hipStream_t s);
hipError_t launch_prep_alm(const LegTables &t, const double2 *alm, hipStream_t s);
hipError_t launch_legendre(const LegTables &t, hipStream_t s);
// Bluestein tables of the lengths hlist[0..nh): chirp[choff + a] =
// e^{i pi a^2/h} (a < h), vhat[vhoff + .] = FFT_P(conj chirp, cyclic) / P.
hipError_t launch_bluestein_prep(const double2 *twiddle, double2 *chirp, double2 *vhat,
                                 const int *hlist, const int64_t *choff, const int64_t *vhoff,
                                 const int *plist, int nh, hipStream_t s);
// LDS slots of the general ring kernel for a ring (n pixels, Bluestein P or
// 0) at mmax M, staging for size class maxp; the size class of a ring (a
// power of two >= 1024, P and n).
int ring_lds_slots(int n, int P, int M, int maxp = kMaxFFT);
int ring_class(int n, int P);
hipError_t ring_set_lds_limit(int slots);
// Register-FFT ring kernels (ring16.hip) for FFT lengths P = 16 .. 4096.
bool ring16_size_ok(int P);
int ring16_lds_slots(int kind, int P);
int64_t ring16_blocks(int kind, int P, int K, int count);
hipError_t ring16_set_lds_limits();
hipError_t launch_ring16(int kind, int P, const LegTables &t, const RingTables &rt, int first,
                         int count, double *map, hipStream_t s);
hipError_t launch_batch_ring16(int kind, int P, const BatchDev &d, int first, int count,
                               hipStream_t s);
int adjoint_groups(const LegTables &t);
int adjoint_groups_within(const LegTables &t, size_t cap);
size_t adjoint_scratch_elems();
double2 *adjoint_scratch(int device);
hipError_t adjoint_set_lds_limit(int slots);
hipError_t launch_adjoint(const LegTables &t, const RingTables &rt, const double *map,
                          double2 *alm, double2 *partial, int G, hipStream_t s);
hipError_t launch_ring_synth(const LegTables &t, const RingTables &rt, double *map,
                             hipStream_t s);

}  // namespace wm

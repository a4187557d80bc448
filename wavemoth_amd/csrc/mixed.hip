// Mixed-precision Legendre stage: the contraction q = sum_l Lambda_l a_l on
// the matrix cores, the recurrence (and its seeds) in FP64.
//
// Per m the contraction is a GEMM: q[ring, col] (cols = Re/Im of the maps,
// interleaved) = sum over l of one parity of Lambda[l, ring] a[l, col].  It
// replaces the reference's register-blocked dgemm-like kernels
// (src/legendre_transform.c.in:389-451, 453-618) for many maps.
//
// A workgroup covers 64 ring pairs of one m and up to 4 column blocks of 32
// columns (16 maps each).  Two producer waves (one per parity, one ring per
// lane) step the FP64 recurrence two l at a time (the x^2 form) and write
// each 64-l super-chunk of Lambda, in the scheme's input format, into a
// double-buffered LDS tile that every column block reads; consumer waves
// (column block, parity) run the MFMAs against the B operand (the prescaled
// a_lm, pre-split at prescale time).  One workgroup barrier per
// super-chunk.  MFMA accumulates in FP32; every kFlush super-chunks the
// FP32 sums are added into FP64 accumulators (the FP64 reduction across
// l-chunks), and q is stored once at the end, rounded to FP32 for the
// mixed plans' FP32 ring stage.
//
// Schemes (products of input parts, FP32 accumulation):
//   0 FP32   v_mfma_f32_16x16x4_f32 on fp32 Lambda and a.
//   1 BF16x3 Lambda = h + l, a = h + l (bf16 parts): hh + hl + lh.
//   2 BF16x6 three bf16 parts each: hh + hm + mh + hl + mm + lh.
//   3 FP16x3 Lambda and a split into two fp16 parts each (~22 bits):
//            hh + hl + lh.  fp16's range is kept by power-of-two scales:
//            Lambda (the true value, stepped by the normalised recurrence)
//            by 2^e with |Lambda| <= sqrt((2 lmax + 1) / 4 pi), a_lm per
//            (m, column) by a power of two from its maximum over m's
//            l-range (found per execute), both undone when the FP32 sums
//            are added into FP64.
// BF16 schemes use v_mfma_f32_16x16x32_bf16, FP16x3 v_mfma_f32_16x16x32_f16.
#include <cmath>
#include <type_traits>

#include "wm_internal.h"

namespace wm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
// native 16-byte vector: arrays of it stay in registers (HIP's uint4 struct
// in an array held across a loop is kept in scratch memory)
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));

#define WM_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const WM_CONST T *cview(const T *p) {
  return (const WM_CONST T *)p;
}

constexpr int kMixCols = 32;                   // real columns per column block
constexpr int kMixCB = 4;                      // column blocks per workgroup (at most)
constexpr int kSC = 64;                        // l per super-chunk
constexpr int kKK = kSC / 2;                   // same-parity l per super-chunk
constexpr int kFlush = 2;                      // super-chunks per FP32 accumulation

template <int S>
struct MixScheme;
template <>
struct MixScheme<0> {
  static constexpr int parts = 1;
  static constexpr int nprod = 1;
  __device__ static constexpr int pa(int) { return 0; }
  __device__ static constexpr int pb(int) { return 0; }
};
template <>
struct MixScheme<1> {
  static constexpr int parts = 2;
  static constexpr int nprod = 3;
  __device__ static constexpr int pa(int i) { return i == 2 ? 1 : 0; }
  __device__ static constexpr int pb(int i) { return i == 1 ? 1 : 0; }
};
template <>
struct MixScheme<2> {
  static constexpr int parts = 3;
  static constexpr int nprod = 6;
  // (0,0) (0,1) (1,0) (0,2) (1,1) (2,0)
  __device__ static constexpr int pa(int i) { return i < 2 ? 0 : (i == 2 ? 1 : (i == 3 ? 0 : (i == 4 ? 1 : 2))); }
  __device__ static constexpr int pb(int i) { return i == 0 ? 0 : (i == 1 ? 1 : (i == 2 ? 0 : (i == 3 ? 2 : (i == 4 ? 1 : 0)))); }
};

template <>
struct MixScheme<3> {
  static constexpr int parts = 2;
  static constexpr int nprod = 3;
  __device__ static constexpr int pa(int i) { return i == 2 ? 1 : 0; }
  __device__ static constexpr int pb(int i) { return i == 1 ? 1 : 0; }
};

int mix_parts(int scheme) { return scheme == 0 ? 1 : (scheme == 2 ? 3 : 2); }

// ------------------------------------------------------------- prescale

// B operand of chunk c (16 l from m + 16 (c - off[m]/16)), parity p, the
// i-th l of that parity (l = l_c + 2 i + p), column col (= 2 map + Re/Im):
//   FP32:  bf[((c * 2 + p) * 8 + i) * NCp + col]
//   BF16:  part q at bq + q * part_stride, [((c * 2 + p) * NCp + col) * 8 + i]
// Values are A_l a_lm (the FP64 path's prescale), zero past lmax and for
// padded columns.  Three trailing zero chunks let a super-chunk of the last
// m read past its end.

// Value of B at (l = t-index j of chunk c, map k, Re/Im): A_l a_lm (FP16:
// the scaled a_lm, the producer steps the normalised Lambda), zero past
// lmax and for padded maps.  csr / csi: the FP16 scales of (m, k, Re) and
// (m, k, Im), used when scheme == 3.
__device__ __forceinline__ void mix_prep_pair(const LegTables &t, const double2 *__restrict__ alm,
                                              int64_t j, int m, int k, int scheme, double csr,
                                              double csi, double *vr, double *vi) {
  // unconditional loads (clamped l and k) so that the loads of all calls
  // of a thread are in flight together
  const double A = t.ascale[j];
  const int l = m + (int)(j - t.off[m]);
  const int lc = l <= t.L ? l : t.L, kc = k < t.K ? k : t.K - 1;
  const int64_t idx = (int64_t)m * (2 * t.L - m + 3) / 2 + (lc - m);
  const double2 a = alm[idx * t.K + kc];
  const bool on = k < t.K && A != 0.0;
  *vr = on ? (scheme == 3 ? csr : A) * a.x : 0.0;
  *vi = on ? (scheme == 3 ? csi : A) * a.y : 0.0;
}

// 16-bit parts of v: FP16 hi + lo (scheme 3), else BF16 hi, mid, lo.
// Formed from f = (float)v: the parts carry 22 (FP16) or 24 (BF16) bits,
// within f's 24, and each residual is exact in FP32, so the conversions
// stay single hardware instructions.
__device__ __forceinline__ void mix_split16(double v, int scheme, unsigned short *h0,
                                            unsigned short *h1, unsigned short *h2) {
  const float f = (float)v;
  if (scheme == 3) {
    const _Float16 a = (_Float16)f;
    const float r = f - (float)a;
    *h0 = __builtin_bit_cast(unsigned short, a);
    *h1 = __builtin_bit_cast(unsigned short, (_Float16)r);
    *h2 = 0;
  } else {
    const __bf16 a = (__bf16)f;
    const float r1 = f - (float)a;
    const __bf16 b = (__bf16)r1;
    const float r2 = r1 - (float)b;
    *h0 = __builtin_bit_cast(unsigned short, a);
    *h1 = __builtin_bit_cast(unsigned short, b);
    *h2 = __builtin_bit_cast(unsigned short, (__bf16)r2);
  }
}

// FP32: one chunk per block; thread u -> (l slot li = u / NK, map k = u % NK)
// reads one a_lm (lanes over maps: coalesced) and writes its two columns.
__global__ __launch_bounds__(256) void k_mix_prep_f32(LegTables t, const double2 *__restrict__ alm,
                                                      float *__restrict__ bf, int NCp) {
  const int64_t c = blockIdx.x;
  const int m = t.chunk_m[c];
  const int NK = NCp / 2;
  for (int u = threadIdx.x; u < kChunk * NK; u += blockDim.x) {
    const int li = u / NK, k = u - li * NK;  // li = 2 i + p
    const int i = li >> 1, p = li & 1;
    double vr, vi;
    mix_prep_pair(t, alm, c * kChunk + li, m, k, 0, 1.0, 1.0, &vr, &vi);
    *reinterpret_cast<float2 *>(bf + ((c * 2 + p) * 8 + i) * NCp + 2 * k) =
        make_float2((float)vr, (float)vi);
  }
}

// 16-bit parts: thread u -> (chunk c = blockIdx.x * CPB + u / (2 NK), parity
// p, map k) reads the 8 a_lm of (c, p, k) and writes 16 bytes per part and
// column.
// BF16 schemes (S = 1, 2).
template <int S>
__global__ __launch_bounds__(256) void k_mix_prep_16(LegTables t, const double2 *__restrict__ alm,
                                                     unsigned short *__restrict__ bq, int NCp,
                                                     int64_t part_stride, int64_t nchunks) {
  const int NK = NCp / 2, per = 2 * NK;
  const int CPB = blockDim.x / per > 0 ? blockDim.x / per : 1;
  for (int u = threadIdx.x; u < CPB * per; u += blockDim.x) {
    const int cl = u / per, r = u - cl * per;
    const int p = r / NK, k = r - p * NK;
    const int64_t c = (int64_t)blockIdx.x * CPB + cl;
    if (c >= nchunks) continue;
    const int m = t.chunk_m[c];
    unsigned short hr[3][8], hi[3][8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      double vr, vi;
      mix_prep_pair(t, alm, c * kChunk + 2 * i + p, m, k, S, 1.0, 1.0, &vr, &vi);
      mix_split16(vr, S, &hr[0][i], &hr[1][i], &hr[2][i]);
      mix_split16(vi, S, &hi[0][i], &hi[1][i], &hi[2][i]);
    }
    constexpr int nparts = MixScheme<S>::parts;
    const int64_t at = ((c * 2 + p) * NCp + 2 * k) * 8;
#pragma unroll
    for (int q = 0; q < nparts; ++q) {
      uint4 wr, wi;
      wr.x = hr[q][0] | (unsigned)hr[q][1] << 16;
      wr.y = hr[q][2] | (unsigned)hr[q][3] << 16;
      wr.z = hr[q][4] | (unsigned)hr[q][5] << 16;
      wr.w = hr[q][6] | (unsigned)hr[q][7] << 16;
      wi.x = hi[q][0] | (unsigned)hi[q][1] << 16;
      wi.y = hi[q][2] | (unsigned)hi[q][3] << 16;
      wi.z = hi[q][4] | (unsigned)hi[q][5] << 16;
      wi.w = hi[q][6] | (unsigned)hi[q][7] << 16;
      uint4 *o = reinterpret_cast<uint4 *>(bq + q * part_stride + at);
      o[0] = wr;
      o[1] = wi;
    }
  }
}

// FP16: one block per m.  The block first finds max |a_lm| of each column
// (map, Re/Im) over m's l-range, picks the power of two 2^e that puts it in
// [2^13, 2^14) (fp16's range with headroom; e = 0 for an all-zero column),
// records e in cexp[m][col] for the contraction to undo, and writes the
// scaled a_lm's two fp16 parts for m's chunks as k_mix_prep_16 does (m's
// rows are read twice; the second read mostly comes from HBM again).
__global__ __launch_bounds__(256) void k_mix_prep_f16(LegTables t, const double2 *__restrict__ alm,
                                                      unsigned short *__restrict__ bq, int NCp,
                                                      int64_t part_stride, int *__restrict__ cexp) {
  extern __shared__ unsigned long long smx[];  // [NCp]: max |a| bits, then the scale
  const int m = blockIdx.x, K = t.K, L = t.L, NK = NCp / 2;
  const int nl = L - m + 1;
  const int64_t row0 = (int64_t)m * (2 * L - m + 3) / 2;  // a_lm row of (l = m, m)
  for (int c = threadIdx.x; c < NCp; c += blockDim.x) smx[c] = 0ull;
  __syncthreads();
  const int cols = NK < 256 ? NK : 256, RP = 256 / cols;
  const int rr = threadIdx.x / cols, cc = threadIdx.x - rr * cols;
  for (int kb = 0; kb < NK; kb += cols) {
    const int k = kb + cc;
    if (rr < RP && k < K) {
      unsigned long long mr = 0ull, mi = 0ull;
      for (int l = rr; l < nl; l += RP) {
        const double2 a = alm[(row0 + l) * K + k];
        const unsigned long long br = __builtin_bit_cast(unsigned long long, fabs(a.x));
        const unsigned long long bi = __builtin_bit_cast(unsigned long long, fabs(a.y));
        mr = br > mr ? br : mr;
        mi = bi > mi ? bi : mi;
      }
      if (mr) atomicMax(&smx[2 * k], mr);
      if (mi) atomicMax(&smx[2 * k + 1], mi);
    }
  }
  __syncthreads();
  for (int c = threadIdx.x; c < NCp; c += blockDim.x) {
    const double mx = __builtin_bit_cast(double, smx[c]);
    const int e = mx > 0.0 ? 13 - ilogb(mx) : 0;
    cexp[(int64_t)m * NCp + c] = e;
    smx[c] = __builtin_bit_cast(unsigned long long, ldexp(1.0, e));
  }
  __syncthreads();
  const int nch = (L - m + kChunk) / kChunk, per = 2 * NK;
  const int64_t c0 = t.off[m] / kChunk;
  for (int u = threadIdx.x; u < nch * per; u += blockDim.x) {
    const int cl = u / per, r = u - cl * per;
    const int p = r / NK, k = r - p * NK;
    const int64_t c = c0 + cl;
    const double csr = __builtin_bit_cast(double, smx[2 * k]);
    const double csi = __builtin_bit_cast(double, smx[2 * k + 1]);
    unsigned short hr[3][8], hi[3][8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      double vr, vi;
      mix_prep_pair(t, alm, c * kChunk + 2 * i + p, m, k, 3, csr, csi, &vr, &vi);
      mix_split16(vr, 3, &hr[0][i], &hr[1][i], &hr[2][i]);
      mix_split16(vi, 3, &hi[0][i], &hi[1][i], &hi[2][i]);
    }
    const int64_t at = ((c * 2 + p) * NCp + 2 * k) * 8;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      uint4 wr, wi;
      wr.x = hr[q][0] | (unsigned)hr[q][1] << 16;
      wr.y = hr[q][2] | (unsigned)hr[q][3] << 16;
      wr.z = hr[q][4] | (unsigned)hr[q][5] << 16;
      wr.w = hr[q][6] | (unsigned)hr[q][7] << 16;
      wi.x = hi[q][0] | (unsigned)hi[q][1] << 16;
      wi.y = hi[q][2] | (unsigned)hi[q][3] << 16;
      wi.z = hi[q][4] | (unsigned)hi[q][5] << 16;
      wi.w = hi[q][6] | (unsigned)hi[q][7] << 16;
      uint4 *o = reinterpret_cast<uint4 *>(bq + q * part_stride + at);
      o[0] = wr;
      o[1] = wi;
    }
  }
}

// ------------------------------------------------------------- contraction

// Workgroup over 64 ring pairs of one m and CB column blocks of 32 columns:
// waves 0 and 1 are producers (parity 0 / 1: one ring per lane, the FP64
// recurrence, conversion to the scheme's parts, the A tile into LDS, plus
// staging of the B rows), waves 2 + 2 c + p consumers (column block c,
// parity p: MFMA over the tiles, FP32 sums, FP64 reduction).  The Lambda
// tile is produced once and read by every column block; tiles and B rows
// are double-buffered, one workgroup barrier per super-chunk.
template <int S, int CB>
struct MixLds {
  static constexpr int P = MixScheme<S>::parts;
  // A tile per buffer, parity and part: 64 rings x (32 values + pad)
  // FP32 rows of 34 floats: the MFMA A reads (row 16 rt + i16, column
  // 4 s + g) then fall on distinct banks for 32 and for 64 banks
  static constexpr int a_row = (S == 0) ? (kKK + 2) : (kKK + 8);  // elements
  static constexpr int a_elem = (S == 0) ? 4 : 2;                   // bytes
  static constexpr int a_part = 64 * a_row * a_elem;                // bytes
  static constexpr int a_par = P * a_part;
  static constexpr int a_buf = 2 * a_par;
  // recurrence coefficients per buffer and parity: d_{l+2}, e_{l+2} and
  // (FP16) A_l of the 32 l of the parity, as doubles
  static constexpr int c_par = 100 * 8;  // + the 4 chunk rescale factors
  // FP32 scheme: B rows staged per buffer, [4 chunks][2 parities][8 l]
  // [32 CB cols + 16] floats (the pad puts the 4 l of one MFMA step on
  // distinct banks); the 16-bit schemes read B straight into registers
  static constexpr int b_row = kMixCols * CB + 16;  // row stride (floats)
  static constexpr int b_buf = (S == 0) ? 64 * b_row * 4 : 0;
  static constexpr int b_u128 = 4 * 2 * 8 * kMixCols * CB * 4 / 16;
  static constexpr int a_off = 0;
  static constexpr int b_off = 2 * a_buf;
  static constexpr int c_off = b_off + 2 * b_buf;
  static constexpr int bytes = c_off + 2 * 2 * c_par;
};

template <int S, int CB>
__global__ __launch_bounds__(64 * (2 + 2 * CB)) void k_mix(
    LegTables t, const void *bmat, int NCp, int ncb, int bpm, int64_t part_stride,
    double lam_scale, const int *cexp, const double *dtil, const double *etil, const double *gtil) {
  using Sc = MixScheme<S>;
  using Ly = MixLds<S, CB>;
  constexpr int P = Sc::parts;
  constexpr int NT = 64 * (2 + 2 * CB);
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];

  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x;
  const int s8 = b >> 3;
  const int m = (b & 7) + 8 * (s8 / bpm);
  if (m > t.M) return;
  const int item = s8 % bpm;
  const int cg = item / t.W, wblk = item - cg * t.W;  // column group, ring block
  const int L = t.L, R = t.R, Mp1 = t.M + 1, K = t.K;
  const int ring0 = wblk * 64;
  const bool producer = wv < 2;
  const int cw = wv - 2;                 // consumer index
  const int pw = producer ? wv : (cw & 1);  // parity of this wave
  const int cb = cg * CB + (cw >> 1);       // column block of a consumer
  const bool cons_on = !producer && cb < ncb;
  const int i16 = lane & 15, g = lane >> 4;
  const int colg0 = cg * CB * kMixCols;     // first column of the group

  int l0 = __builtin_amdgcn_readfirstlane(t.wstart[(int64_t)m * t.W + wblk]);
  const int lsmax = __builtin_amdgcn_readfirstlane(t.wseedmax[(int64_t)m * t.W + wblk]);
  const int64_t offm = t.off[m];

  double qd[4][2][4];  // consumers: [row tile][col tile][r]
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) qd[rt][ct][r] = 0.0;

  if (l0 <= L) {
    const int nsc = (L - l0) / kSC + 1;  // super-chunks, block-uniform
    // ---------------------------------------------------------- producer
    const WM_CONST double *ct_ = cview(t.ctil + (offm - m));
    const int nch = (L - m + kChunk) / kChunk;  // chunks of m's padded l-range
    const int rr = ring0 + lane;
    const bool rok = producer && rr < R;
    const double x = rok ? t.ring_x[rr] : 0.0;
    const double x2 = x * x;
    const int sl = rok ? t.seed_l[(int64_t)m * R + rr] : kSeedNone;
    double u = 0.0, v = 0.0;  // L~ at the next two l of parity pw
    // this lane's start: three-term steps from its seed (L~_{sl-1}, L~_{sl})
    // to the parity's pair (L~_sl, L~_sl+2) or (L~_sl+1, L~_sl+3), computed
    // up front so that the step loop never waits on memory
    double us = 0.0, vs = 0.0;
    if (sl <= L) {
      const double2 sv = t.seed_v[(int64_t)m * R + rr];
      const double t1 = fma(x, sv.y, -(ct_[sl] * sv.x));
      const double t2 = fma(x, t1, -(ct_[sl + 1] * sv.y));
      const double t3 = fma(x, t2, -(ct_[sl + 2] * t1));
      us = pw ? t1 : sv.y;
      vs = pw ? t3 : t2;
      if constexpr (S == 3) {
        // FP16: the producer steps lam_scale Lambda_l itself (normalised x^2
        // form), so the monic start values take their chunk-relative A_l
        // (sl is chunk-aligned, so sl .. sl + 3 share sl's chunk)
        const int64_t j0 = offm - m + sl + pw;
        us *= lam_scale * t.ascale[j0];
        vs *= lam_scale * t.ascale[j0 + 2];
      }
    }

    // 16-bit schemes: B fragments of a consumer (its 32 columns, its
    // parity) straight from global memory into registers, two super-chunks
    // ahead of their use (each B element is used by one wave of the
    // workgroup, so it gains nothing from LDS): lane (i16, g) holds the 8 l
    // of chunk g for column 16 ct + i16, 16 bytes per part and ct.  Columns
    // past the last block read column 0.
    const int64_t chunk0 = (offm + l0 - m) / kChunk;  // first chunk of super-chunk 0
    int colc[2];
#pragma unroll
    for (int ct = 0; ct < 2; ++ct) {
      const int col = cb * kMixCols + 16 * ct + i16;
      colc[ct] = col < NCp ? col : 0;
    }
    typedef u32x4 BFrag[P][2];
    auto load_b = [&](BFrag &f, int sc) {
      const unsigned short *b0 =
          (const unsigned short *)bmat + ((chunk0 + 4 * sc + g) * 2 + pw) * (int64_t)NCp * 8;
#pragma unroll
      for (int q = 0; q < P; ++q)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
          f[q][ct] = *(const u32x4 *)(b0 + q * part_stride + colc[ct] * 8);
    };
    // FP32 scheme: B rows of a super-chunk moved into LDS by the 128 CB
    // consumer lanes, loaded into registers one super-chunk ahead of their
    // store into an LDS buffer (macros, so that the register array is never
    // addressed indirectly)
    constexpr int NC = 128 * CB;          // consumer lanes
    constexpr int BU = Ly::b_u128 / NC;   // 16-byte vectors per consumer lane
    auto b_offsets = [&](int x, int64_t *src, int *dst) {
      const int e = threadIdx.x - 128 + NC * x;
      constexpr int U = 8 * CB;  // rows (chunk, parity, i) of 32 CB floats
      const int row = e / U, cu = e - row * U;
      const int col = colg0 + 4 * cu < NCp ? colg0 + 4 * cu : 0;
      *src = ((int64_t)row * NCp + col) * 4;
      *dst = (row * Ly::b_row + 4 * cu) * 4;
    };
    const int64_t bchunk = (int64_t)16 * NCp * 4;  // bytes between chunks
    const char *bsc0 = (const char *)bmat + chunk0 * bchunk;
#define WM_LOAD_B(breg, sc)                                                                 \
  do {                                                                                      \
    const char *p_ = bsc0 + (int64_t)(sc) * 4 * bchunk;                                     \
    _Pragma("unroll") for (int x = 0; x < BU; ++x) breg[x] = *(const u32x4 *)(p_ + bso[x]); \
  } while (0)
#define WM_STORE_B(breg, buf)                                                               \
  do {                                                                                      \
    unsigned char *d_ = lds + Ly::b_off + (buf) * Ly::b_buf;                                \
    _Pragma("unroll") for (int x = 0; x < BU; ++x) *(u32x4 *)(d_ + bdo[x]) = breg[x];       \
  } while (0)

    // Coefficients of the 32 l of this wave's parity in a super-chunk: lane
    // k < 32 holds d_{l+2}, lane 32 + k e_{l+2} (l = lsc + pw + 2k) (FP16:
    // alpha_{l+2}, beta_{l+2}, and lane k also gamma_{l+2}); loaded one
    // super-chunk ahead, then put in the wave's LDS slot so that the step
    // loop reads them without waiting on memory.  Indices past the tables
    // are clamped (those chunks are unused).
    double creg0 = 0.0, creg1 = 0.0, creg2 = 0.0;
    auto load_coef = [&](int lsc) {
      const int k = lane & 31;
      const int64_t j = offm - m + lsc + pw + 2 * k;  // table index of l
      const int64_t jd = j + 2 < t.nlpad + 15 ? j + 2 : t.nlpad + 15;
      // FP16: alpha_{l+2}, beta_{l+2} (lanes < 32 / >= 32) and gamma_{l+2}
      // of the normalised step, else d_{l+2}, e_{l+2} of the monic one
      creg0 = lane < 32 ? dtil[jd] : etil[jd];
      if constexpr (S == 3) creg1 = gtil[jd];
      // chunk rescale factors of the 4 chunks (lanes 0 .. 3)
      const int64_t jc = (offm - m + lsc) / kChunk + (lane & 3);
      creg2 = t.cscale[jc < t.nlpad / kChunk ? jc : t.nlpad / kChunk - 1];
    };
    auto store_coef = [&](int buf) {
      double *cb_ = (double *)(lds + Ly::c_off + (buf * 2 + pw) * Ly::c_par);
      cb_[lane] = creg0;
      if constexpr (S == 3)
        if (lane < 32) cb_[64 + lane] = creg1;
      if (lane < 4) cb_[96 + lane] = creg2;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };

    // Lambda of super-chunk at lsc (this lane's ring, the wave's parity)
    // into A buffer buf
    auto produce = [&](int lsc, int buf) {
      unsigned char *abase = lds + Ly::a_off + buf * Ly::a_buf + pw * Ly::a_par;
      const double *cf = (const double *)(lds + Ly::c_off + (buf * 2 + pw) * Ly::c_par);
      // coefficient sets of two chunks in registers: the next chunk's are
      // read from LDS while this chunk's steps run
      constexpr int NA = (S == 3) ? 8 : 1;
      double cd[2][8], ce[2][8], ca[2][NA], cs[2];
      auto read_set = [&](int cc, int x) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          cd[x][i] = cf[8 * cc + i];
          ce[x][i] = cf[32 + 8 * cc + i];
          if constexpr (S == 3) ca[x][i] = cf[64 + 8 * cc + i];
        }
        cs[x] = cf[96 + cc];
      };
      read_set(0, 0);
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const int x = cc & 1;
        if (cc + 1 < 4) read_set(cc + 1, x ^ 1);
        const int lc = lsc + kChunk * cc;
        float f[8];
        if ((lc - m) / kChunk < nch) {
          if (lc <= lsmax && sl == lc) {
            u = us;
            v = vs;
          }
#pragma unroll
          for (int i = 0; i < 8; ++i) {  // l = lc + pw + 2 i
            f[i] = (float)u;
            // FP16: w = (alpha x^2 + beta) v - gamma u on lam_scale Lambda;
            // else w = (x^2 - d) v - e u on the monic L~
            const double w = (S == 3) ? fma(fma(cd[x][i], x2, ce[x][i]), v, -(ca[x][i] * u))
                                      : fma(x2 - cd[x][i], v, -(ce[x][i] * u));
            u = v;
            v = w;
          }
          if constexpr (S != 3) {
            u *= cs[x];
            v *= cs[x];
          }
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) f[i] = 0.f;
        }
        // the 8 values of this chunk: A[q][ring][8 cc + i]
        if constexpr (S == 0) {
          typedef float f32x2 __attribute__((ext_vector_type(2)));
          float *dst = (float *)abase + lane * Ly::a_row + 8 * cc;  // 8-byte aligned
#pragma unroll
          for (int i = 0; i < 8; i += 2) *(f32x2 *)(dst + i) = (f32x2){f[i], f[i + 1]};
        } else {
          // parts packed two 16-bit values per dword: wh (high), wm (middle),
          // wl (low)
          unsigned wh[4], wm[4], wl[4];
#pragma unroll
          for (int i = 0; i < 8; i += 2) {
            if constexpr (S == 3) {
              // packed conversions (round toward zero for the high part; the
              // low part takes the exact residual, < 1 ulp of the high part,
              // each formed by one mixed-precision FMA f - h * 1.0 reading
              // its fp16 half of h directly)
              typedef __fp16 h16x2 __attribute__((ext_vector_type(2)));
              const h16x2 h = __builtin_amdgcn_cvt_pkrtz(f[i], f[i + 1]);
              const unsigned hb = __builtin_bit_cast(unsigned, h);
              float r0, r1;
              asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]"
                  : "=v"(r0) : "v"(hb), "v"(f[i]));
              asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"
                  : "=v"(r1) : "v"(hb), "v"(f[i + 1]));
              wh[i >> 1] = hb;
              wm[i >> 1] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(r0, r1));
            } else {
              const __bf16 h0 = (__bf16)f[i], h1 = (__bf16)f[i + 1];
              const float r0 = f[i] - (float)h0, r1 = f[i + 1] - (float)h1;
              const __bf16 m0 = (__bf16)r0, m1 = (__bf16)r1;
              wh[i >> 1] = __builtin_bit_cast(unsigned short, h0) |
                           ((unsigned)__builtin_bit_cast(unsigned short, h1) << 16);
              wm[i >> 1] = __builtin_bit_cast(unsigned short, m0) |
                           ((unsigned)__builtin_bit_cast(unsigned short, m1) << 16);
              if constexpr (S == 2) {
                const __bf16 q0 = (__bf16)(r0 - (float)m0), q1 = (__bf16)(r1 - (float)m1);
                wl[i >> 1] = __builtin_bit_cast(unsigned short, q0) |
                             ((unsigned)__builtin_bit_cast(unsigned short, q1) << 16);
              }
            }
          }
          unsigned char *dst = abase + (lane * Ly::a_row + 8 * cc) * 2;
          *(u32x4 *)dst = (u32x4){wh[0], wh[1], wh[2], wh[3]};
          *(u32x4 *)(dst + Ly::a_part) = (u32x4){wm[0], wm[1], wm[2], wm[3]};
          if constexpr (S == 2)
            *(u32x4 *)(dst + 2 * Ly::a_part) = (u32x4){wl[0], wl[1], wl[2], wl[3]};
        }
      }
    };

    // ---------------------------------------------------------- consumer
    // undoes the FP16 scheme's scales of this lane's two columns (exact)
    double unscale[2] = {1.0, 1.0};
    if constexpr (S == 3) {
      if (cons_on) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
          const int col = cb * kMixCols + 16 * ct + i16;
          unscale[ct] = ldexp(1.0 / lam_scale, -cexp[(int64_t)m * NCp + (col < NCp ? col : 0)]);
        }
      }
    }
    f32x4 acc[4][2];
    auto zero_acc = [&]() {
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) acc[rt][ct] = (f32x4){0.f, 0.f, 0.f, 0.f};
    };
    auto flush = [&]() {
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int ct = 0; ct < 2; ++ct)
#pragma unroll
          for (int r = 0; r < 4; ++r) qd[rt][ct][r] += unscale[ct] * (double)acc[rt][ct][r];
      zero_acc();
    };
    const int cl = (cw >> 1) * kMixCols;  // first column of this consumer in the group
    // fresh: the first product overwrites acc (the previous sums were
    // flushed), so no separate zeroing
    auto consume = [&](int buf, const BFrag &bfr, auto fresh) {
      const unsigned char *abase = lds + Ly::a_off + buf * Ly::a_buf + pw * Ly::a_par;
      if constexpr (S == 0) {
        const float *af = (const float *)abase;
        const float *bf = (const float *)(lds + Ly::b_off + buf * Ly::b_buf);
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          float a[4], bv[2];
#pragma unroll
          for (int rt = 0; rt < 4; ++rt) a[rt] = af[(16 * rt + i16) * Ly::a_row + 4 * s + g];
#pragma unroll
          for (int ct = 0; ct < 2; ++ct)
            bv[ct] = bf[((((s >> 1) * 2 + pw) * 8) + 4 * (s & 1) + g) * Ly::b_row + cl +
                        16 * ct + i16];
#pragma unroll
          for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
              acc[rt][ct] =
                  __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt], bv[ct], acc[rt][ct], 0, 0, 0);
        }
      } else {
        bf16x8 bv[P][2];
#pragma unroll
        for (int q = 0; q < P; ++q)
#pragma unroll
          for (int ct = 0; ct < 2; ++ct) bv[q][ct] = __builtin_bit_cast(bf16x8, bfr[q][ct]);
        // row tiles in two halves: the A fragments of a half first (one LDS
        // wait), then its products with the 4 accumulators interleaved, so
        // that consecutive MFMAs are independent
#pragma unroll
        for (int h2 = 0; h2 < 4; h2 += 2) {
        bf16x8 a[4][P];
#pragma unroll
        for (int rt = h2; rt < h2 + 2; ++rt)
#pragma unroll
          for (int q = 0; q < P; ++q)
            a[rt][q] = *(const bf16x8 *)(abase + q * Ly::a_part +
                                         ((16 * rt + i16) * Ly::a_row + 8 * g) * 2);
#pragma unroll
        for (int pi = 0; pi < Sc::nprod; ++pi)
#pragma unroll
          for (int rt = h2; rt < h2 + 2; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
              const f32x4 c = (decltype(fresh)::value && pi == 0) ? (f32x4){0.f, 0.f, 0.f, 0.f}
                                                                  : acc[rt][ct];
              if constexpr (S == 3)
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                    __builtin_bit_cast(f16x8, a[rt][Sc::pa(pi)]),
                    __builtin_bit_cast(f16x8, bv[Sc::pb(pi)][ct]), c, 0, 0, 0);
              else
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[rt][Sc::pa(pi)], bv[Sc::pb(pi)][ct], c, 0, 0, 0);
            }
        }
      }
    };

    // ---------------------------------------------------------- pipeline
    // two loops with the same barrier sequence (one before the first
    // super-chunk, one after each), so that each role keeps only its own
    // registers live
    if (producer) {
      load_coef(l0);
      store_coef(0);
      if (nsc > 1) load_coef(l0 + kSC);
      produce(l0, 0);
      __syncthreads();
      for (int sc = 0; sc < nsc; ++sc) {
        if (sc + 1 < nsc) {
          store_coef((sc + 1) & 1);
          if (sc + 2 < nsc) load_coef(l0 + kSC * (sc + 2));
          produce(l0 + kSC * (sc + 1), (sc + 1) & 1);
        }
        __syncthreads();
      }
    } else if constexpr (S == 0) {
      int64_t bso[BU];
      int bdo[BU];
#pragma unroll
      for (int x = 0; x < BU; ++x) b_offsets(x, &bso[x], &bdo[x]);
      u32x4 breg[BU];
      WM_LOAD_B(breg, 0);
      WM_STORE_B(breg, 0);
      if (nsc > 1) WM_LOAD_B(breg, 1);
      zero_acc();
      __syncthreads();
      int nfl = 0;
      BFrag none;
      for (int sc = 0; sc < nsc; ++sc) {
        // B of super-chunk sc + 1 into the buffer read last in sc - 1
        if (sc + 1 < nsc) {
          WM_STORE_B(breg, (sc + 1) & 1);
          if (sc + 2 < nsc) WM_LOAD_B(breg, sc + 2);
        }
        if (cons_on) {
          consume(sc & 1, none, std::false_type());
          if (++nfl == kFlush) {
            flush();
            nfl = 0;
          }
        }
        __syncthreads();
      }
      if (cons_on) flush();
    } else {
      // super-chunks in pairs: B fragments of even / odd super-chunks in
      // b0 / b1, each reloaded (two super-chunks ahead) right after its
      // use; the pair's FP32 sums are flushed once (kFlush = 2)
      static_assert(kFlush == 2, "the 16-bit consumer loop flushes every two super-chunks");
      BFrag b0, b1;
      if (cons_on) {
        load_b(b0, 0);
        if (nsc > 1) load_b(b1, 1);
      }
      __syncthreads();
      for (int sc = 0; sc < nsc; sc += 2) {
        if (cons_on) {
          consume(0, b0, std::true_type());
          if (sc + 2 < nsc) load_b(b0, sc + 2);
        }
        __syncthreads();
        if (sc + 1 < nsc) {
          if (cons_on) {
            consume(1, b1, std::false_type());
            if (sc + 3 < nsc) load_b(b1, sc + 3);
          }
          __syncthreads();
        }
        if (cons_on) {
#pragma unroll
          for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
#pragma unroll
              for (int r = 0; r < 4; ++r)
                qd[rt][ct][r] += unscale[ct] * (double)acc[rt][ct][r];
        }
      }
    }
#undef WM_LOAD_B
#undef WM_STORE_B
  }

  // ---- store (consumers): column col = 2 map + (Re, Im) of q(ring, pw, map, m),
  //      rounded to FP32 once after the FP64 l-chunk reduction when t.q32
  if (cons_on) {
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ring = ring0 + 16 * rt + 4 * g + r;
        if (ring < R) {
#pragma unroll
          for (int ct = 0; ct < 2; ++ct) {
            const int col = cb * kMixCols + 16 * ct + i16;
            if (col < 2 * K) {
              const int64_t at = 2 * q_index(ring, pw, col >> 1, m, K, Mp1) + (col & 1);
              if (t.q32)
                ((float *)t.qbuf)[at] = (float)qd[rt][ct][r];
              else
                ((double *)t.qbuf)[at] = qd[rt][ct][r];
            }
          }
        }
      }
  }
}

// ------------------------------------------- FP16x3 with register Lambda

// k_mixr: the FP16x3 contraction without producer waves.  The Lambda
// operand is generated in registers, in the A layout of
// v_mfma_f32_32x32x16_f16 (lane: row = lane % 32, 8 consecutive K values
// for lane half h = lane / 32).
//
// A workgroup covers 128 ring pairs of one m and one parity and 128
// columns: 4 ring groups of 32 ring pairs, each with 2 waves over the
// columns (CT column tiles of 32 per wave).  A block's l-range of
// 4 ns chunks (16 l each, 8 per parity) splits into 4 segments of ns
// chunks; K-step j takes chunk j of every segment.  The lane halves of the
// two waves of a ring group step the FP64 recurrence of the same ring over
// the 4 segments (wave c, half h: segment 2 c + h), each wave builds the A
// fragment of its 2 segments and exchanges it with the other through LDS,
// and each runs the MFMAs of all 4 segments against its own columns, so
// no recurrence step is computed twice.  Segments 1..3 start from
// checkpoints of the recurrence state written by k_mix_ckpt with the same
// steps.  The B rows and coefficients of each K-step are staged through
// LDS, loaded two K-steps ahead; while the MFMAs of K-step j run, the lanes
// build the A fragment of K-step j + 1 (one workgroup barrier per K-step).  FP32
// MFMA sums over kRFlush K-steps (128 l of the parity) are added into FP64
// accumulators.
constexpr int kRRings = 32;   // ring pairs per wave (one 32-row tile)
constexpr int kRGroups = 4;   // ring groups per workgroup
constexpr int kRBlock = kRRings * kRGroups;  // ring pairs per workgroup
constexpr int kRSeg = 4;      // l-segments per ring (2 per wave)
constexpr int kRWaves = 2 * kRGroups;
constexpr int kRFlush = 4;    // K-steps per FP32 accumulation

typedef float f32x16 __attribute__((ext_vector_type(16)));

template <int CT>
struct MixRLds {
  static constexpr int cols = 2 * CT * 32;  // columns of the workgroup
  // B: [segment][part][cols][8 halves]
  static constexpr int b_seg = 2 * cols * 16;
  static constexpr int b_buf = kRSeg * b_seg;
  // coefficients: [segment][8 steps][alpha, beta, gamma, pad] doubles
  static constexpr int c_buf = kRSeg * 8 * 4 * 8;
  // A fragments: [wave][part][lane] 16 bytes
  static constexpr int a_wave = 2 * 64 * 16;
  static constexpr int a_buf = kRWaves * a_wave;
  static constexpr int b_off = 0;
  static constexpr int c_off = 2 * b_buf;
  static constexpr int a_off = c_off + 4 * c_buf;
  static constexpr int bytes = a_off + 2 * a_buf;
};

// Start of a workgroup's l-range: the smallest seed l of its ring pairs
// (from the 64-ring wave table), and the number ns of K-steps (chunks per
// segment).
__device__ __forceinline__ int mixr_block_l0(const LegTables &t, int m, int bb) {
  constexpr int per = kRBlock / 64;  // wave-table entries per workgroup block
  static_assert(per * 64 == kRBlock, "workgroup blocks are whole 64-ring table entries");
  int l0 = t.wstart[(int64_t)m * t.W + per * bb];
#pragma unroll
  for (int k = 1; k < per; ++k) {
    if (per * bb + k < t.W) {
      const int lk = t.wstart[(int64_t)m * t.W + per * bb + k];
      l0 = lk < l0 ? lk : l0;
    }
  }
  return l0;
}
__device__ __forceinline__ int mixr_segment_steps(int L, int l0) {
  return ((L - l0) / kChunk + kRSeg) / kRSeg;  // ceil(chunks / kRSeg)
}

// FP16 scheme: lam_scale Lambda at (sl + p, sl + p + 2) from the seed
// (the monic start values of the seed chunk, scaled by their A_l).
__device__ __forceinline__ void mix16_start(const LegTables &t, int m, int r, double x, int sl,
                                            int p, double lam_scale, double *us, double *vs) {
  const int64_t offm = t.off[m];
  const double *ct_ = t.ctil + (offm - m);
  const double2 sv = t.seed_v[(int64_t)m * t.R + r];
  const double t1 = fma(x, sv.y, -(ct_[sl] * sv.x));
  const double t2 = fma(x, t1, -(ct_[sl + 1] * sv.y));
  const double t3 = fma(x, t2, -(ct_[sl + 2] * t1));
  const int64_t j0 = offm - m + sl + p;
  *us = (p ? t1 : sv.y) * (lam_scale * t.ascale[j0]);
  *vs = (p ? t3 : t2) * (lam_scale * t.ascale[j0 + 2]);
}

// Recurrence coefficients of the step that yields l + 4 from (l, l + 2):
// the table entries of l + 2, zero past m's padded l-range (the values
// there multiply zero B rows and only have to stay finite).
__device__ __forceinline__ double mix16_coef(const double *tab, int64_t offm, int m, int lp,
                                             int l) {
  const int jj = l + 2 - m;
  return jj < lp ? tab[offm + jj] : 0.0;
}

// Coefficient tid < 96 of K-step `step` (segment tid / 24, recurrence step
// (tid % 24) / 3, alpha / beta / gamma = tid % 3) of a workgroup.
__device__ __forceinline__ double mixr_stage_coef(const double *atab, const double *btab,
                                                  const double *gtab, int64_t offm, int m,
                                                  int lp, int l0, int ns, int p, int step,
                                                  int tid) {
  const int g = tid / 24, r2 = tid - g * 24;
  const int i = r2 / 3, which = r2 - i * 3;
  const int l = l0 + kChunk * (g * ns + step) + p + 2 * i;
  const double *tab = which == 0 ? atab : (which == 1 ? btab : gtab);
  return mix16_coef(tab, offm, m, lp, l);
}

// Checkpoints: state (lam_scale Lambda_ls, lam_scale Lambda_{ls + 2}) of
// ring r, parity p at the first l = ls of segments 1 .. kRSeg - 1 of r's
// workgroup (zero when the seed starts at or after it), at
// ckpt[((m * 2 + p) * R + r) * (kRSeg - 1) + segment - 1].  One thread
// per (m, p, r).
__global__ __launch_bounds__(256) void k_mix_ckpt(LegTables t, const double *__restrict__ atab,
                                                  const double *__restrict__ btab,
                                                  const double *__restrict__ gtab,
                                                  double lam_scale, double2 *__restrict__ ckpt) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int R = t.R, L = t.L;
  if (id >= (int64_t)(t.M + 1) * 2 * R) return;
  const int r = (int)(id % R);
  const int mp = (int)(id / R), p = mp & 1, m = mp >> 1;
  double2 out[kRSeg - 1];
#pragma unroll
  for (int s = 0; s < kRSeg - 1; ++s) out[s] = make_double2(0.0, 0.0);
  const int l0 = mixr_block_l0(t, m, r / kRBlock);
  const int sl = t.seed_l[(int64_t)m * R + r];
  if (l0 <= L && sl <= L) {
    const int ns = mixr_segment_steps(L, l0);
    const int llast = l0 + kChunk * ns * (kRSeg - 1);
    if (sl < llast) {
      const double x = t.ring_x[r], x2 = x * x;
      const int64_t offm = t.off[m];
      const int lp = (L - m + kChunk) / kChunk * kChunk;
      double u, v;
      mix16_start(t, m, r, x, sl, p, lam_scale, &u, &v);
      for (int lc = sl; lc < llast; lc += kChunk) {
#pragma unroll
        for (int i = 0; i < 8; ++i) {
          const int l = lc + p + 2 * i;
          const double a = mix16_coef(atab, offm, m, lp, l);
          const double b = mix16_coef(btab, offm, m, lp, l);
          const double g = mix16_coef(gtab, offm, m, lp, l);
          const double w = fma(fma(a, x2, b), v, -(g * u));
          u = v;
          v = w;
        }
        const int ln = lc + kChunk;  // state (u, v) now holds l = ln + p, ln + p + 2
        const int sg = (ln - l0) / (kChunk * ns);
        if (ln == l0 + kChunk * ns * sg && sg >= 1 && sg < kRSeg) {
#pragma unroll
          for (int s = 1; s < kRSeg; ++s)
            if (s == sg) out[s - 1] = make_double2(u, v);
        }
      }
    }
  }
#pragma unroll
  for (int s = 0; s < kRSeg - 1; ++s) ckpt[id * (kRSeg - 1) + s] = out[s];
}

template <int CT>
__global__ __launch_bounds__(64 * kRWaves) void k_mixr(
    LegTables t, const unsigned short *__restrict__ bq, int NCp, int ncg, int bpm,
    int64_t part_stride, double lam_scale, const int *__restrict__ cexp,
    const double *__restrict__ atab, const double *__restrict__ btab,
    const double *__restrict__ gtab, const double2 *__restrict__ ckpt) {
  using Ly = MixRLds<CT>;
  constexpr int NT = 64 * kRWaves;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x;
  const int s8 = b >> 3;
  const int m = (b & 7) + 8 * (s8 / bpm);
  if (m > t.M) return;
  const int L = t.L, R = t.R, K = t.K, Mp1 = t.M + 1;
  const int W2 = (R + kRBlock - 1) / kRBlock;
  const int item = s8 % bpm;  // (column group, parity, ring block)
  const int cg = item / (2 * W2);
  const int rem = item - cg * 2 * W2;
  const int p = rem / W2, bb = rem - p * W2;
  const int n = lane & 31, h = lane >> 5;
  const int rg = wv % kRGroups, chh = wv / kRGroups;
  const int wpart = rg + kRGroups * (1 - chh);  // the other wave of this ring group
  const int ring0 = bb * kRBlock + rg * kRRings;
  const int colg0 = cg * Ly::cols;           // first column of the workgroup
  const int colw0 = colg0 + chh * CT * 32;   // first column of this wave
  const int l0 = __builtin_amdgcn_readfirstlane(mixr_block_l0(t, m, bb));
  const int64_t offm = t.off[m];

  double qd[CT][16];
#pragma unroll
  for (int ct = 0; ct < CT; ++ct)
#pragma unroll
    for (int e = 0; e < 16; ++e) qd[ct][e] = 0.0;

  if (l0 <= L) {
    const int ns = mixr_segment_steps(L, l0);
    const int lp = (L - m + kChunk) / kChunk * kChunk;
    const int64_t c0 = (offm + l0 - m) / kChunk;
    const int64_t cend = (offm + lp) / kChunk;
    // this lane's ring (A row) and segment
    const int sg = 2 * chh + h;
    const int rr = ring0 + n;
    const bool rok = rr < R;
    const double x = rok ? t.ring_x[rr] : 0.0;
    const double x2 = x * x;
    const int sl = rok ? t.seed_l[(int64_t)m * R + rr] : kSeedNone;
    double us = 0.0, vs = 0.0, u = 0.0, v = 0.0;
    if (sl <= L) mix16_start(t, m, rr, x, sl, p, lam_scale, &us, &vs);
    if (sg > 0 && sl < l0 + kChunk * ns * sg && sl <= L) {
      const double2 c = ckpt[(((int64_t)m * 2 + p) * R + rr) * (kRSeg - 1) + sg - 1];
      u = c.x;
      v = c.y;
    }
    // FP16 scales of this lane's columns, undone at the FP64 flush
    double unscale[CT];
#pragma unroll
    for (int ct = 0; ct < CT; ++ct) {
      const int col = colw0 + ct * 32 + n;
      unscale[ct] = ldexp(1.0 / lam_scale, -cexp[(int64_t)m * NCp + (col < NCp ? col : 0)]);
    }

    // Stage j = B rows of chunk j of every segment (zero past m's chunks
    // and past the last column; bbuf j % 2) and the coefficients of K-step
    // j + 1 (cbuf (j + 1) % 4).  At the end of K-step j the register set
    // holding stage j + 1 goes to LDS and is reloaded with stage j + 2, so
    // each load has a whole K-step to arrive.  A(j) of both waves of a ring
    // group sits in abuf j % 2.  Macros, so that no closure object holds
    // the register arrays.
    constexpr int NVEC = kRSeg * 2 * Ly::cols;  // 16-byte vectors per stage
    static_assert(NVEC % NT == 0, "every thread stages the same number of vectors");
    constexpr int NV = NVEC / NT;
    constexpr int NC = kRSeg * 24;              // coefficients per stage
    // per-thread staging state, fixed for the block (a K-step only adds
    // its index): the chunk of K-step 0 and the element offset of each
    // vector; chunks past m's range and padded columns read the B
    // operand's trailing zero chunk instead (no branch)
    const int64_t zch = t.nlpad / kChunk;
    const int64_t cstride = (int64_t)2 * NCp * 8;  // elements per chunk
    int64_t sbase[NV];
    int64_t schb[NV];
#pragma unroll
    for (int x_ = 0; x_ < NV; ++x_) {
      const int vi = threadIdx.x + NT * x_;
      const int g_ = vi / (2 * Ly::cols), r2 = vi - g_ * (2 * Ly::cols);
      const int q_ = r2 / Ly::cols, c_ = r2 - q_ * Ly::cols;
      const int col = colg0 + c_;
      sbase[x_] = q_ * part_stride + ((int64_t)p * NCp + (col < NCp ? col : 0)) * 8;
      // a padded column never reads a real chunk
      schb[x_] = col < NCp ? c0 + (int64_t)g_ * ns : cend;
    }
    const double *ctab = atab;
    int cjj = 0;  // table offset (past m's start) of the coefficient at K-step 0
    if (threadIdx.x < NC) {
      const int g_ = threadIdx.x / 24, r2 = threadIdx.x - g_ * 24;
      const int i_ = r2 / 3, which = r2 - i_ * 3;
      ctab = which == 0 ? atab : (which == 1 ? btab : gtab);
      cjj = l0 + kChunk * g_ * ns + p + 2 * i_ + 2 - m;
    }
    u32x4 breg[NV];
    double creg = 0.0;
    bool cok = false;
#define WM_R_COEF(step) \
  mixr_stage_coef(atab, btab, gtab, offm, m, lp, l0, ns, p, step, threadIdx.x)
#define WM_R_CSLOT(tid) ((((tid) / 24) * 8 + ((tid) % 24) / 3) * 4 + (tid) % 3)
#define WM_R_GLOAD(jg)                                                                       \
  do {                                                                                       \
    _Pragma("unroll") for (int x_ = 0; x_ < NV; ++x_) {                                      \
      const int64_t ch = schb[x_] + (jg);                                                    \
      breg[x_] = *(const u32x4 *)(bq + sbase[x_] + (ch < cend ? ch : zch) * cstride);        \
    }                                                                                        \
    if (threadIdx.x < NC) {                                                                  \
      /* raw value now, zeroing at the LDS store: a select right after */                    \
      /* the load would wait for it here */                                                  \
      const int jj_ = cjj + kChunk * ((jg) + 1);                                             \
      cok = jj_ < lp;                                                                        \
      creg = ctab[offm + (cok ? jj_ : 0)];                                                   \
    }                                                                                        \
  } while (0)
#define WM_R_LSTORE(BB, CB)                                                                  \
  do {                                                                                       \
    unsigned char *bd = lds + Ly::b_off + (BB) * Ly::b_buf + threadIdx.x * 16;               \
    _Pragma("unroll") for (int x_ = 0; x_ < NV; ++x_)* (u32x4 *)(bd + NT * 16 * x_) =         \
        breg[x_];                                                                            \
    if (threadIdx.x < NC)                                                                    \
      ((double *)(lds + Ly::c_off + (CB) * Ly::c_buf))[WM_R_CSLOT(threadIdx.x)] =            \
          cok ? creg : 0.0;                                                                  \
  } while (0)
    // A(jk) from cbuf CB: this lane's 8 recurrence steps, the fp16 parts
    // (high part by round-toward-zero packing, low part the exact residual
    // by one mixed-precision FMA reading the fp16 half), written to abuf AB
#define WM_R_ABUILD(jk, CB, AB)                                                              \
  do {                                                                                       \
    const int lc = l0 + kChunk * (sg * ns + (jk));                                           \
    if (lc == sl) {                                                                          \
      u = us;                                                                                \
      v = vs;                                                                                \
    }                                                                                        \
    const double *cf = (const double *)(lds + Ly::c_off + (CB) * Ly::c_buf) + sg * 32;       \
    float f[8];                                                                              \
    _Pragma("unroll") for (int i_ = 0; i_ < 8; ++i_) {                                       \
      f[i_] = (float)u;                                                                      \
      const double2 ab = *(const double2 *)(cf + 4 * i_);                                    \
      const double gm = cf[4 * i_ + 2];                                                      \
      const double w = fma(fma(ab.x, x2, ab.y), v, -(gm * u));                               \
      u = v;                                                                                 \
      v = w;                                                                                 \
    }                                                                                        \
    unsigned wh[4], wl[4];                                                                   \
    _Pragma("unroll") for (int i_ = 0; i_ < 8; i_ += 2) {                                    \
      typedef __fp16 h16x2 __attribute__((ext_vector_type(2)));                              \
      const h16x2 hh = __builtin_amdgcn_cvt_pkrtz(f[i_], f[i_ + 1]);                         \
      const unsigned hb = __builtin_bit_cast(unsigned, hh);                                  \
      float r0, r1;                                                                          \
      asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r0) : "v"(hb),           \
          "v"(f[i_]));                                                                       \
      asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]"                  \
          : "=v"(r1) : "v"(hb), "v"(f[i_ + 1]));                                             \
      wh[i_ >> 1] = hb;                                                                      \
      wl[i_ >> 1] = __builtin_bit_cast(unsigned, __builtin_amdgcn_cvt_pkrtz(r0, r1));        \
    }                                                                                        \
    ahr[AB] = (u32x4){wh[0], wh[1], wh[2], wh[3]};                                           \
    alr[AB] = (u32x4){wl[0], wl[1], wl[2], wl[3]};                                           \
    unsigned char *ad = lds + Ly::a_off + (AB) * Ly::a_buf + wv * Ly::a_wave + lane * 16;    \
    *(u32x4 *)ad = ahr[AB];                                                                  \
    *(u32x4 *)(ad + 64 * 16) = alr[AB];                                                      \
  } while (0)
    // MFMAs of K-step j: for both waves' A fragments (abuf AB) the B
    // fragments of their 2 segments and this wave's CT column tiles (bbuf
    // BB), 3 MFMAs per (fragment, tile), consecutive MFMAs on different
    // tiles; FRESH: the first product overwrites acc
#define WM_R_MMA(BB, AB, FRESH)                                                              \
  do {                                                                                       \
    _Pragma("unroll") for (int kh = 0; kh < 2; ++kh) {                                       \
      const int csrc = kh == 0 ? chh : 1 - chh;                                              \
      const unsigned char *as_ = lds + Ly::a_off + (AB) * Ly::a_buf + wpart * Ly::a_wave +   \
                                 lane * 16;                                                  \
      const f16x8 ah = __builtin_bit_cast(f16x8, kh == 0 ? ahr[AB] : *(const u32x4 *)as_);   \
      const f16x8 al =                                                                       \
          __builtin_bit_cast(f16x8, kh == 0 ? alr[AB] : *(const u32x4 *)(as_ + 64 * 16));    \
      const unsigned char *bs = lds + Ly::b_off + (BB) * Ly::b_buf +                         \
                                (2 * csrc + h) * Ly::b_seg + (chh * CT * 32 + n) * 16;       \
      f16x8 bh[CT], bl[CT];                                                                  \
      _Pragma("unroll") for (int ct = 0; ct < CT; ++ct) {                                    \
        bh[ct] = *(const f16x8 *)(bs + (ct * 32) * 16);                                      \
        bl[ct] = *(const f16x8 *)(bs + (Ly::cols + ct * 32) * 16);                           \
      }                                                                                      \
      _Pragma("unroll") for (int ct = 0; ct < CT; ++ct) acc[ct] =                            \
          __builtin_amdgcn_mfma_f32_32x32x16_f16(                                            \
              ah, bh[ct], ((FRESH) && kh == 0) ? (f32x16){} : acc[ct], 0, 0, 0);             \
      _Pragma("unroll") for (int ct = 0; ct < CT; ++ct) acc[ct] =                            \
          __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl[ct], acc[ct], 0, 0, 0);              \
      _Pragma("unroll") for (int ct = 0; ct < CT; ++ct) acc[ct] =                            \
          __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh[ct], acc[ct], 0, 0, 0);              \
    }                                                                                        \
  } while (0)

    f32x16 acc[CT];
    u32x4 ahr[2], alr[2];  // this wave's A(j) in registers (set j % 2)
    // prologue: stage 0 and the coefficients of K-step 0 into LDS, A(0)
    WM_R_GLOAD(0);
    if (threadIdx.x < NC)
      ((double *)(lds + Ly::c_off))[WM_R_CSLOT(threadIdx.x)] = WM_R_COEF(0);
    WM_R_LSTORE(0, 1);
    WM_R_GLOAD(1);
    __syncthreads();
    WM_R_ABUILD(0, 0, 0);
    __syncthreads();
    static_assert(kRFlush == 4, "buffer indices are static within a flush group of 4");
    for (int j0 = 0; j0 < ns; j0 += kRFlush) {
#pragma unroll
      for (int jj = 0; jj < kRFlush; ++jj) {
        const int j = j0 + jj;
        if (j < ns) {
          WM_R_MMA(jj & 1, jj & 1, jj == 0);
          WM_R_ABUILD(j + 1, (jj + 1) & 3, (jj + 1) & 1);
#pragma unroll
          for (int q_ = 0; q_ < 6 * CT; ++q_) {
            __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one MFMA
            __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);  // then VALU
          }
          // stage j + 1 (loaded at the end of K-step j - 1) into LDS, then
          // the loads of stage j + 2 (the scheduling fence keeps the store
          // from being hoisted above this K-step's work)
          __builtin_amdgcn_sched_barrier(0);
          WM_R_LSTORE((jj + 1) & 1, (jj + 2) & 3);
          WM_R_GLOAD(j + 2);
          __syncthreads();
        }
      }
#pragma unroll
      for (int ct = 0; ct < CT; ++ct)
#pragma unroll
        for (int e = 0; e < 16; ++e) qd[ct][e] += unscale[ct] * (double)acc[ct][e];
    }
#undef WM_R_COEF
#undef WM_R_CSLOT
#undef WM_R_GLOAD
#undef WM_R_LSTORE
#undef WM_R_ABUILD
#undef WM_R_MMA
  }

  // store: acc entry e = 4 i + jj holds row 8 i + 4 h + jj, column n
#pragma unroll
  for (int ct = 0; ct < CT; ++ct) {
    const int col = colw0 + ct * 32 + n;
    if (col < 2 * K) {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int ring = ring0 + 8 * (e >> 2) + 4 * h + (e & 3);
        if (ring < R) {
          const int64_t at = 2 * q_index(ring, p, col >> 1, m, K, Mp1) + (col & 1);
          ((float *)t.qbuf)[at] = (float)qd[ct][e];
        }
      }
    }
  }
}

// ------------------------------------------------------------- launchers

int64_t mix_bmat_elems(const LegTables &t, int NCp) {
  return (t.nlpad / kChunk + 3) * 2 * (int64_t)NCp * 8;
}

hipError_t launch_mix_prep(const LegTables &t, const double2 *alm, void *bmat, int NCp,
                           int scheme, int64_t part_stride, int *cexp, hipStream_t s) {
  const int64_t nchunks = t.nlpad / kChunk;
  unsigned short *bq = (unsigned short *)bmat;
  if (scheme == 0) {
    hipLaunchKernelGGL(k_mix_prep_f32, dim3((unsigned)nchunks), dim3(256), 0, s, t, alm,
                       (float *)bmat, NCp);
  } else if (scheme == 3) {
    hipLaunchKernelGGL(k_mix_prep_f16, dim3(t.M + 1), dim3(256), (size_t)NCp * 8, s, t, alm, bq,
                       NCp, part_stride, cexp);
  } else {
    const int per = NCp;  // (parity, map) threads per chunk
    const int cpb = 256 / per > 0 ? 256 / per : 1;
    const dim3 grid((unsigned)((nchunks + cpb - 1) / cpb));
    if (scheme == 1)
      hipLaunchKernelGGL(k_mix_prep_16<1>, grid, dim3(256), 0, s, t, alm, bq, NCp, part_stride,
                         nchunks);
    else
      hipLaunchKernelGGL(k_mix_prep_16<2>, grid, dim3(256), 0, s, t, alm, bq, NCp, part_stride,
                         nchunks);
  }
  return hipGetLastError();
}

template <int S, int CB>
static hipError_t launch_mix_t(const LegTables &t, const void *bmat, int NCp,
                               int64_t part_stride, double lam_scale,
                               const int *cexp, const double *dtil,
                               const double *etil, const double *gtil, hipStream_t s) {
  const int ncb = NCp / kMixCols;
  const int ncg = (ncb + CB - 1) / CB;
  const int bpm = t.W * ncg;
  const int mgroups = (t.M + 1 + 7) / 8;
  const unsigned grid = (unsigned)(8 * mgroups * bpm);
  constexpr size_t lds = MixLds<S, CB>::bytes;
  hipLaunchKernelGGL((k_mix<S, CB>), dim3(grid), dim3(64 * (2 + 2 * CB)), lds, s, t, bmat, NCp,
                     ncb, bpm, part_stride, lam_scale, cexp, dtil, etil, gtil);
  return hipGetLastError();
}

template <int S>
static hipError_t launch_mix_s(const LegTables &t, const void *bmat, int NCp,
                               int64_t part_stride, double lam_scale,
                               const int *cexp, const double *dtil,
                               const double *etil, const double *gtil, hipStream_t s) {
  switch (NCp / kMixCols) {
    case 1: return launch_mix_t<S, 1>(t, bmat, NCp, part_stride, lam_scale, cexp, dtil, etil, gtil, s);
    case 2: return launch_mix_t<S, 2>(t, bmat, NCp, part_stride, lam_scale, cexp, dtil, etil, gtil, s);
    case 3: return launch_mix_t<S, 3>(t, bmat, NCp, part_stride, lam_scale, cexp, dtil, etil, gtil, s);
    default:
      return launch_mix_t<S, kMixCB>(t, bmat, NCp, part_stride, lam_scale, cexp, dtil, etil, gtil, s);
  }
}

template <int CT>
static hipError_t launch_mixr_t(const LegTables &t, const void *bmat, int NCp,
                                int64_t part_stride, double lam_scale, const int *cexp,
                                const double *atab, const double *btab, const double *gtab,
                                const double2 *ckpt, hipStream_t s) {
  using Ly = MixRLds<CT>;
  const int ncg = (NCp + Ly::cols - 1) / Ly::cols;
  const int W2 = (t.R + kRBlock - 1) / kRBlock;
  const int bpm = ncg * 2 * W2;
  const int mgroups = (t.M + 1 + 7) / 8;
  const unsigned grid = (unsigned)(8 * mgroups * bpm);
  hipLaunchKernelGGL((k_mixr<CT>), dim3(grid), dim3(64 * kRWaves), Ly::bytes, s, t,
                     (const unsigned short *)bmat, NCp, ncg, bpm, part_stride, lam_scale, cexp,
                     atab, btab, gtab, ckpt);
  return hipGetLastError();
}

hipError_t launch_mixr(const LegTables &t, const void *bmat, int NCp, int64_t part_stride,
                       double lam_scale, const int *cexp, const double *atab,
                       const double *btab, const double *gtab, const double2 *ckpt,
                       hipStream_t s) {
  // two waves of CT column tiles per 32 ring pairs: 64 columns per
  // workgroup up to 64 columns, else groups of 128
  if (NCp <= 64)
    return launch_mixr_t<1>(t, bmat, NCp, part_stride, lam_scale, cexp, atab, btab, gtab, ckpt, s);
  return launch_mixr_t<2>(t, bmat, NCp, part_stride, lam_scale, cexp, atab, btab, gtab, ckpt, s);
}

int64_t mix_ckpt_elems(const LegTables &t) {
  return (int64_t)(t.M + 1) * 2 * t.R * (kRSeg - 1);
}

hipError_t launch_mix_ckpt(const LegTables &t, const double *atab, const double *btab,
                           const double *gtab, double lam_scale, double2 *ckpt, hipStream_t s) {
  const int64_t n = (int64_t)(t.M + 1) * 2 * t.R;  // one thread per (m, parity, ring)
  hipLaunchKernelGGL(k_mix_ckpt, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t, atab,
                     btab, gtab, lam_scale, ckpt);
  return hipGetLastError();
}

hipError_t mix_set_lds_limits() {
  // the FP16 prescale's column maxima: NCp <= 8192 (nmaps <= 4096)
  hipError_t e = hipFuncSetAttribute((const void *)k_mix_prep_f16,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, 8192 * 8);
#define WM_SET(S, CB)                                                                      \
  if (e == hipSuccess)                                                                     \
    e = hipFuncSetAttribute((const void *)k_mix<S, CB>,                                    \
                            hipFuncAttributeMaxDynamicSharedMemorySize,                    \
                            (int)(MixLds<S, CB>::bytes));
#define WM_SET_S(S) WM_SET(S, 1) WM_SET(S, 2) WM_SET(S, 3) WM_SET(S, 4)
  WM_SET_S(0) WM_SET_S(1) WM_SET_S(2) WM_SET_S(3)
#undef WM_SET_S
#undef WM_SET
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_mixr<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)MixRLds<1>::bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_mixr<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)MixRLds<2>::bytes);
  return e;
}

hipError_t launch_mix(const LegTables &t, const void *bmat, int NCp, int scheme,
                      int64_t part_stride, double lam_scale, const int *cexp,
                      const double *dtil, const double *etil, const double *gtil, hipStream_t s) {
  switch (scheme) {
    case 0: return launch_mix_s<0>(t, bmat, NCp, part_stride, lam_scale, cexp, dtil, etil, gtil, s);
    case 1: return launch_mix_s<1>(t, bmat, NCp, part_stride, lam_scale, cexp, dtil, etil, gtil, s);
    case 2: return launch_mix_s<2>(t, bmat, NCp, part_stride, lam_scale, cexp, dtil, etil, gtil, s);
    default:
      return launch_mix_s<3>(t, bmat, NCp, part_stride, lam_scale, cexp, dtil, etil, gtil, s);
  }
}

// Power-of-two scale of Lambda for the FP16 scheme: |Lambda_lm| is at most
// sqrt((2 lmax + 1) / (4 pi)); the scaled maximum stays below 2^15.
double mix_lambda_scale(int lmax) {
  const double bound = std::sqrt((2.0 * lmax + 1.0) / (4.0 * 3.14159265358979323846));
  return std::ldexp(1.0, (int)std::floor(std::log2(32768.0 / bound)));
}

int mix_cols_per_block() { return kMixCols; }
int mix_rings_per_lane_table() { return 1; }  // 64 rings per workgroup

}  // namespace wm

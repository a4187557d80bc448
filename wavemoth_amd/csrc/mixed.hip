// Mixed-precision Legendre stage: the contraction q = sum_l Lambda_l a_l on
// the matrix cores, the recurrence (and its seeds) in FP64.
//
// Per m the contraction is a GEMM: q[ring, col] (cols = Re/Im of the maps,
// interleaved) = sum over l of one parity of Lambda[l, ring] a[l, col].  It
// replaces the reference's register-blocked dgemm-like kernels
// (src/legendre_transform.c.in:389-451, 453-618) for many maps.
//
// Workgroup = 4 waves = 128 ring pairs x 32 columns (16 maps) of one m.
// Waves 2j and 2j+1 take rings 64j .. 64j+63 (one per lane) and the even /
// odd l, stepping the FP64 recurrence two l at a time (the x^2 form; same
// monic scaling, seeds and chunk rescale as the FP64 kernel).  Every 64 l
// (a super-chunk) each lane writes its 32 values, converted to the scheme's
// input format, into a wave-private LDS tile; the wave then reads MFMA
// A-fragments from it.  The
// B operand (prescaled a_lm, pre-split at prescale time) is staged in LDS
// per workgroup, prefetched one super-chunk ahead.  MFMA accumulates in
// FP32; every kFlush super-chunks the FP32 sums are added into FP64
// accumulators (the FP64 reduction across l-chunks), which are written to
// the same q buffer the FP64 path fills.
//
// Schemes (products of input parts, FP32 accumulation):
//   0 FP32   v_mfma_f32_16x16x4_f32 on fp32 Lambda and a.
//   1 BF16x3 Lambda = h + l, a = h + l (bf16 parts): hh + hl + lh.
//   2 BF16x6 three bf16 parts each: hh + hm + mh + hl + mm + lh.
//   3 FP16x3 Lambda and a split into two fp16 parts each (~22 bits):
//            hh + hl + lh.  fp16's range is kept by power-of-two scales:
//            Lambda (the true value, A_l times the monic one) by 2^e with
//            |Lambda| <= sqrt((2 lmax + 1) / 4 pi), each a_lm column by a
//            power of two from its maximum (found per execute), both undone
//            when the FP32 sums are added into FP64.
// BF16 schemes use v_mfma_f32_16x16x32_bf16, FP16x3 v_mfma_f32_16x16x32_f16.
#include <cmath>

#include "wm_internal.h"

namespace wm {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));

#define WM_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const WM_CONST T *cview(const T *p) {
  return (const WM_CONST T *)p;
}

constexpr int kMixWaves = 4;
constexpr int kMixRW = 32;                     // rings per wave
constexpr int kMixRB = kMixWaves * kMixRW;     // rings per workgroup (= 64 * RL with RL 2)
constexpr int kMixCols = 32;                   // real columns per workgroup
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

// Scale of a_lm column c for the FP16 scheme: 2^(13 - floor(log2 max|a|)),
// so the column's largest value lands in [2^13, 2^14).
__device__ __forceinline__ double col_scale(const unsigned long long *colmax, int c) {
  const double mx = __builtin_bit_cast(double, colmax[c]);
  return mx > 0.0 ? ldexp(1.0, 13 - ilogb(mx)) : 1.0;
}

// Per-column max |a| (bit patterns of non-negative doubles order as
// integers): block-level max in LDS, then one global atomic per column.
__global__ __launch_bounds__(256) void k_mix_colmax_clear(unsigned long long *colmax, int NCp) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < NCp) colmax[c] = 0ull;
}

__global__ __launch_bounds__(256) void k_mix_colmax(const double *__restrict__ alm,
                                                    int64_t nrows, int C2,
                                                    unsigned long long *colmax) {
  extern __shared__ unsigned long long smax[];
  for (int c = threadIdx.x; c < C2; c += blockDim.x) smax[c] = 0ull;
  __syncthreads();
  const int64_t per = (nrows + gridDim.x - 1) / gridDim.x;
  const int64_t r0 = blockIdx.x * per;
  const int64_t r1 = r0 + per < nrows ? r0 + per : nrows;
  for (int64_t e = r0 * C2 + threadIdx.x; e < r1 * C2; e += blockDim.x) {
    const int c = (int)(e % C2);
    atomicMax(&smax[c], __builtin_bit_cast(unsigned long long, fabs(alm[e])));
  }
  __syncthreads();
  for (int c = threadIdx.x; c < C2; c += blockDim.x)
    if (smax[c]) atomicMax(&colmax[c], smax[c]);
}

__device__ __forceinline__ unsigned short f16_bits(double v) {
  return __builtin_bit_cast(unsigned short, (_Float16)(float)v);
}
__device__ __forceinline__ double f16_value(unsigned short b) {
  return (double)(float)__builtin_bit_cast(_Float16, b);
}

// ------------------------------------------------------------- prescale

// B operand of chunk c (16 l from m + 16 (c - off[m]/16)), parity p, the
// i-th l of that parity (l = l_c + 2 i + p), column col (= 2 map + Re/Im):
//   FP32:  bf[((c * 2 + p) * 8 + i) * NCp + col]
//   BF16:  part q at bq + q * part_stride, [((c * 2 + p) * NCp + col) * 8 + i]
// Values are A_l a_lm (the FP64 path's prescale), zero past lmax and for
// padded columns.  Three trailing zero chunks let a super-chunk of the last
// m read past its end.
__device__ __forceinline__ unsigned short bf16_bits(double v) {
  const __bf16 h = (__bf16)(float)v;
  return __builtin_bit_cast(unsigned short, h);
}
__device__ __forceinline__ double bf16_value(unsigned short b) {
  return (double)__builtin_bit_cast(float, (unsigned int)b << 16);
}

__global__ __launch_bounds__(256) void k_mix_prep(LegTables t, const double2 *__restrict__ alm,
                                                  void *bmat, int NCp, int scheme,
                                                  int64_t part_stride,
                                                  const unsigned long long *colmax) {
  const int64_t total = t.nlpad * NCp;
  for (int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; g < total;
       g += (int64_t)gridDim.x * blockDim.x) {
    // thread order follows the output layout (innermost index fastest)
    int64_t c;
    int p, i, col;
    if (scheme == 0) {
      col = (int)(g % NCp);
      const int64_t r = g / NCp;  // (c * 2 + p) * 8 + i
      i = (int)(r & 7);
      p = (int)((r >> 3) & 1);
      c = r >> 4;
    } else {
      i = (int)(g & 7);
      const int64_t r = g >> 3;  // (c * 2 + p) * NCp + col
      col = (int)(r % NCp);
      p = (int)((r / NCp) & 1);
      c = r / NCp >> 1;
    }
    const int64_t j = c * kChunk + 2 * i + p;
    const int m = t.chunk_m[c];
    const int l = m + (int)(j - t.off[m]);
    const double A = t.ascale[j];
    const int k = col >> 1;
    double v = 0.0;
    if (k < t.K && A != 0.0) {
      const int64_t idx = (int64_t)m * (2 * t.L - m + 3) / 2 + (l - m);
      const double2 a = alm[idx * t.K + k];
      // FP16: the true a, column-scaled (Lambda carries A_l); else A_l a
      v = (scheme == 3 ? col_scale(colmax, col) : A) * ((col & 1) ? a.y : a.x);
    }
    if (scheme == 3) {
      unsigned short *bq = (unsigned short *)bmat;
      const unsigned short h = f16_bits(v);
      bq[g] = h;
      bq[part_stride + g] = f16_bits(v - f16_value(h));
      continue;
    }
    if (scheme == 0) {
      ((float *)bmat)[g] = (float)v;
      continue;
    }
    unsigned short *bq = (unsigned short *)bmat;
    const unsigned short h = bf16_bits(v);
    const double r1 = v - bf16_value(h);
    const unsigned short mid = bf16_bits(r1);
    bq[g] = h;
    bq[part_stride + g] = mid;
    if (scheme == 2) {
      const double r2 = r1 - bf16_value(mid);
      bq[2 * part_stride + g] = bf16_bits(r2);
    }
  }
}

// ------------------------------------------------------------- contraction

template <int S>
struct MixLds {
  // A tile per wave and part: 64 rings x (32 values of the wave's parity + pad)
  static constexpr int a_row = (S == 0) ? (kKK + 4) : (kKK + 8);  // elements
  static constexpr int a_elem = (S == 0) ? 4 : 2;                   // bytes
  static constexpr int a_part = 64 * a_row;                         // elements
  static constexpr int a_wave = MixScheme<S>::parts * a_part * a_elem;
  // B tile: per part, 4 chunks x 2 parities x 8 l x 32 cols
  static constexpr int b_part = 4 * 2 * 8 * kMixCols * ((S == 0) ? 4 : 2);
  static constexpr int bytes = kMixWaves * a_wave + MixScheme<S>::parts * b_part;
};

// Workgroup = 4 waves over 128 ring pairs x 32 columns of one m: wave w
// takes rings 64 (w >> 1) .. +63 (one per lane) and the l of parity w & 1,
// advanced two at a time by L~_{l+2} = (x^2 - d_l) L~_l - e_l L~_{l-2}
// (d_l = c~_l + c~_{l+1}, e_l = c~_l c~_{l-1}; the x^2 form of
// src/legendre_transform.c.in:680-716), so every coefficient is
// wave-uniform and no lane repeats another's work.
template <int S>
__global__ __launch_bounds__(256, 2) void k_mix(LegTables t, const void *bmat, int NCp, int ncb,
                                                int bpm, int64_t part_stride, double lam_scale,
                                                const unsigned long long *colmax,
                                                const double *dtil, const double *etil) {
  using Sc = MixScheme<S>;
  using Ly = MixLds<S>;
  constexpr int P = Sc::parts;
  extern __shared__ __attribute__((aligned(16))) unsigned char lds[];

  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x;
  const int s8 = b >> 3;
  const int m = (b & 7) + 8 * (s8 / bpm);
  if (m > t.M) return;
  const int item = s8 % bpm;
  const int cb = item / t.W, wblk = item - cb * t.W;
  const int L = t.L, R = t.R, Mp1 = t.M + 1, K = t.K;
  const int pw = wv & 1;                            // parity of this wave
  const int ring0 = wblk * kMixRB + (wv >> 1) * 64;  // first ring of this wave

  // output mapping of the 16x16 MFMA tiles: row 4 g + r, column i16
  const int i16 = lane & 15, g = lane >> 4;
  const int col0 = cb * kMixCols;

  int l0 = __builtin_amdgcn_readfirstlane(t.wstart[(int64_t)m * t.W + wblk]);
  const int lsmax = __builtin_amdgcn_readfirstlane(t.wseedmax[(int64_t)m * t.W + wblk]);

  double qd[4][2][4];  // [row tile][col tile][r]
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int ct = 0; ct < 2; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) qd[rt][ct][r] = 0.0;

  if (l0 <= L) {
    const int64_t offm = t.off[m];
    const WM_CONST double *ct_ = cview(t.ctil + (offm - m));
    const WM_CONST double *dt_ = cview(dtil + (offm - m));
    const WM_CONST double *et_ = cview(etil + (offm - m));
    const WM_CONST double *cs_ = cview(t.cscale + offm / kChunk);
    const WM_CONST double *as_ = cview(t.ascale + (offm - m));  // A_l, 0 past lmax
    const int nch = (L - m + kChunk) / kChunk;  // chunks of m's padded l-range

    const int rr = ring0 + lane;
    const bool rok = rr < R;
    const double x = rok ? t.ring_x[rr] : 0.0;
    const double x2 = x * x;
    const int sl = rok ? t.seed_l[(int64_t)m * R + rr] : kSeedNone;
    double u = 0.0, v = 0.0;  // L~ at the next two l of parity pw

    unsigned char *abase = lds + wv * Ly::a_wave;
    unsigned char *bbase = lds + kMixWaves * Ly::a_wave;

    // B staging: thread -> 16 (BF16/FP16, per part) or 32 (FP32) bytes per
    // super-chunk, prefetched into b0..b2 one super-chunk ahead
    const int tid = threadIdx.x;
    const int brow = (S == 0) ? (tid >> 2) : (tid >> 5);  // FP32: (cc, p, i); 16-bit: (cc, p)
    const int bcol = (S == 0) ? 8 * (tid & 3) : (tid & 31);
    const char *bsrc0;
    if constexpr (S == 0)
      bsrc0 = (const char *)((const float *)bmat +
                             ((int64_t)(brow >> 4) * 16 + (brow & 15)) * NCp + col0 + bcol);
    else
      bsrc0 = (const char *)((const unsigned short *)bmat +
                             ((int64_t)(brow >> 1) * 2 * NCp + (brow & 1) * NCp + col0 + bcol) * 8);
    // byte stride between consecutive chunks of the B layout
    const int64_t bchunk = (S == 0) ? (int64_t)16 * NCp * 4 : (int64_t)2 * NCp * 8 * 2;
    const int64_t bpart = part_stride * 2;
    unsigned char *bdst = (S == 0) ? bbase + (brow * kMixCols + bcol) * 4
                                   : bbase + (brow * kMixCols + bcol) * 16;
    uint4 b0, b1, b2;
#define WM_LOAD_B(lsc)                                                              \
  do {                                                                              \
    const char *src_ = bsrc0 + (offm / kChunk + ((lsc) - m) / kChunk) * bchunk;     \
    b0 = *(const uint4 *)src_;                                                      \
    if constexpr (S == 0) b1 = *(const uint4 *)(src_ + 16);                         \
    if constexpr (S >= 1) b1 = *(const uint4 *)(src_ + bpart);                      \
    if constexpr (S == 2) b2 = *(const uint4 *)(src_ + 2 * bpart);                  \
  } while (0)
#define WM_STORE_B()                                                                \
  do {                                                                              \
    *(uint4 *)bdst = b0;                                                            \
    if constexpr (S == 0) *(uint4 *)(bdst + 16) = b1;                               \
    if constexpr (S >= 1) *(uint4 *)(bdst + Ly::b_part) = b1;                       \
    if constexpr (S == 2) *(uint4 *)(bdst + 2 * Ly::b_part) = b2;                   \
  } while (0)

    // undoes the FP16 scheme's scales of this lane's two columns (exact)
    double unscale[2] = {1.0, 1.0};
    if constexpr (S == 3) {
#pragma unroll
      for (int ct = 0; ct < 2; ++ct) {
        const int col = col0 + 16 * ct + i16;
        unscale[ct] = 1.0 / (lam_scale * col_scale(colmax, col < NCp ? col : 0));
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
    zero_acc();

    WM_LOAD_B(l0);
    int nsc = 0;
    for (; l0 <= L; l0 += kSC) {
      __syncthreads();  // previous super-chunk's B reads are done
      WM_STORE_B();
      __syncthreads();
      if (l0 + kSC <= L) WM_LOAD_B(l0 + kSC);

      // ---- producer: 4 chunks of 16 l; this lane's ring, the wave's parity
#pragma unroll
      for (int cc = 0; cc < 4; ++cc) {
        const int lc = l0 + kChunk * cc;
        float f[8];
        if ((lc - m) / kChunk < nch) {
          if (lc <= lsmax && sl == lc) {
            // three-term steps from the seed (L~_{lc-1}, L~_{lc}) to this
            // parity's pair: (L~_lc, L~_lc+2) or (L~_lc+1, L~_lc+3)
            const double2 sv = t.seed_v[(int64_t)m * R + rr];
            const double t1 = fma(x, sv.y, -(ct_[lc] * sv.x));
            const double t2 = fma(x, t1, -(ct_[lc + 1] * sv.y));
            const double t3 = fma(x, t2, -(ct_[lc + 2] * t1));
            u = pw ? t1 : sv.y;
            v = pw ? t3 : t2;
          }
#pragma unroll
          for (int i = 0; i < 8; ++i) {
            const int l = lc + pw + 2 * i;  // wave-uniform
            if constexpr (S == 3)
              f[i] = (float)(lam_scale * (u * as_[l]));
            else
              f[i] = (float)u;
            const double w = fma(x2 - dt_[l + 2], v, -(et_[l + 2] * u));
            u = v;
            v = w;
          }
          const double sc = cs_[(lc - m) / kChunk];
          u *= sc;
          v *= sc;
        } else {
#pragma unroll
          for (int i = 0; i < 8; ++i) f[i] = 0.f;
        }
        // write the 8 values of this chunk: A[q][ring][8 cc + i]
        if constexpr (S == 0) {
          float *dst = (float *)abase + lane * Ly::a_row + 8 * cc;
          *(f32x4 *)dst = (f32x4){f[0], f[1], f[2], f[3]};
          *(f32x4 *)(dst + 4) = (f32x4){f[4], f[5], f[6], f[7]};
        } else {
          // parts packed two 16-bit values per dword: wh (high), wm (middle),
          // wl (low)
          unsigned wh[4], wm[4], wl[4];
#pragma unroll
          for (int i = 0; i < 8; i += 2) {
            if constexpr (S == 3) {
              const _Float16 h0 = (_Float16)f[i], h1 = (_Float16)f[i + 1];
              const _Float16 m0 = (_Float16)(f[i] - (float)h0);
              const _Float16 m1 = (_Float16)(f[i + 1] - (float)h1);
              wh[i >> 1] = __builtin_bit_cast(unsigned short, h0) |
                           ((unsigned)__builtin_bit_cast(unsigned short, h1) << 16);
              wm[i >> 1] = __builtin_bit_cast(unsigned short, m0) |
                           ((unsigned)__builtin_bit_cast(unsigned short, m1) << 16);
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
          unsigned short *dst = (unsigned short *)abase + lane * Ly::a_row + 8 * cc;
          *(uint4 *)dst = make_uint4(wh[0], wh[1], wh[2], wh[3]);
          *(uint4 *)(dst + Ly::a_part) = make_uint4(wm[0], wm[1], wm[2], wm[3]);
          if constexpr (S == 2)
            *(uint4 *)(dst + 2 * Ly::a_part) = make_uint4(wl[0], wl[1], wl[2], wl[3]);
        }
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");

      // ---- consumer: MFMA over the super-chunk (this wave's parity)
      if constexpr (S == 0) {
        const float *af = (const float *)abase;
        const float *bf = (const float *)bbase;
#pragma unroll
        for (int s = 0; s < 8; ++s) {
          float a[4], bv[2];
#pragma unroll
          for (int rt = 0; rt < 4; ++rt) a[rt] = af[(16 * rt + i16) * Ly::a_row + 4 * s + g];
#pragma unroll
          for (int ct = 0; ct < 2; ++ct)
            bv[ct] = bf[((((s >> 1) * 2 + pw) * 8) + 4 * (s & 1) + g) * kMixCols + 16 * ct + i16];
#pragma unroll
          for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct)
              acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[rt], bv[ct], acc[rt][ct], 0, 0, 0);
        }
      } else {
        bf16x8 a[P][4], bv[P][2];
#pragma unroll
        for (int q = 0; q < P; ++q) {
#pragma unroll
          for (int rt = 0; rt < 4; ++rt)
            a[q][rt] = *(const bf16x8 *)((const unsigned short *)abase + q * Ly::a_part +
                                         (16 * rt + i16) * Ly::a_row + 8 * g);
#pragma unroll
          for (int ct = 0; ct < 2; ++ct)
            bv[q][ct] = *(const bf16x8 *)(bbase + q * Ly::b_part +
                                          (((g * 2 + pw) * kMixCols) + 16 * ct + i16) * 16);
        }
#pragma unroll
        for (int pi = 0; pi < Sc::nprod; ++pi)
#pragma unroll
          for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < 2; ++ct) {
              if constexpr (S == 3)
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                    __builtin_bit_cast(f16x8, a[Sc::pa(pi)][rt]),
                    __builtin_bit_cast(f16x8, bv[Sc::pb(pi)][ct]), acc[rt][ct], 0, 0, 0);
              else
                acc[rt][ct] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    a[Sc::pa(pi)][rt], bv[Sc::pb(pi)][ct], acc[rt][ct], 0, 0, 0);
            }
      }
      if (++nsc == kFlush) {
        flush();
        nsc = 0;
      }
    }
    flush();
#undef WM_LOAD_B
#undef WM_STORE_B
  }

  // ---- store: column col = 2 map + (Re, Im) of q(ring, pw, map, m)
  double *qd_out = (double *)t.qbuf;
#pragma unroll
  for (int rt = 0; rt < 4; ++rt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ring = ring0 + 16 * rt + 4 * g + r;
      if (ring < R) {
#pragma unroll
        for (int ct = 0; ct < 2; ++ct) {
          const int col = col0 + 16 * ct + i16;
          if (col < 2 * K)
            qd_out[2 * q_index(ring, pw, col >> 1, m, K, Mp1) + (col & 1)] = qd[rt][ct][r];
        }
      }
    }
}

// ------------------------------------------------------------- launchers

int64_t mix_bmat_elems(const LegTables &t, int NCp) {
  return (t.nlpad / kChunk + 3) * 2 * (int64_t)NCp * 8;
}

hipError_t launch_mix_prep(const LegTables &t, const double2 *alm, void *bmat, int NCp,
                           int scheme, int64_t part_stride, unsigned long long *colmax,
                           hipStream_t s) {
  if (scheme == 3) {
    hipLaunchKernelGGL(k_mix_colmax_clear, dim3((NCp + 255) / 256), dim3(256), 0, s, colmax,
                       NCp);
    const int64_t nalm = (int64_t)(t.M + 1) * (t.L + 1) - (int64_t)t.M * (t.M + 1) / 2;
    const int C2 = 2 * t.K;
    hipLaunchKernelGGL(k_mix_colmax, dim3(1024), dim3(256), (size_t)C2 * 8, s,
                       (const double *)alm, nalm, C2, colmax);
  }
  const int64_t total = t.nlpad * NCp;
  const int64_t nb = (total + 255) / 256;
  hipLaunchKernelGGL(k_mix_prep, dim3((unsigned)(nb < 1048576 ? nb : 1048576)), dim3(256), 0,
                     s, t, alm, bmat, NCp, scheme, part_stride, (const unsigned long long *)colmax);
  return hipGetLastError();
}

template <int S>
static hipError_t launch_mix_t(const LegTables &t, const void *bmat, int NCp,
                               int64_t part_stride, double lam_scale,
                               const unsigned long long *colmax, const double *dtil,
                               const double *etil, hipStream_t s) {
  const int ncb = NCp / kMixCols;
  const int bpm = t.W * ncb;
  const int mgroups = (t.M + 1 + 7) / 8;
  const unsigned grid = (unsigned)(8 * mgroups * bpm);
  hipLaunchKernelGGL(k_mix<S>, dim3(grid), dim3(256), (size_t)MixLds<S>::bytes, s, t, bmat, NCp,
                     ncb, bpm, part_stride, lam_scale, colmax, dtil, etil);
  return hipGetLastError();
}

hipError_t mix_set_lds_limits() {
  hipError_t e = hipFuncSetAttribute((const void *)k_mix<0>,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, MixLds<0>::bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_mix<1>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            MixLds<1>::bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_mix<2>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            MixLds<2>::bytes);
  if (e == hipSuccess)
    e = hipFuncSetAttribute((const void *)k_mix<3>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            MixLds<3>::bytes);
  return e;
}

hipError_t launch_mix(const LegTables &t, const void *bmat, int NCp, int scheme,
                      int64_t part_stride, double lam_scale, const unsigned long long *colmax,
                      const double *dtil, const double *etil, hipStream_t s) {
  switch (scheme) {
    case 0: return launch_mix_t<0>(t, bmat, NCp, part_stride, lam_scale, colmax, dtil, etil, s);
    case 1: return launch_mix_t<1>(t, bmat, NCp, part_stride, lam_scale, colmax, dtil, etil, s);
    case 2: return launch_mix_t<2>(t, bmat, NCp, part_stride, lam_scale, colmax, dtil, etil, s);
    default:
      return launch_mix_t<3>(t, bmat, NCp, part_stride, lam_scale, colmax, dtil, etil, s);
  }
}

// Power-of-two scale of Lambda for the FP16 scheme: |Lambda_lm| is at most
// sqrt((2 lmax + 1) / (4 pi)); the scaled maximum stays below 2^15.
double mix_lambda_scale(int lmax) {
  const double bound = std::sqrt((2.0 * lmax + 1.0) / (4.0 * 3.14159265358979323846));
  return std::ldexp(1.0, (int)std::floor(std::log2(32768.0 / bound)));
}

int mix_cols_per_block() { return kMixCols; }
int mix_rings_per_lane_table() { return kMixRB / 64; }

}  // namespace wm

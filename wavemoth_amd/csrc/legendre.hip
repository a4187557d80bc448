// Legendre stage of the a_lm -> map synthesis on gfx950, in FP64.
//
// For every (m, ring pair r) the engine needs the even and odd partial sums
//   E[r][m] = sum_{l-m even} Lambda_l^m(x_r) a_lm,
//   O[r][m] = sum_{l-m odd}  Lambda_l^m(x_r) a_lm,
// which the reference computes per (m, parity) with the x^2 two-step
// recurrence from stored seeds (src/legendre_transform.c.in:680-747,
// src/wavemoth.c:958-1028) and keeps in work_q[m][odd][ring].
//
// Here Lambda is generated on the fly with the one-step recurrence written
// in monic form.  With Lambda_{l+1} = a_l x Lambda_l - b_l Lambda_{l-1}
// (the normalised recurrence of libpshtlight/ylmgen_c.c:604-619, 643-732)
// and Lambda_l = A_l L~_l, A_{l+1} = a_l A_l, the scaled values obey
//   L~_{l+1} = x L~_l - c~_l L~_{l-1},   c~_l = (l^2 - m^2) / (4 l^2 - 1),
// i.e. one multiply and one FMA per (l, ring) instead of three operations.
// A_l is reset to 1 at the start of every kChunk-long l-chunk; the state
// is multiplied by the chunk's end product once per chunk and the
// contraction uses a~_lm = A_l a_lm, prescaled once per execute.
// Both parities come out of the same sweep (odd and even l alternate).
//
// Seeds: Lambda_mm = (-1)^m sqrt(prod (2i+1)/(2i)) / sqrt(4 pi) sin^m theta
// underflows FP64 for large m near the poles, so the seed kernel carries a
// separate binary exponent, runs the recurrence through the negligible
// region and records, per (m, ring), the chunk-aligned l where |Lambda|
// first reaches 2^-107 together with the two recurrence states there.  This
// is the GPU form of Ylmgen's firstl and the reference's stripify start
// rows (wavemoth/lib.pyx:301-368, 397-487); terms skipped before that
// point are below 1e-32 in magnitude.
#include "wm_internal.h"

namespace wm {

// -------------------------------------------------------------- seeds

// One thread per (m, ring).  Writes seed_l / seed_v.
__global__ __launch_bounds__(256) void k_seeds(LegTables t) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int R = t.R, L = t.L;
  if (gid >= (int64_t)(t.M + 1) * R) return;
  const int m = (int)(gid / R), r = (int)(gid - (int64_t)m * R);
  const double x = t.ring_x[r];

  // log2 |Lambda_mm| = mfac[m] + m log2 sin(theta), in double-double so
  // the seed is good to ~1 ulp even when the exponent is ~ -10^4.
  const double2 mf = t.mfac[m];
  const double2 ls = t.ring_ls[r];
  const double dm = (double)m;
  double p = dm * ls.x;
  double pe = fma(dm, ls.x, -p) + dm * ls.y;
  double hi = mf.x + p;
  double bb = hi - mf.x;
  double lo = (mf.x - (hi - bb)) + (p - bb) + mf.y + pe;
  double E = floor(hi);
  double frac = (hi - E) + lo;
  if (frac >= 1.0) { E += 1.0; frac -= 1.0; }
  if (frac < 0.0) { E -= 1.0; frac += 1.0; }
  int iE = (int)E;
  double cur = exp2(frac);
  if (m & 1) cur = -cur;
  double prev = 0.0;

  const int64_t offm = t.off[m];
  const double *ct = t.ctil + (offm - m);
  const double *cs = t.cscale + offm / kChunk;
  int found = kSeedNone;
  double sp = 0.0, sc = 0.0;
  int c = 0;
  for (int l0 = m; l0 <= L; l0 += kChunk, ++c) {
    const double cp = prev, cc = cur;
    const int cE = iE;
#pragma unroll
    for (int j = 0; j < kChunk; ++j) {
      const double nx = fma(x, cur, -ct[l0 + j] * prev);
      prev = cur;
      cur = nx;
    }
    const double S = cs[c];
    prev *= S;
    cur *= S;
    const double big = fmax(fabs(prev), fabs(cur));
    if (!(big > 0.0)) break;
    const int e = ilogb(big);
    if (e + iE >= -107) {
      found = l0;
      sp = ldexp(cp, cE);
      sc = ldexp(cc, cE);
      break;
    }
    prev = ldexp(prev, -e);
    cur = ldexp(cur, -e);
    iE += e;
  }
  t.seed_l[gid] = found;
  t.seed_v[gid] = make_double2(sp, sc);
}

// One thread per (m, wave): the wave-uniform first l and the last l at
// which some lane of the wave still has to pick up its seed.
__global__ __launch_bounds__(256) void k_wave_bounds(LegTables t) {
  const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (gid >= (int64_t)(t.M + 1) * t.W) return;
  const int m = (int)(gid / t.W), w = (int)(gid - (int64_t)m * t.W);
  int lo = kSeedNone, hi = -1;
  const int rings = 64 * t.RL;
  for (int i = 0; i < rings; ++i) {
    const int r = w * rings + i;
    if (r >= t.R) break;
    const int s = t.seed_l[(int64_t)m * t.R + r];
    if (s < lo) lo = s;
    if (s <= t.L && s > hi) hi = s;
  }
  t.wstart[gid] = lo;
  t.wseedmax[gid] = hi;
}

// ------------------------------------------------------- a_lm prescale

// map slots of atil: nmaps rounded up to whole blocks of KB
__host__ __device__ inline int legendre_padded_maps_of(int K, int KB) {
  return (K + KB - 1) / KB * KB;
}

// atil[kb][j][k'] = A_l * a_{l m}[kb*KB + k'] on the padded l-grid, zero
// past lmax and for map slots past nmaps: map-block-major so that a wave
// of map block kb reads KB consecutive complex values per l.  One entry:
// padded l-entry j, map slot kk < KB nkb.
__device__ __forceinline__ void prep_entry(const LegTables &t, const double2 *__restrict__ alm,
                                           int64_t j, int kk, int KB) {
  const int K = t.K, L = t.L;
  const int m = t.chunk_m[j / kChunk];
  const int64_t jj = j - t.off[m];
  const int l = m + (int)jj;
  const bool live = l <= L && kk < K;
  const double sc = t.ascale[j];
  const int64_t idx = (int64_t)m * (2 * L - m + 3) / 2 + (l <= L ? jj : 0);
  const double2 a = alm[idx * K + (kk < K ? kk : K - 1)];
  const int kb = kk / KB;
  t.atil[((int64_t)kb * t.nlpad + j) * KB + (kk - kb * KB)] =
      live ? make_double2(a.x * sc, a.y * sc) : make_double2(0.0, 0.0);
}

// The single-plan prescale: one thread per (padded l-entry j, map slot kk),
// maps fastest, so that a wave reads the consecutive a_lm of one l (and of
// JB = 256 / Kp consecutive l) and writes KB-map runs of atil.
__global__ __launch_bounds__(256) void k_prep_alm(LegTables t, const double2 *__restrict__ alm,
                                                  int KB, int nkb) {
  const int Kp = KB * nkb;
  const int JB = Kp >= 256 ? 1 : 256 / Kp;
  const int64_t j0 = (int64_t)blockIdx.x * JB;
  for (int u = threadIdx.x; u < JB * Kp; u += blockDim.x) {
    const int jl = u / Kp, kk = u - jl * Kp;
    const int64_t j = j0 + jl;
    if (j >= t.nlpad) break;  // jl only grows with u
    prep_entry(t, alm, j, kk, KB);
  }
}

// One wave's share of the sweep: ring chunk w (rings w*64*RL + lane + 64*i,
// i < RL) of m, maps kb*KB .. kb*KB+KB-1.  Used by the grid kernel below
// and by the batched work-queue kernel (batch.hip).
// Tables the sweep only reads are accessed through constant-address-space
// views: their loads are then known not to be clobbered by the q stores
// (which matters inside the persistent loop of the batched kernel), and
// wave-uniform ones go through the scalar unit.
#define WM_CONST __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const WM_CONST T *const_view(const T *p) {
  return (const WM_CONST T *)p;
}

// The sweep proper.
template <int RL, int KB>
__device__ __forceinline__ void legendre_sweep(
    int L, int R, int Mp1, int K, int m, int w, int kbase, int lane, int l0, int lsmax,
    const WM_CONST double *ct, const WM_CONST double *at,
    const WM_CONST double *cs, const double *__restrict__ ring_x,
    const int *__restrict__ seed_l, const double2 *__restrict__ seed_v,
    double2 *__restrict__ qbuf) {
  double x[RL], prev[RL], cur[RL];
  int sl[RL], rr[RL];
  double acc[RL][2][KB][2];
#pragma unroll
  for (int i = 0; i < RL; ++i) {
    rr[i] = w * 64 * RL + lane + 64 * i;
    const bool ok = rr[i] < R;
    x[i] = ok ? ring_x[rr[i]] : 0.0;
    sl[i] = ok ? seed_l[(int64_t)m * R + rr[i]] : kSeedNone;
    prev[i] = 0.0;
    cur[i] = 0.0;
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
      for (int k = 0; k < KB; ++k) acc[i][p][k][0] = acc[i][p][k][1] = 0.0;
  }

  for (; l0 <= L; l0 += kChunk) {
    if (l0 <= lsmax) {
#pragma unroll
      for (int i = 0; i < RL; ++i) {
        if (sl[i] == l0) {
          const double2 v = seed_v[(int64_t)m * R + rr[i]];
          prev[i] = v.x;
          cur[i] = v.y;
        }
      }
    }
#pragma unroll
    for (int j = 0; j < kChunk; j += 2) {
      const int l = l0 + j;
      const double c0 = ct[l], c1 = ct[l + 1];
      double2 a0[KB], a1[KB];
#pragma unroll
      for (int k = 0; k < KB; ++k) {
        const int64_t i0 = 2 * ((int64_t)l * KB + k), i1 = i0 + 2 * KB;
        a0[k] = make_double2(at[i0], at[i0 + 1]);
        a1[k] = make_double2(at[i1], at[i1 + 1]);
      }
#pragma unroll
      for (int i = 0; i < RL; ++i) {
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          acc[i][0][k][0] = fma(cur[i], a0[k].x, acc[i][0][k][0]);
          acc[i][0][k][1] = fma(cur[i], a0[k].y, acc[i][0][k][1]);
        }
        const double n1 = fma(x[i], cur[i], -(c0 * prev[i]));
        prev[i] = cur[i];
        cur[i] = n1;
#pragma unroll
        for (int k = 0; k < KB; ++k) {
          acc[i][1][k][0] = fma(cur[i], a1[k].x, acc[i][1][k][0]);
          acc[i][1][k][1] = fma(cur[i], a1[k].y, acc[i][1][k][1]);
        }
        const double n2 = fma(x[i], cur[i], -(c1 * prev[i]));
        prev[i] = cur[i];
        cur[i] = n2;
      }
    }
    const double S = cs[(l0 - m) / kChunk];
#pragma unroll
    for (int i = 0; i < RL; ++i) {
      prev[i] *= S;
      cur[i] *= S;
    }
  }

#pragma unroll
  for (int i = 0; i < RL; ++i) {
    if (rr[i] < R) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
#pragma unroll
        for (int k = 0; k < KB; ++k)
          if (kbase + k < K)
            qbuf[q_index(rr[i], p, kbase + k, m, K, Mp1)] =
                make_double2(acc[i][p][k][0], acc[i][p][k][1]);
      }
    }
  }
}

// One wave's share of the sweep: ring chunk w (rings w*64*RL + lane + 64*i,
// i < RL) of m, maps kb*KB .. kb*KB+KB-1.  Used by the grid kernel below
// and by the batched work-queue kernel.
template <int RL, int KB>
__device__ __forceinline__ void legendre_item(const LegTables &t, int m, int w, int kb,
                                              int lane) {
  const int l0 = __builtin_amdgcn_readfirstlane(t.wstart[(int64_t)m * t.W + w]);
  const int lsmax = __builtin_amdgcn_readfirstlane(t.wseedmax[(int64_t)m * t.W + w]);
  const int64_t offm = t.off[m];
  legendre_sweep<RL, KB>(
      t.L, t.R, t.M + 1, t.K, m, w, kb * KB, lane, l0, lsmax, const_view(t.ctil + (offm - m)),
      const_view((const double *)(t.atil + ((int64_t)kb * t.nlpad + offm - m) * KB)),
      const_view(t.cscale + offm / kChunk), t.ring_x, t.seed_l, t.seed_v, t.qbuf);
}

// Block = 4 waves.  Block b runs on XCD b % 8; every m's blocks are kept
// on one XCD (m = xcd + 8 * i) so the per-m coefficient rows are fetched
// into one L2 only, and each XCD walks its m's from cheap-to-start (small
// m, long l-range) to short ones.
template <int RL, int KB>
__global__ __launch_bounds__(256) void k_legendre(LegTables t, int nkb, int bpm) {
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int b = blockIdx.x;
  const int s = b >> 3;
  const int m = (b & 7) + 8 * (s / bpm);
  if (m > t.M) return;
  const int item = (s % bpm) * 4 + wv;
  if (item >= t.W * nkb) return;
  const int kb = item / t.W, w = item - kb * t.W;
  legendre_item<RL, KB>(t, m, w, kb, lane);
}

// ------------------------------------------------------------ launchers

static inline unsigned nblocks(int64_t n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t launch_seeds(const LegTables &t, hipStream_t s) {
  const int64_t n = (int64_t)(t.M + 1) * t.R;
  hipLaunchKernelGGL(k_seeds, dim3(nblocks(n, 256)), dim3(256), 0, s, t);
  return hipGetLastError();
}

hipError_t launch_wave_bounds(const LegTables &t, hipStream_t s) {
  const int64_t n = (int64_t)(t.M + 1) * t.W;
  hipLaunchKernelGGL(k_wave_bounds, dim3(nblocks(n, 256)), dim3(256), 0, s, t);
  return hipGetLastError();
}

// Maps per pass (KB) from nmaps; rings per lane (RL) is a plan choice
// (heuristic below, or measured by the autotuner).  Two rings per lane and
// at most four maps a pass: each wave then reads a coefficient row for
// twice the rings, and the 8-map pass's registers (3 waves per SIMD at two
// rings) are avoided; measured at Nside 2048, lmax 4096, the sweep takes
// 140 ms at 64 maps against 222 ms with one ring and 8 maps per pass
// (18.3 against 27.9 at 8 maps, 7.8 against 11.9 at 3).
int legendre_map_block(int K) { return K <= 4 ? K : 4; }
int legendre_default_rl(int K) { return 2; }
int legendre_padded_maps(int K) {
  const int KB = legendre_map_block(K);
  return ((K + KB - 1) / KB) * KB;
}

hipError_t launch_prep_alm(const LegTables &t, const double2 *alm, hipStream_t s) {
  const int KB = t.KB;
  const int nkb = (t.K + KB - 1) / KB, Kp = KB * nkb;
  const int JB = Kp >= 256 ? 1 : 256 / Kp;
  hipLaunchKernelGGL(k_prep_alm, dim3((unsigned)((t.nlpad + JB - 1) / JB)), dim3(256), 0, s, t,
                     alm, KB, nkb);
  return hipGetLastError();
}

template <int RL, int KB>
static hipError_t launch_leg_t(const LegTables &t, hipStream_t s) {
  const int nkb = (t.K + KB - 1) / KB;
  const int bpm = (t.W * nkb + 3) / 4;
  const int mgroups = (t.M + 1 + 7) / 8;
  const unsigned grid = (unsigned)(8 * mgroups * bpm);
  hipLaunchKernelGGL((k_legendre<RL, KB>), dim3(grid), dim3(256), 0, s, t, nkb, bpm);
  return hipGetLastError();
}

template <int RL>
static hipError_t launch_leg_rl(const LegTables &t, hipStream_t s) {
  switch (t.KB) {
    case 1: return launch_leg_t<RL, 1>(t, s);
    case 2: return launch_leg_t<RL, 2>(t, s);
    case 3: return launch_leg_t<RL, 3>(t, s);
    case 4: return launch_leg_t<RL, 4>(t, s);
    default: return launch_leg_t<RL, 8>(t, s);
  }
}

// ---------------------------------------------------------------- batch

__global__ void k_batch_reset(int *counters) {
  if (threadIdx.x < kBatchCounters) counters[threadIdx.x] = 0;
}

// Prescale of every request's a_lm: thread g owns entry u = g - prep_off[r]
// of request r, u = j Kp + kk (padded l-entry j, map slot kk of Kp = KB nkb
// slots; maps fastest, as the single-plan prescale).
__global__ __launch_bounds__(256) void k_batch_prep(BatchDev b) {
  __shared__ int s_r0;
  const int64_t g0 = (int64_t)blockIdx.x * blockDim.x;
  if (threadIdx.x == 0) {
    int lo = 0, hi = b.nreq - 1;  // prep_off[lo] <= g0 < prep_off[lo + 1]
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (b.prep_off[mid] <= g0) lo = mid; else hi = mid - 1;
    }
    s_r0 = lo;
  }
  __syncthreads();
  const int64_t g = g0 + threadIdx.x;
  if (g >= b.prep_total) return;
  int lo = s_r0;  // the block's first request; few blocks span more than one
  while (g >= b.prep_off[lo + 1]) ++lo;
  const LegTables &t = b.tabs[lo];
  const int Kp = legendre_padded_maps_of(t.K, t.KB);
  const int64_t u = g - b.prep_off[lo];
  const int64_t j = u / Kp;
  prep_entry(t, b.alm[lo], j, (int)(u - j * Kp), t.KB);
}

#define WM_BATCH_SHAPES(X) \
  X(1, 1) X(1, 2) X(1, 3) X(1, 4) X(1, 8) X(2, 1) X(2, 2) X(2, 3) X(2, 4) X(2, 8) X(4, 1)

bool batch_shape_supported(int RL, int KB) { return batch_shape_index(RL, KB) >= 0; }

int batch_shape_index(int RL, int KB) {
  int i = 0;
#define WM_IDX(R_, K_)                     \
  if (RL == R_ && KB == K_) return i;      \
  ++i;
  WM_BATCH_SHAPES(WM_IDX)
#undef WM_IDX
  return -1;
}

// Persistent sweep of one shape: each wave takes items (request, m, ring
// chunk, map block) of the shape's queue until it runs dry; the host
// orders each queue costliest request first.  One kernel per shape, so
// each has the registers of its own shape (as the single-plan sweep).
template <int R_, int K_>
__global__ __launch_bounds__(256) void k_batch_legendre(BatchDev b, int first, int count,
                                                        int ctr) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    int it = 0;
    if (lane == 0) it = atomicAdd(&b.counters[ctr], 1);
    it = __builtin_amdgcn_readfirstlane(__shfl(it, 0));
    if (it >= count) break;
    const WM_CONST BatchLegItem *ip = const_view(b.litems + first + it);
    const int req = ip->req, m = ip->m, w = ip->w, kb = ip->kb;
    const WM_CONST LegTables *tp = const_view(b.tabs + req);
    LegTables t;
    t.L = tp->L;
    t.M = tp->M;
    t.K = tp->K;
    t.R = tp->R;
    t.W = tp->W;
    t.nlpad = tp->nlpad;
    t.off = tp->off;
    t.ctil = tp->ctil;
    t.cscale = tp->cscale;
    t.ring_x = tp->ring_x;
    t.seed_l = tp->seed_l;
    t.seed_v = tp->seed_v;
    t.wstart = tp->wstart;
    t.wseedmax = tp->wseedmax;
    t.atil = tp->atil;
    t.qbuf = tp->qbuf;
    legendre_item<R_, K_>(t, m, w, kb, lane);
  }
}

hipError_t launch_batch_reset(const BatchDev &b, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, b.counters);
  return hipGetLastError();
}

hipError_t launch_batch_prep(const BatchDev &b, hipStream_t s) {
  if (b.prep_total == 0) return hipSuccess;
  hipLaunchKernelGGL(k_batch_prep, dim3(nblocks(b.prep_total, 256)), dim3(256), 0, s, b);
  return hipGetLastError();
}

hipError_t batch_legendre_occupancy(int RL, int KB, int *blocks_per_cu) {
#define WM_OCC(R_, K_)                                                                   \
  if (RL == R_ && KB == K_)                                                              \
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu,                   \
                                                        k_batch_legendre<R_, K_>, 256, 0);
  WM_BATCH_SHAPES(WM_OCC)
#undef WM_OCC
  return hipErrorInvalidValue;
}

hipError_t launch_batch_legendre(const BatchDev &b, int RL, int KB, int first, int count,
                                 int nb, hipStream_t s) {
  const int ctr = kBatchSweepCtr + batch_shape_index(RL, KB);
#define WM_LAUNCH(R_, K_)                                                                \
  if (RL == R_ && KB == K_) {                                                            \
    hipLaunchKernelGGL((k_batch_legendre<R_, K_>), dim3(nb), dim3(256), 0, s, b, first,  \
                       count, ctr);                                                      \
    return hipGetLastError();                                                            \
  }
  WM_BATCH_SHAPES(WM_LAUNCH)
#undef WM_LAUNCH
  return hipErrorInvalidValue;
}

hipError_t launch_legendre(const LegTables &t, hipStream_t s) {
  switch (t.RL) {
    case 1: return launch_leg_rl<1>(t, s);
    case 2: return launch_leg_rl<2>(t, s);
    case 4: return launch_leg_rl<4>(t, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace wm

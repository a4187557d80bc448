// Adjoint of the synthesis (map -> a_lm): b = Y^T g.
//
// The synthesis evaluates, exactly at every pixel (the aliased c2r of the
// ring stage is an exact trigonometric-polynomial evaluation),
//   f(theta_p, phi_p) = sum_{l, m} c_m Re(a_lm Lambda_l^m(theta_p) e^{i m phi_p}),
// c_0 = 1, c_m = 2 (m > 0).  With the real inner products <f, g> = sum f g and
// <a, b> = sum Re(conj(a) b), its adjoint is
//   b_lm = c_m sum_p g_p Lambda_l^m(theta_p) e^{-i m phi_p},
// the unweighted analysis.  b is also the gradient d<f, g>/dRe(a) +
// i d<f, g>/dIm(a) (torch's convention for complex leaves).  The reference's
// only GPU code computes the transposed Legendre stage of this direction
// (wavemoth/cuda/legendre_transform.cu.in:297-679).
//
// Stages:
//  1. k_ring_adjoint: per ring pair, F_m(ring) = e^{-i m phi0} G^[m mod n]
//     with G^ the forward real DFT of the ring (conjugate-symmetric for
//     bins past n/2), then q_even = F_north + F_south, q_odd = F_north -
//     F_south (the equator ring feeds both), into the [R][2][M+1][K] buffer.
//  2. k_legendre_adj: b~_l = sum_rings L~_l(x_r) q_{parity(l-m)}(r) with the
//     same monic recurrence and seeds as the synthesis; lanes evaluate rings,
//     a 16 x 64 LDS tile per wave transposes them so lanes then own l.
//  3. k_adj_finish: b_lm = c_m A_l sum_groups b~ into the m-major layout.
#include <mutex>

#include "wm_fft16.h"
#include "wm_host.h"

namespace wm {

using namespace fft;

// MAXP: the plan's longest transform (Bluestein length, or twice a power-of-
// two ring's h); the per-thread loops and FFT passes are sized for it.
template <int MAXP>
__global__ __launch_bounds__(NT) void k_ring_adjoint(LegTables t, RingTables rt, int first,
                                                     int lds_slots,
                                                     const double *__restrict__ map) {
  constexpr int kHalf = (MAXP / 2 + 1 + NT - 1) / NT;  // h-range entries per thread
  constexpr int kFull = (MAXP + NT - 1) / NT;          // P-range entries per thread
  extern __shared__ double2 lds[];
  double2 *twl = lds + lds_slots - kTwSlots;
  const RingPair rp = rt.pairs[rt.order[first + blockIdx.x]];
  const int n = rp.n, h = n / 2, K = t.K, M = t.M, P = rp.P;
  const bool south = rp.pix_s >= 0;
  const double2 *chirp = rt.chirp + rp.chirp_off;
  const double2 *vhat = rt.vhat + rp.vhat_off;
  load_twiddles(twl, rt.twiddle);
  const int nhemi = south ? 2 : 1;
  for (int k = 0; k < K; ++k) {
    double2 *qe = t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);  // element m at [m * G]
    double2 *qo = t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
    for (int hemi = 0; hemi < nhemi; ++hemi) {
      int tid = threadIdx.x;
      asm volatile("" : "+v"(tid));
      const double *g = map + (hemi ? rp.pix_s : rp.pix_n) * K;
      // z_p = g_{2p} + i g_{2p+1}; Z = DFT_h(z), computed as conj(IDFT(conj z))
      // for Bluestein lengths
#pragma unroll
      for (int i = 0; i < kHalf; ++i) {
        const int p = tid + NT * i;
        if (p < h) {
          const double2 z = make_double2(g[(int64_t)(2 * p) * K + k], g[(int64_t)(2 * p + 1) * K + k]);
          lds[pidx(p)] = (P == 0) ? z : cmul(cconj(z), chirp[p]);
        }
      }
      if (P == 0) {
        __syncthreads();
        lds_fft<false, MAXP>(lds, h, twl);
      } else {
#pragma unroll
        for (int i = 0; i < kFull; ++i) {
          const int a = h + tid + NT * i;
          if (a < P) lds[pidx(a)] = make_double2(0.0, 0.0);
        }
        __syncthreads();
        lds_fft<false, MAXP>(lds, P, twl);
#pragma unroll
        for (int i = 0; i < kFull; ++i) {
          const int a = tid + NT * i;
          if (a < P) lds[pidx(a)] = cmul(lds[pidx(a)], vhat[a]);
        }
        __syncthreads();
        lds_fft<true, MAXP>(lds, P, twl);
#pragma unroll
        for (int i = 0; i < kHalf; ++i) {
          const int p = tid + NT * i;
          if (p < h) lds[pidx(p)] = cconj(cmul(lds[pidx(p)], chirp[p]));
        }
        __syncthreads();
      }
      // real-DFT split: G^_k = (Z_k + conj Z_{h-k})/2 - (i/2) e^{-2 pi i k/n}
      // (Z_k - conj Z_{h-k}), k in [0, h], Z_h = Z_0.  Each thread reads and
      // rewrites only its own pair (a, h - a), so the split runs in place.
      {
        const double2 step = expipi(-2 * NT, n);
        double2 om = expipi(-2 * (int64_t)tid, n);
#pragma unroll
        for (int i = 0; i < kHalf; ++i) {
          const int a = tid + NT * i;
          if (a <= h / 2) {
            const int c = h - a;
            const double2 za = lds[pidx(a)], zc = lds[pidx(c % h)];
            const double2 e = cadd(za, cconj(zc)), d = csub(za, cconj(zc));
            // -(i/2) om d
            const double2 od = cmul(om, d);
            const double2 ga = make_double2(0.5 * (e.x + od.y), 0.5 * (e.y - od.x));
            // partner: e' = conj(e), d' = -conj(d), om' = e^{-2 pi i c/n} = -conj(om)
            const double2 oc = make_double2(-om.x, om.y);
            const double2 dc = make_double2(-d.x, d.y);
            const double2 odc = cmul(oc, dc);
            const double2 gc = make_double2(0.5 * (e.x + odc.y), 0.5 * (-e.y - odc.x));
            lds[pidx(a)] = ga;
            lds[pidx(c)] = gc;
          }
          om = cmul(om, step);
        }
      }
      __syncthreads();
      // F_m = e^{-i pi s m / n} G^[m mod n] (conjugate past n/2), into q
      {
        const double2 step = expipi(-(int64_t)rp.s * NT, n);
        double2 w = expipi(-(int64_t)rp.s * tid, n);
        for (int m0 = 0; m0 <= M; m0 += NT * kHalf) {
#pragma unroll
          for (int i = 0; i < kHalf; ++i) {
            const int m = m0 + tid + NT * i;
            if (m <= M) {
              if (i == 0 && m0 > 0) w = expipi(-(int64_t)rp.s * m, n);
              const int kk = m % n;
              const double2 v = (kk <= h) ? lds[pidx(kk)] : cconj(lds[pidx(n - kk)]);
              const double2 f = cmul(v, w);
              const int64_t o = (int64_t)m * q_group(K);
              if (hemi == 0) {
                qe[o] = f;
                qo[o] = f;
              } else {
                qe[o] = cadd(qe[o], f);
                qo[o] = csub(qo[o], f);
              }
            }
            w = cmul(w, step);
          }
        }
      }
      __syncthreads();
    }
  }
}

// The ring adjoint of power-of-two rings with mmax <= h (the forward belt
// kernel's rings) on the register FFT: a job of T = h/16 threads per (ring
// pair, map), 512-thread blocks.  Per hemisphere: z_p = g_{2p} + i g_{2p+1}
// for p = j + T rho, Z = DFT_h(z) in registers (canonical layout), Z into the
// job's buffer and back for the partner Z_{h-k}, then for the thread's own
// k = j + T rho
//   G^_k = (Z_k + conj Z_{h-k})/2 - (i/2) e^{-2 pi i k/n} (Z_k - conj Z_{h-k})
// (Z_h = Z_0; thread 0 also forms G^_h), F_m = e^{-i pi s m/n} G^_m for
// m = k <= mmax, and q_e = F_N + F_S, q_o = F_N - F_S (the southern pass
// adds to what the same thread wrote for the northern ring).
constexpr int kAdjRB = 2 * fft16::kThreads;  // threads per block

__host__ __device__ constexpr size_t adj16_bytes(int P) {
  return (size_t)kTwSlots * sizeof(double2) +
         (size_t)(kAdjRB / (P / 16)) * fft16::job_slots(P) * sizeof(double2);
}

template <int P>
__global__ __launch_bounds__(kAdjRB) void k_ring_adjoint16(LegTables t, RingTables rt, int first,
                                                           int count,
                                                           const double *__restrict__ map) {
  using fft16::pad;
  constexpr int T = P / 16, J = kAdjRB / T;
  extern __shared__ double2 lds[];
  double2 *twl = lds;
  for (int u = threadIdx.x; u < kTwSlots; u += kAdjRB)  // read after the FFT's first barrier
    twl[u] = (u < kTwC) ? rt.twiddle[u * 64] : rt.twiddle[u - kTwC];
  const int jb = threadIdx.x / T;
  int j = threadIdx.x - jb * T;
  double2 *buf = lds + kTwSlots + jb * fft16::job_slots(P);
  const int K = t.K, M = t.M;
  const int64_t g = (int64_t)blockIdx.x * J + jb;
  const int pi = (int)(g / K), k = (int)(g - (int64_t)(g / K) * K);
  const bool act = pi < count;
  const RingPair rp = rt.pairs[rt.order[first + (act ? pi : 0)]];
  const int n = rp.n;
  const bool south = rp.pix_s >= 0;
  double2 *qe = t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  double2 *qo = t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
#pragma unroll 1
  for (int hemi = 0; hemi < 2; ++hemi) {  // every job takes both passes (block barriers)
    const bool on = act && (hemi == 0 || south);
    const double *gp = map + (int64_t)(hemi && south ? rp.pix_s : rp.pix_n) * K + k;
    double2 v[16];
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int64_t p = j + T * rho;
      v[rho] = on ? make_double2(gp[2 * p * K], gp[(2 * p + 1) * K]) : make_double2(0.0, 0.0);
    }
    fft16::fft<P, false>(v, buf, twl, j);
    asm volatile("" : "+v"(j));
    __syncthreads();  // the transform's last reads of buf done
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) buf[pad(j + T * rho)] = v[rho];
    __syncthreads();
    const double2 wstep = expipi(-(int64_t)rp.s * T, n), ostep = expipi(-2 * (int64_t)T, n);
    double2 w = expipi(-(int64_t)rp.s * j, n), om = expipi(-2 * (int64_t)j, n);
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int kk = j + T * rho;
      const double2 za = v[rho], zc = buf[pad((P - kk) & (P - 1))];
      const double2 e = cadd(za, cconj(zc)), d = csub(za, cconj(zc));
      const double2 od = cmul(om, d);
      const double2 gk = make_double2(0.5 * (e.x + od.y), 0.5 * (e.y - od.x));
      if (on && kk <= M) {
        const double2 f = cmul(gk, w);
        const int64_t o = (int64_t)kk * qs;
        if (hemi == 0) {
          qe[o] = f;
          qo[o] = f;
        } else {
          qe[o] = cadd(qe[o], f);
          qo[o] = csub(qo[o], f);
        }
      }
      w = cmul(w, wstep);
      om = cmul(om, ostep);
    }
    if (j == 0 && on && P <= M) {
      // G^_h from Z_0 (partner of k = 0): e^{-2 pi i h/n} = -1
      const double2 z0 = v[0];
      const double2 e = cadd(z0, cconj(z0)), d = csub(z0, cconj(z0));
      const double2 gh = make_double2(0.5 * (e.x - d.y), 0.5 * (e.y + d.x));
      const double2 f = cmul(gh, expipi(-(int64_t)rp.s * P, n));
      const int64_t o = (int64_t)P * qs;
      if (hemi == 0) {
        qe[o] = f;
        qo[o] = f;
      } else {
        qe[o] = cadd(qe[o], f);
        qo[o] = csub(qo[o], f);
      }
    }
  }
}

#define WM_ADJ16_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096)

// The ring adjoint of the forward cap kernel's rings (h = n/2 not a power
// of two, n > mmax) on the register FFT: Z = DFT_h(z) as conj(IDFT(conj z))
// by Bluestein's chirp z-transform of length P = next_pow2(2h - 1) (the
// plan's non-split chirp and vhat), u_a = conj(z_a) chirp_a (a < h), then
// FFT_P, times vhat, inverse FFT_P, Z_p = conj(chirp_p u_p); the split of
// k_ring_adjoint16 for G^_0 .. G^_h into the job's buffer, and F_m =
// e^{-i pi s m/n} G^[m] (m <= h) or conj G^[n - m] (h < m <= mmax < n).
template <int P>
__global__ __launch_bounds__(kAdjRB) void k_ring_adjoint_cap16(LegTables t, RingTables rt,
                                                               int first, int count,
                                                               const double *__restrict__ map) {
  using fft16::pad;
  constexpr int T = P / 16, J = kAdjRB / T;
  extern __shared__ double2 lds[];
  double2 *twl = lds;
  for (int u = threadIdx.x; u < kTwSlots; u += kAdjRB)  // read after the FFT's first barrier
    twl[u] = (u < kTwC) ? rt.twiddle[u * 64] : rt.twiddle[u - kTwC];
  const int jb = threadIdx.x / T;
  int j = threadIdx.x - jb * T;
  double2 *buf = lds + kTwSlots + jb * fft16::job_slots(P);
  const int K = t.K, M = t.M;
  const int64_t g = (int64_t)blockIdx.x * J + jb;
  const int pi = (int)(g / K), k = (int)(g - (int64_t)(g / K) * K);
  const bool act = pi < count;
  const RingPair rp = rt.pairs[rt.order[first + (act ? pi : 0)]];
  const int n = rp.n, h = n / 2;
  const bool south = rp.pix_s >= 0;
  const double2 *chirp = rt.chirp + rp.chirp_off;  // [h]
  const double2 *vhat = rt.vhat + rp.vhat_off;     // [P]
  double2 *qe = t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  double2 *qo = t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
#pragma unroll 1
  for (int hemi = 0; hemi < 2; ++hemi) {  // every job takes both passes (block barriers)
    const bool on = act && (hemi == 0 || south);
    const double *gp = map + (int64_t)(hemi && south ? rp.pix_s : rp.pix_n) * K + k;
    double2 v[16];
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int a = j + T * rho;
      double2 x = make_double2(0.0, 0.0);
      if (on && a < h) {
        const double2 z = make_double2(gp[2 * (int64_t)a * K], gp[(2 * (int64_t)a + 1) * K]);
        x = cmul(cconj(z), chirp[a]);
      }
      v[rho] = x;
    }
    fft16::fft<P, false>(v, buf, twl, j);
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) v[rho] = cmul(v[rho], vhat[j + T * rho]);
    fft16::fft<P, true>(v, buf, twl, j);
    asm volatile("" : "+v"(j));
    __syncthreads();  // the transform's last reads of buf done
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int p = j + T * rho;
      if (p < h) buf[pad(p)] = cconj(cmul(v[rho], chirp[p]));  // Z_p
    }
    __syncthreads();
    // G^_k, k = j + T rho <= h (k = h from Z_0), kept in registers until
    // every partner read is done, then into buf
    double2 gk[16];
    {
      const double2 ostep = expipi(-2 * (int64_t)T, n);
      double2 om = expipi(-2 * (int64_t)j, n);
#pragma unroll
      for (int rho = 0; rho < 16; ++rho) {
        const int kk = j + T * rho;
        double2 gv = make_double2(0.0, 0.0);
        if (kk <= h) {
          const double2 za = buf[pad(kk < h ? kk : 0)], zc = buf[pad(kk == 0 || kk == h ? 0 : h - kk)];
          const double2 e = cadd(za, cconj(zc)), d = csub(za, cconj(zc));
          const double2 od = cmul(om, d);
          gv = make_double2(0.5 * (e.x + od.y), 0.5 * (e.y - od.x));
        }
        gk[rho] = gv;
        om = cmul(om, ostep);
      }
    }
    __syncthreads();  // partner reads done
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int kk = j + T * rho;
      if (kk <= h) buf[pad(kk)] = gk[rho];
    }
    __syncthreads();
    if (on) {
      for (int m = j; m <= M; m += T) {
        const double2 gv = m <= h ? buf[pad(m)] : cconj(buf[pad(n - m)]);
        const double2 f = cmul(gv, expipi(-(int64_t)rp.s * m, n));
        const int64_t o = (int64_t)m * qs;
        if (hemi == 0) {
          qe[o] = f;
          qo[o] = f;
        } else {
          qe[o] = cadd(qe[o], f);
          qo[o] = csub(qo[o], f);
        }
      }
    }
    __syncthreads();  // F reads of buf done before the next pass writes it
  }
}

#define WM_ADJ_CAP_SIZES(X) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)


// b~ partial sums.  Each wave is one work item (m, ring group g) and runs
// alone (no workgroup barrier; 4 waves per workgroup only share the launch):
// the group's 64-ring chunks c = g, g + G, ... one after the other, for each
// the l-range in 16-l windows.  Lanes evaluate their ring (monic
// recurrence, 16-l chunk rescale) into the wave's 16 x 65 LDS tile, which
// the matrix cores then multiply with the chunk's q (k_legendre_adj).  Rows
// partial[g][off_m + l - m][k] belong to this wave only.
constexpr int kAdjWaves = 4;  // waves per workgroup of k_legendre_adj

__host__ __device__ constexpr int adj_wave_lds_doubles() {
  return 16 * 65;  // the wave's Lambda tile
}

__device__ __forceinline__ void wave_sync_lds() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

typedef double adj_d4 __attribute__((ext_vector_type(4)));

// One pass over maps kb .. kb + KN - 1 (KN <= 4).  A 16-l window is the
// FP64 matrix-core product D (16 l x 16 columns) = Lambda (16 l x 64 rings)
// . Q (64 rings x 16 columns), Q's columns the real and imaginary parts of
// the pass's maps for the even parity (columns 0 .. 2 KN - 1) and the odd
// parity (2 KN .. 4 KN - 1); row l keeps the columns of its parity.  Sixteen
// v_mfma_f64_16x16x4f64 per window, four rings each: lane i holds
// A[i % 16][i / 16] = Lambda_{l0 + i % 16}(ring 4 t + i / 16) from the tile,
// B[i / 16][i % 16] = Q[4 t + i / 16][i % 16] (the chunk's 16 B values per
// lane stay in registers for all its windows), and D[i / 16 + 4 j][i % 16]
// in result register j.  The sum over the chunk's 64 rings happens inside
// the product: no cross-lane reduction.
template <int KN>
__global__ __launch_bounds__(64 * kAdjWaves) void k_legendre_adj(LegTables t, int G, int kb,
                                                                 double2 *partial) {
  static_assert(KN >= 1 && KN <= 4, "both parities of the pass's maps in 16 columns");
  extern __shared__ double adj_lds[];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int L = t.L, R = t.R, M = t.M, K = t.K;
  const int64_t item = (int64_t)blockIdx.x * kAdjWaves + wv;
  const int m = (int)(item / G), g = (int)(item - (int64_t)(item / G) * G);
  if (m > M) return;  // the whole wave
  double *tile = adj_lds + (size_t)wv * adj_wave_lds_doubles();  // [16][65]
  const int64_t offm = t.off[m];
  const int len = L - m + 1;
  const int lp = (len + kChunk - 1) / kChunk * kChunk;
  const double *ct = t.ctil + (offm - m);
  const double *cs = t.cscale + offm / kChunk;
  double *prow = (double *)(partial + ((int64_t)g * t.nlpad + offm) * K + kb);
  const int64_t rstride = 2 * (int64_t)K;  // doubles per l row of the partial sums
  for (int j = lane; j < lp * 2 * KN; j += 64) {
    const int l = j / (2 * KN), u = j - l * 2 * KN;
    prow[l * rstride + u] = 0.0;
  }
  // this lane's B column and what it stands for
  const int col = lane & 15, kq = lane >> 4;
  const int cpar = col < 2 * KN ? 0 : 1;          // parity of the column
  const int ck = (col - 2 * KN * cpar) >> 1;      // map of the column (valid if col < 4 KN)
  const bool cval = col < 4 * KN;
  const int nchunks = (R + 63) / 64;
  // A wave takes its chunks g, g + G, ... two at a time: both chunks'
  // recurrences run window by window and their products add into one
  // accumulator, so each pass stores its rows once.  The first pass stores
  // them outright, later passes (more than two chunks per wave) add.
  bool first = true;
  for (int c0 = g; c0 < nchunks; c0 += 2 * G) {
    double bq[2][16];
    double x[2], prev[2] = {0.0, 0.0}, cur[2] = {0.0, 0.0};
    int sl[2], lmin[2];
    bool have[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int c = c0 + h * G;
      have[h] = c < nchunks;  // wave-uniform
      const int r = c * 64 + lane;
      const bool ok = have[h] && r < R;
      x[h] = ok ? t.ring_x[r] : 0.0;
      sl[h] = ok ? t.seed_l[(int64_t)m * R + r] : kSeedNone;
      int lm = sl[h];
      for (int d = 32; d >= 1; d >>= 1) {
        const int o = __shfl_xor(lm, d);
        lm = o < lm ? o : lm;
      }
      lmin[h] = __builtin_amdgcn_readfirstlane(lm);
      if (lmin[h] > L) have[h] = false;
#pragma unroll
      for (int tq = 0; tq < 16; ++tq) {
        const int rg = c * 64 + 4 * tq + kq;
        double v = 0.0;
        if (have[h] && cval && rg < R) {
          const double2 q = t.qbuf[q_index(rg, cpar, kb + ck, m, K, M + 1)];
          v = (col & 1) ? q.y : q.x;
        }
        bq[h][tq] = v;
      }
    }
    if (!have[0] && !have[1]) continue;
    const int lstart = have[0] ? (have[1] && lmin[1] < lmin[0] ? lmin[1] : lmin[0]) : lmin[1];
    for (int l0 = lstart; l0 <= L; l0 += kChunk) {
      adj_d4 acc = {0.0, 0.0, 0.0, 0.0};
      const double S = cs[(l0 - m) / kChunk];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if (!have[h] || l0 < lmin[h]) continue;  // wave-uniform
        // lanes = rings: L~ at l0 .. l0 + 15 into the tile (seeds sit on
        // the 16-l grid of m)
        const int r = (c0 + h * G) * 64 + lane;
        if (sl[h] == l0) {
          const double2 v = t.seed_v[(int64_t)m * R + r];
          prev[h] = v.x;
          cur[h] = v.y;
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          tile[j * 65 + lane] = cur[h];
          const double nx = fma(x[h], cur[h], -(ct[l0 + j] * prev[h]));
          prev[h] = cur[h];
          cur[h] = nx;
        }
        prev[h] *= S;
        cur[h] *= S;
        wave_sync_lds();
#pragma unroll
        for (int tq = 0; tq < 16; ++tq)
          acc = __builtin_amdgcn_mfma_f64_16x16x4f64(tile[col * 65 + 4 * tq + kq], bq[h][tq],
                                                     acc, 0, 0, 0);
        wave_sync_lds();  // tile reads done before the next chunk writes it
      }
      if (cval) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int l = l0 + kq + 4 * j;
          if (l <= L && ((l - m) & 1) == cpar) {
            double *dst = prow + (int64_t)(l - m) * rstride + 2 * ck + (col & 1);
            *dst = first ? acc[j] : *dst + acc[j];
          }
        }
      }
    }
    first = false;
  }
}

// b_lm = c_m A_l sum_g partial[g], into the m-major complex layout; one
// thread per (row, map), so that each group's partial sums are read
// contiguously across the threads.
__global__ __launch_bounds__(256) void k_adj_finish(LegTables t, int G,
                                                    const double2 *__restrict__ partial,
                                                    double2 *__restrict__ alm) {
  const int64_t u = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int K = t.K;
  if (u >= t.nlpad * K) return;
  const int64_t j = u / K;
  const int k = (int)(u - j * K);
  const int m = t.chunk_m[j / kChunk];
  const int64_t jj = j - t.off[m];
  const int l = m + (int)jj;
  if (l > t.L) return;
  const double s = t.ascale[j] * (m == 0 ? 1.0 : 2.0);
  const int64_t idx = (int64_t)m * (2 * t.L - m + 3) / 2 + jj;
  double2 acc = make_double2(0.0, 0.0);
  for (int g = 0; g < G; ++g) acc = cadd(acc, partial[(int64_t)g * t.nlpad * K + u]);
  acc.x *= s;
  acc.y = (m == 0) ? 0.0 : acc.y * s;
  alm[idx * K + k] = acc;
}

int adjoint_groups(const LegTables &t) {
  // one wave per (m, group): groups up to one per 64-ring chunk, while the
  // waves stay below ~64 per CU and the partial sums below 2 GiB
  const int nchunks = (t.R + 63) / 64;
  const double row_bytes = (double)t.nlpad * t.K * sizeof(double2);
  int G = 1;
  while (2 * G <= nchunks && (int64_t)(t.M + 1) * 2 * G <= 16384 &&
         row_bytes * 2 * G <= 2.0 * 1024 * 1024 * 1024)
    G *= 2;
  return G;
}

// Partial sums of the adjoint entry that allocates nothing
// (wavemoth_execute_adjoint_device_noalloc): a fixed array in device
// memory, held by each GPU the library runs on until the process ends.
// Every call on a GPU writes its rows before reading them (each wave of
// k_legendre_adj zeroes its own rows first), so the contents left by an
// earlier call never matter; calls that overlap on one GPU race on it.
constexpr size_t kAdjScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) double2 scratch[kAdjScratchBytes / sizeof(double2)];

size_t adjoint_scratch_elems() { return kAdjScratchBytes / sizeof(double2); }

// Device address of `scratch` on the current device (which must be
// `device`), looked up once per device; nullptr with the error set on
// failure.
double2 *adjoint_scratch(int device) {
  constexpr int kMaxDevices = 64;
  static std::mutex mu;
  static double2 *addr[kMaxDevices] = {nullptr};
  if (device < 0 || device >= kMaxDevices) {
    set_err("adjoint scratch: device index out of range");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      set_err(std::string("adjoint scratch: hipGetSymbolAddress: ") + hipGetErrorString(e));
      return nullptr;
    }
    addr[device] = (double2 *)ptr;
  }
  return addr[device];
}

// Ring groups for partial sums that must fit `cap` complex entries: the
// largest power of two G <= adjoint_groups(t) with G nlpad nmaps <= cap, or
// 0 when one group does not fit.
int adjoint_groups_within(const LegTables &t, size_t cap) {
  const size_t per = (size_t)t.nlpad * (size_t)t.K;
  int G = adjoint_groups(t);
  while (G > 0 && (size_t)G * per > cap) G /= 2;
  return G;
}

hipError_t launch_adjoint(const LegTables &t, const RingTables &rt, const double *map,
                          double2 *alm, double2 *partial, int G, hipStream_t s) {
  for (int i = 0; i < rt.nlaunch; ++i) {
    if (rt.count[i] == 0) continue;
    if (rt.kind[i] == kRingBelt16 && rt.fftP[i] <= 4096) {
      // power-of-two rings with mmax <= h: the register-FFT adjoint
      const int P = rt.fftP[i], J = kAdjRB / (P / 16);
      const unsigned nb = (unsigned)(((int64_t)rt.count[i] * t.K + J - 1) / J);
      switch (P) {
#define WM_ADJ16_CASE(PP)                                                                   \
  case PP:                                                                                  \
    hipLaunchKernelGGL(k_ring_adjoint16<PP>, dim3(nb), dim3(kAdjRB), adj16_bytes(PP), s, t,  \
                       rt, rt.first[i], rt.count[i], map);                                  \
    break;
        WM_ADJ16_SIZES(WM_ADJ16_CASE)
#undef WM_ADJ16_CASE
        default:
          return hipErrorInvalidValue;
      }
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      continue;
    }
    if (rt.kind[i] == kRingCap16 && 2 * rt.fftP[i] >= 32 && 2 * rt.fftP[i] <= 8192) {
      // the cap kernel's rings: Bluestein of length P = 2 P2 on the register FFT
      const int P = 2 * rt.fftP[i], J = kAdjRB / (P / 16);
      const unsigned nb = (unsigned)(((int64_t)rt.count[i] * t.K + J - 1) / J);
      switch (P) {
#define WM_ADJCAP_CASE(PP)                                                                  \
  case PP:                                                                                  \
    hipLaunchKernelGGL(k_ring_adjoint_cap16<PP>, dim3(nb), dim3(kAdjRB), adj16_bytes(PP), s, \
                       t, rt, rt.first[i], rt.count[i], map);                               \
    break;
        WM_ADJ_CAP_SIZES(WM_ADJCAP_CASE)
#undef WM_ADJCAP_CASE
        default:
          return hipErrorInvalidValue;
      }
      const hipError_t e = hipGetLastError();
      if (e != hipSuccess) return e;
      continue;
    }
    // every pair of the group fits: adj_lds[i] >= ring_lds_slots of each
    const size_t lds = (size_t)rt.adj_lds[i] * sizeof(double2);
    switch (rt.adj_maxp) {
#define WM_RADJ(MP)                                                                         \
  case MP:                                                                                  \
    hipLaunchKernelGGL(k_ring_adjoint<MP>, dim3(rt.count[i]), dim3(NT), lds, s, t, rt,      \
                       rt.first[i], rt.adj_lds[i], map);                                    \
    break;
      WM_RADJ(1024) WM_RADJ(2048) WM_RADJ(4096) WM_RADJ(8192)
#undef WM_RADJ
      default:
        return hipErrorInvalidValue;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  const int64_t waves = (int64_t)(t.M + 1) * G;
  const unsigned nblk = (unsigned)((waves + kAdjWaves - 1) / kAdjWaves);
  const size_t alds = (size_t)kAdjWaves * adj_wave_lds_doubles() * sizeof(double);
  // passes of up to 4 maps (both parities of a pass fill the 16 columns)
  for (int kb = 0; kb < t.K; kb += 4) {
    const int kn = (t.K - kb) < 4 ? (t.K - kb) : 4;
    switch (kn) {
#define WM_ADJ_CASE(KN)                                                                   \
  case KN:                                                                                \
    hipLaunchKernelGGL(k_legendre_adj<KN>, dim3(nblk), dim3(64 * kAdjWaves), alds, s, t, G, kb, \
                       partial);                                                          \
    break;
      WM_ADJ_CASE(1) WM_ADJ_CASE(2) WM_ADJ_CASE(3) WM_ADJ_CASE(4)
#undef WM_ADJ_CASE
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_adj_finish, dim3((unsigned)((t.nlpad * t.K + 255) / 256)), dim3(256), 0,
                     s, t, G, partial, alm);
  return hipGetLastError();
}

hipError_t adjoint_set_lds_limit(int slots) {
  // The limit only grows, so plans made earlier keep a valid launch size.
  static std::mutex mu;
  static int cur = 0;
  std::lock_guard<std::mutex> lock(mu);
  if (slots <= cur) return hipSuccess;
  hipError_t e = hipSuccess;
  for (const void *f : {(const void *)k_ring_adjoint<1024>, (const void *)k_ring_adjoint<2048>,
                        (const void *)k_ring_adjoint<4096>, (const void *)k_ring_adjoint<8192>})
    if (e == hipSuccess)
      e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)((size_t)slots * sizeof(double2)));
  // the register-FFT adjoint's fixed sizes (set once; they do not grow)
#define WM_ADJ16_SET(PP)                                                                    \
  if (e == hipSuccess)                                                                      \
    e = hipFuncSetAttribute((const void *)k_ring_adjoint16<PP>,                             \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)adj16_bytes(PP));
  WM_ADJ16_SIZES(WM_ADJ16_SET)
#undef WM_ADJ16_SET
#define WM_ADJCAP_SET(PP)                                                                   \
  if (e == hipSuccess)                                                                      \
    e = hipFuncSetAttribute((const void *)k_ring_adjoint_cap16<PP>,                         \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)adj16_bytes(PP));
  WM_ADJ_CAP_SIZES(WM_ADJCAP_SET)
#undef WM_ADJCAP_SET
  if (e == hipSuccess) cur = slots;
  return e;
}

}  // namespace wm

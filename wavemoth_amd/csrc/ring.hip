// Ring stage of the synthesis: even/odd fold, phi0 phase, aliasing into the
// rings' half spectra and the unnormalised complex-to-real DFTs.  One
// workgroup per ring pair (the northern ring and its southern mirror share
// the Legendre sums E, O), everything else in LDS.
//
// Semantics follow src/wavemoth.c:1083-1191: for hemisphere sign
// s_h = +1 (north) / -1 (south), G_m = (E_m + s_h O_m) e^{i m phi0}; bin
// (m mod n) receives G_m and bin ((n - m) mod n) receives conj(G_m) when
// m > 0, each only if the bin is <= n/2; then an FFTW-style c2r of length n
// (imaginary parts of the DC and Nyquist bins ignored, no normalisation).
//
// On the HEALPix grid phi0 = pi s / n with s in {0, 1} for every ring, so
// with m = b + j n the phase is e^{i pi s b/n} (-1)^{s j}: one complex
// factor per bin multiplies the signed alias sum
//   X_b = e^{i pi s b/n} [ sum_j (-1)^{sj} g_{b+jn} + sum_{j>=1} (-1)^{sj} conj g_{jn-b} ].
//
// The length-n c2r is a complex DFT of length h = n/2:
//   Z_k = (X_k + conj X_{h-k}) + i e^{2 pi i k/n} (X_k - conj X_{h-k}),
//   z = IDFT_h(Z),  f[2p] = Re z_p,  f[2p+1] = Im z_p.
// Power-of-two h (the belt, 2N+1 rings) uses a radix-8/4/2 Stockham FFT in
// LDS; other h (polar caps, h = 2i) use Bluestein's chirp-z convolution with
// a power-of-two FFT of P >= 2h-1 points, the chirp and the spectrum of its
// conjugate tabulated per h at plan time.  LDS arrays are padded by one
// slot per 8 (pidx) so the Stockham writes of the first passes do not pile
// onto one bank.
#include <mutex>

#include "wm_fft.h"

namespace wm {

using namespace fft;

// Stores pixels 2p, 2p + 1 of one map: contiguous for a single map, else
// strided by the number of maps (the (npix, K) layout).
__device__ __forceinline__ void store_pix2(double *out, int p, double2 z, int ostride) {
  if (ostride == 1) {
    *reinterpret_cast<double2 *>(out + 2 * p) = z;
  } else {
    out[(int64_t)(2 * p) * ostride] = z.x;
    out[(int64_t)(2 * p + 1) * ostride] = z.y;
  }
}

// Map k of one ring pair (index into rt.pairs) by the whole block; `lds`
// holds lds_slots complex slots, the last kTwSlots of them the twiddle
// table (already loaded).  Pixel p of the ring goes to outk[p * ostride]
// (outk = map + k, ostride = nmaps: the (npix, nmaps) layout).
// Used by k_ring_pairs and the batched kernel.  Q: the element type of q
// (float2 for the mixed-precision plans, widened to double on load).
// MAXP: the size class (>= the ring's Bluestein P and n = 2h, see
// ring_class); the per-thread bin, pair, output and staging counts follow
// from it, the twiddle table stays the kMaxFFT one.
template <typename Q, int MAXP = kMaxFFT>
__device__ __forceinline__ void ring_pair_body(const LegTables &t, const RingTables &rt, int pair, int k,
                               double2 *lds, int lds_slots, double *__restrict__ outk,
                               int ostride) {
  constexpr int kBins = (MAXP / 2 + 1 + NT - 1) / NT;   // bins 0 .. h per thread
  constexpr int kPairs = (MAXP / 4 + 1 + NT - 1) / NT;  // Z pairs per thread
  constexpr int kOut = (MAXP / 2 + NT - 1) / NT;        // output pairs per thread
  constexpr int kFull = (MAXP + NT - 1) / NT;           // FFT points per thread
  constexpr int kStage = kBins;                         // staged m per thread
  constexpr int kStageM = kStage * NT;                  // m staged per round
  double2 *twl = lds + lds_slots - kTwSlots;
  const RingPair rp = rt.pairs[pair];
  const int n = rp.n, h = n / 2, B = h + 1, K = t.K, M = t.M, P = rp.P;
  const bool south = rp.pix_s >= 0;
  // q of map k: element m at qe[m * qs] (maps of a group interleaved)
  const Q *qe = (const Q *)t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  const Q *qo = (const Q *)t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
  const double2 *chirp = rt.chirp + rp.chirp_off;
  const double2 *vhat = rt.vhat + rp.vhat_off;
  const int xoff = (P == 0) ? 0 : (P - h - 1);
  // alias sums: G groups of j per bin when the ring is short
  const int G = ring_groups(B);
  const int J = (M + h) / n + 2;   // j runs over [0, J)
  const int JG = (J + G - 1) / G;  // j-iterations per group

  const int nhemi = south ? 2 : 1;
  {
    for (int hemi = 0; hemi < nhemi; ++hemi) {
      // Opaque copy of the thread index: keeps the per-bin LDS/global
      // addresses from being hoisted out of the map/hemisphere loops (they
      // would stay live across the FFTs and spill).
      int tid = threadIdx.x;
      asm volatile("" : "+v"(tid));
      const double hs = hemi ? -1.0 : 1.0;  // E + O north, E - O south
      double *out = outk + (hemi ? rp.pix_s : rp.pix_n) * ostride;
      // 1.-2. g_m = E_m +/- O_m staged in LDS, kStage*NT values of m at a
      // time (one round of independent loads), then the signed alias sums
      // per bin: slot (i) -> bin b = slot % B, group g = slot / B; direct
      // terms m = b + j n, conjugate terms m = j n - b.
      double2 xv[kBins];
#pragma unroll
      for (int i = 0; i < kBins; ++i) xv[i] = make_double2(0.0, 0.0);
      for (int mb = 0; mb <= M; mb += kStageM) {
        const int me = (mb + kStageM <= M + 1) ? mb + kStageM : M + 1;
#pragma unroll
        for (int i = 0; i < kStage; ++i) {
          const int m = mb + tid + NT * i;
          if (m < me) {
            const double2 e = to_d(qe[m * qs]), o = to_d(qo[m * qs]);
            lds[pidx(m - mb)] = make_double2(fma(hs, o.x, e.x), fma(hs, o.y, e.y));
          }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < kBins; ++i) {
          const int slot = tid + NT * i;
          const int g = slot / B, b = slot - g * B;
          if (g < G) {
            for (int jj = 0; jj < JG; ++jj) {
              const int j = g + jj * G;
              const double sg = (rp.s && (j & 1)) ? -1.0 : 1.0;
              const int m1 = b + j * n, m2 = j * n - b;
              if (m1 >= mb && m1 < me) {
                const double2 v = lds[pidx(m1 - mb)];
                xv[i].x = fma(sg, v.x, xv[i].x);
                xv[i].y = fma(sg, v.y, xv[i].y);
              }
              if (j >= 1 && m2 >= mb && m2 < me) {
                const double2 v = lds[pidx(m2 - mb)];
                xv[i].x = fma(sg, v.x, xv[i].x);
                xv[i].y = fma(-sg, v.y, xv[i].y);
              }
            }
          }
        }
        __syncthreads();
      }
      if (G > 1) {
        // reduce the G partial sums of each bin in a fixed order
        if (tid < G * B) lds[tid] = xv[0];
        __syncthreads();
        if (tid < B) {
          double2 acc = make_double2(0.0, 0.0);
          for (int g = 0; g < G; ++g) acc = cadd(acc, lds[g * B + tid]);
          xv[0] = acc;
        }
        __syncthreads();
      }
      // 3. the alias sums (unphased) into LDS
#pragma unroll
      for (int i = 0; i < kBins; ++i) {
        const int b = tid + NT * i;
        if (b < B) lds[pidx(xoff + b)] = xv[i];
      }
      __syncthreads();
      // 4. per pair (a, h - a): phases e^{i pi s b / n} (angle addition for a,
      // e^{i pi s / 2} conj(.) for h - a), c2r semantics for DC / Nyquist,
      // then the Z pair with twiddle e^{2 pi i a / n} (angle addition).
      {
        const double2 step = expipi(2 * NT, n);
        double2 om = expipi(2 * tid, n);
        const double2 pstep = expipi((int64_t)rp.s * NT, n);
        double2 w = expipi((int64_t)rp.s * tid, n);
        const double2 wh = rp.s ? make_double2(0.0, 1.0) : make_double2(1.0, 0.0);
#pragma unroll
        for (int i = 0; i < kPairs; ++i) {
          const int a = tid + NT * i;
          if (a <= h / 2) {
            const int c = h - a;
            double2 xa = cmul(lds[pidx(xoff + a)], w);
            if (a == 0) xa.y = 0.0;
            double2 xc = cmul(lds[pidx(xoff + c)], c == h ? wh : cmul(wh, cconj(w)));
            if (c == h) xc.y = 0.0;
            const double2 za =
                cadd(cadd(xa, cconj(xc)), cmul_i(cmul(om, csub(xa, cconj(xc)))));
            const bool both = (a != 0) && (a != c);
            double2 zc = make_double2(0.0, 0.0);
            if (both) {
              const double2 oc = make_double2(-om.x, om.y);  // e^{2 pi i c/n}
              zc = cadd(cadd(xc, cconj(xa)), cmul_i(cmul(oc, csub(xc, cconj(xa)))));
            }
            if (P == 0) {
              lds[pidx(a)] = za;
              if (both) lds[pidx(c)] = zc;
            } else {
              lds[pidx(a)] = cmul(za, chirp[a]);
              if (both) lds[pidx(c)] = cmul(zc, chirp[c]);
            }
          }
          om = cmul(om, step);
          w = cmul(w, pstep);
        }
      }
      __syncthreads();
      // 5. transforms and output
      if (P == 0) {
        lds_fft<true, MAXP>(lds, h, twl);
#pragma unroll
        for (int i = 0; i < kOut; ++i) {
          const int p = tid + NT * i;
          if (p < h) {
            store_pix2(out, p, lds[pidx(p)], ostride);
          }
        }
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
        for (int i = 0; i < kOut; ++i) {
          const int p = tid + NT * i;
          if (p < h) {
            store_pix2(out, p, cmul(lds[pidx(p)], chirp[p]), ostride);
          }
        }
      }
      __syncthreads();
    }
  }
}

// Block -> (ring pair, map).  The G = q_group(K) maps of one q group share
// its 128-byte lines, so their G workgroups are put on one XCD (hardware
// dispatches block b to XCD b % 8) at consecutive dispatch slots: block
// b = 8 (sigma G + j) + x runs map j of group gamma = 8 sigma + x, where
// gamma enumerates (ring pair i, map group).
template <typename Q, int MAXP>
__global__ __launch_bounds__(NT) void k_ring_pairs(LegTables t, RingTables rt, int first,
                                                   int count, int lds_slots,
                                                   double *__restrict__ map) {
  extern __shared__ double2 lds[];
  const int K = t.K, G = q_group(K), nG = q_groups(K);
  const int b = blockIdx.x, x = b & 7, s = b >> 3;
  const int sigma = s / G, j = s - sigma * G;
  const int gamma = sigma * 8 + x;
  const int i = gamma / nG, k = (gamma - i * nG) * G + j;
  if (i >= count || k >= K) return;
  load_twiddles(lds + lds_slots - kTwSlots, rt.twiddle);
  ring_pair_body<Q, MAXP>(t, rt, rt.order[first + i], k, lds, lds_slots, map + k, K);
}

// Ring stage of a batch in the general kernel: block i runs item
// ritems[first + i] (request, ring pair, map; pairs of size class MAXP or
// below) as the single-plan kernel runs that pair and map.  One item per
// block, in the host's order (costliest first), rather than a work-queue
// loop: with the pair body inlined into a loop the compiler kept
// thread-invariant values live across it (248 VGPRs, 96-368 bytes of
// spills per lane).
template <int MAXP>
__global__ __launch_bounds__(NT) void k_batch_rings(BatchDev b, int first, int slots) {
  extern __shared__ double2 lds[];
  const BatchRingItem item = b.ritems[first + blockIdx.x];
  const LegTables &t = b.tabs[item.req];
  load_twiddles(lds + slots - kTwSlots, b.twiddle);
  ring_pair_body<double2, MAXP>(t, b.rtabs[item.req], item.pair, item.k, lds, slots,
                                b.map[item.req] + item.k, t.K);
}

hipError_t batch_rings_set_lds_limit(int slots) {
  // The limit only grows, so plans made earlier keep a valid launch size.
  static std::mutex mu;
  static int cur = 0;
  std::lock_guard<std::mutex> lock(mu);
  if (slots <= cur) return hipSuccess;
  hipError_t e = hipSuccess;
  for (const void *f : {(const void *)k_batch_rings<1024>, (const void *)k_batch_rings<2048>,
                        (const void *)k_batch_rings<4096>, (const void *)k_batch_rings<8192>})
    if (e == hipSuccess)
      e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)((size_t)slots * sizeof(double2)));
  if (e == hipSuccess) cur = slots;
  return e;
}

#define WM_BR_CASES(X) X(1024) X(2048) X(4096) X(8192)

hipError_t launch_batch_rings(const BatchDev &b, int maxp, int first, int count, int slots,
                              hipStream_t s) {
  if (count <= 0) return hipSuccess;
#define WM_BR_LAUNCH(MP)                                                                   \
  if (maxp == MP) {                                                                        \
    hipLaunchKernelGGL(k_batch_rings<MP>, dim3((unsigned)count), dim3(NT),                 \
                       (size_t)slots * sizeof(double2), s, b, first, slots);               \
    return hipGetLastError();                                                              \
  }
  WM_BR_CASES(WM_BR_LAUNCH)
#undef WM_BR_LAUNCH
  return hipErrorInvalidValue;
}

// Plan-time Bluestein tables, one block per distinct h:
// chirp[a] = e^{i pi a^2 / h} (a < h), and vhat = FFT_P(v) / P with
// v[d] = conj(chirp[|d|]) placed cyclically (d in (-h, h)).
__global__ __launch_bounds__(NT) void k_bluestein_prep(const double2 *__restrict__ twiddle,
                                                       double2 *__restrict__ chirp_all,
                                                       double2 *__restrict__ vhat_all,
                                                       const int *__restrict__ hlist,
                                                       const int64_t *__restrict__ choff,
                                                       const int64_t *__restrict__ vhoff,
                                                       const int *__restrict__ plist) {
  extern __shared__ double2 lds[];
  const int tid = threadIdx.x;
  const int h = hlist[blockIdx.x], P = plist[blockIdx.x];
  double2 *twl = lds + pidx(kMaxFFT - 1) + 1;
  load_twiddles(twl, twiddle);
  double2 *chirp = chirp_all + choff[blockIdx.x];
  double2 *vhat = vhat_all + vhoff[blockIdx.x];
  for (int a = tid; a < P; a += NT) lds[pidx(a)] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int a = tid; a < h; a += NT) {
    const int64_t a2 = (int64_t)a * a;
    const double2 c = expipi(a2 % (2 * (int64_t)h), h);
    chirp[a] = c;
    lds[pidx(a)] = cconj(c);
    if (a > 0) lds[pidx(P - a)] = cconj(c);
  }
  __syncthreads();
  lds_fft<false>(lds, P, twl);
  const double inv = 1.0 / (double)P;
  for (int a = tid; a < P; a += NT) vhat[a] = cscale(lds[pidx(a)], inv);
}

int ring_class(int n, int P) {
  int c = 1024;
  while (c < (P ? P : n)) c <<= 1;
  return c;
}

int ring_lds_slots(int n, int P, int M, int maxp) {
  const int h = n / 2;
  int need = (P == 0) ? pidx(h) + 1 : pidx(P - 1) + 1;
  const int G = ring_groups(h + 1);
  if (G > 1 && G * (h + 1) > need) need = G * (h + 1);
  const int stage_m = (maxp / 2 + 1 + NT - 1) / NT * NT;  // m staged per round
  const int ms = (M < stage_m - 1) ? M : stage_m - 1;
  if (pidx(ms) + 1 > need) need = pidx(ms) + 1;  // staged g_m
  return need + kTwSlots;
}

hipError_t launch_bluestein_prep(const double2 *twiddle, double2 *chirp, double2 *vhat,
                                 const int *hlist, const int64_t *choff, const int64_t *vhoff,
                                 const int *plist, int nh, hipStream_t s) {
  if (nh <= 0) return hipSuccess;
  const size_t lds = (size_t)(pidx(kMaxFFT - 1) + 1 + kTwSlots) * sizeof(double2);
  hipError_t e = hipFuncSetAttribute((const void *)k_bluestein_prep,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bluestein_prep, dim3(nh), dim3(NT), lds, s, twiddle, chirp, vhat, hlist,
                     choff, vhoff, plist);
  return hipGetLastError();
}

hipError_t ring_set_lds_limit(int slots) {
  // The limit only grows, so plans made earlier keep a valid launch size.
  static std::mutex mu;
  static int cur = 0;
  std::lock_guard<std::mutex> lock(mu);
  if (slots <= cur) return hipSuccess;
  const int bytes = (int)((size_t)slots * sizeof(double2));
  hipError_t e = hipSuccess;
#define WM_RP(MP)                                                                          \
  for (const void *f : {(const void *)k_ring_pairs<double2, MP>,                           \
                        (const void *)k_ring_pairs<float2, MP>})                           \
    if (e == hipSuccess) e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  WM_RP(1024) WM_RP(2048) WM_RP(4096) WM_RP(8192)
#undef WM_RP
  if (e == hipSuccess) cur = slots;
  return e;
}

hipError_t launch_ring_synth(const LegTables &t, const RingTables &rt, double *map,
                             hipStream_t s) {
  for (int i = 0; i < rt.nlaunch; ++i) {
    // the long path (launch_long_rings) runs the kRingLong groups
    if (rt.count[i] == 0 || rt.kind[i] == kRingLong) continue;
    hipError_t e;
    if (rt.kind[i] != kRingGeneral) {
      e = launch_ring16(rt.kind[i], rt.fftP[i], t, rt, rt.first[i], rt.count[i], map, s);
    } else {
      // a general group's fftP is its size class (ring_class of its pairs)
      const size_t lds = (size_t)rt.lds[i] * sizeof(double2);
      const int64_t groups = (int64_t)rt.count[i] * q_groups(t.K);
      const int64_t nb = (groups + 7) / 8 * 8 * q_group(t.K);
      switch (rt.fftP[i]) {
#define WM_RP(MP)                                                                           \
  case MP:                                                                                  \
    if (t.q32)                                                                              \
      hipLaunchKernelGGL((k_ring_pairs<float2, MP>), dim3((unsigned)nb), dim3(NT), lds, s, t, \
                         rt, rt.first[i], rt.count[i], rt.lds[i], map);                     \
    else                                                                                    \
      hipLaunchKernelGGL((k_ring_pairs<double2, MP>), dim3((unsigned)nb), dim3(NT), lds, s,   \
                         t, rt, rt.first[i], rt.count[i], rt.lds[i], map);                  \
    break;
        WM_RP(1024) WM_RP(2048) WM_RP(4096) WM_RP(8192)
#undef WM_RP
        default:
          return hipErrorInvalidValue;
      }
      e = hipGetLastError();
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace wm

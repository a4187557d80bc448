// Ring stage of the synthesis: even/odd fold, phi0 phase, aliasing into the
// ring's half spectrum and the unnormalised complex-to-real DFT, one
// workgroup per ring, everything in LDS.
//
// Semantics follow src/wavemoth.c:1083-1191: for ring pair r and
// hemisphere s (+1 north, -1 south), G_m = (E_m + s O_m) e^{i m phi0};
// bin (m mod n) receives G_m and bin ((n - m) mod n) receives conj(G_m)
// (m > 0), each only when the bin is <= n/2; then an FFTW-style c2r of
// length n (imaginary parts of the DC and Nyquist bins ignored, no
// normalisation).
//
// The length-n c2r is computed as a complex DFT of length h = n/2:
//   Z_k = (X_k + conj X_{h-k}) + i e^{i pi k/h} (X_k - conj X_{h-k}),
//   z = IDFT_h(Z),  f[2p] = Re z_p,  f[2p+1] = Im z_p.
// h is a power of two for the 2N+1 belt rings (radix-4/2 Stockham in LDS);
// the polar-cap lengths h = 2i use Bluestein's chirp-z convolution with a
// power-of-two FFT of P >= 2h-1 points, the chirp and the spectrum of its
// conjugate being tabulated per h at plan time.
#include "wm_internal.h"

namespace wm {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by +i / -i
__device__ __forceinline__ double2 cmul_i(double2 a) { return make_double2(-a.y, a.x); }
__device__ __forceinline__ double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }

// In-place FFT of length P (power of two, <= kMaxFFT) held in LDS, by the
// whole block (kRingThreads threads).  Stockham autosort, radix 4 while
// P/Ns is divisible by 4, then one radix-2 stage.  Each stage reads all
// its butterfly inputs into registers, syncs, writes, syncs.
// INV = false: e^{-2 pi i jk/P};  INV = true: e^{+2 pi i jk/P}.
template <bool INV>
__device__ void lds_fft(double2 *buf, int P, const double2 *__restrict__ tw) {
  const int tid = threadIdx.x;
  constexpr int NT = kRingThreads;
  constexpr int B4 = (kMaxFFT / 4 + NT - 1) / NT;
  constexpr int B2 = (kMaxFFT / 2 + NT - 1) / NT;
  int Ns = 1;
  while (Ns < P) {
    if (((P / Ns) & 3) == 0) {
      const int nb = P >> 2;
      const int tstep = kMaxFFT / (Ns * 4);
      double2 v[B4][4];
#pragma unroll
      for (int it = 0; it < B4; ++it) {
        const int j = tid + it * NT;
        if (j < nb) {
          const int k = j & (Ns - 1);
#pragma unroll
          for (int r = 0; r < 4; ++r) v[it][r] = buf[j + r * nb];
          if (k) {
#pragma unroll
            for (int r = 1; r < 4; ++r) {
              double2 w = tw[k * r * tstep];
              if (INV) w.y = -w.y;
              v[it][r] = cmul(v[it][r], w);
            }
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int it = 0; it < B4; ++it) {
        const int j = tid + it * NT;
        if (j < nb) {
          const int k = j & (Ns - 1);
          const double2 a0 = cadd(v[it][0], v[it][2]);
          const double2 a1 = csub(v[it][0], v[it][2]);
          const double2 a2 = cadd(v[it][1], v[it][3]);
          const double2 a3 = csub(v[it][1], v[it][3]);
          const int d = (j - k) * 4 + k;
          buf[d] = cadd(a0, a2);
          buf[d + 2 * Ns] = csub(a0, a2);
          if (INV) {
            buf[d + Ns] = cadd(a1, cmul_i(a3));
            buf[d + 3 * Ns] = cadd(a1, cmul_mi(a3));
          } else {
            buf[d + Ns] = cadd(a1, cmul_mi(a3));
            buf[d + 3 * Ns] = cadd(a1, cmul_i(a3));
          }
        }
      }
      __syncthreads();
      Ns <<= 2;
    } else {
      const int nb = P >> 1;
      const int tstep = kMaxFFT / (Ns * 2);
      double2 v[B2][2];
#pragma unroll
      for (int it = 0; it < B2; ++it) {
        const int j = tid + it * NT;
        if (j < nb) {
          const int k = j & (Ns - 1);
          v[it][0] = buf[j];
          v[it][1] = buf[j + nb];
          if (k) {
            double2 w = tw[k * tstep];
            if (INV) w.y = -w.y;
            v[it][1] = cmul(v[it][1], w);
          }
        }
      }
      __syncthreads();
#pragma unroll
      for (int it = 0; it < B2; ++it) {
        const int j = tid + it * NT;
        if (j < nb) {
          const int k = j & (Ns - 1);
          const int d = (j - k) * 2 + k;
          buf[d] = cadd(v[it][0], v[it][1]);
          buf[d + Ns] = csub(v[it][0], v[it][1]);
        }
      }
      __syncthreads();
      Ns <<= 1;
    }
  }
}

// e^{i pi num/den} with num reduced exactly modulo 2 den first.
__device__ __forceinline__ double2 expipi(int64_t num, int64_t den) {
  int64_t t = num % (2 * den);
  if (t < 0) t += 2 * den;
  double s, c;
  sincospi((double)t / (double)den, &s, &c);
  return make_double2(c, s);
}

// Half-spectrum bin b of one ring for map k (see the header comment).
__device__ double2 gather_bin(int b, int n, int M, const double2 *__restrict__ qe,
                              const double2 *__restrict__ qo, int K, int k, double sgn,
                              int ph_num, int ph_den) {
  double2 acc = make_double2(0.0, 0.0);
  for (int m = b; m <= M; m += n) {
    const double2 e = qe[(int64_t)m * K + k], o = qo[(int64_t)m * K + k];
    const double2 g = make_double2(fma(sgn, o.x, e.x), fma(sgn, o.y, e.y));
    acc = cadd(acc, cmul(g, expipi((int64_t)m * ph_num, ph_den)));
  }
  for (int m = n - b; m <= M; m += n) {
    const double2 e = qe[(int64_t)m * K + k], o = qo[(int64_t)m * K + k];
    const double2 g = make_double2(fma(sgn, o.x, e.x), fma(sgn, o.y, e.y));
    acc = cadd(acc, cconj(cmul(g, expipi((int64_t)m * ph_num, ph_den))));
  }
  return acc;
}

__global__ __launch_bounds__(kRingThreads) void k_ring_synth(LegTables t, RingTables rt,
                                                             double *__restrict__ map) {
  extern __shared__ double2 lds[];
  const int tid = threadIdx.x;
  constexpr int NT = kRingThreads;
  const RingInfo ri = rt.info[rt.order[blockIdx.x]];
  const int n = ri.n, h = n / 2, K = t.K, M = t.M;
  const double sgn = (double)ri.sgn;
  const double2 *qe = t.qbuf + ((int64_t)ri.r * 2 + 0) * (M + 1) * K;
  const double2 *qo = t.qbuf + ((int64_t)ri.r * 2 + 1) * (M + 1) * K;
  const int P = ri.P;
  double *out = map + ri.pix_off * K;

  for (int k = 0; k < K; ++k) {
    // X region: [0, h] for power-of-two h, [P-h-1, P) for Bluestein.
    double2 *X = (P == 0) ? lds : lds + (P - h - 1);
    for (int b = tid; b <= h; b += NT) {
      double2 v = gather_bin(b, n, M, qe, qo, K, k, sgn, ri.ph_num, ri.ph_den);
      if (b == 0 || b == h) v.y = 0.0;
      X[b] = v;
    }
    __syncthreads();
    if (P == 0) {
      for (int a = tid; a <= h / 2; a += NT) {
        const int c = h - a;
        const double2 xa = X[a], xc = X[c];
        const double2 wa = expipi(a, h);
        const double2 za = cadd(cadd(xa, cconj(xc)), cmul_i(cmul(wa, csub(xa, cconj(xc)))));
        double2 zc = make_double2(0.0, 0.0);
        const bool two = (a != 0) && (a != c);
        if (two) {
          const double2 wc = expipi(c, h);
          zc = cadd(cadd(xc, cconj(xa)), cmul_i(cmul(wc, csub(xc, cconj(xa)))));
        }
        lds[a] = za;
        if (two) lds[c] = zc;
      }
      __syncthreads();
      lds_fft<true>(lds, h, rt.twiddle);
      for (int p = tid; p < h; p += NT) {
        const double2 z = lds[p];
        out[(int64_t)(2 * p) * K + k] = z.x;
        out[(int64_t)(2 * p + 1) * K + k] = z.y;
      }
    } else {
      const double2 *chirp = rt.chirp + ri.chirp_off;
      const double2 *vhat = rt.vhat + ri.vhat_off;
      for (int a = tid; a < h; a += NT) {
        const double2 xa = X[a], xc = X[h - a];
        const double2 wa = expipi(a, h);
        const double2 za = cadd(cadd(xa, cconj(xc)), cmul_i(cmul(wa, csub(xa, cconj(xc)))));
        lds[a] = cmul(za, chirp[a]);
      }
      __syncthreads();
      for (int a = h + tid; a < P; a += NT) lds[a] = make_double2(0.0, 0.0);
      __syncthreads();
      lds_fft<false>(lds, P, rt.twiddle);
      for (int a = tid; a < P; a += NT) lds[a] = cmul(lds[a], vhat[a]);
      __syncthreads();
      lds_fft<true>(lds, P, rt.twiddle);
      for (int p = tid; p < h; p += NT) {
        const double2 z = cmul(lds[p], chirp[p]);
        out[(int64_t)(2 * p) * K + k] = z.x;
        out[(int64_t)(2 * p + 1) * K + k] = z.y;
      }
    }
    __syncthreads();
  }
}

// Plan-time Bluestein tables, one block per distinct h:
// chirp[a] = e^{i pi a^2 / h} (a < h), and vhat = FFT_P(v) / P with
// v[d] = conj(chirp[|d|]) placed cyclically (d in (-h, h)).
__global__ __launch_bounds__(kRingThreads) void k_bluestein_prep(RingTables rt,
                                                                 const int *__restrict__ hlist,
                                                                 const int64_t *__restrict__ choff,
                                                                 const int64_t *__restrict__ vhoff,
                                                                 const int *__restrict__ plist) {
  extern __shared__ double2 lds[];
  const int tid = threadIdx.x;
  constexpr int NT = kRingThreads;
  const int h = hlist[blockIdx.x], P = plist[blockIdx.x];
  double2 *chirp = rt.chirp + choff[blockIdx.x];
  double2 *vhat = rt.vhat + vhoff[blockIdx.x];
  for (int a = tid; a < P; a += NT) lds[a] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int a = tid; a < h; a += NT) {
    const int64_t a2 = (int64_t)a * a;
    const double2 c = expipi(a2 % (2 * (int64_t)h), h);
    chirp[a] = c;
    lds[a] = cconj(c);
    if (a > 0) lds[P - a] = cconj(c);
  }
  __syncthreads();
  lds_fft<false>(lds, P, rt.twiddle);
  const double inv = 1.0 / (double)P;
  for (int a = tid; a < P; a += NT) vhat[a] = make_double2(lds[a].x * inv, lds[a].y * inv);
}

hipError_t launch_bluestein_prep(const RingTables &rt, const int *hlist, const int64_t *choff,
                                 const int64_t *vhoff, const int *plist, int nh, hipStream_t s) {
  if (nh <= 0) return hipSuccess;
  const size_t lds = (size_t)kMaxFFT * sizeof(double2);
  hipError_t e = hipFuncSetAttribute((const void *)k_bluestein_prep,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bluestein_prep, dim3(nh), dim3(kRingThreads), lds, s, rt, hlist, choff,
                     vhoff, plist);
  return hipGetLastError();
}

hipError_t launch_ring_synth(const LegTables &t, const RingTables &rt, double *map,
                             hipStream_t s) {
  const size_t lds = (size_t)rt.max_lds_complex * sizeof(double2);
  hipError_t e = hipFuncSetAttribute((const void *)k_ring_synth,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_ring_synth, dim3(rt.nrings), dim3(kRingThreads), lds, s, t, rt, map);
  return hipGetLastError();
}

}  // namespace wm

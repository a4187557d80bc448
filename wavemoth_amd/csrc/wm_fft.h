// Device helpers of the ring stage and its adjoint: complex arithmetic,
// the LDS-resident Stockham FFT (radix 8/4/2, one padding slot per 8) and
// the two-level LDS twiddle table.  Included by ring.hip and adjoint.hip.
#pragma once

#include "wm_internal.h"

namespace wm {
namespace fft {

constexpr int NT = kRingThreads;
constexpr int kBins = (kMaxFFT / 2 + 1 + NT - 1) / NT;  // bins per thread
constexpr int kPairs = (kMaxFFT / 4 + 1 + NT - 1) / NT; // Z pairs per thread
constexpr int kOut = (kMaxFFT / 2 + NT - 1) / NT;       // output pairs per thread
constexpr int kFull = (kMaxFFT + NT - 1) / NT;          // FFT points per thread
constexpr int kStage = (kMaxFFT / 2 + 1 + NT - 1) / NT; // staged m per thread
constexpr int kStageM = kStage * NT;                    // m staged per round
constexpr int kHalf = (kMaxFFT / 2 + 1 + NT - 1) / NT;  // half-spectrum entries per thread

__host__ static __device__ __forceinline__ int pidx(int d) { return d + (d >> 3); }

// Number of j-groups that share the alias sum of one bin (B bins per ring),
// so that short rings still spread their (M+1) loads over the block.
__host__ static __device__ __forceinline__ int ring_groups(int B) {
  int G = 1;
  while (2 * G * B <= NT) G *= 2;
  return G;
}

static __device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
static __device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
static __device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(fma(a.x, b.x, -a.y * b.y), fma(a.x, b.y, a.y * b.x));
}
static __device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
static __device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// multiply by +i / -i
static __device__ __forceinline__ double2 cmul_i(double2 a) { return make_double2(-a.y, a.x); }
static __device__ __forceinline__ double2 cmul_mi(double2 a) { return make_double2(a.y, -a.x); }

// The same operations on FP32 complex values (the ring stage of the
// mixed-precision plans).
static __device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
static __device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
static __device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
static __device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
static __device__ __forceinline__ float2 cmul_i(float2 a) { return make_float2(-a.y, a.x); }
static __device__ __forceinline__ float2 cmul_mi(float2 a) { return make_float2(a.y, -a.x); }

// Complex value type C (double2 or float2): its real type, a constant from
// doubles, conversions from and to double2.
template <typename C> struct cx;
template <> struct cx<double2> {
  using R = double;
  static __device__ __forceinline__ double2 mk(double a, double b) { return make_double2(a, b); }
  static __device__ __forceinline__ double2 of(double2 a) { return a; }
};
template <> struct cx<float2> {
  using R = float;
  static __device__ __forceinline__ float2 mk(double a, double b) {
    return make_float2((float)a, (float)b);
  }
  static __device__ __forceinline__ float2 of(double2 a) {
    return make_float2((float)a.x, (float)a.y);
  }
};
static __device__ __forceinline__ double2 to_d(double2 a) { return a; }
static __device__ __forceinline__ double2 to_d(float2 a) {
  return make_double2((double)a.x, (double)a.y);
}

template <bool INV, typename C>
static __device__ __forceinline__ void dft4(C &a, C &b, C &c, C &d) {
  const C t0 = cadd(a, c), t1 = csub(a, c), t2 = cadd(b, d), t3 = csub(b, d);
  a = cadd(t0, t2);
  c = csub(t0, t2);
  if (INV) {
    b = cadd(t1, cmul_i(t3));
    d = cadd(t1, cmul_mi(t3));
  } else {
    b = cadd(t1, cmul_mi(t3));
    d = cadd(t1, cmul_i(t3));
  }
}

// In-register 8-point DFT, y_k = sum_r v_r e^{-/+ 2 pi i r k / 8}.
template <bool INV, typename C>
static __device__ __forceinline__ void dft8(C v[8]) {
  constexpr double c = 0.70710678118654752440;
  dft4<INV>(v[0], v[2], v[4], v[6]);
  dft4<INV>(v[1], v[3], v[5], v[7]);
  // after dft4: e_k in v[0], v[2], v[4], v[6] (k = 0..3), o_k likewise odd
  const C w1 = INV ? cx<C>::mk(c, c) : cx<C>::mk(c, -c);
  const C w3 = INV ? cx<C>::mk(-c, c) : cx<C>::mk(-c, -c);
  const C o1 = cmul(v[3], w1);
  const C o2 = INV ? cmul_i(v[5]) : cmul_mi(v[5]);
  const C o3 = cmul(v[7], w3);
  const C e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6], o0 = v[1];
  v[0] = cadd(e0, o0);
  v[4] = csub(e0, o0);
  v[1] = cadd(e1, o1);
  v[5] = csub(e1, o1);
  v[2] = cadd(e2, o2);
  v[6] = csub(e2, o2);
  v[3] = cadd(e3, o3);
  v[7] = csub(e3, o3);
}

// Twiddle table in LDS: e^{-2 pi i t / kMaxFFT} = C[t >> 6] * F[t & 63],
// C[u] = e^{-2 pi i 64 u / kMaxFFT} (kMaxFFT/64 entries), F[v] =
// e^{-2 pi i v / kMaxFFT} (64 entries).
constexpr int kTwC = kMaxFFT / 64;
constexpr int kTwSlots = kTwC + 64;

static __device__ __forceinline__ double2 twiddle(const double2 *twl, int t) {
  return cmul(twl[t >> 6], twl[kTwC + (t & 63)]);
}

// One Stockham pass of radix R over P <= MAXP points held in LDS (padded).
template <int R, bool INV, int MAXP = kMaxFFT>
static __device__ __forceinline__ void stockham_pass(double2 *buf, int P, int Ns,
                                              const double2 *twl) {
  constexpr int B = (MAXP / R + NT - 1) / NT;
  const int tid = threadIdx.x;
  const int nb = P / R;
  const int tstep = kMaxFFT / (Ns * R);
  double2 v[B][R];
#pragma unroll
  for (int it = 0; it < B; ++it) {
    const int j = tid + it * NT;
    if (j < nb) {
      const int k = j & (Ns - 1);
#pragma unroll
      for (int r = 0; r < R; ++r) v[it][r] = buf[pidx(j + r * nb)];
      if (k) {
        double2 w1 = twiddle(twl, k * tstep);
        if (INV) w1.y = -w1.y;
        double2 w = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[it][r] = cmul(v[it][r], w);
          if (r + 1 < R) w = cmul(w, w1);
        }
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int it = 0; it < B; ++it) {
    const int j = tid + it * NT;
    if (j < nb) {
      const int k = j & (Ns - 1);
      if (R == 8) {
        dft8<INV>(v[it]);
      } else if (R == 4) {
        dft4<INV>(v[it][0], v[it][1], v[it][2], v[it][3]);
      } else {
        const double2 a = v[it][0];
        v[it][0] = cadd(a, v[it][1]);
        v[it][1] = csub(a, v[it][1]);
      }
      const int d = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) buf[pidx(d + r * Ns)] = v[it][r];
    }
  }
  __syncthreads();
}

// In-place power-of-two FFT (P <= MAXP <= kMaxFFT) in padded LDS by the
// whole block.  INV = false: e^{-2 pi i jk/P};  INV = true: e^{+2 pi i jk/P}.
// Expects the data visible to all threads on entry; leaves it so on exit.
template <bool INV, int MAXP = kMaxFFT>
static __device__ __forceinline__ void lds_fft(double2 *buf, int P, const double2 *twl) {
  int Ns = 1;
  while (Ns < P) {
    const int rem = P / Ns;
    if (rem >= 8) {
      stockham_pass<8, INV, MAXP>(buf, P, Ns, twl);
      Ns *= 8;
    } else if (rem == 4) {
      stockham_pass<4, INV, MAXP>(buf, P, Ns, twl);
      Ns *= 4;
    } else {
      stockham_pass<2, INV, MAXP>(buf, P, Ns, twl);
      Ns *= 2;
    }
  }
}

// Fills the LDS twiddle table from the global one (block-wide; syncs).
static __device__ __forceinline__ void load_twiddles(double2 *twl, const double2 *__restrict__ tw) {
  for (int u = threadIdx.x; u < kTwSlots; u += NT)
    twl[u] = (u < kTwC) ? tw[u * 64] : tw[u - kTwC];
  __syncthreads();
}

// e^{i pi num/den}, num reduced exactly modulo 2 den first.
static __device__ __forceinline__ double2 expipi(int64_t num, int64_t den) {
  int64_t t = num % (2 * den);
  if (t < 0) t += 2 * den;
  double s, c;
  sincospi((double)t / (double)den, &s, &c);
  return make_double2(c, s);
}

}  // namespace fft
}  // namespace wm

// Register-resident FFT of the ring stage: each thread of a T = P/16 thread
// job holds 16 points in registers, passes of radix 16 (the last one radix
// 16, 8, 4 or 2) exchange through one padded LDS buffer per job.
//
// The point type C is double2 or float2 (FP32 data, half the LDS); the
// twiddle table and the twiddle powers stay in double either way.
//
// Canonical layout: thread j of a job holds positions j + T rho in v[rho],
// rho = 0..15.  A Stockham pass of radix R over stride Ns reads butterfly
// j' = j + T q (q < 16/R) inputs x[j' + (P/R) r], twiddles input r by
// W^{r (j' mod Ns)}, W = e^{-/+ 2 pi i/(Ns R)}, and writes output r to
// y[(j' - k) R + k + Ns r].  The first pass (Ns = 1, R = 16) reads the
// canonical layout and the last pass (Ns R = P) writes it, so a forward
// transform, a pointwise product and an inverse transform chain in
// registers.  Every pass but the last costs one LDS round trip and two
// barriers; the barriers are block-wide, so all jobs of a block run the
// same pass sequence.
#pragma once

#include "wm_fft.h"

namespace wm {
namespace fft16 {

using fft::cadd;
using fft::cmul;
using fft::csub;

constexpr int kThreads = 256;  // threads per block of the register-FFT kernels

// one padding slot per 16: the stride-16 writes of the first pass land on
// distinct banks
__host__ __device__ constexpr int pad(int d) { return d + (d >> 4); }
// LDS slots of one job's buffer for an FFT of P points plus one extra bin
__host__ __device__ constexpr int job_slots(int P) { return pad(P) + 2; }

// 16-point DFT in registers, y_k = sum_r v_r e^{-/+ 2 pi i r k/16} (INV: +).
// Radix 4 x 4: r = r0 + 4 r1, k = k1 + 4 k2; y_k ends in v[4 (k & 3) + (k >> 2)].
template <bool INV, typename C>
__device__ __forceinline__ void dft16(C v[16]) {
  constexpr double c1 = 0.92387953251128675613;  // cos(pi/8)
  constexpr double s1 = 0.38268343236508977173;  // sin(pi/8)
  constexpr double r2 = 0.70710678118654752440;
  constexpr double sg = INV ? 1.0 : -1.0;
  using X = fft::cx<C>;
#pragma unroll
  for (int r0 = 0; r0 < 4; ++r0) fft::dft4<INV>(v[r0], v[r0 + 4], v[r0 + 8], v[r0 + 12]);
  // slot r0 + 4 k1 *= w^{r0 k1}, w = e^{-/+ 2 pi i/16}
  const C w1 = X::mk(c1, sg * s1), w2 = X::mk(r2, sg * r2), w3 = X::mk(s1, sg * c1),
          w6 = X::mk(-r2, sg * r2), w9 = X::mk(-c1, -sg * s1);
  v[5] = cmul(v[5], w1);
  v[9] = cmul(v[9], w2);
  v[13] = cmul(v[13], w3);
  v[6] = cmul(v[6], w2);
  v[10] = INV ? fft::cmul_i(v[10]) : fft::cmul_mi(v[10]);
  v[14] = cmul(v[14], w6);
  v[7] = cmul(v[7], w3);
  v[11] = cmul(v[11], w6);
  v[15] = cmul(v[15], w9);
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1)
    fft::dft4<INV>(v[4 * k1], v[4 * k1 + 1], v[4 * k1 + 2], v[4 * k1 + 3]);
}

// Radix-R DFT of x[0..R) in place, natural output order.
template <int R, bool INV, typename C>
__device__ __forceinline__ void dft_r(C *x) {
  if constexpr (R == 16) {
    C t[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) t[r] = x[r];
    dft16<INV>(t);
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = t[4 * (k & 3) + (k >> 2)];
  } else if constexpr (R == 8) {
    fft::dft8<INV>(x);
  } else if constexpr (R == 4) {
    fft::dft4<INV>(x[0], x[1], x[2], x[3]);
  } else {
    const C a = x[0];
    x[0] = cadd(a, x[1]);
    x[1] = csub(a, x[1]);
  }
}

// One pass (Ns, R) of NV independent P-point transforms of the same
// thread layout: transform t in v[t] with LDS buffer buf[t]; twl the
// two-level twiddle table, j the thread's index in its job.  The twiddle
// powers are formed once for all NV transforms.
template <int P, int Ns, bool INV, int NV, typename C>
__device__ __forceinline__ void pass(C (&v)[NV][16], C *const (&buf)[NV], const double2 *twl,
                                     int j) {
  constexpr int T = P / 16;
  constexpr int R = (P / Ns >= 16) ? 16 : P / Ns;
  constexpr int Q = 16 / R;   // butterflies per thread
  constexpr int NB = P / R;   // butterflies per job
  C x[NV][Q][R];
  if constexpr (Ns == 1) {
#pragma unroll
    for (int t = 0; t < NV; ++t)
#pragma unroll
      for (int r = 0; r < 16; ++r) x[t][0][r] = v[t][r];
  } else {
#pragma unroll
    for (int t = 0; t < NV; ++t)
#pragma unroll
      for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) x[t][q][r] = buf[t][pad(j + T * q + NB * r)];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int k = (j + T * q) & (Ns - 1);
      double2 w = fft::twiddle(twl, k * (kMaxFFT / (Ns * R)));
      if (INV) w.y = -w.y;
      double2 p = w;
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const C pc = fft::cx<C>::of(p);
#pragma unroll
        for (int t = 0; t < NV; ++t) x[t][q][r] = cmul(x[t][q][r], pc);
        if (r + 1 < R) p = cmul(p, w);
      }
    }
  }
#pragma unroll
  for (int t = 0; t < NV; ++t)
#pragma unroll
    for (int q = 0; q < Q; ++q) dft_r<R, INV>(x[t][q]);
  if constexpr (Ns * R == P) {
#pragma unroll
    for (int t = 0; t < NV; ++t)
#pragma unroll
      for (int q = 0; q < Q; ++q)
#pragma unroll
        for (int r = 0; r < R; ++r) v[t][q + Q * r] = x[t][q][r];
  } else {
    __syncthreads();  // every read of buf (this pass or earlier users) done
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int jq = j + T * q;
      const int k = jq & (Ns - 1);
      const int d = (jq - k) * R + k;
#pragma unroll
      for (int t = 0; t < NV; ++t)
#pragma unroll
        for (int r = 0; r < R; ++r) buf[t][pad(d + Ns * r)] = x[t][q][r];
    }
    __syncthreads();
    pass<P, Ns * R, INV, NV>(v, buf, twl, j);
  }
}

// P-point transform (16 <= P <= 8192, power of two) of the canonical
// layout, in place: INV = false e^{-2 pi i jk/P}, INV = true e^{+...};
// unnormalised.  Starts with a barrier before its first LDS write, so
// the caller may still read buf up to the call.  Every thread of the
// block must call it.
template <int P, bool INV, typename C>
__device__ __forceinline__ void fft(C v[16], C *buf, const double2 *twl, int j) {
  static_assert(P >= 16 && P <= kMaxFFT && (P & (P - 1)) == 0, "P");
  C(&vv)[1][16] = *reinterpret_cast<C(*)[1][16]>(v);
  C *const bb[1] = {buf};
  pass<P, 1, INV, 1>(vv, bb, twl, j);
}

// Two independent transforms of the same size and direction together
// (buffers buf0, buf1): half the barriers of two calls of fft.
template <int P, bool INV, typename C>
__device__ __forceinline__ void fft2(C (&v)[2][16], C *buf0, C *buf1, const double2 *twl,
                                     int j) {
  static_assert(P >= 16 && P <= kMaxFFT && (P & (P - 1)) == 0, "P");
  C *const bb[2] = {buf0, buf1};
  pass<P, 1, INV, 2>(v, bb, twl, j);
}

}  // namespace fft16
}  // namespace wm

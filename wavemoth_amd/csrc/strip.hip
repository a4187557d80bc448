// Per-strip Legendre transform: the low-level entry the reference's
// wavemoth/lib.pyx:220-280 (legendre_transform) exposes and
// wavemoth/test/test_legendre_transforms.py:22-72 drives.
//
// For a strip of nx rings and nk rows of one parity (l = lmin, lmin + 2,
// ...), starting values P[i] = Lambda_lmin(x_i), Pp1[i] = Lambda_lmin+2(x_i)
// and the recurrence coefficients auxdata[3 (k - 2) + {0, 1, 2}] =
// {alpha_k, beta_k, gamma_k}:
//   P_k = (x_i^2 + alpha_k) beta_k P_{k-1} + gamma_k P_{k-2},
//   y[i][j] = sum_k P_k a[2 k nvecs + j]   (`a` in every-other-row layout).
// Semantics of src/legendre_transform.c.in:680-716 (the scalar kernel) and
// of its SSE twin :640-678, which computes the same sums.
//
// On the GPU one thread owns one output y[i][j]: j is fastest across the
// lanes, so the loads of a row of `a` and the stores of y coalesce, and the
// threads of one ring repeat its (cheap) recurrence.  The operations run in
// the reference's order without contraction into FMAs, so the result is
// bitwise the one of the scalar kernel compiled without FMA contraction.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>

#include "wavemoth.h"
#include "wm_host.h"
#include "wm_internal.h"

namespace wm {

__global__ __launch_bounds__(256) void k_strip_transform(int64_t nx, int64_t nk, int64_t nvecs,
                                                         const double *__restrict__ a,
                                                         double *__restrict__ y,
                                                         const double *__restrict__ x2,
                                                         const double *__restrict__ aux,
                                                         const double *__restrict__ P0,
                                                         const double *__restrict__ P1) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * nvecs) return;
  const int64_t i = t / nvecs, j = t - i * nvecs;
  const double xs = x2[i];
  double pp = P0[i], p = P1[i];
  double acc = pp * a[j];
  acc += p * a[2 * nvecs + j];
  for (int64_t k = 2; k < nk; ++k) {
    const double alpha = aux[3 * (k - 2) + 0];
    const double beta = aux[3 * (k - 2) + 1];
    const double gamma = aux[3 * (k - 2) + 2];
    const double pv = (xs + alpha) * beta * p + gamma * pp;
    pp = p;
    p = pv;
    acc += pv * a[2 * k * nvecs + j];
  }
  y[t] = acc;
}

static bool strip_args_ok(size_t nx, size_t nk, size_t nvecs) {
  if (nk < 2) {
    set_err("legendre transform: nk must be >= 2");
    return false;
  }
  if (nvecs == 0 || nvecs % 2 != 0) {
    set_err("legendre transform: nvecs not divisible by 2");
    return false;
  }
  if (nx > ((size_t)1 << 40) / nvecs) {
    set_err("legendre transform: strip too large");
    return false;
  }
  return true;
}

static hipError_t launch_strip(size_t nx, size_t nk, size_t nvecs, const double *a, double *y,
                               const double *x2, const double *aux, const double *P,
                               const double *Pp1, hipStream_t s) {
  const int64_t n = (int64_t)nx * (int64_t)nvecs;
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_strip_transform, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     (int64_t)nx, (int64_t)nk, (int64_t)nvecs, a, y, x2, aux, P, Pp1);
  return hipGetLastError();
}

// The Legendre stage from a resource file's strips: thread (item, map
// block) with item -> (strip, ring pair) by binary search over the strips'
// prefix column counts.  It sums P_k(x) a_{l_k} over the strip's rows for
// up to 8 maps, P_k from the stored start rows by the x^2 recurrence
// (src/legendre_transform.c.in:680-716) or read from a dense block, which
// is what src/wavemoth.c:958-1009 (pull_a_through_legendre_block) feeds its
// kernels; rows before the strip's first are zero there and here.
__global__ __launch_bounds__(256) void k_res_legendre(LegTables t, const double2 *__restrict__ alm,
                                                      const double *__restrict__ img,
                                                      const ResStrip *__restrict__ strips,
                                                      const int64_t *__restrict__ prefix,
                                                      int nstrips, int64_t items,
                                                      double2 *__restrict__ chunk) {
  const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= items) return;
  const int K = t.K, L = t.L, Mp1 = t.M + 1;
  const int kb0 = blockIdx.y * 8;
  const int kn = (K - kb0) < 8 ? (K - kb0) : 8;
  int lo = 0, hi = nstrips - 1;  // last strip with prefix <= it
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= it) lo = mid;
    else hi = mid - 1;
  }
  const ResStrip s = strips[lo];
  const int c = (int)(it - prefix[lo]);
  const int m = s.m;
  const int64_t base = (int64_t)m * (2 * L - m + 3) / 2;  // a_lm row of (l = m, m)
  const int64_t l0 = (int64_t)s.odd + 2 * s.row;          // l - m of the first row
  double2 acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = make_double2(0.0, 0.0);
  // (arguments evaluated once, before the loop over maps)
#define WM_RES_ADD(row, P)                                                    \
  do {                                                                        \
    const double p_ = (P);                                                    \
    const double2 *a_ = alm + (base + l0 + 2 * (int64_t)(row)) * K + kb0;    \
    _Pragma("unroll") for (int kk_ = 0; kk_ < 8; ++kk_) {                     \
      if (kk_ < kn) {                                                         \
        const double2 v_ = a_[kk_];                                           \
        acc[kk_].x = fma(p_, v_.x, acc[kk_].x);                               \
        acc[kk_].y = fma(p_, v_.y, acc[kk_].y);                               \
      }                                                                       \
    }                                                                         \
  } while (0)
  if (s.x2 < 0) {
    const double *D = img + s.dense + (int64_t)s.nks * c;
    for (int k = 0; k < s.nks; ++k) WM_RES_ADD(k, D[k]);
  } else {
    const double x2 = img[s.x2 + c];
    double pp = img[s.p0 + c], p = img[s.p1 + c];
    WM_RES_ADD(0, pp);
    WM_RES_ADD(1, p);
    const double *aux = img + s.aux;
    for (int k = 2; k < s.nks; ++k) {
      const double pv = (x2 + aux[3 * (k - 2)]) * aux[3 * (k - 2) + 1] * p +
                        aux[3 * (k - 2) + 2] * pp;
      pp = p;
      p = pv;
      WM_RES_ADD(k, pv);
    }
  }
#undef WM_RES_ADD
  if (s.chunk >= 0) {
    double2 *dst = chunk + (s.chunk + c) * K + kb0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (k < kn) dst[k] = acc[k];
    return;
  }
  const int ring = s.cstart + c;
#pragma unroll
  for (int k = 0; k < 8; ++k)
    if (k < kn) t.qbuf[q_index(ring, s.odd, kb0 + k, m, K, Mp1)] = acc[k];
}

// The transposed interpolations of one tree depth of the compressed
// matrices: thread (output row, map block of 8), row -> op by binary
// search over the ops' first rows.  The T and B blocks of a pair feed the
// same rows, so one thread forms both terms and writes the sum once (the
// reference's second scatter adds into the first's output).
__global__ __launch_bounds__(256) void k_res_interp(LegTables t, const double *__restrict__ img,
                                                    const ResInterp *__restrict__ ops,
                                                    const int64_t *__restrict__ prefix,
                                                    const int32_t *__restrict__ codes, int op_lo,
                                                    int op_hi, int64_t item_lo, int64_t item_hi,
                                                    double2 *__restrict__ chunk) {
  const int64_t it = item_lo + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (it >= item_hi) return;
  const int K = t.K, Mp1 = t.M + 1;
  const int kb0 = blockIdx.y * 8;
  const int kn = (K - kb0) < 8 ? (K - kb0) : 8;
  int lo = op_lo, hi = op_hi - 1;  // last op with prefix <= it
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (prefix[mid] <= it) lo = mid;
    else hi = mid - 1;
  }
  const ResInterp o = ops[lo];
  const int p = (int)(it - prefix[lo]);
  double2 acc[8];
#pragma unroll
  for (int k = 0; k < 8; ++k) acc[k] = make_double2(0.0, 0.0);
#pragma unroll
  for (int side = 0; side < 2; ++side) {
    const int code = codes[(side ? o.codeB : o.codeT) + p];
    const int kx = side ? o.kB : o.kT;
    const double2 *x = chunk + (side ? o.inB : o.inT) * K + kb0;
    if (code >= 0) {
#pragma unroll
      for (int k = 0; k < 8; ++k)
        if (k < kn) {
          const double2 v = x[(int64_t)code * K + k];
          acc[k].x += v.x;
          acc[k].y += v.y;
        }
    } else {
      const double *A = img + (side ? o.ipB : o.ipT) + (int64_t)(-1 - code) * kx;
      for (int r = 0; r < kx; ++r) {
        const double a = A[r];
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (k < kn) {
            const double2 v = x[(int64_t)r * K + k];
            acc[k].x = fma(a, v.x, acc[k].x);
            acc[k].y = fma(a, v.y, acc[k].y);
          }
      }
    }
  }
  const bool left = p < o.nl;
  const int64_t row = left ? o.outL + p : o.outR + (p - o.nl);
  if (o.to_rings) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (k < kn) t.qbuf[q_index((int)row, o.odd, kb0 + k, o.m, K, Mp1)] = acc[k];
  } else {
    double2 *dst = chunk + row * K + kb0;
#pragma unroll
    for (int k = 0; k < 8; ++k)
      if (k < kn) dst[k] = acc[k];
  }
}

__global__ __launch_bounds__(256) void k_zero_q(double2 *q, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) q[i] = make_double2(0.0, 0.0);
}

hipError_t launch_res_legendre(const LegTables &t, const double2 *alm, const double *img,
                               const ResStrip *strips, const int64_t *prefix, int nstrips,
                               int64_t items, double2 *chunk, bool zero_q, hipStream_t s) {
  if (zero_q) {
    const int64_t n = q_elems(t.R, t.K, t.M + 1);
    hipLaunchKernelGGL(k_zero_q, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, t.qbuf, n);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (items == 0) return hipSuccess;
  const dim3 grid((unsigned)((items + 255) / 256), (unsigned)((t.K + 7) / 8));
  hipLaunchKernelGGL(k_res_legendre, grid, dim3(256), 0, s, t, alm, img, strips, prefix, nstrips,
                     items, chunk);
  return hipGetLastError();
}

hipError_t launch_res_interp(const LegTables &t, const double *img, const ResInterp *ops,
                             const int64_t *prefix, const int32_t *codes, int op_lo, int op_hi,
                             int64_t item_lo, int64_t item_hi, double2 *chunk, hipStream_t s) {
  if (item_hi <= item_lo || op_hi <= op_lo) return hipSuccess;
  const dim3 grid((unsigned)((item_hi - item_lo + 255) / 256), (unsigned)((t.K + 7) / 8));
  hipLaunchKernelGGL(k_res_interp, grid, dim3(256), 0, s, t, img, ops, prefix, codes, op_lo,
                     op_hi, item_lo, item_hi, chunk);
  return hipGetLastError();
}

}  // namespace wm

using namespace wm;

extern "C" void wavemoth_legendre_transform_auxdata(size_t m, size_t lmin, size_t nk,
                                                   double *auxdata) {
  // c_l = sqrt((l-m+1)(l-m+2)(l+m+1)(l+m+2) / ((2l+1)(2l+3)^2(2l+5))),
  // d_l = (2l(l+1) - 2m^2 - 1) / ((2l-1)(2l+3)), l = lmin + 2k; then for
  // k >= 2 alpha = -d_{k-1}, beta = 1 / c_{k-1}, gamma = -c_{k-2} / c_{k-1}.
  // The integer products are exact in 64 bits for l < ~40000.
  double c1 = 0, c2 = 0, d1 = 0;
  const int64_t mm = (int64_t)m;
  for (size_t k = 0; k < nk; ++k) {
    const int64_t l = (int64_t)lmin + 2 * (int64_t)k;
    const double cn = (double)((l - mm + 1) * (l - mm + 2) * (l + mm + 1) * (l + mm + 2));
    const double cd = (double)((2 * l + 1) * (2 * l + 3) * (2 * l + 3) * (2 * l + 5));
    const double c = std::sqrt(cn / cd);
    const double d = (double)(2 * l * (l + 1) - 2 * mm * mm - 1) /
                     (double)((2 * l - 1) * (2 * l + 3));
    if (k >= 2) {
      auxdata[3 * (k - 2) + 0] = -d1;
      auxdata[3 * (k - 2) + 1] = 1 / c1;
      auxdata[3 * (k - 2) + 2] = -c2 / c1;
    }
    d1 = d;
    c2 = c1;
    c1 = c;
  }
}

extern "C" int wavemoth_legendre_transform_device(size_t nx, size_t nk, size_t nvecs,
                                                  const double *a, double *y,
                                                  const double *x_squared,
                                                  const double *auxdata, const double *P,
                                                  const double *Pp1, void *stream) {
  clear_err();
  if (!strip_args_ok(nx, nk, nvecs)) return -1;
  if (nx > 0 && (!a || !y || !x_squared || !P || !Pp1 || (nk > 2 && !auxdata))) {
    set_err("legendre transform: NULL argument");
    return -1;
  }
  const hipError_t e =
      launch_strip(nx, nk, nvecs, a, y, x_squared, auxdata, P, Pp1, (hipStream_t)stream);
  if (e != hipSuccess) {
    set_err(std::string("legendre transform launch: ") + hipGetErrorString(e));
    return -1;
  }
  return 0;
}

extern "C" void wavemoth_legendre_transform(size_t nx, size_t nk, size_t nvecs, double *a,
                                            double *y, double *x_squared, double *auxdata,
                                            double *P, double *Pp1) {
  clear_err();
  if (!strip_args_ok(nx, nk, nvecs)) return;
  if (nx == 0) return;
  DevArena arena;
  const size_t na = 2 * nk * nvecs, ny = nx * nvecs, nau = nk > 2 ? 3 * (nk - 2) : 1;
  double *da = nullptr, *dy = nullptr, *dx = nullptr, *dau = nullptr, *dp = nullptr,
         *dp1 = nullptr;
  auto fail = [&](const char *what, hipError_t e) {
    set_err(std::string("legendre transform ") + what + ": " + hipGetErrorString(e));
    arena.release();
  };
  if (!dalloc(arena, &da, na, "a") || !dalloc(arena, &dy, ny, "y") ||
      !dalloc(arena, &dx, nx, "x_squared") || !dalloc(arena, &dau, nau, "auxdata") ||
      !dalloc(arena, &dp, nx, "P") || !dalloc(arena, &dp1, nx, "Pp1")) {
    arena.release();
    return;
  }
  hipError_t e;
  if ((e = hipMemcpy(da, a, na * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dx, x_squared, nx * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (nk > 2 && (e = hipMemcpy(dau, auxdata, nau * 8, hipMemcpyHostToDevice)) != hipSuccess) ||
      (e = hipMemcpy(dp, P, nx * 8, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(dp1, Pp1, nx * 8, hipMemcpyHostToDevice)) != hipSuccess)
    return fail("upload", e);
  if ((e = launch_strip(nx, nk, nvecs, da, dy, dx, dau, dp, dp1, nullptr)) != hipSuccess)
    return fail("launch", e);
  if ((e = hipMemcpy(y, dy, ny * 8, hipMemcpyDeviceToHost)) != hipSuccess)
    return fail("download", e);
  arena.release();
}

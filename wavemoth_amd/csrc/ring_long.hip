// Ring stage for rings whose transform exceeds one workgroup's LDS (Nside >
// 4096: belt rings of 2 Nside > 8192 complex points, polar caps whose
// Bluestein length passes 8192).  Same mathematics as the general LDS
// kernel (k_ring_pairs, src/wavemoth.c:1083-1191 semantics): per ring and
// map the alias sums of g = E +/- O over the bins 0 .. h, the Z pairs, then
// the length-h inverse DFT directly (h a power of two) or by Bluestein's
// chirp z-transform (FFT_P, times vhat, inverse FFT_P, times the chirp),
// and the pixels z_p -> (2p, 2p + 1).
//
// The P-point FFTs (P > 8192) run as four-step transforms through an HBM
// scratch buffer, P = A B (A <= B <= 8192, powers of two), n = B a + b,
// k = k1 + A k2:
//   k_long_cols: for each column b, the A-point DFT over a (stride B), times
//                the twiddle e^{-/+2 pi i b k1 / P}, in place;
//   k_long_rows: for each row k1, the B-point DFT over b (contiguous), into
//                the output at k1 + A k2 (natural order), optionally times a
//                per-vector table (Bluestein's vhat).
// Both hold each A- or B-point transform in registers (the radix-16
// register FFT of the ring kernels, one job per column or row).  A launch
// handles a batch of vectors of one length P; vectors are (long ring,
// map) pairs, or Bluestein kernel vectors at plan time.
#include <mutex>

#include "wm_fft16.h"
#include "wm_host.h"

namespace wm {

using namespace fft;
using fft16::job_slots;
using fft16::pad;

namespace {

constexpr int LB = 2 * fft16::kThreads;  // threads per block of the FFT kernels

__device__ __forceinline__ void load_tw_long(double2 *twl, const double2 *__restrict__ tw) {
  for (int u = threadIdx.x; u < kTwSlots; u += LB) twl[u] = (u < kTwC) ? tw[u * 64] : tw[u - kTwC];
}

// LDS of a four-step kernel: the twiddle table, then one buffer per job.
__host__ __device__ constexpr size_t long_lds_bytes(int S) {
  return (size_t)kTwSlots * sizeof(double2) +
         (size_t)(LB / (S / 16)) * job_slots(S) * sizeof(double2);
}

// Column pass of vectors x + t P (t < nvec): job (t, b) of b < B.
template <int A, bool INV>
__global__ __launch_bounds__(LB) void k_long_cols(const double2 *__restrict__ tw,
                                                  double2 *__restrict__ x, int64_t nvec, int P,
                                                  int B) {
  constexpr int T = A / 16, J = LB / T;
  extern __shared__ double2 lds[];
  double2 *twl = lds;
  load_tw_long(twl, tw);  // first read after the transform's first barrier
  const int jb = threadIdx.x / T;
  int j = threadIdx.x - jb * T;
  const int64_t g = (int64_t)blockIdx.x * J + jb;
  const int64_t t = g / B;
  const int b = (int)(g - t * B);
  const bool act = t < nvec;
  double2 *xv = x + (act ? t : 0) * P + b;
  double2 v[16];
#pragma unroll
  for (int rho = 0; rho < 16; ++rho)
    v[rho] = act ? xv[(int64_t)B * (j + T * rho)] : make_double2(0.0, 0.0);
  fft16::fft<A, INV>(v, lds + kTwSlots + jb * job_slots(A), twl, j);
  asm volatile("" : "+v"(j));
  if (act) {
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int k1 = j + T * rho;
      const double2 w = expipi((INV ? 2 : -2) * (int64_t)b * k1, P);
      xv[(int64_t)B * k1] = cmul(v[rho], w);
    }
  }
}

// Row pass: job (t, k1) of k1 < A reads x + t P + B k1 (B points) and
// writes y + t P + k1 + A k2; with mul, times mul[t] table entry k (the
// table of vector t starts at mtab + mul_off(t)).
template <int Bn, bool INV>
__global__ __launch_bounds__(LB) void k_long_rows(const double2 *__restrict__ tw,
                                                  const double2 *__restrict__ x,
                                                  double2 *__restrict__ y, int64_t nvec, int P,
                                                  int A, const double2 *__restrict__ mtab,
                                                  const LongVec *__restrict__ vecs, int64_t v0,
                                                  int K, const RingPair *__restrict__ pairs) {
  constexpr int T = Bn / 16, J = LB / T;
  extern __shared__ double2 lds[];
  double2 *twl = lds;
  load_tw_long(twl, tw);
  const int jb = threadIdx.x / T;
  int j = threadIdx.x - jb * T;
  const int64_t g = (int64_t)blockIdx.x * J + jb;
  const int64_t t = g / A;
  const int k1 = (int)(g - t * A);
  const bool act = t < nvec;
  const double2 *xv = x + (act ? t : 0) * P + (int64_t)Bn * k1;
  double2 v[16];
#pragma unroll
  for (int rho = 0; rho < 16; ++rho) v[rho] = act ? xv[j + T * rho] : make_double2(0.0, 0.0);
  fft16::fft<Bn, INV>(v, lds + kTwSlots + jb * job_slots(Bn), twl, j);
  asm volatile("" : "+v"(j));
  if (act) {
    const double2 *mt = nullptr;
    if (mtab) mt = mtab + pairs[vecs[(v0 + t) / K].pair].vhat_off;
    double2 *yv = y + t * P + k1;
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int64_t k = k1 + (int64_t)A * (j + T * rho);
      yv[(int64_t)A * (j + T * rho)] = mt ? cmul(v[rho], mt[k]) : v[rho];
    }
  }
}

// g_m = E_m + hs O_m of one (ring pair, map), m <= M
template <typename Q>
__device__ __forceinline__ double2 gval(const Q *qe, const Q *qo, int64_t qs, int64_t m,
                                        double hs) {
  const double2 e = to_d(qe[m * qs]), o = to_d(qo[m * qs]);
  return make_double2(fma(hs, o.x, e.x), fma(hs, o.y, e.y));
}

// Signed alias sum of bin b (0 <= b <= h): sum_j sg_j g_{b + j n} +
// sum_{j >= 1} sg_j conj g_{j n - b}, sg_j = -1 for odd j when s = 1.
template <typename Q>
__device__ __forceinline__ double2 alias_sum(const Q *qe, const Q *qo, int64_t qs, int M, int n,
                                             int s, int b, double hs) {
  double2 acc = make_double2(0.0, 0.0);
  for (int64_t j = 0; b + j * n <= M || j * n - b <= M; ++j) {
    const double sg = (s && (j & 1)) ? -1.0 : 1.0;
    const int64_t m1 = b + j * n, m2 = j * n - b;
    if (m1 <= M) {
      const double2 v = gval(qe, qo, qs, m1, hs);
      acc.x = fma(sg, v.x, acc.x);
      acc.y = fma(sg, v.y, acc.y);
    }
    if (j >= 1 && m2 >= 0 && m2 <= M) {
      const double2 v = gval(qe, qo, qs, m2, hs);
      acc.x = fma(sg, v.x, acc.x);
      acc.y = fma(-sg, v.y, acc.y);
    }
  }
  return acc;
}

// Input vectors of a batch: vector t = (long ring entry (v0 + t) / K, map
// (v0 + t) % K); entry a < P: Z_a (times the chirp for Bluestein rings),
// zero past h.  Z_a = (x_a + conj x_c) + i e^{2 pi i a/n} (x_a - conj x_c),
// c = h - a, x_b = e^{i pi s b/n} X_b with X the alias sums (imaginary part
// of the DC and Nyquist bins dropped).
template <typename Q>
__global__ __launch_bounds__(256) void k_long_z(LegTables t, RingTables rt,
                                                const LongVec *__restrict__ vecs, int64_t v0,
                                                int64_t nvec, int P, double2 *__restrict__ u) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nvec * P) return;
  const int64_t tv = idx / P;
  const int a = (int)(idx - tv * P);
  const int64_t gv = v0 + tv;
  const LongVec lv = vecs[gv / t.K];
  const int k = (int)(gv % t.K);
  const RingPair rp = rt.pairs[lv.pair];
  const int n = rp.n, h = n / 2, K = t.K, M = t.M;
  if (a >= h) {
    u[idx] = make_double2(0.0, 0.0);
    return;
  }
  const Q *qe = (const Q *)t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  const Q *qo = (const Q *)t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
  const double hs = lv.hemi ? -1.0 : 1.0;
  const int c = h - a;
  double2 xa = cmul(alias_sum(qe, qo, qs, M, n, rp.s, a, hs), expipi((int64_t)rp.s * a, n));
  double2 xc = cmul(alias_sum(qe, qo, qs, M, n, rp.s, c, hs), expipi((int64_t)rp.s * c, n));
  if (a == 0) xa.y = xc.y = 0.0;
  const double2 om = expipi(2 * (int64_t)a, n);
  const double2 cc = cconj(xc);
  double2 z = cadd(cadd(xa, cc), cmul_i(cmul(om, csub(xa, cc))));
  if (rp.P) z = cmul(z, rt.chirp[rp.chirp_off + a]);
  u[idx] = z;
}

// Pixels of a batch: z_p = y_p (times the chirp for Bluestein rings), p < h,
// to pixels 2p, 2p + 1 of the vector's ring and map.
__global__ __launch_bounds__(256) void k_long_out(LegTables t, RingTables rt,
                                                  const LongVec *__restrict__ vecs, int64_t v0,
                                                  int64_t nvec, int P,
                                                  const double2 *__restrict__ y,
                                                  double *__restrict__ map) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nvec * P) return;
  const int64_t tv = idx / P;
  const int p = (int)(idx - tv * P);
  const int64_t gv = v0 + tv;
  const LongVec lv = vecs[gv / t.K];
  const int k = (int)(gv % t.K);
  const RingPair rp = rt.pairs[lv.pair];
  if (p >= rp.n / 2) return;
  double2 z = y[idx];
  if (rp.P) z = cmul(z, rt.chirp[rp.chirp_off + p]);
  const int K = t.K;
  double *out = map + k + (lv.hemi ? rp.pix_s : rp.pix_n) * K;
  out[(int64_t)(2 * p) * K] = z.x;
  out[(int64_t)(2 * p + 1) * K] = z.y;
}

// Bluestein kernel vectors of a plan-time batch: vector t is entry e0 + t
// of (h, chirp offset, vhat offset): chirp[a] = e^{i pi a^2/h} (a < h) into
// the table, v[d] = conj(chirp[|d|]) placed cyclically into u.
__global__ __launch_bounds__(256) void k_long_chirp(const int *__restrict__ hs,
                                                    const int64_t *__restrict__ choff, int e0,
                                                    int64_t nvec, int P, double2 *__restrict__ chirp,
                                                    double2 *__restrict__ u) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nvec * P) return;
  const int64_t tv = idx / P;
  const int a = (int)(idx - tv * P);
  const int h = hs[e0 + tv];
  const int d = a < h ? a : (P - a < h ? P - a : -1);
  double2 v = make_double2(0.0, 0.0);
  if (d >= 0) {
    const double2 c = expipi(((int64_t)d * d) % (2 * (int64_t)h), h);
    if (a < h) chirp[choff[e0 + tv] + a] = c;
    v = cconj(c);
  }
  u[idx] = v;
}

// vhat = y / P into the table at its offset
__global__ __launch_bounds__(256) void k_long_vhat(const int64_t *__restrict__ vhoff, int e0,
                                                   int64_t nvec, int P,
                                                   const double2 *__restrict__ y,
                                                   double2 *__restrict__ vhat) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nvec * P) return;
  const int64_t tv = idx / P;
  const int a = (int)(idx - tv * P);
  vhat[vhoff[e0 + tv] + a] = cscale(y[idx], 1.0 / (double)P);
}

#define WM_LONG_SIZES(X) X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)

int log2i(int v) {
  int l = 0;
  while ((1 << l) < v) ++l;
  return l;
}

// P = A B, A = 2^floor(log2 P / 2) <= B
void long_split(int P, int *A, int *B) {
  *A = 1 << (log2i(P) / 2);
  *B = P / *A;
}

hipError_t cols(const double2 *tw, double2 *x, int64_t nvec, int P, bool inv, hipStream_t s) {
  int A, B;
  long_split(P, &A, &B);
  const int64_t jobs = nvec * B;
#define WM_COLS(S)                                                                            \
  if (A == S) {                                                                               \
    constexpr int J = LB / (S / 16);                                                          \
    const unsigned nb = (unsigned)((jobs + J - 1) / J);                                       \
    if (inv)                                                                                  \
      hipLaunchKernelGGL((k_long_cols<S, true>), dim3(nb), dim3(LB), long_lds_bytes(S), s, tw, \
                         x, nvec, P, B);                                                      \
    else                                                                                      \
      hipLaunchKernelGGL((k_long_cols<S, false>), dim3(nb), dim3(LB), long_lds_bytes(S), s,    \
                         tw, x, nvec, P, B);                                                  \
    return hipGetLastError();                                                                 \
  }
  WM_LONG_SIZES(WM_COLS)
#undef WM_COLS
  return hipErrorInvalidValue;
}

hipError_t rows(const double2 *tw, const double2 *x, double2 *y, int64_t nvec, int P, bool inv,
                const double2 *mtab, const LongVec *vecs, int64_t v0, int K,
                const RingPair *pairs, hipStream_t s) {
  int A, B;
  long_split(P, &A, &B);
  const int64_t jobs = nvec * A;
#define WM_ROWS(S)                                                                              \
  if (B == S) {                                                                                 \
    constexpr int J = LB / (S / 16);                                                            \
    const unsigned nb = (unsigned)((jobs + J - 1) / J);                                         \
    if (inv)                                                                                    \
      hipLaunchKernelGGL((k_long_rows<S, true>), dim3(nb), dim3(LB), long_lds_bytes(S), s, tw,   \
                         x, y, nvec, P, A, mtab, vecs, v0, K, pairs);                          \
    else                                                                                        \
      hipLaunchKernelGGL((k_long_rows<S, false>), dim3(nb), dim3(LB), long_lds_bytes(S), s, tw,  \
                         x, y, nvec, P, A, mtab, vecs, v0, K, pairs);                          \
    return hipGetLastError();                                                                   \
  }
  WM_LONG_SIZES(WM_ROWS)
#undef WM_ROWS
  return hipErrorInvalidValue;
}

unsigned grid256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

bool long_size_ok(int P) {
  int A, B;
  long_split(P, &A, &B);
  return P > kMaxFFT && (P & (P - 1)) == 0 && A >= 16 && B <= kMaxFFT;
}

hipError_t long_set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
#define WM_SET(S)                                                                           \
  for (const void *f : {(const void *)k_long_cols<S, false>, (const void *)k_long_cols<S, true>, \
                        (const void *)k_long_rows<S, false>, (const void *)k_long_rows<S, true>}) \
    if (err == hipSuccess)                                                                  \
      err = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,              \
                                (int)long_lds_bytes(S));
    WM_LONG_SIZES(WM_SET)
#undef WM_SET
  });
  return err;
}

hipError_t launch_long_rings(const LegTables &t, const RingTables &rt, const LongRings &lr,
                             double *map, hipStream_t s) {
  const int K = t.K;
  for (const LongGroup &g : lr.groups) {
    const int64_t nvec_all = (int64_t)g.count * K;
    const int64_t per = lr.elems / g.P;  // vectors per batch
    if (per < 1) return hipErrorInvalidValue;
    for (int64_t v0 = 0; v0 < nvec_all; v0 += per) {
      const int64_t nv = nvec_all - v0 < per ? nvec_all - v0 : per;
      const int64_t gv0 = (int64_t)g.first * K + v0;
      const int64_t n = nv * g.P;
      hipError_t e;
      if (t.q32)
        hipLaunchKernelGGL(k_long_z<float2>, dim3(grid256(n)), dim3(256), 0, s, t, rt, lr.vecs,
                           gv0, nv, g.P, lr.u);
      else
        hipLaunchKernelGGL(k_long_z<double2>, dim3(grid256(n)), dim3(256), 0, s, t, rt,
                           lr.vecs, gv0, nv, g.P, lr.u);
      if ((e = hipGetLastError()) != hipSuccess) return e;
      const double2 *res = lr.v;
      if (!g.blue) {
        // the inverse DFT of Z (h = P points)
        if ((e = cols(rt.twiddle, lr.u, nv, g.P, true, s)) != hipSuccess) return e;
        if ((e = rows(rt.twiddle, lr.u, lr.v, nv, g.P, true, nullptr, lr.vecs, gv0, K, rt.pairs,
                      s)) != hipSuccess)
          return e;
      } else {
        // Bluestein: FFT, times vhat, inverse FFT
        if ((e = cols(rt.twiddle, lr.u, nv, g.P, false, s)) != hipSuccess) return e;
        if ((e = rows(rt.twiddle, lr.u, lr.v, nv, g.P, false, rt.vhat, lr.vecs, gv0, K, rt.pairs,
                      s)) != hipSuccess)
          return e;
        if ((e = cols(rt.twiddle, lr.v, nv, g.P, true, s)) != hipSuccess) return e;
        if ((e = rows(rt.twiddle, lr.v, lr.u, nv, g.P, true, nullptr, lr.vecs, gv0, K, rt.pairs,
                      s)) != hipSuccess)
          return e;
        res = lr.u;
      }
      hipLaunchKernelGGL(k_long_out, dim3(grid256(n)), dim3(256), 0, s, t, rt, lr.vecs, gv0, nv,
                         g.P, res, map);
      if ((e = hipGetLastError()) != hipSuccess) return e;
    }
  }
  return hipSuccess;
}

hipError_t launch_long_bluestein_prep(const double2 *twiddle, double2 *chirp, double2 *vhat,
                                      const int *d_h, const int64_t *d_choff,
                                      const int64_t *d_vhoff, int nent, int P, double2 *u,
                                      double2 *v, int64_t elems, hipStream_t s) {
  const int64_t per = elems / P;
  if (per < 1) return hipErrorInvalidValue;
  for (int64_t e0 = 0; e0 < nent; e0 += per) {
    const int64_t nv = nent - e0 < per ? nent - e0 : per;
    const int64_t n = nv * P;
    hipLaunchKernelGGL(k_long_chirp, dim3(grid256(n)), dim3(256), 0, s, d_h, d_choff, (int)e0,
                       nv, P, chirp, u);
    hipError_t e;
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = cols(twiddle, u, nv, P, false, s)) != hipSuccess) return e;
    if ((e = rows(twiddle, u, v, nv, P, false, nullptr, nullptr, 0, 1, nullptr, s)) != hipSuccess)
      return e;
    hipLaunchKernelGGL(k_long_vhat, dim3(grid256(n)), dim3(256), 0, s, d_vhoff, (int)e0, nv, P,
                       v, vhat);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace wm

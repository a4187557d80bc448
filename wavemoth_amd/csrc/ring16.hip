// Ring stage, register-FFT kernels.  Same semantics as ring.hip
// (src/wavemoth.c:1083-1191): G_m = E_m +/- O_m, phase e^{i m phi0},
// bins (m mod n) and conjugates ((n - m) mod n) up to n/2, FFTW-style c2r of
// length n (imaginary parts of DC and Nyquist ignored, unnormalised), with
// the same Z-pair reduction of the c2r to a complex DFT of h = n/2.
//
// kind 1, belt: h a power of two and mmax <= h (the equatorial belt): one
//   job per (ring pair, map), both rings from one read of E and O, one
//   P = h point inverse FFT per ring.
// kind 2, caps: h = 2i not a power of two, 4i > mmax (no wrap: bin b
//   receives at most g_b and the conjugate term (-1)^s conj g_{n-b}, so the
//   fold is two passes over LDS with one writer per bin): the length-h
//   inverse DFT splits into the DFTs of the even and odd Z entries,
//   z_p = A_p + e^{2 pi i p/h} B_p, z_{p+i} = A_p - e^{2 pi i p/h} B_p, each a
//   length-i Bluestein convolution of P = next_pow2(2i - 1) <= 4096 points
//   (chirp c_k = e^{i pi k^2/i}, kernel spectrum vhat tabulated per i at
//   plan time).  Splitting keeps every transform within 4096 points.
// kind 3, wrapping rings (n <= mmax): the fold sums the row blocks of every
//   bin in registers, then as kind 2.
//
// One job = T = P/16 threads holding 16 points each (wm_fft16.h); a block of
// 512 threads runs 512/T belt jobs or 512/(2T) cap jobs.  Pixel rows hold
// all maps of a pixel (K doubles), so the stores of one map are strided and
// the kernels put several maps of a row into each store instruction: FP64
// belt jobs go in pairs that exchange half-waves (two maps per lane); FP32
// belt jobs, and FP32 cap jobs of Bluestein length 4096, hold two maps each
// at K >= 2 (one 16-byte store per row, the paired job writing the next two
// maps of the same rows); other FP32 cap blocks whose jobs are consecutive
// maps of one pair store all of them together from LDS.
//
// Every kernel is instantiated for two point types C: double2 (the FP64
// plans) and float2 (the mixed-precision plans, whose contraction stores q
// in FP32).  With float2, q, the folds, LDS and the FFT arithmetic are FP32,
// and the LDS one FP64 ring takes holds both rings of the pair, which are
// transformed together; phase factors, twiddles and their powers are formed
// in double and rounded once; pixels are stored as double.  Jobs of the G
// maps of one q group (they share the lines of q) and the blocks holding
// them are placed on one XCD at consecutive dispatch slots.
#include <mutex>

#include "wm_fft16.h"

namespace wm {

using namespace fft;
using fft16::pad;
using fft16::job_slots;
constexpr int RB = 2 * fft16::kThreads;  // threads per block of both kernels

__host__ __device__ inline int pow2_ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Job (pair index, map) of job slot jb of block b.  U = pow2_ceil(q_group(K))
// jobs per unit (pair, q group), J jobs per block.  Block b runs on XCD
// x = b mod 8; the units of XCD x are pairs x, x + 8, ..., each with all its
// q groups, so that every line of a pair's pixel rows (all maps of a pixel)
// is written through one L2.  Returns false for a padding job.
__device__ __forceinline__ bool ring16_job(int b, int J, int jb, int K, int count, int *pi,
                                           int *k) {
  const int G = q_group(K), nG = q_groups(K), U = pow2_ceil(G);
  const int x = b & 7, s = b >> 3;
  int sg, u;  // unit index on XCD x, job in unit
  if (J <= U) {
    const int BU = U / J;  // blocks per unit, consecutive on XCD x
    sg = s / BU;
    u = (s - sg * BU) * J + jb;
  } else {
    const int UB = J / U;  // units per block
    sg = s * UB + jb / U;
    u = jb & (U - 1);
  }
  const int q = sg / nG;
  *pi = q * 8 + x;
  *k = (sg - q * nG) * G + u;
  return *pi < count && u < G && *k < K;
}

// Entry e (< 8) of the complex values a thread of a cap_seq job parks in
// its own map column `col` (stride K doubles) of a ring's pixel rows: an FP64
// value in rows (2 e) T + tt and (2 e + 1) T + tt, an FP32 one in row
// e T + tt, so consecutive threads use consecutive rows.  Rows stay below
// 16 T = 8192, and the rings of cap_seq jobs (Bluestein halves of 8192
// points: i = n/4 >= 2049) have n >= 8196.
template <int T, typename C>
__device__ __forceinline__ void park_put(double *col, int K, int tt, int e, C v) {
  if constexpr (sizeof(C) == 16) {
    col[(int64_t)((2 * e) * T + tt) * K] = v.x;
    col[(int64_t)((2 * e + 1) * T + tt) * K] = v.y;
  } else {
    *reinterpret_cast<float2 *>(col + (int64_t)(e * T + tt) * K) = make_float2(v.x, v.y);
  }
}
template <int T, typename C>
__device__ __forceinline__ C park_get(const double *col, int K, int tt, int e) {
  if constexpr (sizeof(C) == 16) {
    return cx<C>::mk(col[(int64_t)((2 * e) * T + tt) * K], col[(int64_t)((2 * e + 1) * T + tt) * K]);
  } else {
    const float2 f = *reinterpret_cast<const float2 *>(col + (int64_t)(e * T + tt) * K);
    return cx<C>::mk(f.x, f.y);
  }
}

// Cap and wrapping-ring jobs whose 2T threads would exceed a block (P =
// 8192): one T-thread job per block runs the two halves in turn through one
// buffer (cap_block_seq).
__host__ __device__ constexpr bool cap_seq(int kind, int P) {
  return kind != kRingBelt16 && P / 8 > RB;
}

// Blocks of a launch: belt jobs are T = P/16 threads, cap jobs 2T threads
// (T for cap_seq), all in 512-thread blocks.
__host__ __device__ constexpr int ring16_jobs(int kind, int P) {
  return kind == kRingBelt16 ? RB / (P / 16) : cap_seq(kind, P) ? 1 : RB / (P / 8);
}

int64_t ring16_blocks(int kind, int P, int K, int count) {
  const int J = ring16_jobs(kind, P);
  const int G = q_group(K), U = pow2_ceil(G);
  const int64_t ux = (int64_t)(count + 7) / 8 * q_groups(K);  // units per XCD
  if (J <= U) return 8 * ux * (U / J);
  const int UB = J / U;
  return 8 * ((ux + UB - 1) / UB);
}


// LDS of one block: the twiddle table (double2) first, then the jobs'
// buffers of cs-byte points
// buffers per job: FP64 belt 1, FP64 cap 2 (A, B); FP32 twice as many, so
// that both rings of a pair are transformed together (same bytes); a
// cap_seq job 1 in either precision
__host__ __device__ constexpr int ring16_bufs(int kind, int P, int cs) {
  return cap_seq(kind, P) ? 1 : (kind == kRingBelt16 ? 1 : 2) * (cs == 8 ? 2 : 1);
}
__host__ __device__ constexpr size_t ring16_bytes(int kind, int P, int cs) {
  return (size_t)kTwSlots * sizeof(double2) +
         (size_t)ring16_bufs(kind, P, cs) * ring16_jobs(kind, P) * job_slots(P) * cs;
}
int ring16_lds_slots(int kind, int P) {
  return (int)(ring16_bytes(kind, P, sizeof(double2)) / sizeof(double2));
}


// Z_a = (X_a + conj X_c) + i e^{2 pi i a/n} (X_a - conj X_c), c = h - a
template <typename C>
__device__ __forceinline__ C zpair(C xa, C xc, double2 om) {
  const C cc = cconj(xc);
  return cadd(cadd(xa, cc), cmul_i(cmul(cx<C>::of(om), csub(xa, cc))));
}

// pixels 2p, 2p + 1 of one map of the (npix, K) map
template <typename C>
__device__ __forceinline__ void put2(double *out, int p, C z, int K) {
  if (K == 1) {
    *reinterpret_cast<double2 *>(out + 2 * p) = to_d(z);
  } else {
    out[(int64_t)(2 * p) * K] = z.x;
    out[(int64_t)(2 * p + 1) * K] = z.y;
  }
}

// Lanes 32 .. 63 of x trade places with lanes 0 .. 31 of y.
__device__ __forceinline__ void swap_halves(double &x, double &y) {
  const unsigned long long xb = __builtin_bit_cast(unsigned long long, x);
  const unsigned long long yb = __builtin_bit_cast(unsigned long long, y);
  const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)xb, (unsigned)yb, false, false);
  const auto hi =
      __builtin_amdgcn_permlane32_swap((unsigned)(xb >> 32), (unsigned)(yb >> 32), false, false);
  x = __builtin_bit_cast(double, (unsigned long long)lo[0] | (unsigned long long)hi[0] << 32);
  y = __builtin_bit_cast(double, (unsigned long long)lo[1] | (unsigned long long)hi[1] << 32);
}

// pixels 2p, 2p + 1 (p = j + T rho) of the 16 values of one thread
template <int T, typename C>
__device__ __forceinline__ void put16(double *out, int j, const C v[16], int K) {
  if (K == 1) {
    double2 *o = reinterpret_cast<double2 *>(out) + j;
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) o[T * rho] = to_d(v[rho]);
  } else if (T >= 64 && 2 * T <= RB && (K & 1) == 0) {
    // paired jobs (job_lane): lane l and lane l + 32 hold the same pixels of
    // maps k (even) and k + 1 (out = map + k resp. map + k + 1); after the
    // exchange lane l holds pixel 2p of both maps and lane l + 32 pixel
    // 2p + 1, one 16-byte store each.  Every lane of the wave takes part
    // (both jobs of a pair are live or padding together when K is even).
    const int hi = (threadIdx.x >> 5) & 1;
    double *o = out - hi + (int64_t)(2 * j + hi) * K;
    const int64_t st = (int64_t)(2 * T) * K;
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      double x = v[rho].x, y = v[rho].y;
      swap_halves(x, y);
      *reinterpret_cast<double2 *>(o) = make_double2(x, y);
      o += st;
    }
  } else {
    double *o = out + (int64_t)(2 * j) * K;
    const int64_t st = (int64_t)(2 * T) * K;
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      o[0] = v[rho].x;
      o[K] = v[rho].y;
      o += st;
    }
  }
}

// Z pairs of the X values in buf (bins 0 .. h), the inverse FFT, the pixels.
template <int P, typename C>
__device__ __forceinline__ void belt_fft_out(C *buf, const double2 *twl, int j, int n,
                                             bool on, int K, double *__restrict__ out) {
  constexpr int T = P / 16, h = P;
  C v[16];
  {
    const double2 ostep = expipi(2 * T, n);
    double2 om = expipi(2 * j, n);
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int a = j + T * rho;
      v[rho] = zpair(buf[pad(a)], buf[pad(h - a)], om);
      om = cmul(om, ostep);
    }
  }
  fft16::fft<P, true>(v, buf, twl, j);
  if (on) put16<T>(out, j, v, K);
}

// Both rings of a belt pair, one map, by the T threads of a job: E and O
// are read once; X^N = w (E + O) goes through LDS while X^S = w (E - O)
// waits in registers.  The q rows past mmax are read (the q buffer has a
// tail of kQTail rows) and discarded, so that all loads are unconditional
// and in flight together.
template <int P, typename C>
__device__ __forceinline__ void belt_pair(const RingPair &rp, const C *__restrict__ qe,
                                          const C *__restrict__ qo, int64_t qs, int M,
                                          int K, C *buf, const double2 *twl, int j,
                                          bool act, bool south, double *__restrict__ outn,
                                          double *__restrict__ outs) {
  constexpr int T = P / 16, h = P;
  using X = cx<C>;
  const int n = rp.n;
  // 1. X_a = w_a (E_a +/- O_a), w_a = e^{i pi s a/n}, a = j + T rho; DC
  //    keeps its real part
  C xs[16];
  {
    const double2 wstep = expipi((int64_t)rp.s * T, n);
    double2 w = expipi((int64_t)rp.s * j, n);
    const C *pe = qe + j * qs, *po = qo + j * qs;
    const int64_t st = T * qs;
    C e[16], o[16];
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      e[rho] = pe[rho * st];
      o[rho] = po[rho * st];
    }
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      const int a = j + T * rho;
      const bool in = act && a <= M;
      C xn = in ? cmul(cadd(e[rho], o[rho]), X::of(w)) : X::mk(0.0, 0.0);
      xs[rho] = in ? cmul(csub(e[rho], o[rho]), X::of(w)) : X::mk(0.0, 0.0);
      if (a == 0) xn.y = xs[rho].y = 0.0;
      buf[pad(a)] = xn;
      w = cmul(w, wstep);
    }
  }
  // Nyquist bin h: e^{i pi s/2} (g_h + (-1)^s conj g_h), real part
  C xsh = X::mk(0.0, 0.0);
  if (j == 0) {
    C gn = X::mk(0.0, 0.0), gs = gn;
    if (act && h <= M) {
      const C e = qe[h * qs], o = qo[h * qs];
      gn = cadd(e, o);
      gs = csub(e, o);
    }
    const double sg = rp.s ? -1.0 : 1.0;
    const C wh = rp.s ? X::mk(0.0, 1.0) : X::mk(1.0, 0.0);
    C xn = cmul(X::mk(fma(sg, (double)gn.x, (double)gn.x), fma(-sg, (double)gn.y, (double)gn.y)), wh);
    xsh = cmul(X::mk(fma(sg, (double)gs.x, (double)gs.x), fma(-sg, (double)gs.y, (double)gs.y)), wh);
    xn.y = xsh.y = 0.0;
    buf[pad(h)] = xn;
  }
  if constexpr (sizeof(C) == 8) {
    // FP32: X^S into the job's second buffer, both rings transformed together
    C *bufs = buf + job_slots(P);
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) bufs[pad(j + T * rho)] = xs[rho];
    if (j == 0) bufs[pad(h)] = xsh;
    __syncthreads();
    C v[2][16];
    {
      const double2 ostep = expipi(2 * T, n);
      double2 om = expipi(2 * j, n);
#pragma unroll
      for (int rho = 0; rho < 16; ++rho) {
        const int a = j + T * rho;
        v[0][rho] = zpair(buf[pad(a)], buf[pad(h - a)], om);
        v[1][rho] = zpair(bufs[pad(a)], bufs[pad(h - a)], om);
        om = cmul(om, ostep);
      }
    }
    fft16::fft2<P, true>(v, buf, bufs, twl, j);
    if (act) put16<T>(outn, j, v[0], K);
    if (act && south) put16<T>(outs, j, v[1], K);
    return;
  }
  __syncthreads();
  // 2. northern ring
  belt_fft_out<P>(buf, twl, j, n, act, K, outn);
  // 3. southern ring: X^S through the same buffer
  __syncthreads();  // last FFT pass reads done
#pragma unroll
  for (int rho = 0; rho < 16; ++rho) buf[pad(j + T * rho)] = xs[rho];
  if (j == 0) buf[pad(h)] = xsh;
  __syncthreads();
  asm volatile("" : "+v"(j));  // no address of the northern pass kept live
  belt_fft_out<P>(buf, twl, j, n, act && south, K, outs);
}

// Job slot and thread-in-job of thread tid for T-thread jobs.  Jobs of
// T >= 64 threads go in pairs whose waves hold 32 lanes of each: lane l
// and lane l + 32 then handle the same positions of two adjacent maps,
// whose q rows share a line and whose pixels share a map row, so each
// load and store request serves both.  Smaller jobs sit side by side in a
// wave already; a job of a whole block (T = RB, P = 8192) has no partner.
template <int T>
__device__ __forceinline__ void job_lane(int tid, int *jb, int *j) {
  if constexpr (T >= 64 && 2 * T <= RB) {
    constexpr int WPP = T / 32;  // waves per job pair
    const int l = tid & 63, w = tid >> 6;
    const int pr = w / WPP, wi = w - pr * WPP;
    *jb = 2 * pr + (l >> 5);
    *j = wi * 32 + (l & 31);
  } else {
    *jb = tid / T;
    *j = tid % T;
  }
}

// Job slot and thread-in-job for the FP32 two-map belt jobs (T >= 64):
// the two jobs of a pair alternate lanes (lane 2 i job 2 pr, lane 2 i + 1
// job 2 pr + 1, both at j = wi 32 + i), so the maps k .. k + 1 and
// k + 2 .. k + 3 of a pixel row are written by adjacent lanes of one store
// instruction (32 contiguous bytes, one request where the 16-byte stores of
// lanes l and l + 32 went out separately) and their q entries are read by
// adjacent lanes too.  Jobs of fewer than 64 threads keep job_lane's
// layout (side by side in a wave).
template <int T>
__device__ __forceinline__ void job_lane_il(int tid, int *jb, int *j) {
  if constexpr (T >= 64 && 2 * T <= RB) {
    constexpr int WPP = T / 32;  // waves per job pair
    const int l = tid & 63, w = tid >> 6;
    const int pr = w / WPP, wi = w - pr * WPP;
    *jb = 2 * pr + (l & 1);
    *j = wi * 32 + (l >> 1);
  } else {
    job_lane<T>(tid, jb, j);
  }
}

// pixels 2p, 2p + 1 (p = j + T rho) of maps k and (two) k + 1 of one ring;
// out = map + k + ring offset.  With K even the two maps of a pixel row are
// one aligned 16-byte store.
template <int T, typename C>
__device__ __forceinline__ void put16x2(double *out, int j, const C v0[16], const C v1[16],
                                        int K, bool two) {
  double *o = out + (int64_t)(2 * j) * K;
  const int64_t st = (int64_t)(2 * T) * K;
  if (two && (K & 1) == 0) {
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      *reinterpret_cast<double2 *>(o) = make_double2(v0[rho].x, v1[rho].x);
      *reinterpret_cast<double2 *>(o + K) = make_double2(v0[rho].y, v1[rho].y);
      o += st;
    }
  } else {
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) {
      o[0] = v0[rho].x;
      o[K] = v0[rho].y;
      if (two) {
        o[1] = v1[rho].x;
        o[K + 1] = v1[rho].y;
      }
      o += st;
    }
  }
}

// FP32 belt at K >= 2: a job is (ring pair, maps k = 2 kp and k + 1); the two
// maps of a ring are transformed together (fft2), the northern ring first,
// while X^S of both maps waits in registers.  Paired jobs (maps k .. k + 3)
// share each wave in alternate lanes (job_lane_il), so a store instruction
// writes 32 contiguous bytes of each of its pixel rows.  The second map's q
// is the next entry of the q group (k is
// even; past the last map it reads the group's padding or the q tail).
template <int P, typename C>
__device__ __forceinline__ void belt_block_maps2(const LegTables &t, const RingTables &rt,
                                                 int first, int count,
                                                 double *__restrict__ map, int b, C *lds,
                                                 const double2 *twl) {
  constexpr int T = P / 16, J = RB / T, S = job_slots(P), h = P;
  using X = cx<C>;
  int jb, j;
  job_lane_il<T>(threadIdx.x, &jb, &j);
  C *buf0 = lds + jb * S * ring16_bufs(kRingBelt16, P, sizeof(C)), *buf1 = buf0 + S;
  const int K = t.K, M = t.M;
  int pi, kp;
  const bool act = ring16_job(b, J, jb, (K + 1) / 2, count, &pi, &kp);
  if (!act) {
    pi = 0;
    kp = 0;
  }
  const int k = 2 * kp;
  const bool two = act && k + 1 < K;
  const RingPair rp = rt.pairs[rt.order[first + pi]];
  const bool south = rp.pix_s >= 0;
  const int n = rp.n;
  const C *qe = (const C *)t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  const C *qo = (const C *)t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
  // 1. X_a = w_a (E_a +/- O_a) of both maps; northern into buf0 / buf1
  C xs0[16], xs1[16];
  {
    const double2 wstep = expipi((int64_t)rp.s * T, n);
    double2 w = expipi((int64_t)rp.s * j, n);
    const C *pe = qe + j * qs, *po = qo + j * qs;
    const int64_t st = T * qs;
#pragma unroll
    for (int h8 = 0; h8 < 16; h8 += 8) {
      C e0[8], e1[8], o0[8], o1[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int64_t at = (h8 + q) * st;
        e0[q] = pe[at];
        e1[q] = pe[at + 1];
        o0[q] = po[at];
        o1[q] = po[at + 1];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int rho = h8 + q, a = j + T * rho;
        const bool in = act && a <= M, in1 = two && a <= M;
        const C wc = X::of(w);
        C xn0 = in ? cmul(cadd(e0[q], o0[q]), wc) : X::mk(0.0, 0.0);
        C xn1 = in1 ? cmul(cadd(e1[q], o1[q]), wc) : X::mk(0.0, 0.0);
        xs0[rho] = in ? cmul(csub(e0[q], o0[q]), wc) : X::mk(0.0, 0.0);
        xs1[rho] = in1 ? cmul(csub(e1[q], o1[q]), wc) : X::mk(0.0, 0.0);
        if (a == 0) xn0.y = xn1.y = xs0[rho].y = xs1[rho].y = 0.0;
        buf0[pad(a)] = xn0;
        buf1[pad(a)] = xn1;
        w = cmul(w, wstep);
      }
    }
  }
  // Nyquist bin h: e^{i pi s/2} (g_h + (-1)^s conj g_h), real part
  C xsh0 = X::mk(0.0, 0.0), xsh1 = xsh0;
  if (j == 0) {
    const double sg = rp.s ? -1.0 : 1.0;
    const C wh = rp.s ? X::mk(0.0, 1.0) : X::mk(1.0, 0.0);
    auto nyq = [&](C g) {
      C x = cmul(X::mk(fma(sg, (double)g.x, (double)g.x), fma(-sg, (double)g.y, (double)g.y)), wh);
      x.y = 0.0;
      return x;
    };
    C e0 = X::mk(0.0, 0.0), e1 = e0, o0 = e0, o1 = e0;
    if (act && h <= M) {
      e0 = qe[h * qs];
      o0 = qo[h * qs];
      if (two) {
        e1 = qe[h * qs + 1];
        o1 = qo[h * qs + 1];
      }
    }
    buf0[pad(h)] = nyq(cadd(e0, o0));
    buf1[pad(h)] = nyq(cadd(e1, o1));
    xsh0 = nyq(csub(e0, o0));
    xsh1 = nyq(csub(e1, o1));
  }
  double *outn = map + k + (int64_t)rp.pix_n * K;
  double *outs = map + k + (int64_t)(south ? rp.pix_s : 0) * K;
#pragma unroll 1
  for (int ring = 0; ring < 2; ++ring) {
    if (ring == 1) {
      // southern ring: X^S of both maps through the same buffers
      __syncthreads();  // last FFT pass reads done
#pragma unroll
      for (int rho = 0; rho < 16; ++rho) {
        buf0[pad(j + T * rho)] = xs0[rho];
        buf1[pad(j + T * rho)] = xs1[rho];
      }
      if (j == 0) {
        buf0[pad(h)] = xsh0;
        buf1[pad(h)] = xsh1;
      }
    }
    __syncthreads();
    C v[2][16];
    {
      const double2 ostep = expipi(2 * T, n);
      double2 om = expipi(2 * j, n);
#pragma unroll
      for (int rho = 0; rho < 16; ++rho) {
        const int a = j + T * rho;
        v[0][rho] = zpair(buf0[pad(a)], buf0[pad(h - a)], om);
        v[1][rho] = zpair(buf1[pad(a)], buf1[pad(h - a)], om);
        om = cmul(om, ostep);
      }
    }
    fft16::fft2<P, true>(v, buf0, buf1, twl, j);
    if (act && (ring == 0 || south)) put16x2<T>(ring ? outs : outn, j, v[0], v[1], K, two);
  }
}

// Block b of a belt launch over pairs order[first .. first + count).
template <int P, typename C>
__device__ __forceinline__ void belt_block(const LegTables &t, const RingTables &rt, int first,
                                           int count, double *__restrict__ map, int b,
                                           C *lds, const double2 *twl) {
  constexpr int T = P / 16, J = RB / T, S = job_slots(P);
  if constexpr (sizeof(C) == 8) {
    if (t.K >= 2) {  // block-uniform
      belt_block_maps2<P>(t, rt, first, count, map, b, lds, twl);
      return;
    }
  }
  int jb, j;
  job_lane<T>(threadIdx.x, &jb, &j);
  C *buf = lds + jb * S * ring16_bufs(kRingBelt16, P, sizeof(C));
  int pi, k;
  const bool act = ring16_job(b, J, jb, t.K, count, &pi, &k);
  if (!act) {
    pi = 0;
    k = 0;
  }
  const RingPair rp = rt.pairs[rt.order[first + pi]];
  const int K = t.K, M = t.M;
  const bool south = rp.pix_s >= 0;
  const C *qe = (const C *)t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  const C *qo = (const C *)t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
  belt_pair<P>(rp, qe, qo, qs, M, K, buf, twl, j, act, south, map + k + rp.pix_n * K,
               map + k + (south ? rp.pix_s : 0) * K);
}

// Cap jobs are 2T threads per (ring pair, map), the first T threads
// transforming the even Z entries (A), the other T the odd ones (B), each
// in its own LDS buffer, so that no thread holds one transform's result
// through the other's.  Both rings of the pair come from one read of E and
// O; the Z entries of the southern ring wait in registers (8 per thread)
// while the northern ring is transformed.

// F_b = P_b + (-1)^s conj P_{n-b} (0 < b <= h), F_0 = P_0 + conj(P_0 - g0),
// in place over the P_rho (rho < n) in buf
template <int P, typename C>
__device__ __forceinline__ void wrap_fstep(C *buf, int n, int s, int tt, C g0) {
  using Rt = typename cx<C>::R;
  constexpr int T2 = P / 8;
  const int h = n / 2;
  const Rt sg = s ? -1 : 1;
  for (int b = tt; b <= h; b += T2) {
    const C pb = buf[pad(b)];
    const C pc = b == 0 ? csub(pb, g0) : buf[pad(n - b)];
    const Rt c = b == 0 ? 1 : sg;
    C f;
    f.x = fma(c, pc.x, pb.x);
    f.y = fma(-c, pc.y, pb.y);
    buf[pad(b)] = f;
  }
}

// Fold of a ring pair whose m-range wraps (n <= mmax), into F[0 .. h] at
// buf[pad(.)] (buf spans the job's two buffers, >= n padded slots):
//   P_rho = sum_w (-1)^{s w} g_{w n + rho}  (rho < n, blocks of n rows),
//   F_b = P_b + (-1)^s conj P_{n-b}  (0 < b <= h),
//   F_0 = P_0 + conj(P_0 - g_0),
// the direct terms m = b + w n and conjugate terms m = w n - b of every bin
// regrouped by row block.  Thread tt owns bins rho = tt + 2T q (q < 16;
// n < 2P = 16 (2T)) and sums their row blocks in order in registers, for
// g = E + O (north, folded into buf here) and g = E - O (south, returned
// in ps[] and *g0s for wrap_fold_south).  Rows of bins rho >= n are not
// read: those lanes load the row of bin tt again.
template <int P, typename C>
__device__ __forceinline__ void wrap_fold2(const C *__restrict__ qe, const C *__restrict__ qo,
                                           int64_t qs, int M, int n, int s, C *buf, int tt,
                                           bool on_n, bool on_s, C ps[16], C *g0s) {
  using Rt = typename cx<C>::R;
  using X = cx<C>;
  constexpr int T2 = P / 8;
  C an[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) an[q] = ps[q] = X::mk(0.0, 0.0);
  C g0n = X::mk(0.0, 0.0);
  *g0s = X::mk(0.0, 0.0);
  for (int w0 = 0; w0 <= M; w0 += n) {
    const Rt se = (s && ((w0 / n) & 1)) ? -1 : 1;
    const int64_t m0 = w0 + tt <= M ? w0 + tt : M;
#pragma unroll
    for (int h8 = 0; h8 < 16; h8 += 8) {
      C e[8], o[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int rho = tt + T2 * (h8 + q), m = w0 + rho;
        const int64_t mc = (rho < n && m <= M) ? m : m0;
        e[q] = qe[mc * qs];
        o[q] = qo[mc * qs];
      }
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int rho = tt + T2 * (h8 + q), m = w0 + rho;
        if (rho < n && m <= M) {
          const Rt ex = se * e[q].x, ey = se * e[q].y;
          an[h8 + q].x += fma(se, o[q].x, ex);
          an[h8 + q].y += fma(se, o[q].y, ey);
          ps[h8 + q].x += fma(-se, o[q].x, ex);
          ps[h8 + q].y += fma(-se, o[q].y, ey);
        }
      }
      if (h8 == 0 && w0 == 0 && tt == 0) {
        g0n = cadd(e[0], o[0]);
        *g0s = csub(e[0], o[0]);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int rho = tt + T2 * q;
    if (rho < n) buf[pad(rho)] = on_n ? an[q] : X::mk(0.0, 0.0);
    if (!on_s) ps[q] = X::mk(0.0, 0.0);
  }
  __syncthreads();
  wrap_fstep<P>(buf, n, s, tt, on_n ? g0n : X::mk(0.0, 0.0));
  __syncthreads();
}

// The southern fold of wrap_fold2 into buf.  The caller has passed a
// barrier after its last read of buf.
template <int P, typename C>
__device__ __forceinline__ void wrap_fold_south(C *buf, int n, int s, int tt, const C ps[16],
                                                C g0s) {
  constexpr int T2 = P / 8;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    const int rho = tt + T2 * q;
    if (rho < n) buf[pad(rho)] = ps[q];
  }
  __syncthreads();
  wrap_fstep<P>(buf, n, s, tt, g0s);
  __syncthreads();
}

// Fold of a cap ring pair without wrap (n > mmax): F_b = g_b
// (b <= min(h, mmax)), then F_{n-m} += (-1)^s conj g_m for h <= m <= mmax
// (one writer per bin in each pass); h < P and n > mmax give mmax - h < h,
// so 8 rows per thread cover each pass.  g = E + O (north) into bufN,
// g = E - O (south) into bufS.
template <int P, typename C>
__device__ __forceinline__ void cap_fold2(const C *__restrict__ qe, const C *__restrict__ qo,
                                          int64_t qs, int M, int n, int s, C *bufN, C *bufS,
                                          int tt, bool on_n, bool on_s) {
  using Rt = typename cx<C>::R;
  constexpr int T2 = P / 8;
  const int h = n / 2;
  const Rt sg = s ? -1 : 1;
  C e[16], o[16];
#pragma unroll
  for (int rho = 0; rho < 8; ++rho) {
    const int b = tt + T2 * rho, m = h + b;
    const int64_t bc = b <= M ? b : M, mc = m <= M ? m : M;
    e[rho] = qe[bc * qs];
    o[rho] = qo[bc * qs];
    e[8 + rho] = qe[mc * qs];
    o[8 + rho] = qo[mc * qs];
  }
  using X = cx<C>;
  C cn[8], cs[8];
#pragma unroll
  for (int rho = 0; rho < 8; ++rho) {
    const int b = tt + T2 * rho, m = h + b;
    const bool db = b <= M, dm = m <= M;
    if (b <= h) {
      bufN[pad(b)] = (on_n && db) ? cadd(e[rho], o[rho]) : X::mk(0.0, 0.0);
      bufS[pad(b)] = (on_s && db) ? csub(e[rho], o[rho]) : X::mk(0.0, 0.0);
    }
    const C c = cadd(e[8 + rho], o[8 + rho]), d = csub(e[8 + rho], o[8 + rho]);
    C cc, dd;
    cc.x = sg * c.x;
    cc.y = -sg * c.y;
    dd.x = sg * d.x;
    dd.y = -sg * d.y;
    cn[rho] = (on_n && dm) ? cc : X::mk(0.0, 0.0);
    cs[rho] = (on_s && dm) ? dd : X::mk(0.0, 0.0);
  }
  __syncthreads();
#pragma unroll
  for (int rho = 0; rho < 8; ++rho) {
    const int m = h + tt + T2 * rho;
    if (m <= M) {
      C &fn = bufN[pad(n - m)];
      fn = cadd(fn, cn[rho]);
      C &fs = bufS[pad(n - m)];
      fs = cadd(fs, cs[rho]);
    }
  }
  __syncthreads();
}

// Z entries of one cap ring from its fold F[0 .. h] in bufF: entry
// b = 2 kappa + half (kappa = j + T rho < i < P/2, so rho < 8) times the
// chirp c_kappa; phase w_b = e^{i pi s b/n}, partner phase
// e^{i pi s/2} conj(w_b), twiddle e^{2 pi i b/n}; DC / Nyquist real.
template <int P, typename C>
__device__ __forceinline__ void cap_z(const RingPair &rp, const double2 *__restrict__ chirp,
                                      const C *bufF, int tt, C z[8]) {
  using X = cx<C>;
  constexpr int T = P / 16;
  const int n = rp.n, h = n / 2, i = h / 2;
  const int half = tt / T, j = tt - half * T;
  const double2 wh = rp.s ? make_double2(0.0, 1.0) : make_double2(1.0, 0.0);
  const double2 wstep = expipi((int64_t)rp.s * 2 * T, n), ostep = expipi(4 * T, n);
  double2 w = expipi((int64_t)rp.s * (2 * j + half), n), om = expipi(2 * (2 * j + half), n);
#pragma unroll
  for (int rho = 0; rho < 8; ++rho) {
    const int kap = j + T * rho;
    const bool in = kap < i;
    const C c = X::of(chirp[in ? kap : 0]);
    const int b = in ? 2 * kap + half : 0;
    C xa = cmul(bufF[pad(b)], X::of(w));
    C xc = cmul(bufF[pad(h - b)], X::of(b == 0 ? wh : cmul(wh, cconj(w))));
    if (b == 0) xa.y = xc.y = 0;
    z[rho] = in ? cmul(zpair(xa, xc, om), c) : X::mk(0.0, 0.0);
    w = cmul(w, wstep);
    om = cmul(om, ostep);
  }
}

// One cap ring of one map from its Z entries: the two Bluestein
// convolutions, the pixels.  tt = thread in the job (0 .. 2T), bufA / bufB
// the job's two buffers (free: the transform starts with a barrier).
template <int P, typename C>
__device__ __forceinline__ void cap_xform(const RingPair &rp, int K,
                                          const double2 *__restrict__ chirp,
                                          const double2 *__restrict__ vhat, const C z[8],
                                          C *bufA, C *bufB, const double2 *twl, int tt, bool on,
                                          double *__restrict__ out) {
  using X = cx<C>;
  constexpr int T = P / 16, T2 = 2 * T;
  const int n = rp.n, h = n / 2, i = h / 2;
  const int half = tt / T, j = tt - half * T;
  C v[16];
#pragma unroll
  for (int rho = 0; rho < 8; ++rho) {
    v[rho] = z[rho];
    v[8 + rho] = X::mk(0.0, 0.0);
  }
  // Bluestein: y = IFFT(vhat . FFT(c . Z)) in the half's own buffer
  C *buf = half ? bufB : bufA;
  fft16::fft<P, false>(v, buf, twl, j);
  asm volatile("" : "+v"(tt));
#pragma unroll
  for (int rho = 0; rho < 16; ++rho) v[rho] = cmul(v[rho], X::of(vhat[j + T * rho]));
  fft16::fft<P, true>(v, buf, twl, j);
  asm volatile("" : "+v"(tt));
  // A_p = c_p y_p into bufA, e^{2 pi i p/h} c_p y_p (B) into bufB, p < i
  __syncthreads();  // last-pass reads of buf done
  {
    const double2 ostep = expipi(4 * T, n);
    double2 om = half ? expipi(4 * j, n) : make_double2(1.0, 0.0);
    const double2 st = half ? ostep : make_double2(1.0, 0.0);
#pragma unroll
    for (int rho = 0; rho < 8; ++rho) {
      const int p = j + T * rho;
      if (p < i) buf[pad(p)] = cmul(v[rho], X::of(cmul(chirp[p], om)));
      om = cmul(om, st);
    }
  }
  __syncthreads();
  // z_p = A_p + B_p, z_{p+i} = A_p - B_p
  if (on) {
#pragma unroll
    for (int rho = 0; rho < 4; ++rho) {
      const int p = tt + T2 * rho;
      if (p < i) {
        const C a = bufA[pad(p)], b = bufB[pad(p)];
        put2(out, p, cadd(a, b), K);
        put2(out, p + i, csub(a, b), K);
      }
    }
  }
}

// Both cap rings of a pair from their Z entries together (FP32): half 0
// transforms the A (even) entries of both rings in AN, AS, half 1 the B
// entries in BN, BS; then (own) the pixels of both rings.  Without own the
// results stay in the buffers for cap_store_block.
template <int P, typename C>
__device__ __forceinline__ void cap_xform2(const RingPair &rp, int K,
                                           const double2 *__restrict__ chirp,
                                           const double2 *__restrict__ vhat, const C zn[8],
                                           const C zs[8], C *AN, C *BN, C *AS, C *BS,
                                           const double2 *twl, int tt, bool onn, bool ons,
                                           bool own, double *__restrict__ outn,
                                           double *__restrict__ outs) {
  using X = cx<C>;
  constexpr int T = P / 16, T2 = 2 * T;
  const int n = rp.n, h = n / 2, i = h / 2;
  const int half = tt / T, j = tt - half * T;
  C v[2][16];
#pragma unroll
  for (int rho = 0; rho < 8; ++rho) {
    v[0][rho] = zn[rho];
    v[1][rho] = zs[rho];
    v[0][8 + rho] = v[1][8 + rho] = X::mk(0.0, 0.0);
  }
  C *b0 = half ? BN : AN, *b1 = half ? BS : AS;
  fft16::fft2<P, false>(v, b0, b1, twl, j);
  asm volatile("" : "+v"(tt));
#pragma unroll
  for (int rho = 0; rho < 16; ++rho) {
    const C vh = X::of(vhat[j + T * rho]);
    v[0][rho] = cmul(v[0][rho], vh);
    v[1][rho] = cmul(v[1][rho], vh);
  }
  fft16::fft2<P, true>(v, b0, b1, twl, j);
  asm volatile("" : "+v"(tt));
  __syncthreads();  // last-pass reads of the buffers done
  {
    const double2 ostep = expipi(4 * T, n);
    double2 om = half ? expipi(4 * j, n) : make_double2(1.0, 0.0);
    const double2 st = half ? ostep : make_double2(1.0, 0.0);
#pragma unroll
    for (int rho = 0; rho < 8; ++rho) {
      const int p = j + T * rho;
      if (p < i) {
        const C f = X::of(cmul(chirp[p], om));
        b0[pad(p)] = cmul(v[0][rho], f);
        b1[pad(p)] = cmul(v[1][rho], f);
      }
      om = cmul(om, st);
    }
  }
  __syncthreads();
  if (!own) return;
#pragma unroll
  for (int rho = 0; rho < 4; ++rho) {
    const int p = tt + T2 * rho;
    if (p < i) {
      if (onn) {
        const C a = AN[pad(p)], b = BN[pad(p)];
        put2(outn, p, cadd(a, b), K);
        put2(outn, p + i, csub(a, b), K);
      }
      if (ons) {
        const C a = AS[pad(p)], b = BS[pad(p)];
        put2(outs, p, cadd(a, b), K);
        put2(outs, p + i, csub(a, b), K);
      }
    }
  }
}

// Pixels of both rings of every job of block b, stored by the whole block
// (FP32 caps with 2 <= J <= U).  The J jobs of such a block are maps
// k0 .. k0 + J - 1 of one pair (a job is padding only if every later job
// is), so with the map fastest over the lanes each pixel row gets J
// adjacent doubles per request instead of one per lane.  The caller has
// passed the barrier after the buffers' last write.
template <int P, typename C>
__device__ __forceinline__ void cap_store_block(const LegTables &t, const RingTables &rt,
                                                int first, int count, double *__restrict__ map,
                                                int b, const C *lds) {
  constexpr int T = P / 16, J = RB / (2 * T), S = job_slots(P), NP = RB / J;
  const int K = t.K;
  int pi0, k0;
  if (!ring16_job(b, J, 0, K, count, &pi0, &k0)) return;
  const int jj = threadIdx.x % J, pp = threadIdx.x / J;
  int pj, kj;
  if (!ring16_job(b, J, jj, K, count, &pj, &kj)) return;
  const RingPair rp = rt.pairs[rt.order[first + pi0]];
  const int i = rp.n / 4;
  const bool south = rp.pix_s >= 0;
  const C *AN = lds + ring16_bufs(kRingCap16, P, sizeof(C)) * jj * S, *BN = AN + S;
  const C *AS = BN + S, *BS = AS + S;
  double *outn = map + kj + (int64_t)rp.pix_n * K;
  double *outs = map + kj + (int64_t)(south ? rp.pix_s : 0) * K;
#pragma unroll
  for (int rho = 0; rho < 4; ++rho) {  // i <= P/2 = 4 NP
    const int p = pp + NP * rho;
    if (p < i) {
      const C a = AN[pad(p)], bb = BN[pad(p)];
      put2(outn, p, cadd(a, bb), K);
      put2(outn, p + i, csub(a, bb), K);
      if (south) {
        const C c = AS[pad(p)], d = BS[pad(p)];
        put2(outs, p, cadd(c, d), K);
        put2(outs, p + i, csub(c, d), K);
      }
    }
  }
}

// pixels 2p, 2p + 1 of maps k (za) and, if two, k + 1 (zb); out = map + k +
// ring offset.  With K even the two maps of a row are one 16-byte store.
template <typename C>
__device__ __forceinline__ void put2x2(double *out, int p, C za, C zb, int K, bool two) {
  double *o = out + (int64_t)(2 * p) * K;
  if (two && (K & 1) == 0) {
    *reinterpret_cast<double2 *>(o) = make_double2(za.x, zb.x);
    *reinterpret_cast<double2 *>(o + K) = make_double2(za.y, zb.y);
  } else {
    o[0] = za.x;
    o[K] = za.y;
    if (two) {
      o[1] = zb.x;
      o[K + 1] = zb.y;
    }
  }
}

// FP32 polar caps without wrap at P = 4096 (one job per block) and K >= 2:
// a job is (ring pair, maps k = 2 kp and k + 1).  Per ring (north, then
// south): both maps are folded into the four buffers (north k, south k,
// north k + 1, south k + 1), the ring's Z entries of both maps are formed,
// the two maps are transformed together (cap_xform2 with the maps in place
// of the rings) and stored, the two maps of a pixel row in one 16-byte
// store.  The second map's q is the next entry of the q group, as in
// belt_block_maps2.
template <int P, typename C>
__device__ __forceinline__ void cap_block_maps2(const LegTables &t, const RingTables &rt,
                                                int first, int count,
                                                double *__restrict__ map, int b, C *lds,
                                                const double2 *twl) {
  constexpr int T = P / 16, T2 = 2 * T, S = job_slots(P);
  static_assert(RB == T2, "one cap job per block");
  int tt = threadIdx.x;
  C *b0 = lds, *b1 = b0 + S, *b2 = b1 + S, *b3 = b2 + S;
  const int K = t.K, M = t.M;
  int pi, kp;
  const bool act = ring16_job(b, 1, 0, (K + 1) / 2, count, &pi, &kp);
  if (!act) {
    pi = 0;
    kp = 0;
  }
  const int k = 2 * kp;
  const bool two = act && k + 1 < K;
  const RingPair rp = rt.pairs[rt.order[first + pi]];
  const bool south = rp.pix_s >= 0;
  const int n = rp.n, i = n / 4;
  const C *qe = (const C *)t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  const C *qo = (const C *)t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
  const double2 *chirp = rt.chirp2 + rp.chirp2_off;  // [i]
  const double2 *vhat = rt.vhat2 + rp.vhat2_off;     // [P]
  double *outn = map + k + (int64_t)rp.pix_n * K;
  double *outs = map + k + (int64_t)(south ? rp.pix_s : 0) * K;
  // With two maps, the southern ring's Z entries are formed from the first
  // fold and parked in the southern ring's own pixel rows 8 tt .. 8 tt + 7
  // (maps k and k + 1, one float2 each; n > 4096 rows), which only this
  // thread reads back and which the southern pass's pixel stores overwrite
  // after the transform's barriers.  A one-map job folds again from q.
  static_assert(sizeof(C) == 8, "parked Z entries are FP32");
  const bool park = two && south;  // block-uniform (one job per block)
  float2 *parked = reinterpret_cast<float2 *>(outs + (int64_t)(8 * tt) * K);
#pragma unroll 1
  for (int ring = 0; ring < 2; ++ring) {
    if (ring == 1) {
      if (!south) break;  // block-uniform
      __syncthreads();    // pixel reads of the northern ring done
    }
    C za[8], zb[8];
    if (ring == 0 || !park) {
      // the folds of both rings and maps
      cap_fold2<P>(qe, qo, qs, M, n, rp.s, b0, b1, tt, act, act && south);
      cap_fold2<P>(qe + 1, qo + 1, qs, M, n, rp.s, b2, b3, tt, two, two && south);
      cap_z<P>(rp, chirp, ring ? b1 : b0, tt, za);
      cap_z<P>(rp, chirp, ring ? b3 : b2, tt, zb);
      if (ring == 0 && park) {
        C sa[8], sb[8];
        cap_z<P>(rp, chirp, b1, tt, sa);
        cap_z<P>(rp, chirp, b3, tt, sb);
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          float2 *row = reinterpret_cast<float2 *>(reinterpret_cast<double *>(parked) +
                                                   (int64_t)e * K);
          row[0] = make_float2((float)sa[e].x, (float)sa[e].y);
          row[1] = make_float2((float)sb[e].x, (float)sb[e].y);
        }
      }
    } else {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const float2 *row = reinterpret_cast<const float2 *>(
            reinterpret_cast<const double *>(parked) + (int64_t)e * K);
        const float2 a = row[0], b = row[1];
        za[e] = cx<C>::mk(a.x, a.y);
        zb[e] = cx<C>::mk(b.x, b.y);
      }
    }
    __syncthreads();  // fold reads done before the transforms write the buffers
    asm volatile("" : "+v"(tt));
    cap_xform2<P>(rp, K, chirp, vhat, za, zb, b0, b1, b2, b3, twl, tt, false, false, false,
                  nullptr, nullptr);
    if (act) {
      double *out = ring ? outs : outn;
#pragma unroll
      for (int rho = 0; rho < 4; ++rho) {
        const int p = tt + T2 * rho;
        if (p < i) {
          const C a0 = b0[pad(p)], c0 = b1[pad(p)], a1 = b2[pad(p)], c1 = b3[pad(p)];
          put2x2(out, p, cadd(a0, c0), cadd(a1, c1), K, two);
          put2x2(out, p + i, csub(a0, c0), csub(a1, c1), K, two);
        }
      }
    }
  }
}

// g_m = E_m + O_m (north) or E_m - O_m (south)
template <typename C>
__device__ __forceinline__ C gring(C e, C o, bool south) {
  return south ? csub(e, o) : cadd(e, o);
}

// One ring's fold without wrap (n > mmax) by the T = P/16 threads of a
// cap_seq job: F_b = g_b (b <= min(h, mmax)), then F_{n-m} += (-1)^s
// conj g_m for h <= m <= mmax (distinct bins n - m, one writer each).  h < P
// and mmax - h < h, so bins b = tt + T rho, rho < 16, cover both passes,
// which run 8 rows per thread at a time.  Ends with a barrier.
template <int P, typename C>
__device__ __forceinline__ void cap_fold_seq(const C *__restrict__ qe, const C *__restrict__ qo,
                                             int64_t qs, int M, int n, int s, bool south,
                                             C *buf, int tt, bool on) {
  using Rt = typename cx<C>::R;
  using X = cx<C>;
  constexpr int T = P / 16;
  const int h = n / 2;
  const Rt sg = s ? -1 : 1;
#pragma unroll
  for (int r8 = 0; r8 < 16; r8 += 8) {
    C e[8], o[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int b = tt + T * (r8 + q);
      const int64_t bc = b <= M ? b : M;
      e[q] = qe[bc * qs];
      o[q] = qo[bc * qs];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int b = tt + T * (r8 + q);
      if (b <= h) buf[pad(b)] = (on && b <= M) ? gring(e[q], o[q], south) : X::mk(0.0, 0.0);
    }
  }
  __syncthreads();
#pragma unroll
  for (int r8 = 0; r8 < 16; r8 += 8) {
    C e[8], o[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int m = h + tt + T * (r8 + q);
      const int64_t mc = m <= M ? m : M;
      e[q] = qe[mc * qs];
      o[q] = qo[mc * qs];
    }
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int m = h + tt + T * (r8 + q);
      if (on && m <= M) {
        const C g = gring(e[q], o[q], south);
        C &f = buf[pad(n - m)];
        f.x = fma(sg, g.x, f.x);
        f.y = fma(-sg, g.y, f.y);
      }
    }
  }
  __syncthreads();
}

// NQ bins b = tt + T (q0 + q) of one wrapping ring's fold, summed over the
// row blocks w in order:
//   F_b = sum_w (-1)^{s w} (g_{w n + b} + (-1)^s conj g_{w n + n - b})  (0 < b <= h),
//   F_0 = sum_w (-1)^{s w} (g_{w n} + conj g_{w n}) - conj g_0,
// which regroups P_b + (-1)^s conj P_{n-b} and P_0 + conj(P_0 - g_0) of
// wrap_fold2 (P_rho = sum_w (-1)^{s w} g_{w n + rho}) by bin.
template <int P, int NQ, typename C>
__device__ __forceinline__ void wrap_bins_seq(const C *__restrict__ qe, const C *__restrict__ qo,
                                              int64_t qs, int M, int n, int s, bool south,
                                              C *buf, int tt, int q0, bool on) {
  using Rt = typename cx<C>::R;
  using X = cx<C>;
  constexpr int T = P / 16;
  const int h = n / 2;
  const Rt sp = s ? -1 : 1;
  C acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = X::mk(0.0, 0.0);
  for (int w0 = 0; w0 <= M; w0 += n) {
    const Rt se = (s && ((w0 / n) & 1)) ? -1 : 1;
    C e[NQ], o[NQ], ec[NQ], oc[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = tt + T * (q0 + q);
      const int m = w0 + b, mp = b == 0 ? w0 : w0 + n - b;
      const int64_t mc = m <= M ? m : M, mpc = mp <= M ? mp : M;
      e[q] = qe[mc * qs];
      o[q] = qo[mc * qs];
      ec[q] = qe[mpc * qs];
      oc[q] = qo[mpc * qs];
    }
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int b = tt + T * (q0 + q);
      const int m = w0 + b, mp = b == 0 ? w0 : w0 + n - b;
      if (b <= h && m <= M) {
        const C g = gring(e[q], o[q], south);
        acc[q].x = fma(se, g.x, acc[q].x);
        acc[q].y = fma(se, g.y, acc[q].y);
      }
      if (b <= h && mp <= M) {
        const C g = gring(ec[q], oc[q], south);
        const Rt c = b == 0 ? se : se * sp;
        acc[q].x = fma(c, g.x, acc[q].x);
        acc[q].y = fma(-c, g.y, acc[q].y);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) {
    const int b = tt + T * (q0 + q);
    if (b > h) continue;
    C f = on ? acc[q] : X::mk(0.0, 0.0);
    if (b == 0 && on) {
      const C g0 = gring(qe[0], qo[0], south);
      f.x -= g0.x;
      f.y += g0.y;
    }
    buf[pad(b)] = f;
  }
}

// One wrapping ring's fold (n <= mmax, n <= 2P) by the T = P/16 threads of
// a cap_seq job into buf[pad(b)], b = 0 .. h (h <= P: the bin b = P = 16 T,
// present when n = 2P, is the 17th of thread 0).  Ends with a barrier.
template <int P, typename C>
__device__ __forceinline__ void wrap_fold_seq(const C *__restrict__ qe, const C *__restrict__ qo,
                                              int64_t qs, int M, int n, int s, bool south,
                                              C *buf, int tt, bool on) {
  wrap_bins_seq<P, 8>(qe, qo, qs, M, n, s, south, buf, tt, 0, on);
  wrap_bins_seq<P, 8>(qe, qo, qs, M, n, s, south, buf, tt, 8, on);
  if (tt == 0 && n / 2 == P) wrap_bins_seq<P, 1>(qe, qo, qs, M, n, s, south, buf, tt, 16, on);
  __syncthreads();
}

// Cap and wrapping-ring block at P = 8192 (cap_seq): one job of T = P/16 =
// RB threads per block, one buffer.  Per ring (north, then south, each
// folding its own g from q, so the southern ring reads q again): the fold,
// the Z entries of both halves (A: even entries, B: odd), then the A
// Bluestein convolution, whose results c_p y_p stay in registers, then the
// B one through the same buffer; thread j holds A_p and B_p for the same
// p = j + T rho (rho < 8, p < i <= P/2) and writes z_p = A_p + B_p,
// z_{p+i} = A_p - B_p.
template <int P, bool WRAP, typename C>
__device__ __forceinline__ void cap_block_seq(const LegTables &t, const RingTables &rt,
                                              int first, int count, double *__restrict__ map,
                                              int b, C *buf, const double2 *twl) {
  constexpr int T = P / 16;
  static_assert(T == RB, "one job of a whole block");
  using X = cx<C>;
  int tt = threadIdx.x;
  const int K = t.K, M = t.M;
  int pi, k;
  const bool act = ring16_job(b, 1, 0, K, count, &pi, &k);
  if (!act) {
    pi = 0;
    k = 0;
  }
  const RingPair rp = rt.pairs[rt.order[first + pi]];
  const bool south = rp.pix_s >= 0;
  const int n = rp.n, i = n / 4;
  const C *qe = (const C *)t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  const C *qo = (const C *)t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
  const double2 *chirp = rt.chirp2 + rp.chirp2_off;  // [i]
  const double2 *vhat = rt.vhat2 + rp.vhat2_off;     // [P]
#pragma unroll 1
  for (int ring = 0; ring < 2; ++ring) {
    if (ring == 1 && !south) break;  // block-uniform
    __syncthreads();  // the previous ring's last reads of buf done
    if constexpr (WRAP)
      wrap_fold_seq<P>(qe, qo, qs, M, n, rp.s, ring == 1, buf, tt, act);
    else
      cap_fold_seq<P>(qe, qo, qs, M, n, rp.s, ring == 1, buf, tt, act);
    // The second half's Z entries wait through the first transform, and the
    // first half's result through the second, parked in this ring's own
    // pixel column (park_put: rows below 16 T <= n), which only this thread
    // reads back; the pixel stores come after a barrier that follows every
    // thread's last read-back.  An inactive job (past the last map) has no
    // column and keeps them in registers.
    double *out = map + k + (int64_t)(ring ? rp.pix_s : rp.pix_n) * K;
    C za[8], zb[8];
    cap_z<P>(rp, chirp, buf, tt, za);
    cap_z<P>(rp, chirp, buf, T + tt, zb);
    if (act) {
#pragma unroll
      for (int rho = 0; rho < 8; ++rho) park_put<T>(out, K, tt, rho, zb[rho]);
    }
    asm volatile("" : "+v"(tt));  // no table load hoisted above the barriers
    C v[16], a[8];
#pragma unroll
    for (int rho = 0; rho < 8; ++rho) {
      v[rho] = za[rho];
      v[8 + rho] = X::mk(0.0, 0.0);
    }
    fft16::fft<P, false>(v, buf, twl, tt);  // first barrier: cap_z reads done
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) v[rho] = cmul(v[rho], X::of(vhat[tt + T * rho]));
    fft16::fft<P, true>(v, buf, twl, tt);
#pragma unroll
    for (int rho = 0; rho < 8; ++rho) {
      const int p = tt + T * rho;
      a[rho] = cmul(v[rho], X::of(chirp[p < i ? p : 0]));
      if (act) {
        v[rho] = park_get<T, C>(out, K, tt, rho);
        park_put<T>(out, K, tt, rho, a[rho]);
      } else {
        v[rho] = zb[rho];
      }
      v[8 + rho] = X::mk(0.0, 0.0);
    }
    asm volatile("" : "+v"(tt));
    fft16::fft<P, false>(v, buf, twl, tt);
#pragma unroll
    for (int rho = 0; rho < 16; ++rho) v[rho] = cmul(v[rho], X::of(vhat[tt + T * rho]));
    fft16::fft<P, true>(v, buf, twl, tt);
    if (act) {
#pragma unroll
      for (int rho = 0; rho < 8; ++rho) a[rho] = park_get<T, C>(out, K, tt, rho);
    }
    __syncthreads();  // every parked read-back done before the pixel stores
    if (act) {
      const double2 ostep = expipi(4 * T, n);
      double2 om = expipi(4 * tt, n);
#pragma unroll
      for (int rho = 0; rho < 8; ++rho) {
        const int p = tt + T * rho;
        if (p < i) {
          const C bb = cmul(v[rho], X::of(cmul(chirp[p], om)));
          put2(out, p, cadd(a[rho], bb), K);
          put2(out, p + i, csub(a[rho], bb), K);
        }
        om = cmul(om, ostep);
      }
    }
  }
}

// Block b of a cap (WRAP: wrapping-ring) launch over pairs
// order[first .. first + count).
template <int P, bool WRAP, typename C>
__device__ __forceinline__ void cap_block(const LegTables &t, const RingTables &rt, int first,
                                          int count, double *__restrict__ map, int b,
                                          C *lds, const double2 *twl) {
  constexpr int T = P / 16, J = RB / (2 * T), S = job_slots(P);
  constexpr bool dual = sizeof(C) == 8;  // FP32: both rings together
  if constexpr (dual && !WRAP && J == 1) {
    if (t.K >= 2) {  // block-uniform
      cap_block_maps2<P>(t, rt, first, count, map, b, lds, twl);
      return;
    }
  }
  const int jb = threadIdx.x / (2 * T);
  int tt = threadIdx.x % (2 * T);
  C *bufA = lds + ring16_bufs(kRingCap16, P, sizeof(C)) * jb * S, *bufB = bufA + S;
  C *bufA2 = bufB + S, *bufB2 = bufA2 + S;  // dual only
  int pi, k;
  const bool act = ring16_job(b, J, jb, t.K, count, &pi, &k);
  if (!act) {
    pi = 0;
    k = 0;
  }
  const RingPair rp = rt.pairs[rt.order[first + pi]];
  const int K = t.K, M = t.M;
  const bool south = rp.pix_s >= 0;
  const C *qe = (const C *)t.qbuf + q_index(rp.r, 0, k, 0, K, M + 1);
  const C *qo = (const C *)t.qbuf + q_index(rp.r, 1, k, 0, K, M + 1);
  const int64_t qs = q_group(K);
  // every pair of a cap launch (also the padding jobs' stand-in) has tables
  const double2 *chirp = rt.chirp2 + rp.chirp2_off;  // [i]
  const double2 *vhat = rt.vhat2 + rp.vhat2_off;     // [P]
  C zn[8], zs[8];
  if constexpr (dual) {
    if constexpr (WRAP) {
      C ps[16], g0s;
      wrap_fold2<P>(qe, qo, qs, M, rp.n, rp.s, bufA, tt, act, act && south, ps, &g0s);
      wrap_fold_south<P>(bufA2, rp.n, rp.s, tt, ps, g0s);
    } else {
      cap_fold2<P>(qe, qo, qs, M, rp.n, rp.s, bufA, bufA2, tt, act, act && south);
    }
    cap_z<P>(rp, chirp, bufA, tt, zn);
    cap_z<P>(rp, chirp, bufA2, tt, zs);
    asm volatile("" : "+v"(tt));
    const bool coop = J >= 2 && K > 1 && J <= pow2_ceil(q_group(K));  // block-uniform
    cap_xform2<P>(rp, K, chirp, vhat, zn, zs, bufA, bufB, bufA2, bufB2, twl, tt, act,
                  act && south, !coop, map + k + rp.pix_n * K,
                  map + k + (south ? rp.pix_s : 0) * K);
    if (coop) cap_store_block<P>(t, rt, first, count, map, b, lds);
    return;
  }
  if constexpr (WRAP) {
    C ps[16], g0s;
    wrap_fold2<P>(qe, qo, qs, M, rp.n, rp.s, bufA, tt, act, act && south, ps, &g0s);
    cap_z<P>(rp, chirp, bufA, tt, zn);
    __syncthreads();  // reads of the northern fold done
    wrap_fold_south<P>(bufA, rp.n, rp.s, tt, ps, g0s);
    cap_z<P>(rp, chirp, bufA, tt, zs);
  } else {
    cap_fold2<P>(qe, qo, qs, M, rp.n, rp.s, bufA, bufB, tt, act, act && south);
    cap_z<P>(rp, chirp, bufA, tt, zn);
    cap_z<P>(rp, chirp, bufB, tt, zs);
  }
  asm volatile("" : "+v"(tt));  // no table load hoisted above the barriers
  cap_xform<P>(rp, K, chirp, vhat, zn, bufA, bufB, twl, tt, act, map + k + rp.pix_n * K);
  __syncthreads();  // the northern ring's reads of bufA / bufB done
  asm volatile("" : "+v"(tt));
  cap_xform<P>(rp, K, chirp, vhat, zs, bufA, bufB, twl, tt, act && south,
               map + k + (south ? rp.pix_s : 0) * K);
}

template <int KIND, int P, typename C>
__device__ __forceinline__ void ring16_block(const LegTables &t, const RingTables &rt, int first,
                                             int count, double *__restrict__ map, int b,
                                             C *lds, const double2 *twl) {
  if constexpr (KIND == kRingBelt16)
    belt_block<P>(t, rt, first, count, map, b, lds, twl);
  else if constexpr (cap_seq(KIND, P))
    cap_block_seq<P, KIND == kRingWrap16>(t, rt, first, count, map, b, lds, twl);
  else
    cap_block<P, KIND == kRingWrap16>(t, rt, first, count, map, b, lds, twl);
}

__device__ __forceinline__ void load_twiddles16(double2 *twl, const double2 *__restrict__ tw) {
  for (int u = threadIdx.x; u < kTwSlots; u += RB)
    twl[u] = (u < kTwC) ? tw[u * 64] : tw[u - kTwC];
}

// Waves per SIMD the kernel is built for (the second launch-bounds
// argument): one 512-thread block per CU.  FP32 points would fit the LDS of
// two blocks, but 4 waves per SIMD (<= 128 VGPRs) spill and measured slower.
template <int KIND, typename C> constexpr int ring16_waves_per_eu() { return 2; }

template <int KIND, int P, typename C>
__global__ __launch_bounds__(RB, (ring16_waves_per_eu<KIND, C>())) void k_ring16(LegTables t,
                                                                        RingTables rt, int first,
                                                                        int count,
                                                                        double *__restrict__ map) {
  extern __shared__ double2 lds[];
  double2 *twl = lds;
  load_twiddles16(twl, rt.twiddle);  // first read after a barrier
  ring16_block<KIND, P>(t, rt, first, count, map, blockIdx.x, (C *)(lds + kTwSlots), twl);
}

// Ring kernel of a batch: block i runs item r16items[first + i] (request,
// launch group, block) as the single-plan kernel runs block b (bitwise the
// same maps).  One item per block, in the host's order, rather than a
// work-queue loop: with the block body inlined into a loop the compiler kept
// thread-invariant values live across it and spilled 12-376 bytes per lane.
template <int KIND, int P>
__global__ __launch_bounds__(RB, (ring16_waves_per_eu<KIND, double2>())) void k_batch_ring16(
    BatchDev d, int first) {
  extern __shared__ double2 lds[];
  double2 *twl = lds;
  load_twiddles16(twl, d.twiddle);
  const BatchRing16Item x = d.r16items[first + blockIdx.x];
  const LegTables &t = d.tabs[x.req];
  const RingTables &rt = d.rtabs[x.req];
  ring16_block<KIND, P>(t, rt, rt.first[x.g], rt.count[x.g], d.map[x.req], x.b,
                        lds + kTwSlots, twl);
}

#define WM_RING16_SIZES(X) \
  X(16) X(32) X(64) X(128) X(256) X(512) X(1024) X(2048) X(4096) X(8192)

bool ring16_size_ok(int P) { return P >= 16 && P <= 8192 && (P & (P - 1)) == 0; }

hipError_t ring16_set_lds_limits() {
  static std::once_flag once;
  static hipError_t err = hipSuccess;
  std::call_once(once, [] {
#define WM_SET_K(KIND, P)                                                                  \
  for (const void *f : {(const void *)k_ring16<KIND, P, double2>,                          \
                        (const void *)k_batch_ring16<KIND, P>})                            \
    if (err == hipSuccess)                                                                 \
      err = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,             \
                                (int)ring16_bytes(KIND, P, sizeof(double2)));              \
  if (err == hipSuccess)                                                                   \
    err = hipFuncSetAttribute((const void *)k_ring16<KIND, P, float2>,                     \
                              hipFuncAttributeMaxDynamicSharedMemorySize,                  \
                              (int)ring16_bytes(KIND, P, sizeof(float2)));
#define WM_SET(P) WM_SET_K(kRingBelt16, P) WM_SET_K(kRingCap16, P) WM_SET_K(kRingWrap16, P)
    WM_RING16_SIZES(WM_SET)
#undef WM_SET

#undef WM_SET_K
  });
  return err;
}

// pairs order[first .. first + count) of one launch group (kind, FFT size
// P); t.q32: q in FP32 and the FP32 kernels
hipError_t launch_ring16(int kind, int P, const LegTables &t, const RingTables &rt, int first,
                         int count, double *map, hipStream_t s) {
  if (count <= 0) return hipSuccess;
  // FP32 belt jobs, and FP32 cap jobs of a whole block, hold two maps each at K >= 2
  const bool maps2 = (kind == kRingBelt16 ||
                      (kind == kRingCap16 && !cap_seq(kind, P) && ring16_jobs(kind, P) == 1)) &&
                     t.q32 && t.K >= 2;
  const int64_t nb = ring16_blocks(kind, P, maps2 ? (t.K + 1) / 2 : t.K, count);
  const size_t lds = ring16_bytes(kind, P, t.q32 ? sizeof(float2) : sizeof(double2));
#define WM_CASE_K(KIND, PP)                                                                \
  if (kind == KIND && P == PP) {                                                           \
    if (t.q32)                                                                             \
      hipLaunchKernelGGL((k_ring16<KIND, PP, float2>), dim3((unsigned)nb), dim3(RB), lds,  \
                         s, t, rt, first, count, map);                                     \
    else                                                                                   \
      hipLaunchKernelGGL((k_ring16<KIND, PP, double2>), dim3((unsigned)nb), dim3(RB), lds, \
                         s, t, rt, first, count, map);                                     \
    return hipGetLastError();                                                              \
  }
#define WM_CASE(PP) WM_CASE_K(kRingBelt16, PP) WM_CASE_K(kRingCap16, PP) WM_CASE_K(kRingWrap16, PP)
  WM_RING16_SIZES(WM_CASE)
#undef WM_CASE
#undef WM_CASE_K
  return hipErrorInvalidValue;
}

// Persistent batch launch of class (kind, P): items r16items[first ..
// first + count) through queue head ctr, nblocks workgroups.
hipError_t launch_batch_ring16(int kind, int P, const BatchDev &d, int first, int count,
                               hipStream_t s) {
  if (count <= 0) return hipSuccess;
  const size_t lds = ring16_bytes(kind, P, sizeof(double2));
#define WM_CASE_K(KIND, PP)                                                                \
  if (kind == KIND && P == PP) {                                                           \
    hipLaunchKernelGGL((k_batch_ring16<KIND, PP>), dim3((unsigned)count), dim3(RB), lds, s,  \
                       d, first);                                                          \
    return hipGetLastError();                                                              \
  }
#define WM_CASE(PP) WM_CASE_K(kRingBelt16, PP) WM_CASE_K(kRingCap16, PP) WM_CASE_K(kRingWrap16, PP)
  WM_RING16_SIZES(WM_CASE)
#undef WM_CASE
#undef WM_CASE_K
  return hipErrorInvalidValue;
}

}  // namespace wm

// Host side of libwavemoth_hip.so: the C ABI of include/wavemoth.h.
//
// A plan owns every device buffer and table of one (Nside, lmax, mmax,
// nmaps) synthesis on the HIP device that was current when it was made.
// Plan creation corresponds to the reference's wavemoth_plan_to_healpix
// (src/wavemoth.c:383-615) plus its offline resource precompute
// (wavemoth/lib.pyx:376-592): recurrence coefficient tables, per-(m, ring)
// seeds (computed on the GPU in FP64), ring geometry and the Bluestein
// tables of the ring stage.  Execute runs three kernels on one stream:
// a_lm prescale -> Legendre sweep -> ring synthesis.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>
#include <unistd.h>
#include <memory>

#include "../../include/wavemoth.h"
#include "wm_host.h"
#include "wm_resource.h"

using namespace wm;

static thread_local std::string g_err;
static std::string g_resource_dir;

namespace wm {
void set_err(const std::string &s) { g_err = s; }
void clear_err() { g_err.clear(); }
const char *last_err() { return g_err.c_str(); }
}  // namespace wm

struct _wavemoth_plan {
  int N = 0, L = 0, M = 0, K = 0, nthreads = 0;
  unsigned flags = 0;
  double *h_in = nullptr, *h_out = nullptr;
  int64_t nalm = 0, npix = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  LegTables lt;
  RingTables rt;
  RingHost rhost;
  double2 *d_alm = nullptr;
  double *d_map = nullptr;
  bool phase_shifts = true;
  bool fly = false;
  int rl_choice = 0;  // rings per lane of the Legendre sweep (0 = heuristic)
  // mixed-precision contraction (extension): -1 = FP64 sweep, else scheme
  // 0 FP32 MFMA, 1 BF16x3, 2 BF16x6
  int mix = -1;
  int mix_ncp = 0;          // padded real columns
  int64_t mix_pstride = 0;  // elements per B part
  void *mix_b = nullptr;
  int *mix_cexp = nullptr;  // FP16 scheme: a_lm scale exponents per (m, column)
  // x^2-form recurrence coefficients: d, e of the monic step (FP32 / BF16
  // schemes) or alpha, beta, gamma of the normalised one (FP16)
  double *mix_dtil = nullptr, *mix_etil = nullptr, *mix_gtil = nullptr;
  double mix_lam_scale = 1.0;
  // FP16x3: recurrence checkpoints of the register-Lambda kernel, and the
  // choice between it (default) and the producer/consumer kernel
  double2 *mix_ckpt = nullptr;
  bool mix_regs = true;
  LegCounts cnt;
  int64_t T_exec = 0;
  DevArena arena;
  // adjoint (extension): partial sums over ring groups, allocated on first use
  double2 *adj_partial = nullptr;
  int adj_groups = 0;
  // rev1 resource file given or found in the configured directory: mapped
  // and validated at plan time (the GPU path generates Lambda itself)
  std::unique_ptr<Resource> res;
  // WAVEMOTH_RESOURCE_APPLY: the Legendre stage runs from the file's strips
  // (the file image and its strip table on the device)
  bool res_apply = false;
  double *res_img = nullptr;
  ResStrip *res_strips = nullptr;
  int64_t *res_prefix = nullptr;
  int res_nstrips = 0;
  int64_t res_items = 0;
  // compressed matrices: the chunk buffer of the interpolations' blocks,
  // the interpolations sorted by tree depth (ops res_depth_op[d] ..
  // res_depth_op[d + 1] - 1, output rows res_depth_item[d] ..), their first
  // output rows and the filter position codes
  double2 *res_chunk = nullptr;
  ResInterp *res_ops = nullptr;
  int64_t *res_op_prefix = nullptr;
  int32_t *res_codes = nullptr;
  std::vector<int> res_depth_op;
  std::vector<int64_t> res_depth_item;
  // stage timing (extension): events recorded around prep / sweep / rings
  bool profiling = false;
  std::vector<hipEvent_t> ev_free;
  std::vector<std::array<hipEvent_t, 4>> ev_used;
};

static void release_events(wavemoth_plan p) {
  for (auto &a : p->ev_used)
    for (hipEvent_t e : a) p->ev_free.push_back(e);
  p->ev_used.clear();
}

static void free_plan(wavemoth_plan p) {
  if (!p) return;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  release_events(p);
  for (hipEvent_t e : p->ev_free) (void)hipEventDestroy(e);
  p->arena.release();
  if (p->stream) (void)hipStreamDestroy(p->stream);
  delete p;
}

// double-double from long double
static double2 dd(long double v) {
  double hi = (double)v;
  double lo = (double)(v - (long double)hi);
  return make_double2(hi, lo);
}

// z of the 2N northern rings, equator -> pole (wavemoth/healpix.py:5-18).
static double ring_z(int N, int r) {
  const int i = 2 * N - r;
  if (i <= N) return 1.0 - (double)i * (double)i / (3.0 * (double)N * (double)N);
  return 4.0 / 3.0 - 2.0 * (double)i / (3.0 * (double)N);
}

namespace wm {

bool build_leg_geometry(DevArena &a, LegTables &t, int N, int L, int M, hipStream_t stream,
                        LegCounts *cnt) {
  const int R = 2 * N;
  t.N = N;
  t.L = L;
  t.M = M;
  t.K = 0;
  t.R = R;

  std::vector<int64_t> off(M + 1);
  int64_t tot = 0;
  for (int m = 0; m <= M; ++m) {
    off[m] = tot;
    const int64_t len = L - m + 1;
    tot += (len + kChunk - 1) / kChunk * kChunk;
  }
  t.nlpad = tot;
  std::vector<int> chunk_m(tot / kChunk);
  std::vector<double> ctil(tot), ascale(tot), cscale(tot / kChunk);
  for (int m = 0; m <= M; ++m) {
    const int64_t len = L - m + 1;
    const int64_t lp = (len + kChunk - 1) / kChunk * kChunk;
    const double dm = (double)m;
    for (int64_t c = 0; c < lp / kChunk; ++c) {
      chunk_m[off[m] / kChunk + c] = m;
      double A = 1.0;
      for (int j = 0; j < kChunk; ++j) {
        const int64_t jj = c * kChunk + j;
        const double l = dm + (double)jj;
        ctil[off[m] + jj] = (l * l - dm * dm) / (4.0 * l * l - 1.0);
        ascale[off[m] + jj] = (jj < len) ? A : 0.0;
        const double l1 = l + 1.0;
        A *= std::sqrt((4.0 * l1 * l1 - 1.0) / (l1 * l1 - dm * dm));
      }
      cscale[off[m] / kChunk + c] = A;
    }
  }
  // log2 of sqrt(prod_{i<=m} (2i+1)/(2i)) / sqrt(4 pi)
  std::vector<double2> mfac(M + 1);
  {
    long double acc = -0.5L * log2l(4.0L * 3.141592653589793238462643383279502884L);
    for (int m = 0; m <= M; ++m) {
      if (m > 0) acc += 0.5L * log2l((2.0L * m + 1.0L) / (2.0L * m));
      mfac[m] = dd(acc);
    }
  }
  // Ring cos/sin as the reference sees them: theta = arccos(z) in double
  // (wavemoth/lib.pyx:382-383), x = cos(theta), sin(theta).
  std::vector<double> rx(R);
  std::vector<double2> rls(R);
  for (int r = 0; r < R; ++r) {
    const double th = std::acos(ring_z(N, r));
    rx[r] = std::cos(th);
    rls[r] = dd(log2l((long double)std::sin(th)));
  }

  if (!upload(a, &t.off, off, "off")) return false;
  if (!upload(a, &t.chunk_m, chunk_m, "chunk_m")) return false;
  if (!upload(a, &t.ctil, ctil, "ctil")) return false;
  if (!upload(a, &t.ascale, ascale, "ascale")) return false;
  if (!upload(a, &t.cscale, cscale, "cscale")) return false;
  if (!upload(a, &t.mfac, mfac, "mfac")) return false;
  if (!upload(a, &t.ring_x, rx, "ring_x")) return false;
  if (!upload(a, &t.ring_ls, rls, "ring_ls")) return false;
  const size_t nseed = (size_t)(M + 1) * R;
  if (!dalloc(a, &t.seed_l, nseed, "seed_l")) return false;
  if (!dalloc(a, &t.seed_v, nseed, "seed_v")) return false;
  WM_TRY(launch_seeds(t, stream), "seed kernel");
  WM_TRY(hipStreamSynchronize(stream), "seed stage");

  // Work accounting from the seeds: T_eps and per-(m, odd) element counts.
  std::vector<int> sl(nseed);
  WM_TRY(hipMemcpy(sl.data(), t.seed_l, nseed * sizeof(int), hipMemcpyDeviceToHost),
         "seed_l download");
  cnt->elems.assign((size_t)(M + 1) * 2, 0);
  cnt->T_dense = 0;
  cnt->T_eps = 0;
  for (int m = 0; m <= M; ++m) {
    cnt->T_dense += (int64_t)(L - m + 1) * R;
    for (int r = 0; r < R; ++r) {
      const int s = sl[(size_t)m * R + r];
      if (s > L) continue;
      const int64_t n = L - s + 1;
      cnt->T_eps += n;
      // l in [s, L] with l - m even / odd
      const int64_t ne = (L - m) / 2 - (s - m + 1) / 2 + 1;
      cnt->elems[2 * m] += std::max<int64_t>(ne, 0);
      cnt->elems[2 * m + 1] += n - std::max<int64_t>(ne, 0);
    }
  }
  return true;
}

bool build_leg_waves(DevArena &a, LegTables &t, int RL, hipStream_t stream, int64_t *T_exec) {
  const int L = t.L, M = t.M;
  t.RL = RL;
  t.W = (t.R + 64 * RL - 1) / (64 * RL);
  if (!dalloc(a, &t.wstart, (size_t)(M + 1) * t.W, "wstart")) return false;
  if (!dalloc(a, &t.wseedmax, (size_t)(M + 1) * t.W, "wseedmax")) return false;
  WM_TRY(launch_wave_bounds(t, stream), "wave-bounds kernel");
  // (l, ring) slots the sweep evaluates: every wave runs its 64*RL ring
  // slots from its first seed to the end of its last l-chunk.
  std::vector<int> ws((size_t)(M + 1) * t.W);
  WM_TRY(hipMemcpyAsync(ws.data(), t.wstart, ws.size() * sizeof(int), hipMemcpyDeviceToHost,
                        stream),
         "wstart download");
  WM_TRY(hipStreamSynchronize(stream), "wave bounds");
  int64_t tx = 0;
  for (int m = 0; m <= M; ++m) {
    const int64_t lend = m + (L - m + kChunk) / kChunk * kChunk;  // exclusive
    for (int w = 0; w < t.W; ++w) {
      const int s0 = ws[(size_t)m * t.W + w];
      if (s0 <= L) tx += (lend - s0) * 64 * RL;
    }
  }
  *T_exec = tx;
  return true;
}

}  // namespace wm

static bool build_legendre_tables(wavemoth_plan p) {
  LegTables &t = p->lt;
  if (!build_leg_geometry(p->arena, t, p->N, p->L, p->M, p->stream, &p->cnt)) return false;
  int RL = p->rl_choice > 0 ? p->rl_choice : legendre_default_rl(p->K);
  if (p->mix >= 0) RL = mix_rings_per_lane_table();  // one wave-table entry per workgroup
  if (!build_leg_waves(p->arena, t, RL, p->stream, &p->T_exec)) return false;
  t.K = p->K;
  t.KB = legendre_map_block(p->K);
  if (p->mix >= 0) {
    // the mixed path reads the pre-split B operand instead of atil
    const int cpb = mix_cols_per_block();
    p->mix_ncp = (2 * p->K + cpb - 1) / cpb * cpb;
    p->mix_pstride = mix_bmat_elems(t, p->mix_ncp);
    const size_t esz = p->mix == 0 ? 4 : 2;
    const size_t bytes = (size_t)p->mix_pstride * mix_parts(p->mix) * esz;
    unsigned char *b = nullptr;
    if (!dalloc(p->arena, &b, bytes, "mixed B operand")) return false;
    // the trailing padding chunks are never written by the prescale
    WM_TRY(hipMemsetAsync(b, 0, bytes, p->stream), "mixed B clear");
    WM_TRY(hipStreamSynchronize(p->stream), "mixed B clear");
    p->mix_b = b;
    if (!dalloc(p->arena, &p->mix_cexp, (size_t)(p->M + 1) * p->mix_ncp, "scale exponents"))
      return false;
    {
      // d_l = c~_l + c~_{l+1}, e_l = c~_l c~_{l-1}, c~_l = (l^2 - m^2)/(4 l^2 - 1)
      std::vector<int64_t> offh(p->M + 1);
      WM_TRY(hipMemcpy(offh.data(), t.off, offh.size() * sizeof(int64_t), hipMemcpyDeviceToHost),
             "off download");
      std::vector<double> dt(t.nlpad + 16, 0.0), et(t.nlpad + 16, 0.0), gt;
      auto ctl = [](double l, double m) { return (l * l - m * m) / (4.0 * l * l - 1.0); };
      // A_{l+1} / A_l of Lambda_l = A_l L~_l: r0(l) = sqrt((4(l+1)^2 - 1) / ((l+1)^2 - m^2))
      auto r0 = [](double l, double m) {
        return std::sqrt((4.0 * (l + 1.0) * (l + 1.0) - 1.0) / ((l + 1.0) * (l + 1.0) - m * m));
      };
      const bool norm = p->mix == 3;
      if (norm) gt.assign(t.nlpad + 16, 0.0);
      for (int m = 0; m <= p->M; ++m) {
        const int64_t lp = (p->L - m + kChunk) / kChunk * kChunk;
        const double dm = (double)m;
        for (int64_t jj = 0; jj < lp; ++jj) {
          const double l = dm + (double)jj;
          const double d = ctl(l, dm) + ctl(l + 1.0, dm);
          const double e = (jj == 0) ? 0.0 : ctl(l, dm) * ctl(l - 1.0, dm);
          if (!norm) {
            dt[offh[m] + jj] = d;
            et[offh[m] + jj] = e;
          } else {
            // Lambda_{l+2} = (alpha x^2 + beta) Lambda_l - gamma Lambda_{l-2}:
            // alpha = A_{l+2}/A_l, beta = -alpha d, gamma = e A_{l+2}/A_{l-2}
            // (e = 0 for l < m + 2, where A_{l-2} is undefined)
            const double alpha = r0(l + 1.0, dm) * r0(l, dm);
            dt[offh[m] + jj] = alpha;
            et[offh[m] + jj] = -alpha * d;
            gt[offh[m] + jj] =
                (jj < 2) ? 0.0 : e * alpha * r0(l - 1.0, dm) * r0(l - 2.0, dm);
          }
        }
      }
      if (!upload(p->arena, &p->mix_dtil, dt, "mixed d table")) return false;
      if (!upload(p->arena, &p->mix_etil, et, "mixed e table")) return false;
      if (norm && !upload(p->arena, &p->mix_gtil, gt, "mixed gamma table")) return false;
    }
    p->mix_lam_scale = mix_lambda_scale(p->L);
    WM_TRY(mix_set_lds_limits(), "mixed kernel LDS limit");
    if (p->mix == 3) {
      const char *ev = std::getenv("WAVEMOTH_MIX_PRODUCERS");
      p->mix_regs = !(ev && ev[0] == '1');
    }
    if (p->mix == 3 && p->mix_regs) {
      if (!dalloc(p->arena, &p->mix_ckpt, (size_t)mix_ckpt_elems(t), "recurrence checkpoints"))
        return false;
      WM_TRY(launch_mix_ckpt(t, p->mix_dtil, p->mix_etil, p->mix_gtil, p->mix_lam_scale,
                             p->mix_ckpt, p->stream),
             "checkpoint kernel");
      WM_TRY(hipStreamSynchronize(p->stream), "checkpoint kernel");
    }
    if (!dalloc(p->arena, &t.atil, 1, "atil")) return false;
  } else if (!dalloc(p->arena, &t.atil, (size_t)t.nlpad * legendre_padded_maps(p->K), "atil")) {
    return false;
  }
  // mixed-precision plans store q in FP32 (float2 elements at the same
  // indices); the buffer keeps the FP64 size, which the adjoint uses
  t.q32 = p->mix >= 0 ? 1 : 0;
  if (!dalloc(p->arena, &t.qbuf, (size_t)q_alloc_elems(t.R, p->K, p->M + 1), "qbuf")) return false;
  t.npix = 12 * (int64_t)p->N * p->N;
  return true;
}

static int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

static bool is_pow2(int v) { return v > 0 && (v & (v - 1)) == 0; }

namespace wm {

// Whether the general LDS ring kernel (and its adjoint) holds a ring of
// h = n/2 bins: a power-of-two h (P = 0) as h + 1 <= kMaxFFT/2 + 1 bins,
// otherwise the Bluestein length P <= kMaxFFT.
static bool general_fits(int h, int P) { return P == 0 ? h <= kMaxFFT / 2 : P <= kMaxFFT; }

bool build_ring_geometry(DevArena &a, RingTables &rt, RingHost &host, int N, int M,
                         bool phase_shifts, bool wide, hipStream_t stream) {
  const int nrings = 4 * N - 1, mid = 2 * N - 1, R = 2 * N;
  // Ring length of iring and phi0 = pi * num / den of grid entry iring
  // (src/wavemoth.c:1199-1232); den always equals the ring length.
  auto geom = [&](int iring, int &n, int &num) {
    if (iring <= N - 1) {
      n = 4 * (iring + 1);
      num = 1;
    } else if (iring > 3 * N - 1) {
      n = 4 * (nrings - iring);
      num = 1;
    } else {
      n = 4 * N;
      num = iring % 2;
    }
  };
  std::vector<int64_t> offs(nrings + 1, 0);
  for (int i = 0; i < nrings; ++i) {
    int n, num;
    geom(i, n, num);
    offs[i + 1] = offs[i] + n;
  }
  host.npix = offs[nrings];
  rt.npairs = R;
  std::vector<RingPair> pairs(R);
  std::vector<double> cost(R);
  std::vector<int> need(R);
  std::vector<int> hlist, plist, h2list, p2list;
  std::vector<int64_t> choff, vhoff, ch2off, vh2off;
  std::vector<int> h_index(2 * N + 1, -1), i_index(N + 1, -1);
  std::vector<int> kind(R, kRingGeneral), fftP(R, 0);
  bool all_general = true;  // every ring within the general kernel's limits
  int64_t chtot = 0, vhtot = 0, ch2tot = 0, vh2tot = 0;
  for (int r = 0; r < R; ++r) {
    // The pair is processed with the southern ring's length and phi0, as
    // src/wavemoth.c:532-546 sets ri->phi0 / ri->length from mid_ring + r;
    // it differs from the northern ring's own phi0 only for Nside = 1.
    int n, num;
    geom(mid + r, n, num);
    RingPair &rp = pairs[r];
    rp.pix_n = offs[mid - r];
    rp.pix_s = (r == 0) ? -1 : offs[mid + r];
    rp.n = n;
    rp.s = phase_shifts ? num : 0;
    rp.r = r;
    rp.chirp2_off = rp.vhat2_off = -1;
    rp.P2 = 0;
    const int h = n / 2;
    rp.chirp_off = rp.vhat_off = 0;  // set below for the rings with tables
    if (is_pow2(h)) {
      rp.P = 0;
      cost[r] = (double)h * std::log2((double)h + 1.0) + (double)(M + 1);
    } else {
      rp.P = next_pow2(2 * h - 1);
      cost[r] = 2.0 * (double)rp.P * std::log2((double)rp.P) + (double)(M + 1);
    }
    if (is_pow2(h) && M <= h && ring16_size_ok(h)) {
      // non-aliasing power-of-two ring: one FFT of h points
      kind[r] = kRingBelt16;
      fftP[r] = h;
    } else if (n % 4 == 0) {
      // split Bluestein of the even / odd Z entries (length i = h/2), with
      // the two-pass fold (n > mmax) or the wrapping fold
      const int i = h / 2, P2 = next_pow2(2 * i - 1);
      if (ring16_size_ok(P2)) {
        if (i_index[i] < 0) {
          i_index[i] = (int)h2list.size();
          h2list.push_back(i);
          p2list.push_back(P2);
          ch2off.push_back(ch2tot);
          vh2off.push_back(vh2tot);
          ch2tot += i;
          vh2tot += P2;
        }
        rp.P2 = P2;
        rp.chirp2_off = ch2off[i_index[i]];
        rp.vhat2_off = vh2off[i_index[i]];
        fftP[r] = P2;
        // the two-pass fold covers bins up to P2 - 1: h < P2 unless h is a
        // power of two (P2 = h), which takes the wrapping fold; that one
        // walks (mmax + 1) / n row blocks one after the other, so short
        // rings (P2 < 256) stay with the general kernel, and so do the
        // others at few maps where that kernel holds the ring's FFT
        if (n > M && h < P2) {
          kind[r] = kRingCap16;
        } else if (P2 >= 256 && (wide || !general_fits(h, rp.P))) {
          kind[r] = kRingWrap16;
        }
      }
    }
    need[r] = ring_lds_slots(n, rp.P, M);
    if (!general_fits(h, rp.P)) {
      all_general = false;
      if (kind[r] == kRingGeneral) {
        // longer than the LDS holds: the four-step path
        const int PL = rp.P == 0 ? h : rp.P;
        if (!long_size_ok(PL)) {
          set_err("ring stage: a ring's transform of " + std::to_string(PL) +
                  " points exceeds the supported lengths");
          return false;
        }
        kind[r] = kRingLong;
        fftP[r] = PL;
      }
    }
  }
  // the adjoint transforms every ring in the general kernel's LDS layout:
  // rings longer than it holds (Nside > 2048) leave the plan without one
  int adj_need = 0, adj_len = 0;
  for (int r = 0; r < R; ++r) {
    adj_need = std::max(adj_need, need[r]);
    adj_len = std::max(adj_len, pairs[r].P ? pairs[r].P : pairs[r].n);
  }
  rt.adj_ok = all_general && adj_need <= kLdsSlots ? 1 : 0;
  rt.adj_maxp = std::max(1024, next_pow2(adj_len));
  // Bluestein tables (chirp of h entries, kernel spectrum of P) of each
  // distinct h among the rings that use them: the general kernel's and the
  // long path's, and every ring's when the plan has an adjoint
  for (int r = 0; r < R; ++r) {
    RingPair &rp = pairs[r];
    if (rp.P == 0 || !(kind[r] == kRingGeneral || kind[r] == kRingLong || rt.adj_ok)) continue;
    const int h = rp.n / 2;
    if (h_index[h] < 0) {
      h_index[h] = (int)hlist.size();
      hlist.push_back(h);
      plist.push_back(rp.P);
      choff.push_back(chtot);
      vhoff.push_back(vhtot);
      chtot += h;
      vhtot += rp.P;
    }
    rp.chirp_off = choff[h_index[h]];
    rp.vhat_off = vhoff[h_index[h]];
  }
  // The general kernel's rings: size class and LDS need at that class
  std::vector<int> gneed(R, 0);
  for (int r = 0; r < R; ++r) {
    if (kind[r] != kRingGeneral) continue;
    fftP[r] = ring_class(pairs[r].n, pairs[r].P);
    gneed[r] = ring_lds_slots(pairs[r].n, pairs[r].P, M, fftP[r]);
  }
  // Launch groups: register-FFT caps and wrapping rings by FFT length
  // (largest first), the belt, then the general kernel's rings whose LDS
  // need exceeds 80 KiB (one block per CU) and the rest, each by size class;
  // costliest first within each.
  const int big_slots = (80 * 1024) / (int)sizeof(double2);
  auto key = [&](int r) -> int64_t {
    if (kind[r] == kRingLong)  // one group per (length, Bluestein or not)
      return 5 * 100000 + 2 * (int64_t)std::log2((double)fftP[r]) + (pairs[r].P ? 1 : 0);
    if (kind[r] == kRingCap16) return 0 * 100000 + (8192 - fftP[r]);
    if (kind[r] == kRingWrap16) return 1 * 100000 + (8192 - fftP[r]);
    if (kind[r] == kRingBelt16) return 2 * 100000 + (8192 - fftP[r]);
    return (gneed[r] > big_slots ? 3 : 4) * 100000 + (8192 - fftP[r]);
  };
  std::vector<int> order(R);
  for (int i = 0; i < R; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    if (key(a) != key(b)) return key(a) < key(b);
    return cost[a] > cost[b];
  });
  rt.nlaunch = 0;
  int max_need = 0;
  for (int i = 0; i < R; ++i) {
    const int r = order[i];
    if (rt.nlaunch == 0 || key(r) != key(order[rt.first[rt.nlaunch - 1]])) {
      if (rt.nlaunch == kMaxLaunches) {
        set_err("ring stage: too many ring launch groups");
        return false;
      }
      rt.first[rt.nlaunch] = i;
      rt.count[rt.nlaunch] = 0;
      rt.lds[rt.nlaunch] = 0;
      rt.adj_lds[rt.nlaunch] = 0;
      ++rt.nlaunch;
    }
    const int g = rt.nlaunch - 1;
    rt.count[g]++;
    rt.kind[g] = kind[r];
    rt.fftP[g] = fftP[r];
    rt.lds[g] = kind[r] == kRingGeneral ? std::max(rt.lds[g], gneed[r])
                : kind[r] == kRingLong   ? 0
                                         : ring16_lds_slots(kind[r], fftP[r]);
    rt.adj_lds[g] = std::max(rt.adj_lds[g], need[r]);
    if (kind[r] == kRingGeneral) max_need = std::max(max_need, gneed[r]);
  }
  WM_TRY(ring16_set_lds_limits(), "register-FFT ring kernel LDS limit");
  WM_TRY(ring_set_lds_limit(max_need), "ring kernel LDS limit");
  if (rt.adj_ok) WM_TRY(adjoint_set_lds_limit(adj_need), "ring adjoint kernel LDS limit");
  std::vector<double2> tw(kMaxFFT);
  for (int j = 0; j < kMaxFFT; ++j) {
    const long double a = -2.0L * 3.141592653589793238462643383279502884L * j / kMaxFFT;
    tw[j] = make_double2((double)cosl(a), (double)sinl(a));
  }
  if (!upload(a, &rt.pairs, pairs, "ring pairs")) return false;
  if (!upload(a, &rt.order, order, "ring order")) return false;
  if (!upload(a, &rt.twiddle, tw, "twiddle")) return false;
  if (!dalloc(a, &rt.chirp, (size_t)chtot, "chirp")) return false;
  if (!dalloc(a, &rt.vhat, (size_t)vhtot, "vhat")) return false;
  if (!dalloc(a, &rt.chirp2, (size_t)ch2tot, "split chirp")) return false;
  if (!dalloc(a, &rt.vhat2, (size_t)vh2tot, "split vhat")) return false;
  // the long path: its (ring pair, hemisphere) entries by launch group, and
  // two scratch buffers (at most 2^26 complex entries each, at least one
  // vector of the longest transform)
  LongRings &lr = host.lr;
  lr = LongRings();
  std::vector<LongVec> lvecs;
  int64_t lwant = 0;
  int maxP = 0;
  for (int g = 0; g < rt.nlaunch; ++g) {
    if (rt.kind[g] != kRingLong) continue;
    LongGroup lg{(int)lvecs.size(), 0, rt.fftP[g], pairs[order[rt.first[g]]].P != 0};
    for (int i = rt.first[g]; i < rt.first[g] + rt.count[g]; ++i) {
      const int r = order[i];
      lvecs.push_back({r, 0});
      if (pairs[r].pix_s >= 0) lvecs.push_back({r, 1});
    }
    lg.count = (int)lvecs.size() - lg.first;
    lr.groups.push_back(lg);
    lwant += (int64_t)lg.count * lg.P;
    maxP = std::max(maxP, lg.P);
  }
  for (size_t i = 0; i < hlist.size(); ++i)
    if (plist[i] > kMaxFFT) {
      lwant = std::max(lwant, (int64_t)plist[i]);
      maxP = std::max(maxP, plist[i]);
    }
  if (maxP > 0) {
    WM_TRY(long_set_lds_limits(), "long-ring kernel LDS limit");
    lr.elems = std::max<int64_t>(maxP, std::min<int64_t>(lwant, (int64_t)1 << 26));
    if (!upload(a, &lr.vecs, lvecs, "long-ring entries")) return false;
    if (!dalloc(a, &lr.u, (size_t)lr.elems, "long-ring scratch")) return false;
    if (!dalloc(a, &lr.v, (size_t)lr.elems, "long-ring scratch")) return false;
  }
  if (!hlist.empty()) {
    // short lengths by the LDS prep kernel, long ones by the four-step FFT
    std::vector<int> hs, ps;
    std::vector<int64_t> cs, vs;
    for (size_t i = 0; i < hlist.size(); ++i) {
      if (plist[i] <= kMaxFFT) {
        hs.push_back(hlist[i]);
        ps.push_back(plist[i]);
        cs.push_back(choff[i]);
        vs.push_back(vhoff[i]);
      }
    }
    if (!hs.empty()) {
      int *d_h = nullptr, *d_p = nullptr;
      int64_t *d_co = nullptr, *d_vo = nullptr;
      if (!upload(a, &d_h, hs, "hlist")) return false;
      if (!upload(a, &d_p, ps, "plist")) return false;
      if (!upload(a, &d_co, cs, "choff")) return false;
      if (!upload(a, &d_vo, vs, "vhoff")) return false;
      WM_TRY(launch_bluestein_prep(rt.twiddle, rt.chirp, rt.vhat, d_h, d_co, d_vo, d_p,
                                   (int)hs.size(), stream),
             "bluestein prep");
    }
    std::vector<int> lens;
    for (size_t i = 0; i < hlist.size(); ++i)
      if (plist[i] > kMaxFFT && std::find(lens.begin(), lens.end(), plist[i]) == lens.end())
        lens.push_back(plist[i]);
    for (int P : lens) {
      std::vector<int> hP;
      std::vector<int64_t> cP, vP;
      for (size_t i = 0; i < hlist.size(); ++i)
        if (plist[i] == P) {
          hP.push_back(hlist[i]);
          cP.push_back(choff[i]);
          vP.push_back(vhoff[i]);
        }
      int *d_h = nullptr;
      int64_t *d_co = nullptr, *d_vo = nullptr;
      if (!upload(a, &d_h, hP, "long hlist")) return false;
      if (!upload(a, &d_co, cP, "long choff")) return false;
      if (!upload(a, &d_vo, vP, "long vhoff")) return false;
      WM_TRY(launch_long_bluestein_prep(rt.twiddle, rt.chirp, rt.vhat, d_h, d_co, d_vo,
                                        (int)hP.size(), P, lr.u, lr.v, lr.elems, stream),
             "long bluestein prep");
    }
    WM_TRY(hipStreamSynchronize(stream), "bluestein prep");
  }
  if (!h2list.empty()) {
    int *d_h = nullptr, *d_p = nullptr;
    int64_t *d_co = nullptr, *d_vo = nullptr;
    if (!upload(a, &d_h, h2list, "split hlist")) return false;
    if (!upload(a, &d_p, p2list, "split plist")) return false;
    if (!upload(a, &d_co, ch2off, "split choff")) return false;
    if (!upload(a, &d_vo, vh2off, "split vhoff")) return false;
    WM_TRY(launch_bluestein_prep(rt.twiddle, rt.chirp2, rt.vhat2, d_h, d_co, d_vo, d_p,
                                 (int)h2list.size(), stream),
           "split bluestein prep");
    WM_TRY(hipStreamSynchronize(stream), "split bluestein prep");
  }
  host.pairs = std::move(pairs);
  host.need = std::move(need);
  host.cost = std::move(cost);
  host.kind = std::move(kind);
  return true;
}

}  // namespace wm

static bool build_ring_tables(wavemoth_plan p) {
  if (!build_ring_geometry(p->arena, p->rt, p->rhost, p->N, p->M, p->phase_shifts,
                           p->K >= kRingWideMaps, p->stream))
    return false;
  p->npix = p->rhost.npix;
  return true;
}

// Clears the phi0 phase of every ring pair (phase shifting off).  The
// table is rewritten only after all work queued on the device has finished
// (the plan's stream and any caller stream of wavemoth_execute_device), so
// no execute in flight reads a half-updated table.
static bool refresh_phases(wavemoth_plan p) {
  WM_TRY(hipDeviceSynchronize(), "wait for pending executes");
  std::vector<RingPair> pairs(p->rt.npairs);
  WM_TRY(hipMemcpy(pairs.data(), p->rt.pairs, pairs.size() * sizeof(RingPair),
                   hipMemcpyDeviceToHost),
         "ring table download");
  for (auto &rp : pairs) rp.s = 0;
  WM_TRY(hipMemcpy(p->rt.pairs, pairs.data(), pairs.size() * sizeof(RingPair),
                   hipMemcpyHostToDevice),
         "ring table upload");
  return true;
}

static hipEvent_t take_event(wavemoth_plan p) {
  if (!p->ev_free.empty()) {
    hipEvent_t e = p->ev_free.back();
    p->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Event slot i (0..3) of the current execute, when profiling.
static bool mark(wavemoth_plan p, int i, hipStream_t s) {
  if (!p->profiling) return true;
  if (i == 0) p->ev_used.push_back({nullptr, nullptr, nullptr, nullptr});
  hipEvent_t e = take_event(p);
  if (!e) {
    set_err("hipEventCreate failed");
    return false;
  }
  p->ev_used.back()[i] = e;
  WM_TRY(hipEventRecord(e, s), "hipEventRecord");
  return true;
}

static bool run_legendre(wavemoth_plan p, const double2 *d_alm, hipStream_t s) {
  if (!mark(p, 0, s)) return false;
  if (p->fly) {
    WM_TRY(launch_seeds(p->lt, s), "seed kernel");
    WM_TRY(launch_wave_bounds(p->lt, s), "wave-bounds kernel");
    if (p->mix_ckpt)
      WM_TRY(launch_mix_ckpt(p->lt, p->mix_dtil, p->mix_etil, p->mix_gtil, p->mix_lam_scale,
                             p->mix_ckpt, s),
             "checkpoint kernel");
  }
  if (p->res_apply) {
    if (!mark(p, 1, s)) return false;
    WM_TRY(launch_res_legendre(p->lt, d_alm, p->res_img, p->res_strips, p->res_prefix,
                               p->res_nstrips, p->res_items, p->res_chunk, true, s),
           "resource Legendre kernel");
    for (size_t d = 0; d + 1 < p->res_depth_op.size(); ++d)
      WM_TRY(launch_res_interp(p->lt, p->res_img, p->res_ops, p->res_op_prefix, p->res_codes,
                               p->res_depth_op[d], p->res_depth_op[d + 1],
                               p->res_depth_item[d], p->res_depth_item[d + 1], p->res_chunk,
                               s),
             "resource interpolation kernel");
  } else if (p->mix >= 0) {
    WM_TRY(launch_mix_prep(p->lt, d_alm, p->mix_b, p->mix_ncp, p->mix, p->mix_pstride,
                           p->mix_cexp, s),
           "mixed prescale kernel");
    if (!mark(p, 1, s)) return false;
    if (p->mix_ckpt)
      WM_TRY(launch_mixr(p->lt, p->mix_b, p->mix_ncp, p->mix_pstride, p->mix_lam_scale,
                         p->mix_cexp, p->mix_dtil, p->mix_etil, p->mix_gtil, p->mix_ckpt, s),
             "mixed contraction kernel");
    else
      WM_TRY(launch_mix(p->lt, p->mix_b, p->mix_ncp, p->mix, p->mix_pstride, p->mix_lam_scale,
                        p->mix_cexp, p->mix_dtil, p->mix_etil, p->mix_gtil, s),
             "mixed contraction kernel");
  } else {
    WM_TRY(launch_prep_alm(p->lt, d_alm, s), "a_lm prescale kernel");
    if (!mark(p, 1, s)) return false;
    WM_TRY(launch_legendre(p->lt, s), "legendre kernel");
  }
  if (!mark(p, 2, s)) return false;
  return true;
}

// WAVEMOTH_RESOURCE_APPLY: the file image and its strip table on the device.
static bool build_resource_apply(wavemoth_plan p) {
  if (!p->res_apply) return true;
  const Resource &r = *p->res;
  std::vector<ResStrip> st;
  std::vector<int64_t> prefix;
  int64_t items = 0;
  for (const ResourceStrip &x : r.strips) {
    ResStrip d;
    d.m = x.m;
    d.odd = x.odd;
    d.nks = (int)(x.nk - x.rstart);
    d.cstart = (int)x.cstart;
    d.nx = (int)(x.cstop - x.cstart);
    d.row = x.row0 + x.rstart;
    d.chunk = x.chunk;
    if (x.x2_off >= 0) {
      d.x2 = x.x2_off / 8;
      d.p0 = x.p0_off / 8;
      d.p1 = x.p1_off / 8;
      d.aux = x.aux_off / 8 + 3 * x.rstart;  // the strip's first recurrence step
    } else {
      d.dense = x.dense_off / 8;
    }
    prefix.push_back(items);
    items += d.nx;
    st.push_back(d);
  }
  p->res_nstrips = (int)st.size();
  p->res_items = items;
  // interpolations, stably sorted by depth (a parent's outputs are its
  // children's inputs), with the filters turned into position codes
  std::vector<const ResourceInterp *> ord;
  for (const ResourceInterp &x : r.interps) ord.push_back(&x);
  std::stable_sort(ord.begin(), ord.end(), [](const ResourceInterp *a, const ResourceInterp *b) {
    return a->depth < b->depth;
  });
  std::vector<ResInterp> ops;
  std::vector<int64_t> op_prefix;
  std::vector<int32_t> codes;
  int64_t oitems = 0;
  p->res_depth_op.assign(1, 0);
  p->res_depth_item.assign(1, 0);
  for (size_t i = 0; i < ord.size(); ++i) {
    const ResourceInterp &x = *ord[i];
    if (i > 0 && x.depth != ord[i - 1]->depth) {
      p->res_depth_op.push_back((int)i);
      p->res_depth_item.push_back(oitems);
    }
    ResInterp o;
    o.m = x.m;
    o.odd = x.odd;
    o.kT = (int)x.kT;
    o.kB = (int)x.kB;
    o.nl = (int)x.nl;
    o.nr = (int)x.nr;
    o.to_rings = x.to_rings;
    o.inT = x.inT;
    o.inB = x.inB;
    o.outL = x.outL;
    o.outR = x.outR;
    o.ipT = x.ipT / 8;
    o.ipB = x.ipB / 8;
    const int64_t n = x.nl + x.nr;
    for (int side = 0; side < 2; ++side) {
      (side ? o.codeB : o.codeT) = (int64_t)codes.size();
      const char *mask = r.base + (side ? x.maskB : x.maskT);
      int32_t zeros = 0, ones = 0;
      for (int64_t u = 0; u < n; ++u) codes.push_back(mask[u] == 0 ? zeros++ : -1 - ones++);
    }
    ops.push_back(o);
    op_prefix.push_back(oitems);
    oitems += n;
  }
  // every index the kernels form stays inside the chunk buffer, the rings
  // and the file image
  const int64_t rows = r.chunk_rows, img = (int64_t)(r.size / 8), R = 2 * (int64_t)p->N;
  for (const ResStrip &d : st)
    if (d.chunk >= 0 && (d.chunk + d.nx > rows)) {
      set_err("resource apply: a strip's chunk rows lie outside the chunk buffer");
      return false;
    }
  for (const ResInterp &o : ops) {
    const int64_t n = (int64_t)o.nl + o.nr;
    const bool in_ok = o.inT >= 0 && o.inT + o.kT <= rows && o.inB >= 0 && o.inB + o.kB <= rows;
    const int64_t lim = o.to_rings ? R : rows;
    const bool out_ok = o.outL >= 0 && o.outL + o.nl <= lim && o.outR >= 0 && o.outR + o.nr <= lim;
    const bool ip_ok = o.ipT >= 0 && o.ipT + (n - o.kT) * o.kT <= img && o.ipB >= 0 &&
                       o.ipB + (n - o.kB) * o.kB <= img;
    if (!in_ok || !out_ok || !ip_ok || o.m < 0 || o.m > p->M) {
      set_err("resource apply: an interpolation's rows lie outside its buffers");
      return false;
    }
  }
  if (!ops.empty()) {
    p->res_depth_op.push_back((int)ops.size());
    p->res_depth_item.push_back(oitems);
  } else {
    p->res_depth_op.clear();
    p->res_depth_item.clear();
  }
  if (st.empty() && ops.empty()) return true;
  if (!dalloc(p->arena, &p->res_img, r.size / 8 + 1, "resource image")) return false;
  WM_TRY(hipMemcpy(p->res_img, r.base, r.size, hipMemcpyHostToDevice), "resource upload");
  if (!st.empty() && (!upload(p->arena, &p->res_strips, st, "resource strips") ||
                      !upload(p->arena, &p->res_prefix, prefix, "resource strip prefix")))
    return false;
  if (!ops.empty()) {
    if (!upload(p->arena, &p->res_ops, ops, "resource interpolations") ||
        !upload(p->arena, &p->res_op_prefix, op_prefix, "resource interpolation prefix") ||
        !upload(p->arena, &p->res_codes, codes, "resource filter codes"))
      return false;
  }
  if (r.chunk_rows > 0 &&
      !dalloc(p->arena, &p->res_chunk, (size_t)r.chunk_rows * p->K, "resource chunk buffer"))
    return false;
  return true;
}

static bool run_rings(wavemoth_plan p, double *d_map, hipStream_t s) {
  WM_TRY(launch_ring_synth(p->lt, p->rt, d_map, s), "ring kernel");
  if (!p->rhost.lr.groups.empty())
    WM_TRY(launch_long_rings(p->lt, p->rt, p->rhost.lr, d_map, s), "long-ring kernels");
  if (p->profiling && !p->ev_used.empty() && p->ev_used.back()[2] != nullptr)
    if (!mark(p, 3, s)) return false;
  return true;
}

extern "C" {

void wavemoth_configure(char *resource_dir) {
  g_resource_dir = resource_dir ? resource_dir : "";
}

int wavemoth_query_resourcefile(char *filename, int *out_Nside, int *out_lmax) {
  if (!filename) return -1;
  FILE *fd = std::fopen(filename, "rb");
  if (!fd) return -1;
  int64_t fields[3];
  const size_t got = std::fread(fields, sizeof(fields), 1, fd);
  std::fclose(fd);
  if (got != 1) {
    set_err(std::string("could not read resource header: ") + filename);
    return -1;
  }
  *out_lmax = (int)fields[0];
  *out_Nside = (int)fields[2];
  return 0;
}

int wavemoth_resource_info(const char *filename, int64_t *info, int ninfo) {
  g_err.clear();
  if (!filename || (ninfo > 0 && !info)) {
    set_err("NULL argument");
    return -1;
  }
  std::unique_ptr<Resource> r(resource_open(filename));
  if (!r) return -1;
  const double dev = resource_x2_deviation(*r);
  const int64_t v[11] = {r->lmax, r->mmax, r->Nside, r->present, r->missing, r->trees,
                         r->recurrence_strips, r->dense_blocks, r->elements, r->max_depth,
                         dev < 0 ? -1 : (int64_t)std::ceil(dev * 1e18)};
  for (int i = 0; i < ninfo && i < 11; ++i) info[i] = v[i];
  return 0;
}

wavemoth_plan wavemoth_plan_to_healpix(int Nside, int lmax, int mmax, int nmaps, int nthreads,
                                       double *input, double *output, int ordering,
                                       unsigned flags, char *resource_filename) {
  g_err.clear();
  if (nthreads <= 0) {
    set_err("Require nthreads > 0");
    return nullptr;
  }
  if (ordering != WAVEMOTH_MMAJOR) {
    set_err("only WAVEMOTH_MMAJOR ordering is supported");
    return nullptr;
  }
  // Resource: the file given, else <dir>/rev1/<Nside>.dat when a directory
  // is configured and the file exists (src/wavemoth.c:556-569).  It is
  // validated completely before any device work; Nside < 0 takes the
  // file's.  Without one the plan needs none (Lambda is generated on the
  // device), where the reference would abort.
  Resource *res = nullptr;
  if (resource_filename) {
    res = resource_open(resource_filename);
    if (!res) return nullptr;
  } else if (Nside < 0) {
    set_err("Invalid Nside");
    return nullptr;
  } else if (!g_resource_dir.empty()) {
    const std::string path = resource_default_path(g_resource_dir, Nside);
    if (access(path.c_str(), F_OK) == 0) {
      res = resource_open(path.c_str());
      if (!res) return nullptr;
    }
  }
  if (res && !resource_matches(*res, &Nside, lmax, mmax)) {
    delete res;
    return nullptr;
  }
  std::unique_ptr<Resource> res_owner(res);
  if (Nside < 1 || Nside > kMaxNsidePlan) {
    set_err("Nside must be in [1, " + std::to_string(kMaxNsidePlan) + "]");
    return nullptr;
  }
  if (lmax < 0 || mmax < 0 || mmax > lmax) {
    set_err("require 0 <= mmax <= lmax");
    return nullptr;
  }
  if (nmaps < 1) {
    set_err("nmaps must be >= 1");
    return nullptr;
  }
  if ((int64_t)lmax + kChunk > (1 << 28)) {
    set_err("lmax too large");
    return nullptr;
  }
  wavemoth_plan p = new _wavemoth_plan();
  p->res = std::move(res_owner);
  p->N = Nside;
  p->L = lmax;
  p->M = mmax;
  p->K = nmaps;
  p->nthreads = nthreads;
  p->flags = flags;
  p->h_in = input;
  p->h_out = output;
  p->fly = (flags & WAVEMOTH_SEEDS_ON_THE_FLY) != 0;
  {
    const unsigned mf = flags & (WAVEMOTH_MIXED_F32 | WAVEMOTH_MIXED_BF16X3 |
                                 WAVEMOTH_MIXED_BF16X6 | WAVEMOTH_MIXED_F16X3);
    if (mf == WAVEMOTH_MIXED_F32) p->mix = 0;
    else if (mf == WAVEMOTH_MIXED_BF16X3) p->mix = 1;
    else if (mf == WAVEMOTH_MIXED_BF16X6) p->mix = 2;
    else if (mf == WAVEMOTH_MIXED_F16X3) p->mix = 3;
    else if (mf != 0) {
      set_err("at most one WAVEMOTH_MIXED_* flag");
      delete p;
      return nullptr;
    }
    if (p->mix == 3 && nmaps > 4096) {
      set_err("WAVEMOTH_MIXED_F16X3 supports nmaps <= 4096");
      delete p;
      return nullptr;
    }
  }
  if (flags & WAVEMOTH_RESOURCE_APPLY) {
    const char *why = !p->res ? "WAVEMOTH_RESOURCE_APPLY needs a resource file"
                      : p->mix >= 0
                          ? "WAVEMOTH_RESOURCE_APPLY runs in FP64: no WAVEMOTH_MIXED_* flag"
                          : nullptr;
    if (why) {
      set_err(why);
      delete p;
      return nullptr;
    }
    p->res_apply = true;
  }
  p->nalm = 0;
  for (int m = 0; m <= mmax; ++m) p->nalm += lmax - m + 1;
  if (hipGetDevice(&p->device) != hipSuccess) {
    set_err("no HIP device");
    delete p;
    return nullptr;
  }
  if (hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess) {
    set_err("hipStreamCreate failed");
    delete p;
    return nullptr;
  }
  if (!build_legendre_tables(p) || !build_ring_tables(p) || !build_resource_apply(p)) {
    std::string e = g_err;
    free_plan(p);
    g_err = e;
    return nullptr;
  }
  return p;
}

void wavemoth_destroy_plan(wavemoth_plan plan) { free_plan(plan); }

// Device staging of the host-buffer executes, allocated on first use (a
// plan driven only through device pointers never holds them).
static bool ensure_staging(wavemoth_plan p) {
  if (!p->d_alm && !dalloc(p->arena, &p->d_alm, (size_t)p->nalm * p->K, "alm staging"))
    return false;
  if (!p->d_map && !dalloc(p->arena, &p->d_map, (size_t)p->npix * p->K, "map staging"))
    return false;
  return true;
}

static bool execute_host(wavemoth_plan p, bool leg, bool rings) {
  if (!p) {
    set_err("NULL plan");
    return false;
  }
  if (!ensure_staging(p)) return false;
  hipStream_t s = p->stream;
  if (leg) {
    if (!p->h_in) {
      set_err("plan has no host input");
      return false;
    }
    WM_TRY(hipMemcpyAsync(p->d_alm, p->h_in, (size_t)p->nalm * p->K * sizeof(double2),
                          hipMemcpyHostToDevice, s),
           "input upload");
    if (!run_legendre(p, p->d_alm, s)) return false;
  }
  if (rings) {
    if (!p->h_out) {
      set_err("plan has no host output");
      return false;
    }
    if (!run_rings(p, p->d_map, s)) return false;
    WM_TRY(hipMemcpyAsync(p->h_out, p->d_map, (size_t)p->npix * p->K * sizeof(double),
                          hipMemcpyDeviceToHost, s),
           "output download");
  }
  WM_TRY(hipStreamSynchronize(s), "execute");
  return true;
}

void wavemoth_execute(wavemoth_plan plan) {
  g_err.clear();
  (void)execute_host(plan, true, true);
}

void wavemoth_perform_legendre_transforms(wavemoth_plan plan) {
  g_err.clear();
  (void)execute_host(plan, true, false);
}

void wavemoth_perform_backward_ffts(wavemoth_plan plan) {
  g_err.clear();
  (void)execute_host(plan, false, true);
}

void wavemoth_disable_phase_shifting(wavemoth_plan plan) {
  g_err.clear();
  if (!plan) return;
  plan->phase_shifts = false;
  (void)refresh_phases(plan);
}

int64_t wavemoth_get_legendre_flops(wavemoth_plan plan, int m, int odd) {
  if (!plan || m < 0 || m > plan->M || odd < 0 || odd > 1) return -1;
  return 4 * plan->cnt.elems[(size_t)2 * m + odd];
}

const char *wavemoth_last_error(void) { return last_err(); }

int wavemoth_get_work(wavemoth_plan p, double *out) {
  g_err.clear();
  if (!p || !out) {
    set_err("NULL argument");
    return -1;
  }
  const int R = p->lt.R, M = p->M, K = p->K;
  std::vector<double2> q((size_t)q_elems(R, K, M + 1));
  if (p->lt.q32) {
    // mixed-precision plans keep q in FP32 at the same indices
    std::vector<float2> qf(q.size());
    if (hipMemcpy(qf.data(), p->lt.qbuf, qf.size() * sizeof(float2), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      set_err("work download failed");
      return -1;
    }
    for (size_t u = 0; u < q.size(); ++u) q[u] = make_double2(qf[u].x, qf[u].y);
  } else if (hipMemcpy(q.data(), p->lt.qbuf, q.size() * sizeof(double2),
                       hipMemcpyDeviceToHost) != hipSuccess) {
    set_err("work download failed");
    return -1;
  }
  double2 *o = (double2 *)out;
  for (int r = 0; r < R; ++r)
    for (int odd = 0; odd < 2; ++odd)
      for (int m = 0; m <= M; ++m)
        for (int k = 0; k < K; ++k)
          o[(((size_t)m * 2 + odd) * R + r) * K + k] = q[q_index(r, odd, k, m, K, M + 1)];
  return 0;
}

int wavemoth_execute_device(wavemoth_plan p, const double *d_input, double *d_output,
                            void *stream) {
  g_err.clear();
  if (!p || !d_input || !d_output) {
    set_err("NULL argument");
    return -1;
  }
  if (((uintptr_t)d_input | (uintptr_t)d_output) & 15) {
    set_err("device buffers must be 16-byte aligned");
    return -1;
  }
  hipStream_t s = (hipStream_t)stream;
  if (!run_legendre(p, (const double2 *)d_input, s)) return -1;
  if (!run_rings(p, d_output, s)) return -1;
  return 0;
}

static bool adjoint_available(wavemoth_plan p) {
  if (p->rt.adj_ok) return true;
  set_err("adjoint: ring FFTs longer than " + std::to_string(kMaxFFT) +
          " points (Nside > 2048) are not supported");
  return false;
}

static bool run_adjoint(wavemoth_plan p, const double *d_map, double2 *d_alm, hipStream_t s) {
  if (!adjoint_available(p)) return false;
  if (!p->adj_partial) {
    p->adj_groups = adjoint_groups(p->lt);
    if (!dalloc(p->arena, &p->adj_partial, (size_t)p->adj_groups * p->lt.nlpad * p->K,
                "adjoint partial sums"))
      return false;
  }
  WM_TRY(launch_adjoint(p->lt, p->rt, d_map, d_alm, p->adj_partial, p->adj_groups, s),
         "adjoint kernels");
  return true;
}

int wavemoth_execute_adjoint_device(wavemoth_plan p, const double *d_map, double *d_alm,
                                    void *stream) {
  g_err.clear();
  if (!p || !d_map || !d_alm) {
    set_err("NULL argument");
    return -1;
  }
  if (((uintptr_t)d_map | (uintptr_t)d_alm) & 15) {
    set_err("device buffers must be 16-byte aligned");
    return -1;
  }
  return run_adjoint(p, d_map, (double2 *)d_alm, (hipStream_t)stream) ? 0 : -1;
}

int wavemoth_execute_adjoint_device_noalloc(wavemoth_plan p, const double *d_map,
                                            double *d_alm, void *stream) {
  g_err.clear();
  if (!p || !d_map || !d_alm) {
    set_err("NULL argument");
    return -1;
  }
  if (((uintptr_t)d_map | (uintptr_t)d_alm) & 15) {
    set_err("device buffers must be 16-byte aligned");
    return -1;
  }
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev != p->device) {
    set_err("adjoint without allocation: make the plan's device current first");
    return -1;
  }
  if (!adjoint_available(p)) return -1;
  double2 *scr = adjoint_scratch(dev);
  if (!scr) return -1;
  const size_t cap = adjoint_scratch_elems();
  const int G = adjoint_groups_within(p->lt, cap);
  if (G == 0) {
    set_err("adjoint without allocation: the partial sums of one ring group need " +
            std::to_string((size_t)p->lt.nlpad * p->K * sizeof(double2)) +
            " bytes, above the fixed scratch of " + std::to_string(cap * sizeof(double2)) +
            " bytes (use wavemoth_execute_adjoint_device)");
    return -1;
  }
  const hipError_t e = launch_adjoint(p->lt, p->rt, d_map, (double2 *)d_alm, scr, G,
                                      (hipStream_t)stream);
  if (e != hipSuccess) {
    set_err(std::string("adjoint kernels: ") + hipGetErrorString(e));
    return -1;
  }
  return 0;
}

void wavemoth_execute_adjoint(wavemoth_plan p) {
  g_err.clear();
  if (!p || !p->h_in || !p->h_out) {
    set_err("plan has no host buffers");
    return;
  }
  if (!ensure_staging(p)) return;
  hipStream_t s = p->stream;
  if (hipMemcpyAsync(p->d_map, p->h_out, (size_t)p->npix * p->K * sizeof(double),
                     hipMemcpyHostToDevice, s) != hipSuccess) {
    set_err("map upload failed");
    return;
  }
  if (!run_adjoint(p, p->d_map, p->d_alm, s)) return;
  if (hipMemcpyAsync(p->h_in, p->d_alm, (size_t)p->nalm * p->K * sizeof(double2),
                     hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    set_err("adjoint download failed");
}

int wavemoth_set_profiling(wavemoth_plan p, int on) {
  g_err.clear();
  if (!p) return -1;
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  release_events(p);
  p->profiling = on != 0;
  return 0;
}

int wavemoth_stage_times(wavemoth_plan p, double *avg_ms, int *count) {
  g_err.clear();
  if (!p || !avg_ms || !count) return -1;
  double acc[3] = {0, 0, 0};
  int n = 0;
  for (auto &a : p->ev_used) {
    if (!a[0] || !a[1] || !a[2] || !a[3]) continue;
    if (hipEventSynchronize(a[3]) != hipSuccess) {
      set_err("hipEventSynchronize failed");
      return -1;
    }
    for (int i = 0; i < 3; ++i) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, a[i], a[i + 1]) != hipSuccess) {
        set_err("hipEventElapsedTime failed");
        return -1;
      }
      acc[i] += ms;
    }
    ++n;
  }
  for (int i = 0; i < 3; ++i) avg_ms[i] = n ? acc[i] / n : 0.0;
  *count = n;
  release_events(p);
  return 0;
}

int64_t wavemoth_plan_info(wavemoth_plan p, int key) {
  if (!p) return -1;
  switch (key) {
    case 0: return p->N;
    case 1: return p->L;
    case 2: return p->M;
    case 3: return p->K;
    case 4: return p->npix;
    case 5: return p->cnt.T_dense;
    case 6: return p->cnt.T_eps;
    case 7: return p->flags;
    case 8: return 1;
    case 9: return (int64_t)p->arena.bytes;
    case 10: return p->T_exec;
    case 11: return p->mix;
    case 12: return p->rt.adj_ok ? adjoint_groups(p->lt) : 0;
    case 13: return p->rt.adj_ok ? adjoint_groups_within(p->lt, adjoint_scratch_elems()) : 0;
    case 14:
    case 15:
    case 16:
    case 17: {
      // ring pairs of the general, belt, cap and wrapping-ring launches
      int64_t n = 0;
      for (int g = 0; g < p->rt.nlaunch; ++g)
        if (p->rt.kind[g] == key - 14) n += p->rt.count[g];
      return n;
    }
    case 18: {
      int64_t mx = 0;
      for (int g = 0; g < p->rt.nlaunch; ++g)
        if (p->rt.kind[g] != kRingGeneral && p->rt.kind[g] != kRingLong &&
            p->rt.fftP[g] > mx)
          mx = p->rt.fftP[g];
      return mx;
    }
    case 19: {
      int64_t n = 0;
      for (int g = 0; g < p->rt.nlaunch; ++g)
        if (p->rt.kind[g] == kRingLong) n += p->rt.count[g];
      return n;
    }
    default: return -1;
  }
}

}  // extern "C"

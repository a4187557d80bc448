/* CPU baseline: a vectorised restatement of the reference's SIMD Legendre
   kernel (src/legendre_transform.c.in:63-316 strips of columns sharing one
   recurrence, :389-451 the blocked product Lambda^T a over a block of rows
   kept in L1, :453-616 the driver that alternates the two per K-chunk).

   TEST INFRASTRUCTURE ONLY.  It is the timed leg of bench.py's CPU baseline
   and is checked against or_legendre_stage (the oracle's scalar restatement)
   in tests/test_oracle_fast.py; the product never links or calls it.

   It runs over the oracle plan (sht_oracle.h): per (m, odd, ring) the first
   row k0 and up to four seed values.  Rings are taken S at a time (W
   doubles per vector: 4 for AVX2, 8 for AVX-512), in ring order, so that
   neighbouring rings with close k0 share a strip, as the reference's
   stripify groups columns of a resource file.  A strip starts at its
   smallest k0; each ring joins at its own k0 (rows before it are zero, the
   seeds are injected, dense rings with <= 4 rows take their seeds as the
   rows themselves), and from the row past every ring's seeds on the
   recurrence
       P_k = (x^2 + alpha_k) beta_k P_{k-1} + gamma_k P_{k-2}
   runs vectorised over the strip's rings.  KC rows at a time go to a block
   in L1, then
     - many columns (2*nmaps >= 2W): a register-blocked product over XR rings
       x 2 vectors of columns, P broadcast, a loaded (the reference's
       legendre_matmul_chunk, vectorised over columns);
     - one map (2 columns): P vectors over rings times broadcast a (the
       reference's _nvec2 kernels, vectorised over rings).
   Compiled once per vector width (FAST_W, FAST_SUFFIX); sht_oracle.c picks
   one at run time from the CPU's features. */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <immintrin.h>
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "sht_oracle.h"

#ifndef FAST_W
#error "FAST_W (doubles per vector) must be defined"
#endif
#define W FAST_W
#ifndef SV
#define SV (W == 4 ? 3 : 2)   /* ring vectors per strip */
#endif
#define S (SV * W)            /* rings per strip */
#ifndef KC
#define KC 64                 /* rows of Lambda per block */
#endif
#ifndef XR
#define XR (W == 4 ? 4 : 8)   /* rings per register block (many columns) */
#endif
#ifndef JV
#define JV (W == 4 ? 3 : 2)   /* column vectors per register block */
#endif
_Static_assert(S % XR == 0, "XR must divide the strip");

#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, FAST_SUFFIX)

typedef double vd __attribute__((vector_size(8 * W)));
typedef double vdu __attribute__((vector_size(8 * W), aligned(8)));

/* Broadcast (vbroadcastsd: one load, no shuffle). */
static inline vd bc(double x) {
#if W == 8
  return (vd)_mm512_set1_pd(x);
#else
  return (vd)_mm256_set1_pd(x);
#endif
}

typedef struct {
  int64_t kmin;      /* first row any ring of the strip needs */
  int64_t kinj;      /* first row past every ring's injected seeds */
  int64_t k0[S];
  int64_t ninj[S];   /* seeds injected: 2 (recurrence) or rows (dense) */
  double x2[S] __attribute__((aligned(64)));
  double seed[S][4];
} strip_t;

static void strip_setup(const or_plan *p, int m, int odd, int64_t nk, int r0, strip_t *st) {
  int R = p->R;
  st->kmin = nk;
  st->kinj = 0;
  for (int l = 0; l < S; ++l) {
    int r = r0 + l;
    int64_t k0 = nk, ninj = 0;
    st->x2[l] = 0.0;
    memset(st->seed[l], 0, sizeof(st->seed[l]));
    if (r < R) {
      size_t slot = ((size_t)m * 2 + odd) * R + r;
      int64_t rows = nk - p->k0[slot];
      st->x2[l] = p->x2[r];
      if (rows > 0) {
        k0 = p->k0[slot];
        ninj = rows <= 4 ? rows : 2;
        memcpy(st->seed[l], p->seed + slot * 4, sizeof(st->seed[l]));
      }
    }
    st->k0[l] = k0;
    st->ninj[l] = ninj;
    if (k0 < st->kmin) st->kmin = k0;
    if (k0 + ninj > st->kinj) st->kinj = k0 + ninj;
  }
  if (st->kinj < st->kmin) st->kinj = st->kmin;
}

/* Rows [kc0, kc1) of the strip into pb[(k - kc0) * S + ring]; p0/p1 carry
   P_{k-2}, P_{k-1} per ring from one block to the next.  Row k uses
   aux[3 (k - 2) ..] (src/legendre_transform.c.in:114 auxdata -= 6). */
static void gen_rows(const strip_t *st, const double *aux, int64_t kc0, int64_t kc1,
                     double *pb, double *p0, double *p1) {
  int64_t k = kc0;
  for (; k < kc1 && k < st->kinj; ++k) {
    for (int l = 0; l < S; ++l) {
      int64_t t = k - st->k0[l];
      double v;
      if (t < 0)
        v = 0.0;
      else if (t < st->ninj[l])
        v = st->seed[l][t];
      else {
        const double *ax = aux + 3 * (k - 2);
        v = (st->x2[l] + ax[0]) * ax[1] * p1[l] + ax[2] * p0[l];
      }
      p0[l] = p1[l];
      p1[l] = v;
      pb[(k - kc0) * S + l] = v;
    }
  }
  if (k >= kc1) return;
  vd P0[S / W], P1[S / W], X2[S / W];
  for (int v = 0; v < S / W; ++v) {
    P0[v] = *(const vdu *)(p0 + v * W);
    P1[v] = *(const vdu *)(p1 + v * W);
    X2[v] = *(const vdu *)(st->x2 + v * W);
  }
  double *out = pb + (k - kc0) * S;
  for (; k < kc1; ++k, out += S) {
    const double *ax = aux + 3 * (k - 2);
    vd al = bc(ax[0]), be = bc(ax[1]), ga = bc(ax[2]);
    for (int v = 0; v < S / W; ++v) {
      vd t = (X2[v] + al) * be;
      vd nv = t * P1[v] + ga * P0[v];
      P0[v] = P1[v];
      P1[v] = nv;
      *(vdu *)(out + v * W) = nv;
    }
  }
  for (int v = 0; v < S / W; ++v) {
    *(vdu *)(p0 + v * W) = P0[v];
    *(vdu *)(p1 + v * W) = P1[v];
  }
}

/* Register block of XR rings x NV column vectors over nrows rows of a
   packed column block (row k at pa + k * NV * W): P broadcast, a loaded,
   y[., j ..] += the block's sum. */
#define WM_MM_BLOCK(NV)                                                   \
  for (int x0 = 0; x0 < S; x0 += XR) {                                    \
    vd acc[XR][NV];                                                       \
    for (int x = 0; x < XR; ++x)                                          \
      for (int v = 0; v < NV; ++v) acc[x][v] = bc(0.0);                   \
    const double *pa = ap + j * nk + kc0 * (NV * W);                      \
    const double *pp = pb + x0;                                           \
    for (int64_t k = 0; k < nrows; ++k, pa += NV * W, pp += S) {          \
      vd av[NV];                                                          \
      for (int v = 0; v < NV; ++v) av[v] = *(const vd *)(pa + v * W);     \
      for (int x = 0; x < XR; ++x) {                                      \
        vd b = bc(pp[x]);                                                 \
        for (int v = 0; v < NV; ++v) acc[x][v] += b * av[v];              \
      }                                                                   \
    }                                                                     \
    for (int x = 0; x < XR; ++x)                                          \
      for (int v = 0; v < NV; ++v) {                                      \
        vdu *dst = (vdu *)(y + (x0 + x) * nvecs + j + v * W);             \
        *dst = *dst + acc[x][v];                                          \
      }                                                                   \
  }

/* The column blocks of mm_cols: JV vectors wide, then single vectors, then
   single columns.  Width of the block starting at column j. */
static inline int64_t col_block(int64_t j, int64_t nvecs) {
  int64_t full = nvecs - nvecs % (JV * W);
  if (j < full) return JV * W;
  if (j + W <= nvecs) return W;
  return 1;
}

/* Copy the nk rows of a (row k at a + k * astride) into column blocks, each
   k-major and contiguous: block at column j of width cw holds (k, c) at
   ap[j * nk + k * cw + c].  Rows of a lie 2 * nvecs doubles apart, a power
   of two for power-of-two map counts, which would put a block's rows in a
   few L1 sets; the packed copy streams (the reference packs its A chunks
   for the product, src/legendre_transform.c.in:330-363). */
static void pack_cols(const double *a, int64_t astride, int64_t nk, int64_t nvecs,
                      double *ap) {
  for (int64_t j = 0; j < nvecs;) {
    int64_t cw = col_block(j, nvecs);
    double *dst = ap + j * nk;
    for (int64_t k = 0; k < nk; ++k)
      for (int64_t c = 0; c < cw; ++c) dst[k * cw + c] = a[k * astride + j + c];
    j += cw;
  }
}

/* y[ring][j] += sum_k pb[k - kc0][ring] a[k][j] over rows [kc0, kc0 + nrows)
   of the packed a. */
static void mm_cols(const double *pb, int64_t nrows, const double *ap, int64_t nk,
                    int64_t kc0, int64_t nvecs, double *y) {
  int64_t j = 0;
  for (; j + JV * W <= nvecs; j += JV * W) WM_MM_BLOCK(JV)
  for (; j + W <= nvecs; j += W) WM_MM_BLOCK(1)
  for (; j < nvecs; ++j)
    for (int x = 0; x < S; ++x) {
      const double *pa = ap + j * nk + kc0;
      double acc = 0.0;
      for (int64_t k = 0; k < nrows; ++k) acc += pb[k * S + x] * pa[k];
      y[x * nvecs + j] += acc;
    }
}

/* One map (nvecs == 2): y2[v][c] (ring vector v, column c) += P vectors x
   broadcast a, rings in the vector lanes. */
static void mm_rings2(const double *pb, int64_t nrows, const double *a, int64_t astride,
                      vd *y2) {
  vd acc[S / W][2];
  for (int v = 0; v < S / W; ++v) acc[v][0] = acc[v][1] = bc(0.0);
  for (int64_t k = 0; k < nrows; ++k, a += astride, pb += S) {
    vd ar = bc(a[0]), ai = bc(a[1]);
    for (int v = 0; v < S / W; ++v) {
      vd pv = *(const vdu *)(pb + v * W);
      acc[v][0] += pv * ar;
      acc[v][1] += pv * ai;
    }
  }
  for (int v = 0; v < S / W; ++v) {
    y2[2 * v] += acc[v][0];
    y2[2 * v + 1] += acc[v][1];
  }
}

/* Same contract as or_legendre_stage (sht_oracle.c): work[m][odd][ring][map]
   complex128, or only the sampled m in order when compact. */
void FN(or_legendre_stage_fast)(const or_plan *p, const double *alm, double *work,
                                int m_start, int m_stop, int m_step, int compact) {
  int L = p->lmax, R = p->R, K = p->nmaps;
  int64_t nvecs = 2 * (int64_t)K;
  int nsamp = 0;
  for (int m = m_start; m < m_stop && m <= p->mmax; m += m_step) ++nsamp;
#pragma omp parallel
  {
    double *pb = (double *)aligned_alloc(64, sizeof(double) * KC * S);
    size_t ybytes = (sizeof(double) * S * nvecs + 63) / 64 * 64;
    double *y = (double *)aligned_alloc(64, ybytes);
    double *ap = nvecs == 2 ? NULL
                            : (double *)aligned_alloc(64, (sizeof(double) * (L / 2 + 1) *
                                                               nvecs + 63) / 64 * 64);
    double p0[S] __attribute__((aligned(64))), p1[S] __attribute__((aligned(64)));
    strip_t st;
#pragma omp for schedule(dynamic, 1)
    for (int task = 0; task < 2 * nsamp; ++task) {
      int m = m_start + (task / 2) * m_step, odd = task % 2;
      int64_t nk = or_nk(L, m, odd);
      const double *a = alm + nvecs * ((int64_t)m * (2 * L - m + 3) / 2) + odd * nvecs;
      int64_t astride = 2 * nvecs;
      const double *aux = p->aux + p->aux_off[2 * m + odd];
      size_t mo = compact ? (size_t)(task / 2) : (size_t)m;
      if (ap) pack_cols(a, astride, nk, nvecs, ap);
      for (int r0 = 0; r0 < R; r0 += S) {
        strip_setup(p, m, odd, nk, r0, &st);
        memset(y, 0, sizeof(double) * S * nvecs);
        vd y2[2 * (S / W)];
        for (int v = 0; v < 2 * (S / W); ++v) y2[v] = bc(0.0);
        memset(p0, 0, sizeof(p0));
        memset(p1, 0, sizeof(p1));
        for (int64_t kc0 = st.kmin; kc0 < nk; kc0 += KC) {
          int64_t kc1 = kc0 + KC < nk ? kc0 + KC : nk;
          gen_rows(&st, aux, kc0, kc1, pb, p0, p1);
          if (nvecs == 2)
            mm_rings2(pb, kc1 - kc0, a + kc0 * astride, astride, y2);
          else
            mm_cols(pb, kc1 - kc0, ap, nk, kc0, nvecs, y);
        }
        int nr = R - r0 < S ? R - r0 : S;
        for (int l = 0; l < nr; ++l) {
          double *out = work + 2 * ((mo * 2 + odd) * R + r0 + l) * K;
          if (nvecs == 2) {
            out[0] = y2[2 * (l / W)][l % W];
            out[1] = y2[2 * (l / W) + 1][l % W];
          } else {
            memcpy(out, y + l * nvecs, sizeof(double) * nvecs);
          }
        }
      }
    }
    free(pb);
    free(y);
    free(ap);
  }
}

/* Ring stage fold for the baseline, on a sample of ring pairs: sampled pair
   s is ring r = s * r_step of the northern half and its mirror (the
   reference's perform_backward_ffts_thread, src/wavemoth.c:1083-1176: the
   phase shift by m phi0, the even/odd fold into the top and bottom rings,
   the +m and -m bins wrapped into the half spectrum).  work holds the
   sampled rings only: work[m][odd][s][map] complex.  Pair s writes its top
   ring's half spectra, map-major (nmaps x (n/2 + 1) bins complex, so that
   each map's c2r runs over contiguous bins), from bin soff[s] (in rows of
   nmaps bins) and, unless it is the equator ring, the bottom ring's right
   after. */
void FN(or_ring_fold_sample)(const or_plan *p, const double *work, int phase_shifts,
                             int r_step, int nsamp, const int64_t *soff, double *spec) {
  int N = p->N, K = p->nmaps, M = p->mmax;
  int mid = 2 * N - 1;
#pragma omp parallel for schedule(dynamic, 1)
  for (int s = 0; s < nsamp; ++s) {
    int r = s * r_step;
    int it = mid - r, ib = mid + r;
    int64_t n = p->ring_offsets[ib + 1] - p->ring_offsets[ib];
    int64_t nb = n / 2 + 1;
    double phi0 = phase_shifts ? p->phi0s[ib] : 0.0;
    double *top = spec + 2 * soff[s] * K;
    double *bot = top + 2 * nb * K;
    int two = ib != it;
    memset(top, 0, sizeof(double) * 2 * nb * K * (two ? 2 : 1));
    for (int m = 0; m <= M; ++m) {
      const double *qe = work + 2 * (((size_t)m * 2 + 0) * nsamp + s) * K;
      const double *qo = work + 2 * (((size_t)m * 2 + 1) * nsamp + s) * K;
      double cs = cos(m * phi0), sn = sin(m * phi0);
      int64_t j1 = m % n;
      int64_t j2 = (n - m % n) % n;
      int up1 = j1 <= n / 2, up2 = m != 0 && j2 <= n / 2;
      double *t1 = top + 2 * j1, *t2 = top + 2 * j2;
      double *b1 = bot + 2 * j1, *b2 = bot + 2 * j2;
      for (int k = 0; k < K; ++k) {
        double tr = qe[2 * k] + qo[2 * k], ti = qe[2 * k + 1] + qo[2 * k + 1];
        double br = qe[2 * k] - qo[2 * k], bi = qe[2 * k + 1] - qo[2 * k + 1];
        double ur = tr * cs - ti * sn, ui = tr * sn + ti * cs;
        double vr = br * cs - bi * sn, vi = br * sn + bi * cs;
        int64_t o = 2 * k * nb;
        if (up1) {
          t1[o] += ur;
          t1[o + 1] += ui;
          if (two) {
            b1[o] += vr;
            b1[o + 1] += vi;
          }
        }
        if (up2) {
          t2[o] += ur;
          t2[o + 1] -= ui;
          if (two) {
            b2[o] += vr;
            b2[o + 1] -= vi;
          }
        }
      }
    }
  }
}

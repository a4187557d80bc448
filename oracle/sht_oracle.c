/*
 * CPU oracle for a_lm -> HEALPix map synthesis.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is a plain-C restatement of the
 * reference algorithm (wavemoth) and is used exclusively to check results:
 * by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.
 * The product path (wavemoth_amd/, libwavemoth_hip.so) never links, loads
 * or calls it.
 *
 * What each part restates (paths relative to the reference tree):
 *   or_grid                 src/wavemoth.c:1199-1232 (ring offsets, phi0)
 *   or_ring_z               wavemoth/healpix.py:5-18 (z of the northern rings,
 *                           equator -> pole, i.e. the [::-1] at :13-14)
 *   or_ylm_*                libpshtlight/ylmgen_c.c:419-504, 548-585,
 *                           604-619, 643-732 (normalised Lambda_l^m(theta)
 *                           with 2^90 exponent scaling and the eps cut-off
 *                           `firstl`, including the m_crit/cth_crit skip)
 *   or_auxdata              src/legendre_transform.c.in:723-747
 *   or_legendre_transform   src/legendre_transform.c.in:680-716 (scalar)
 *   or_plan_precompute      wavemoth/lib.pyx:397-487 (per (m, odd) payload:
 *                           Lambda with eps=1e-300, start row at the first
 *                           |Lambda| >= 1e-30, P0/P1 seeds, <=4-row dense
 *                           blocks) restated per column (a one-column strip
 *                           satisfies stripify's constraints, lib.pyx:301-368)
 *   or_legendre_stage       src/wavemoth.c:958-1028 (input indexing, seeds +
 *                           x^2 recurrence per strip) -> work_q[m][odd][ring]
 *   or_ring_spectra         src/wavemoth.c:1083-1191 (even/odd fold, phi0
 *                           phase, aliasing into the c2r half spectrum; the
 *                           c2r itself is numpy.fft.irfft * n in oracle.py)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846

/* ------------------------------------------------------------------ grid */

/* src/wavemoth.c:1199-1232.  phi0s has 4N-1 entries, ring_offsets 4N. */
void or_grid(int N, double *phi0s, int64_t *ring_offsets) {
  int nrings = 4 * N - 1;
  int64_t npix_ring = 0, ipix = 0;
  for (int iring = 0; iring < nrings; ++iring) {
    if (iring <= N - 1) {
      npix_ring += 4;
      phi0s[iring] = OR_PI / (4.0 * (iring + 1));
    } else if (iring > 3 * N - 1) {
      npix_ring -= 4;
      phi0s[iring] = OR_PI / (4.0 * (nrings - iring));
    } else {
      phi0s[iring] = (OR_PI / (4.0 * N)) * (iring % 2);
    }
    ring_offsets[iring] = ipix;
    ipix += npix_ring;
  }
  ring_offsets[nrings] = ipix;
}

/* wavemoth/healpix.py:5-18 with positive_only=True: z of the 2N northern
   rings ordered from the equator (r=0, z=0) towards the pole. */
void or_ring_z(int N, double *z) {
  for (int r = 0; r < 2 * N; ++r) {
    int i = 2 * N - r; /* 1-based ring number counted from the north pole */
    double zi;
    if (i <= N)
      zi = 1.0 - (double)i * (double)i / (3.0 * (double)N * (double)N);
    else
      zi = 4.0 / 3.0 - 2.0 * (double)i / (3.0 * (double)N);
    z[r] = zi;
  }
}

/* ------------------------------------------------ Lambda_l^m (Ylmgen) */

enum { OR_LARGE_EXP2 = 90, OR_MINSCALE = -4, OR_MAXSCALE = 11 };

typedef struct {
  int lmax, mmax;
  double eps, fbig, fsmall;
  double cf[OR_MAXSCALE - OR_MINSCALE + 1];
  double *mfac;   /* log2 of sqrt(prod (2i+1)/(2i)) / sqrt(4 pi), per m */
  double *t1fac;  /* sqrt(4 (l+1)^2 - 1) */
  double *t2fac;  /* 1 / sqrt(j + 1) */
  double *rf0, *rf1; /* recurrence factors for the current m */
  int m_cur;
  int m_crit;
  double cth_crit;
} or_ylm;

or_ylm *or_ylm_new(int lmax, int mmax, double eps) {
  const double inv_sqrt4pi = 0.2820947917738781434740397257803862929220;
  const double inv_ln2 = 1.4426950408889634073599246810018921;
  or_ylm *g = (or_ylm *)calloc(1, sizeof(or_ylm));
  g->lmax = lmax;
  g->mmax = mmax;
  g->eps = eps;
  g->fbig = ldexp(1.0, OR_LARGE_EXP2);
  g->fsmall = ldexp(1.0, -OR_LARGE_EXP2);
  for (int s = 0; s <= OR_MAXSCALE - OR_MINSCALE; ++s)
    g->cf[s] = ldexp(1.0, (s + OR_MINSCALE) * OR_LARGE_EXP2);
  g->mfac = (double *)malloc(sizeof(double) * (mmax + 1));
  g->mfac[0] = 1.0;
  for (int m = 1; m <= mmax; ++m)
    g->mfac[m] = g->mfac[m - 1] * sqrt((2 * m + 1.0) / (2 * m));
  for (int m = 0; m <= mmax; ++m)
    g->mfac[m] = inv_ln2 * log(inv_sqrt4pi * g->mfac[m]);
  g->t1fac = (double *)malloc(sizeof(double) * (lmax + 1));
  for (int l = 0; l <= lmax; ++l)
    g->t1fac[l] = sqrt(4.0 * (l + 1) * (l + 1) - 1.0);
  g->t2fac = (double *)malloc(sizeof(double) * (lmax + mmax + 1));
  for (int j = 0; j <= lmax + mmax; ++j)
    g->t2fac[j] = 1.0 / sqrt(j + 1.0);
  g->rf0 = (double *)malloc(sizeof(double) * (lmax + 1));
  g->rf1 = (double *)malloc(sizeof(double) * (lmax + 1));
  g->m_cur = -1;
  g->m_crit = mmax + 1;
  g->cth_crit = 2.0;
  return g;
}

void or_ylm_free(or_ylm *g) {
  if (!g) return;
  free(g->mfac);
  free(g->t1fac);
  free(g->t2fac);
  free(g->rf0);
  free(g->rf1);
  free(g);
}

/* Forget the m_crit/cth_crit memory (a fresh generator per matrix, as
   wavemoth/legendre.pyx:24-57 creates one per call). */
void or_ylm_reset_crit(or_ylm *g) {
  g->m_crit = g->mmax + 1;
  g->cth_crit = 2.0;
}

static void or_ylm_set_m(or_ylm *g, int m) {
  if (m == g->m_cur) return;
  double f_old = 1.0;
  for (int l = m; l <= g->lmax; ++l) {
    g->rf0[l] = g->t1fac[l] * g->t2fac[l + m] * g->t2fac[l - m];
    g->rf1[l] = g->rf0[l] / f_old;
    f_old = g->rf0[l];
  }
  g->m_cur = m;
}

/* Fills out[0..lmax] with Lambda_l^m(theta); entries with l < firstl are 0.
   Returns firstl (lmax+1 when every value is below eps). */
int or_ylm_column(or_ylm *g, int m, double theta, double *out) {
  const double inv_ln2 = 1.4426950408889634073599246810018921;
  const double ln2 = 0.6931471805599453094172321214581766;
  int lmax = g->lmax;
  double th = theta;
  if (th == 0.0) th = 1e-16;
  if (fabs(th - OR_PI) < 1e-15) th = OR_PI - 1e-15;
  double cth = cos(th), sth = sin(th);
  double logsth = inv_ln2 * log(sth);
  for (int l = 0; l <= lmax; ++l) out[l] = 0.0;

  if (((m >= g->m_crit) && (fabs(cth) >= g->cth_crit)) || ((m > 0) && (sth == 0)))
    return lmax + 1;
  or_ylm_set_m(g, m);
  const double *rf0 = g->rf0, *rf1 = g->rf1;

  double logval = g->mfac[m];
  if (m > 0) logval += m * logsth;
  int scale = (int)(logval / OR_LARGE_EXP2) - OR_MINSCALE;
  double corfac = (scale < 0) ? 0.0 : g->cf[scale];
  double lam_1 = 0.0;
  double lam_2 = exp(ln2 * (logval - (scale + OR_MINSCALE) * OR_LARGE_EXP2));
  if (m & 1) lam_2 = -lam_2;

  int l = m;
  if (scale < 0) {
    for (;;) {
      if (++l > lmax) break;
      lam_1 = cth * lam_2 * rf0[l - 1] - lam_1 * rf1[l - 1];
      if (++l > lmax) break;
      lam_2 = cth * lam_1 * rf0[l - 1] - lam_2 * rf1[l - 1];
      if (fabs(lam_2) > g->fbig) {
        while (fabs(lam_2) > g->fbig) {
          lam_1 *= g->fsmall;
          lam_2 *= g->fsmall;
          ++scale;
        }
        corfac = (scale < 0) ? 0.0 : g->cf[scale];
        if (scale >= 0) break;
      }
    }
  }
  lam_1 *= corfac;
  lam_2 *= corfac;

  if (l <= lmax) {
    for (;;) {
      if (fabs(lam_2) > g->eps) break;
      if (++l > lmax) break;
      lam_1 = cth * lam_2 * rf0[l - 1] - lam_1 * rf1[l - 1];
      if (fabs(lam_1) > g->eps) {
        double x = lam_1;
        lam_1 = lam_2;
        lam_2 = x;
        break;
      }
      if (++l > lmax) break;
      lam_2 = cth * lam_1 * rf0[l - 1] - lam_2 * rf1[l - 1];
    }
  }
  int firstl = l;
  if (l > lmax) {
    g->m_crit = m;
    g->cth_crit = fabs(cth);
    return lmax + 1;
  }
  for (;;) {
    out[l] = lam_2;
    if (++l > lmax) break;
    lam_1 = cth * lam_2 * rf0[l - 1] - lam_1 * rf1[l - 1];
    out[l] = lam_1;
    if (++l > lmax) break;
    lam_2 = cth * lam_1 * rf0[l - 1] - lam_2 * rf1[l - 1];
  }
  return firstl;
}

/* Convenience: one column with a fresh generator. */
int or_lambda_column(int m, int lmax, double theta, double eps, double *out) {
  or_ylm *g = or_ylm_new(lmax, lmax, eps);
  int fl = or_ylm_column(g, m, theta, out);
  or_ylm_free(g);
  return fl;
}

/* ------------------------------------------- x^2 recurrence (reference) */

/* src/legendre_transform.c.in:723-747.  aux has 3*(nk-2) entries. */
void or_auxdata(int64_t m, int64_t lmin, int64_t nk, double *aux) {
  double cp = 0, cpp = 0, dp = 0;
  int64_t l = lmin;
  for (int64_t k = 0; k < nk; ++k, l += 2) {
    double x = (double)((l - m + 1) * (l - m + 2) * (l + m + 1) * (l + m + 2));
    double y = (double)((2 * l + 1) * (2 * l + 3) * (2 * l + 3) * (2 * l + 5));
    double c = sqrt(x / y);
    x = (double)(2 * l * (l + 1) - 2 * m * m - 1);
    y = (double)((2 * l - 1) * (2 * l + 3));
    double d = x / y;
    if (k >= 2) {
      aux[3 * (k - 2) + 0] = -dp;
      aux[3 * (k - 2) + 1] = 1.0 / cp;
      aux[3 * (k - 2) + 2] = -cpp / cp;
    }
    dp = d;
    cpp = cp;
    cp = c;
  }
}

/* src/legendre_transform.c.in:680-716.  `a` is in every-other-row layout:
   row k of the transform is a[2*k*nvecs + j]. */
void or_legendre_transform(int64_t nx, int64_t nk, int64_t nvecs, const double *a,
                           double *y, const double *x_squared, const double *aux,
                           const double *P, const double *Pp1) {
  for (int64_t i = 0; i < nx; ++i) {
    int64_t k = 0;
    double pp = P[i], p = Pp1[i];
    for (int64_t j = 0; j < nvecs; ++j) y[i * nvecs + j] = pp * a[2 * k * nvecs + j];
    ++k;
    for (int64_t j = 0; j < nvecs; ++j) y[i * nvecs + j] += p * a[2 * k * nvecs + j];
    ++k;
    for (; k < nk; ++k) {
      double alpha = aux[3 * (k - 2) + 0];
      double beta = aux[3 * (k - 2) + 1];
      double gamma = aux[3 * (k - 2) + 2];
      double pv = (x_squared[i] + alpha) * beta * p + gamma * pp;
      pp = p;
      p = pv;
      for (int64_t j = 0; j < nvecs; ++j) y[i * nvecs + j] += pv * a[2 * k * nvecs + j];
    }
  }
}

/* --------------------------------------- plan: precompute + execute */

#include "sht_oracle.h"

/* Precompute the per-(m, odd, ring) seeds: the restated resource file.
   Only m in [m_start, m_stop) with stride m_step are filled (bounded
   samples for CPU timing); the rest stay empty. */
or_plan *or_plan_new(int N, int lmax, int mmax, int nmaps, int m_start, int m_stop,
                     int m_step) {
  or_plan *p = (or_plan *)calloc(1, sizeof(or_plan));
  p->N = N;
  p->lmax = lmax;
  p->mmax = mmax;
  p->nmaps = nmaps;
  int R = p->R = 2 * N;
  p->z = (double *)malloc(sizeof(double) * R);
  p->x2 = (double *)malloc(sizeof(double) * R);
  or_ring_z(N, p->z);
  for (int r = 0; r < R; ++r) {
    /* lib.pyx:382,445: thetas from arccos(z), x_squared = cos(theta)^2 */
    double c = cos(acos(p->z[r]));
    p->x2[r] = c * c;
  }
  p->phi0s = (double *)malloc(sizeof(double) * (4 * N - 1));
  p->ring_offsets = (int64_t *)malloc(sizeof(int64_t) * 4 * N);
  or_grid(N, p->phi0s, p->ring_offsets);
  size_t nslots = (size_t)(mmax + 1) * 2 * R;
  p->k0 = (int32_t *)malloc(sizeof(int32_t) * nslots);
  p->seed = (double *)calloc(nslots * 4, sizeof(double));
  for (size_t s = 0; s < nslots; ++s) p->k0[s] = -1;
  p->aux_off = (int64_t *)malloc(sizeof(int64_t) * ((mmax + 1) * 2 + 1));
  int64_t off = 0;
  for (int m = 0; m <= mmax; ++m)
    for (int odd = 0; odd < 2; ++odd) {
      p->aux_off[2 * m + odd] = off;
      int64_t nk = or_nk(lmax, m, odd);
      off += nk > 2 ? 3 * (nk - 2) : 0;
    }
  p->aux_off[2 * (mmax + 1)] = off;
  p->aux = (double *)malloc(sizeof(double) * (off > 0 ? off : 1));

#pragma omp parallel
  {
    or_ylm *g = or_ylm_new(lmax, mmax, 1e-300);
    double *col = (double *)malloc(sizeof(double) * (lmax + 1));
#pragma omp for schedule(dynamic, 1)
    for (int m = m_start; m < m_stop; m += m_step) {
      if (m > mmax) continue;
      for (int odd = 0; odd < 2; ++odd) {
        int64_t nk = or_nk(lmax, m, odd);
        if (nk > 2) or_auxdata(m, m + odd, nk, p->aux + p->aux_off[2 * m + odd]);
      }
      or_ylm_reset_crit(g);
      for (int r = 0; r < R; ++r) {
        or_ylm_column(g, m, acos(p->z[r]), col);
        for (int odd = 0; odd < 2; ++odd) {
          int64_t nk = or_nk(lmax, m, odd);
          size_t slot = ((size_t)m * 2 + odd) * R + r;
          int64_t k0 = nk;
          for (int64_t k = 0; k < nk; ++k)
            if (fabs(col[m + odd + 2 * k]) >= 1e-30) { k0 = k; break; }
          p->k0[slot] = (int32_t)k0;
          double *sd = p->seed + slot * 4;
          int64_t rows = nk - k0;
          for (int64_t t = 0; t < 4 && t < rows; ++t) sd[t] = col[m + odd + 2 * (k0 + t)];
        }
      }
    }
    free(col);
    or_ylm_free(g);
  }
  return p;
}

void or_plan_free(or_plan *p) {
  if (!p) return;
  free(p->z);
  free(p->x2);
  free(p->phi0s);
  free(p->ring_offsets);
  free(p->k0);
  free(p->seed);
  free(p->aux);
  free(p->aux_off);
  free(p);
}

/* Legendre stage (src/wavemoth.c:1011-1028 + 958-1009 + the scalar kernel).
   alm: complex128 m-major [nalm][nmaps] viewed as doubles.
   work: complex128 [M+1][2][R][nmaps] (the reference's get_work layout,
   wavemoth/lib.pyx:158-179).  Only m in [m_start, m_stop) step m_step. */
void or_legendre_stage(const or_plan *p, const double *alm, double *work, int m_start,
                       int m_stop, int m_step, int compact) {
  int L = p->lmax, R = p->R, K = p->nmaps;
  int64_t nvecs = 2 * K;
#pragma omp parallel
  {
    double *y = (double *)malloc(sizeof(double) * nvecs);
#pragma omp for schedule(dynamic, 1)
    for (int m = m_start; m < m_stop; m += m_step) {
      if (m > p->mmax) continue;
      const double *input_m = alm + nvecs * ((int64_t)m * (2 * L - m + 3) / 2);
      for (int odd = 0; odd < 2; ++odd) {
        int64_t nk = or_nk(L, m, odd);
        const double *a = input_m + odd * nvecs; /* row k at a[2*k*nvecs] */
        const double *aux = p->aux + p->aux_off[2 * m + odd];
        for (int r = 0; r < R; ++r) {
          size_t slot = ((size_t)m * 2 + odd) * R + r;
          /* compact: work holds only the sampled m, in sampling order */
          size_t mo = compact ? (size_t)((m - m_start) / m_step) : (size_t)m;
          double *out = work + 2 * ((mo * 2 + odd) * R + r) * K;
          int64_t k0 = p->k0[slot];
          const double *sd = p->seed + slot * 4;
          int64_t rows = nk - k0;
          if (rows <= 0) {
            for (int64_t j = 0; j < nvecs; ++j) out[j] = 0.0;
          } else if (rows <= 4) {
            /* dense block (dgemm path, src/wavemoth.c:968-991) */
            for (int64_t j = 0; j < nvecs; ++j) {
              double acc = 0.0;
              for (int64_t t = 0; t < rows; ++t) acc += sd[t] * a[2 * (k0 + t) * nvecs + j];
              out[j] = acc;
            }
          } else {
            or_legendre_transform(1, rows, nvecs, a + 2 * k0 * nvecs, y, p->x2 + r,
                                  aux + 3 * k0, sd, sd + 1);
            for (int64_t j = 0; j < nvecs; ++j) out[j] = y[j];
          }
        }
      }
    }
    free(y);
  }
}

/* Ring stage up to the c2r (src/wavemoth.c:1083-1191): for each ring of
   the map, fill its half spectrum X[0..n/2] per map, ready for an
   unnormalised c2r.  spec is complex128 laid out ring by ring (north to
   south, ring iring at spec_off[iring]), each bin holding nmaps complex. */
void or_ring_spectra(const or_plan *p, const double *work, int phase_shifts,
                     const int64_t *spec_off, double *spec) {
  int N = p->N, R = p->R, K = p->nmaps, M = p->mmax;
  int mid = 2 * N - 1;
#pragma omp parallel for schedule(dynamic, 4)
  for (int r = 0; r < R; ++r) {
    int it = mid - r, ib = mid + r;
    int64_t n = p->ring_offsets[ib + 1] - p->ring_offsets[ib];
    double phi0 = phase_shifts ? p->phi0s[ib] : 0.0;
    double *top = spec + 2 * spec_off[it] * K;
    double *bot = spec + 2 * spec_off[ib] * K;
    int64_t nb = n / 2 + 1;
    memset(top, 0, sizeof(double) * 2 * nb * K);
    if (ib != it) memset(bot, 0, sizeof(double) * 2 * nb * K);
    for (int m = 0; m <= M; ++m) {
      const double *qe = work + 2 * (((size_t)m * 2 + 0) * R + r) * K;
      const double *qo = work + 2 * (((size_t)m * 2 + 1) * R + r) * K;
      double cs = cos(m * phi0), sn = sin(m * phi0);
      int64_t j1 = m % n;
      int64_t j2 = (n - m) % n;
      if (j2 < 0) j2 += n;
      for (int k = 0; k < K; ++k) {
        double tr = qe[2 * k] + qo[2 * k], ti = qe[2 * k + 1] + qo[2 * k + 1];
        double br = qe[2 * k] - qo[2 * k], bi = qe[2 * k + 1] - qo[2 * k + 1];
        double t1r = tr * cs - ti * sn, t1i = tr * sn + ti * cs;
        double b1r = br * cs - bi * sn, b1i = br * sn + bi * cs;
        if (j1 <= n / 2) {
          top[2 * (j1 * K + k)] += t1r;
          top[2 * (j1 * K + k) + 1] += t1i;
          if (ib != it) {
            bot[2 * (j1 * K + k)] += b1r;
            bot[2 * (j1 * K + k) + 1] += b1i;
          }
        }
        if (m != 0 && j2 <= n / 2) {
          top[2 * (j2 * K + k)] += t1r;
          top[2 * (j2 * K + k) + 1] -= t1i;
          if (ib != it) {
            bot[2 * (j2 * K + k)] += b1r;
            bot[2 * (j2 * K + k) + 1] -= b1i;
          }
        }
      }
    }
  }
}

/* Exact count of (l, m, ring-pair) triples at or after each column's
   start row (both parities), i.e. the trimmed work T_eps of the plan. */
int64_t or_plan_trimmed_triples(const or_plan *p, int m_start, int m_stop, int m_step) {
  int64_t total = 0;
  for (int m = m_start; m < m_stop && m <= p->mmax; m += m_step)
    for (int odd = 0; odd < 2; ++odd) {
      int64_t nk = or_nk(p->lmax, m, odd);
      for (int r = 0; r < p->R; ++r) {
        int64_t k0 = p->k0[((size_t)m * 2 + odd) * p->R + r];
        if (k0 >= 0 && k0 < nk) total += nk - k0;
      }
    }
  return total;
}

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Vector width the CPU supports for the vectorised Legendre stage: 8
   (AVX-512F), 4 (AVX2 + FMA) or 0 (neither). */
int or_fast_width(void) {
  __builtin_cpu_init();
  if (__builtin_cpu_supports("avx512f")) return 8;
  if (__builtin_cpu_supports("avx2") && __builtin_cpu_supports("fma")) return 4;
  return 0;
}

/* The vectorised Legendre stage of sht_fast.c at the given width (0: the
   widest the CPU has).  Returns the width run, or 0 if none is available. */
int or_legendre_stage_fast(const or_plan *p, const double *alm, double *work, int m_start,
                           int m_stop, int m_step, int compact, int width) {
  int have = or_fast_width();
  if (width == 0) width = have;
  if (width > have) return 0;
  if (width == 8)
    or_legendre_stage_fast_w8(p, alm, work, m_start, m_stop, m_step, compact);
  else if (width == 4)
    or_legendre_stage_fast_w4(p, alm, work, m_start, m_stop, m_step, compact);
  else
    return 0;
  return width;
}

/* The sampled ring-stage fold of sht_fast.c at the given width (0: the
   widest the CPU has).  Returns the width run, or 0 if none is available. */
int or_ring_fold_sample(const or_plan *p, const double *work, int phase_shifts, int r_step,
                        int nsamp, const int64_t *soff, double *spec, int width) {
  int have = or_fast_width();
  if (width == 0) width = have;
  if (width > have) return 0;
  if (width == 8)
    or_ring_fold_sample_w8(p, work, phase_shifts, r_step, nsamp, soff, spec);
  else if (width == 4)
    or_ring_fold_sample_w4(p, work, phase_shifts, r_step, nsamp, soff, spec);
  else
    return 0;
  return width;
}

/* Plan of the CPU oracle (test infrastructure): shared by sht_oracle.c
   (the oracle that results are compared with) and sht_fast.c (the CPU
   baseline's restatement). */
#ifndef SHT_ORACLE_H_
#define SHT_ORACLE_H_

#include <stdint.h>

typedef struct {
  int N, lmax, mmax, nmaps, R;
  double *z;      /* [R] */
  double *x2;     /* [R] */
  double *phi0s;  /* [4N-1] */
  int64_t *ring_offsets; /* [4N] */
  /* per (m, odd, ring): first row k0 used, and up to 4 values:
     P0, P1 (recurrence) or the dense Lambda rows when <= 4 rows remain. */
  int32_t *k0;
  double *seed;   /* [(M+1)*2*R*4] */
  double *aux;    /* per (m, odd): 3*(nk-2) entries, offsets in aux_off */
  int64_t *aux_off;
} or_plan;

static inline int64_t or_nk(int L, int m, int odd) {
  int d = L - m - odd;
  return d < 0 ? 0 : (int64_t)(d / 2 + 1);
}

/* Vectorised Legendre stage (sht_fast.c, built once per vector width). */
void or_legendre_stage_fast_w4(const or_plan *p, const double *alm, double *work, int m_start,
                               int m_stop, int m_step, int compact);
void or_legendre_stage_fast_w8(const or_plan *p, const double *alm, double *work, int m_start,
                               int m_stop, int m_step, int compact);
void or_ring_fold_sample_w4(const or_plan *p, const double *work, int phase_shifts, int r_step,
                            int nsamp, const int64_t *soff, double *spec);
void or_ring_fold_sample_w8(const or_plan *p, const double *work, int phase_shifts, int r_step,
                            int nsamp, const int64_t *soff, double *spec);

#endif

/*
 * Test-only driver around the reference's own Ylmgen (libpshtlight), built
 * from the reference sources into oracle/_ref/ by oracle/Makefile.  It
 * fills a (nth, lmax+1) matrix exactly the way
 * wavemoth/legendre.pyx:24-57 (compute_normalized_associated_legendre)
 * does: one generator per call, theta rows in order, zeros before firstl.
 * Used only by tests/ to pin the oracle's restatement of Ylmgen.
 */
#include "ylmgen_c.h"

int ref_lambda_matrix(int m, int lmax, const double *theta, int nth, double eps,
                      double *out, int *firstl) {
  Ylmgen_C ctx;
  if (lmax < m) return -1;
  Ylmgen_init(&ctx, lmax, lmax, 0, 0, eps);
  Ylmgen_set_theta(&ctx, theta, nth);
  for (int row = 0; row < nth; ++row) {
    Ylmgen_prepare(&ctx, row, m);
    Ylmgen_recalc_Ylm(&ctx);
    int fl = ctx.firstl[0];
    firstl[row] = fl;
    for (int l = 0; l <= lmax; ++l)
      out[(long)row * (lmax + 1) + l] = (l >= m && l >= fl) ? ctx.ylm[l] : 0.0;
  }
  Ylmgen_destroy(&ctx);
  return 0;
}

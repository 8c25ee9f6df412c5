// legendre_mixed.hip -- split-precision Legendre contractions (placeholder until built).
#include "dsht_internal.hpp"

namespace dsht {

void mixed_free(dsht_plan* p) { (void)p; }

int mixed_run_fwd(dsht_plan* p, int op, const double* in0, const double* in1, long ldg, double* out0,
                  double* out1, long sl, long sc, int ncols, const double* diag0, const double* diag1,
                  double scale, const double* cw, hipStream_t s) {
  return fail(DSHT_EUNSUPPORTED, "precision form %d not available", p->prec);
}
int mixed_run_inv(dsht_plan* p, int op, const double* in0, const double* in1, long sl, long sc,
                  double* out0, double* out1, long ldg, int ncols, const double* pro, const double* rs,
                  hipStream_t s) {
  return fail(DSHT_EUNSUPPORTED, "precision form %d not available", p->prec);
}

}  // namespace dsht

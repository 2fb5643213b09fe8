// Newton + FGMRES + block-triangular preconditioner (placeholder until the
// Krylov layer lands).
#include <cstring>
#include "mhd_internal.h"

void mhd_solver_free(mhd_ctx *c) { (void)c; }

extern "C" {

void mhd_default_solver_opts(mhd_solver_opts *o) {
  std::memset(o, 0, sizeof(*o));
  o->max_newton = 20;
  o->restart = 15;
  o->max_krylov = 15;
  o->precision = 0;
  o->rtol = 1e-6;
  o->atol = 1e-8;
  o->ls_rtol = 1e-7;
  o->inner_iters = 4;
}

int mhd_newton_solve(mhd_ctx *c, double *x, const mhd_solver_opts *o, mhd_solve_info *info) {
  (void)x; (void)o; (void)info;
  if (c) c->err = "mhd_newton_solve: not available in this build";
  return -5;
}

}

// Runs the library's host symbolic stages (gdx_symbolic.cpp) on host-only
// problems, built with AddressSanitizer and UndefinedBehaviorSanitizer by
// tests/test_sanitized_builds.py.  Host-only problems never reach the device
// stages, which this build leaves out (the definitions below only satisfy the
// linker and would fail loudly if reached).
#include <cstdio>
#include <cstdlib>
#include <string>

#include "gdx_internal.h"

namespace gdx {
int numbering_device(gdx_problem *, const std::vector<int>[3], const int[27][3], int, const int (*)[3], int,
                     const int (*)[3], int) {
  std::abort();
}
int structure_device(gdx_problem *) { std::abort(); }
int build_tiles_device(gdx_problem *, const int[3]) { std::abort(); }
int host_fe(gdx_problem *) { std::abort(); }
void set_error(const std::string &) {}
}  // namespace gdx

struct Case {
  int dim, order, strategy, dir_kind, op_kind;
  long long cells[3];
  int parts[3];
};

static int run(const Case &c) {
  gdx_problem p;
  p.ctx = nullptr;
  gdx_problem_desc &d = p.desc;
  d.dim = c.dim;
  d.order = c.order;
  d.qdeg = 2 * c.order;
  d.strategy = c.strategy;
  for (int k = 0; k < 3; ++k) {
    d.cells[k] = k < c.dim ? c.cells[k] : 1;
    d.parts[k] = k < c.dim ? c.parts[k] : 1;
    d.domain[2 * k] = 0.0;
    d.domain[2 * k + 1] = 1.0;
  }
  d.coef_const = 1.0;
  d.rhs_kind = GDX_RHS_CONST;
  d.rhs_const = 1.0;
  d.dir_kind = c.dir_kind;
  d.op_kind = c.op_kind;
  d.flux_p = 3.0;
  d.flux_beta = 1.0;
  d.lame_lambda = 0.5;
  d.lame_mu = 0.4;
  p.D = c.dim;
  p.K = c.order;
  p.NQ = c.order + 1;
  p.ncomp = c.op_kind == GDX_OP_ELASTICITY ? 3 : 1;
  p.nparts = d.parts[0] * d.parts[1] * d.parts[2];
  const int rc = gdx::symbolic(&p);
  if (rc) return rc;
  gdx::fill_kparams(&p);
  // the tile rebuild a kernel change makes
  const int R[3] = {8, 8, c.dim == 3 ? 4 : 1};
  gdx::build_tiles(&p, R);
  long long nnz = 0, rows = 0;
  for (const gdx::Part &P : p.P) {
    nnz += P.nnz;
    rows += P.nrows;
    if ((long long)P.rowptr.size() != P.nrows + 1 || P.rowptr[P.nrows] != P.nnz) return 100;
    for (long long i = 0; i < P.nnz; ++i)
      if (P.colidx[i] < 0 || P.colidx[i] >= P.ncols) return 101;
  }
  if (nnz != p.nnz_total || rows != p.rows_total || p.npat <= 0) return 102;
  std::printf("case dim %d order %d strategy %d dir %d op %d: %lld rows, %lld entries, %d patterns\n", c.dim, c.order,
              c.strategy, c.dir_kind, c.op_kind, rows, nnz, p.npat);
  return 0;
}

int main() {
  const Case cases[] = {
      {2, 1, GDX_SUB_ASSEMBLED_ROWS, 1, GDX_OP_DIFFUSION, {9, 7, 1}, {2, 3, 1}},
      {2, 2, GDX_FULLY_ASSEMBLED_ROWS, 1, GDX_OP_DIFFUSION, {6, 5, 1}, {2, 2, 1}},
      {3, 1, GDX_SUB_ASSEMBLED_ROWS, 1, GDX_OP_DIFFUSION, {7, 6, 5}, {3, 2, 2}},
      {3, 2, GDX_SUB_ASSEMBLED_ROWS, 1, GDX_OP_DIFFUSION, {6, 5, 4}, {2, 2, 2}},
      {3, 2, GDX_FULLY_ASSEMBLED_ROWS, 0, GDX_OP_DIFFUSION, {5, 4, 6}, {2, 1, 3}},
      {3, 1, GDX_SUB_ASSEMBLED_ROWS, 1, GDX_OP_FLUX, {8, 8, 8}, {2, 2, 1}},
      {3, 2, GDX_SUB_ASSEMBLED_ROWS, GDX_DIR_CLAMPED_X0, GDX_OP_ELASTICITY, {6, 5, 4}, {3, 2, 2}},
      {3, 1, GDX_FULLY_ASSEMBLED_ROWS, 0, GDX_OP_ELASTICITY, {5, 6, 4}, {2, 2, 2}},
  };
  for (const Case &c : cases) {
    const int rc = run(c);
    if (rc) {
      std::printf("case dim %d order %d failed: %d\n", c.dim, c.order, rc);
      return 1;
    }
  }
  std::printf("all host symbolic cases clean\n");
  return 0;
}

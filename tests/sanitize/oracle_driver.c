/* Runs the C oracle (oracle/gdx_oracle.c, included below as one translation
   unit) on a set of cases, built with AddressSanitizer and
   UndefinedBehaviorSanitizer by tests/test_sanitized_builds.py: fresh
   assembly, in-place re-assembly, both strategies, the flux operator at a
   state and linear elasticity, with one and with several threads. */
#include "gdx_oracle.c"

#include <stdio.h>

typedef struct {
  int dim, order, strategy, dir_kind, op_kind;
  i64 cells[3];
  int parts[3];
} ocase;

static int run(const ocase *c, int threads) {
  orc_desc d;
  memset(&d, 0, sizeof d);
  d.dim = c->dim;
  d.order = c->order;
  d.qdeg = 2 * c->order;
  d.strategy = c->strategy;
  for (int k = 0; k < 3; ++k) {
    d.cells[k] = k < c->dim ? c->cells[k] : 1;
    d.parts[k] = k < c->dim ? c->parts[k] : 1;
    d.domain[2 * k] = 0.0;
    d.domain[2 * k + 1] = 1.0;
  }
  double modes[10] = {1, 2, 0, 0.3, 0.1, 0, 1, 3, -0.2, 0.5};
  d.nmodes = c->op_kind == 0 ? 2 : 0;
  d.modes = modes;
  d.coef_const = 1.0;
  d.rhs_const = 1.0;
  d.dir_kind = c->dir_kind;
  d.dir_coef[1] = 1.0;
  d.op_kind = c->op_kind;
  d.flux_p = 3.0;
  d.flux_alpha = 1e-2;
  d.flux_beta = 1.0;
  d.lame_lambda = 0.5;
  d.lame_mu = 0.4;
  d.body_force[2] = -1.0;
  i64 nn = 1;
  for (int k = 0; k < c->dim; ++k) nn *= c->order * c->cells[k] + 1;
  double *state = NULL;
  if (c->op_kind == 1) {
    state = malloc(nn * sizeof(double));
    for (i64 i = 0; i < nn; ++i) state[i] = 0.01 * (double)(i % 17);
    d.state = state;
  }
  orc_set_threads(threads);
  void *h = orc_run(&d, 2);
  free(state);
  if (!h) return 1;
  const int np = orc_nparts(h);
  i64 total = 0;
  for (int r = 0; r < np; ++r) {
    i64 s[8];
    orc_part_sizes(h, r, s);
    total += s[0];
  }
  printf("case dim %d order %d strategy %d op %d threads %d: %d parts, %lld local dofs\n", c->dim, c->order,
         c->strategy, c->op_kind, threads, np, (long long)total);
  orc_free(h);
  return 0;
}

int main(void) {
  const ocase cases[] = {
      {2, 1, 0, 1, 0, {9, 7, 1}, {2, 3, 1}},
      {2, 2, 1, 1, 0, {6, 5, 1}, {2, 2, 1}},
      {3, 1, 0, 1, 0, {7, 6, 5}, {3, 2, 2}},
      {3, 2, 0, 1, 0, {5, 4, 4}, {2, 2, 2}},
      {3, 2, 1, 0, 0, {5, 4, 6}, {2, 1, 3}},
      {3, 1, 0, 1, 1, {6, 6, 6}, {2, 2, 1}},
      {3, 2, 0, 2, 2, {5, 4, 4}, {3, 2, 2}},
      {3, 1, 1, 0, 2, {5, 6, 4}, {2, 2, 2}},
  };
  for (size_t i = 0; i < sizeof cases / sizeof cases[0]; ++i)
    for (int threads = 1; threads <= 3; threads += 2)
      if (run(&cases[i], threads)) {
        printf("case %zu failed\n", i);
        return 1;
      }
  printf("all oracle cases clean\n");
  return 0;
}

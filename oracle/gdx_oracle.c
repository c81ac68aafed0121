/*
 * gdx_oracle.c -- CPU ORACLE (test infrastructure only).
 *
 * A plain-C restatement of GridapDistributed's assembly path for Lagrangian
 * Q1/Q2 spaces on a partitioned Cartesian mesh.  It is the checker for the HIP
 * library and the CPU baseline timed beside it; nothing in the product links
 * or calls it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.
 *
 * What it restates (reference files under GridapDistributed.jl v0.4.17):
 *   partition          src/PArraysExtras.jl:5-84  (_uniform_partition,
 *                      _block_with_constant_size, _local_range)
 *   local model        src/Geometry.jl:390-411 (per-part CartesianDiscreteModel
 *                      over the owned+ghost cell box)
 *   dof numbering      Gridap 0.20 conforming numbering [ext, not in the
 *                      reference tree]: faces dimension-major (vertices
 *                      lexicographic, edges/faces first-encounter over cells in
 *                      the polytope's local face order, cell interiors), free
 *                      dofs numbered in that order, Dirichlet ones skipped.
 *   gids               src/FESpaces.jl:139-261 (generate_gids: owner = max
 *                      owner of touching local cells, ghost-cell fix-up,
 *                      rank-blocked global ids, ghost gids fetched from owners)
 *   triangulation      src/Geometry.jl:660-663,797-847 (owned cells for
 *                      SubAssembledRows, owned+ghost for FullyAssembledRows)
 *   cell integrals     Gridap numeric loop [ext]: Q1 geometric map, Gauss
 *                      quadrature with div(degree,2)+1 points per direction,
 *                      kappa grad(u).grad(v) and f v with Dirichlet lifting
 *                      b_c -= A_c u_D (src/FESpaces.jl:703-715)
 *   COO + compression  src/Algebra.jl:410-418,849-857 (COO counter/alloc),
 *                      :593-646 (_sa_create_from_nz_with_callback),
 *                      :1048-1131 (assemble_coo_with_column_owner!),
 *                      :1191-1257 (_setup_prange_with_ghosts, first-encounter
 *                      ghost columns), :919-962 (ghost rows sorted by local id)
 *   FullyAssembledRows src/Algebra.jl:537-579, src/FESpaces.jl:806-816
 *   RHS                src/Algebra.jl:720-753,786-789,870-907 (_rhs_callback +
 *                      assemble! : ghost values added to owners, then zeroed)
 *   in-place numeric   src/Algebra.jl:306-393,530-535 (LocalView +
 *                      create_from_nz(::PSparseMatrix) = assemble!)
 *
 * Parity status.  Pinned by closed forms, invariants, an independent numpy
 * restatement of the cell integrals (oracle/micro.py) and hand enumeration (the
 * tests under tests/test_oracle_*.py): values, sparsity keyed by lattice node,
 * partition, owners and gids, the Q2 numbering rule as written here, mapped
 * geometry, the reaction term, elasticity, the flux operator.  Parity
 * unpinned (no Gridap or PartitionedArrays to compare with): whether Gridap's
 * own Q2 edge/face enumeration is the first-encounter rule restated here,
 * PartitionedArrays' assembly_neighbors order (senders ascending by rank
 * here), and explicit-zero retention.  These fix the numbering and order of the
 * outputs only.
 * All indices in the outputs are 0-based.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef long long i64;

typedef struct {
  int dim, order, qdeg, strategy; /* strategy 0 SubAssembledRows, 1 FullyAssembledRows */
  i64 cells[3];
  int parts[3];
  double domain[6];
  int nmodes;            /* kappa = coef_const * exp(sum_m a_m cos(2 pi k_m.x + phi_m)) */
  const double *modes;   /* 5 per mode: k1 k2 k3 a phi */
  double coef_const;
  int rhs_kind;          /* 0: f = rhs_const; 1: f = -div(kappa grad uD) for linear uD */
  double rhs_const;
  int dir_kind;          /* 0: no Dirichlet; 1: whole boundary, uD = c0 + c1 x + c2 y + c3 z */
  double dir_coef[4];
  /* op_kind 1: nonlinear flux operator at a state u (test/PLaplacianTests.jl:26-31 form)
       residual r(u,v) = int( grad(v).sigma(grad(u)) - v f ) dOmega, f = rhs_const
       jacobian j(u,du,v) = int( grad(v).dsigma(grad(du),grad(u)) ) dOmega
       sigma(g) = (alpha + beta |g|^2)^((p-2)/2) g
     state: u at every global lattice node (x fastest), NULL = 0; Dirichlet nodes take uD */
  int op_kind;
  double flux_p, flux_alpha, flux_beta;
  const double *state;
  /* kappa per cell (x fastest), NULL = the Fourier form above; f = rhs_const */
  const double *cell_kappa;
  /* op_kind 2: linear elasticity, a vector-valued Lagrangian space with dim
     components per node (cell dofs component-major, each node's components
     numbered consecutively):
       a(u,v) = int( eps(v) : (lambda tr(eps(u)) I + 2 mu eps(u)) ) dOmega
       l(v)   = int( v . body_force ) dOmega
     with dir_kind 0 (no Dirichlet) or 2 (clamped: u = 0 at every node of the
     face x = x0, its edges and vertices included). */
  double lame_lambda, lame_mu;
  double body_force[3];
  /* op_kind 2: (lambda, mu) per cell (x fastest, 2 values a cell), NULL = the
     constants above: sigma = lambda_c tr(eps) I + 2 mu_c eps on cell c */
  const double *cell_lame;
  /* mapped Cartesian grid: coordinates of every mesh vertex (dim values each,
     vertices x fastest over (cells_0 + 1) x (cells_1 + 1) [x (cells_2 + 1)]),
     NULL = the uniform box.  Cells take the multilinear (Q1) map of their
     vertices; a Q2 node sits at that map's image of its reference position */
  const double *vertex_coords;
  /* op_kind 0 with weights_set: a(u,v) = int( diffusion_weight kappa grad(u).grad(v)
     + reaction u v ); without, diffusion_weight = 1 and reaction = 0 */
  int weights_set;
  double diffusion_weight, reaction;
} orc_desc;

typedef struct {
  int rank, pc[3];
  i64 own_lo[3], own_hi[3], loc_lo[3], loc_hi[3];
  i64 lnn[3], n_lnodes;
  i64 *lnode_dof;  /* local node -> free local dof (>=0) or -1-dirichlet index */
  i64 n_ldofs, n_ddofs;
  i64 *ldof_lnode; /* free local dof -> local node */
  int *ldof_owner;
  i64 *ldof_gid;
  i64 n_own;
  i64 *own_ldofs;  /* own index -> ldof */
  /* COO (FE local ids, later global ids) */
  i64 nnz_coo, cap_coo;
  i64 *ci, *cj;
  int *cjo;
  double *cv;
  double *bfe;     /* FE-indexed RHS */
  char *touched;
  /* matrix PRanges */
  i64 n_ghost_rows, n_ghost_cols;
  i64 *row_gid; int *row_owner;  /* own then ghost */
  i64 *col_gid; int *col_owner;
  /* CSR */
  i64 nrows, ncols;
  i64 *rowptr; int *colidx; double *val;
  double *b;       /* on A's rows */
} part_t;

typedef struct {
  orc_desc d;
  double *modes;
  double *state;
  double *cell_kappa;
  double *cell_lame;
  double *vertex_coords;
  int nparts, K, D, nq;
  int nc;          /* components per node: dim for elasticity, else 1 */
  i64 nn[3];       /* global node counts per dim */
  i64 ngdofs;
  part_t *P;
  double qx[8], qw[8];
  double t_setup, t_assemble, t_inplace;
  char err[256];
} orc_ctx;

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* ---------- partition: src/PArraysExtras.jl:72-84, 0-based half-open ---------- */
static void local_range(int pi, int np, i64 n, int ghost, i64 *lo, i64 *hi) {
  i64 l = n / np, rem = n % np;
  i64 p = pi + 1;
  i64 offset = l * (p - 1);
  if (rem >= np - p + 1) {
    l += 1;
    offset += p - (np - rem) - 1;
  }
  i64 s = 1 + offset - ghost, e = l + offset + ghost;
  if (s < 1) s = 1;
  if (e > n) e = n;
  *lo = s - 1;
  *hi = e;
}

/* owner part index (per dim) of global cell c */
static int cell_owner_dim(const orc_ctx *C, int d, i64 c) {
  for (int p = 0; p < C->d.parts[d]; p++) {
    i64 lo, hi;
    local_range(p, C->d.parts[d], C->d.cells[d], 0, &lo, &hi);
    if (c >= lo && c < hi) return p;
  }
  return -1;
}

static int rank_of(const orc_ctx *C, const int pc[3]) {
  return pc[0] + C->d.parts[0] * (pc[1] + C->d.parts[1] * pc[2]);
}

/* ---------- 1D Gauss-Legendre on [0,1] (ascending points) ---------- */
static void gauss_legendre(int n, double *x, double *w) {
  for (int i = 0; i < n; i++) {
    double z = cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; it++) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 1; j <= n; j++) {
        double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j - 1.0) * z * p2 - (j - 1.0) * p3) / j;
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      double z1 = z;
      z = z1 - p1 / pp;
      if (fabs(z - z1) < 1e-16) break;
    }
    /* map the root z of P_n on [-1,1] to [0,1]; the weight halves */
    x[n - 1 - i] = 0.5 * (1.0 + z);
    w[n - 1 - i] = 1.0 / ((1.0 - z * z) * pp * pp);
  }
}

/* 1D Lagrange basis on equispaced nodes j/K, j=0..K */
static double lag1(int K, int a, double x) {
  double r = 1.0;
  for (int j = 0; j <= K; j++)
    if (j != a) r *= (x - (double)j / K) / ((double)(a - j) / K);
  return r;
}
static double dlag1(int K, int a, double x) {
  double s = 0.0;
  for (int m = 0; m <= K; m++) {
    if (m == a) continue;
    double r = 1.0 / ((double)(a - m) / K);
    for (int j = 0; j <= K; j++)
      if (j != a && j != m) r *= (x - (double)j / K) / ((double)(a - j) / K);
    s += r;
  }
  return s;
}

/* ---------- Gridap polytope local faces, doubled-lattice offsets ---------- */
static const int HEX_V[8][3] = {{0,0,0},{2,0,0},{0,2,0},{2,2,0},{0,0,2},{2,0,2},{0,2,2},{2,2,2}};
static const int HEX_E[12][3] = {{1,0,0},{1,2,0},{1,0,2},{1,2,2},{0,1,0},{2,1,0},
                                 {0,1,2},{2,1,2},{0,0,1},{2,0,1},{0,2,1},{2,2,1}};
static const int HEX_F[6][3] = {{1,1,0},{1,1,2},{1,0,1},{1,2,1},{0,1,1},{2,1,1}};
static const int QUAD_V[4][3] = {{0,0,0},{2,0,0},{0,2,0},{2,2,0}};
static const int QUAD_E[4][3] = {{1,0,0},{1,2,0},{0,1,0},{2,1,0}};

/* cell-local node list in Gridap's LagrangianRefFE order; offsets in the
   order-K node lattice.  Returns the count. */
static int cell_nodes(int D, int K, int off[27][3]) {
  int n = 0;
  const int (*V)[3] = D == 3 ? HEX_V : QUAD_V;
  int nv = D == 3 ? 8 : 4;
  for (int v = 0; v < nv; v++) {
    for (int d = 0; d < 3; d++) off[n][d] = V[v][d] * K / 2;
    n++;
  }
  if (K == 1) return n;
  const int (*E)[3] = D == 3 ? HEX_E : QUAD_E;
  int ne = D == 3 ? 12 : 4;
  for (int e = 0; e < ne; e++, n++)
    for (int d = 0; d < 3; d++) off[n][d] = E[e][d];
  if (D == 3)
    for (int f = 0; f < 6; f++, n++)
      for (int d = 0; d < 3; d++) off[n][d] = HEX_F[f][d];
  off[n][0] = 1; off[n][1] = 1; off[n][2] = D == 3 ? 1 : 0;
  n++;
  return n;
}

static inline i64 lidx(const i64 lnn[3], i64 x, i64 y, i64 z) { return x + lnn[0] * (y + lnn[1] * z); }

static int is_dirichlet_gnode(const orc_ctx *C, const i64 g[3]) {
  if (C->d.dir_kind == 0) return 0;
  if (C->d.dir_kind == 2) return g[0] == 0;
  for (int d = 0; d < C->D; d++)
    if (g[d] == 0 || g[d] == C->nn[d] - 1) return 1;
  return 0;
}

static double gnode_coord(const orc_ctx *C, int d, i64 g) {
  double x0 = C->d.domain[2 * d], x1 = C->d.domain[2 * d + 1];
  return x0 + (x1 - x0) * (double)g / (double)(C->nn[d] - 1);
}

/* coordinates of global lattice node g: on the uniform box per direction; on a
   mapped grid the multilinear interpolation of the vertices around it (a vertex
   itself, or the mean of the 2, 4 or 8 vertices of the edge, face or cell whose
   midpoint it is) */
static void node_xyz(const orc_ctx *C, const i64 g[3], double x[3]) {
  const int D = C->D, K = C->K;
  x[0] = x[1] = x[2] = 0.0;
  if (!C->vertex_coords) {
    for (int d = 0; d < D; d++) x[d] = gnode_coord(C, d, g[d]);
    return;
  }
  i64 nv[3] = {1, 1, 1}, v0[3] = {0, 0, 0};
  int two[3] = {0, 0, 0};
  for (int d = 0; d < D; d++) {
    nv[d] = C->d.cells[d] + 1;
    v0[d] = g[d] / K;
    two[d] = g[d] % K != 0;
  }
  const int nb = 1 << (two[0] + two[1] + two[2]);
  for (int c = 0; c < 8; c++) {
    int s[3] = {c & 1, (c >> 1) & 1, (c >> 2) & 1};
    if ((s[0] && !two[0]) || (s[1] && !two[1]) || (s[2] && !two[2])) continue;
    i64 v = (v0[0] + s[0]) + nv[0] * ((v0[1] + s[1]) + nv[1] * (v0[2] + s[2]));
    for (int d = 0; d < D; d++) x[d] += C->vertex_coords[D * v + d] / nb;
  }
}

/* ---------- local dof numbering (Gridap conforming numbering, restated) ---------- */
static void number_local_dofs(orc_ctx *C, part_t *P) {
  int D = C->D, K = C->K;
  i64 *lnn = P->lnn;
  i64 n = P->n_lnodes;
  i64 *list = malloc(sizeof(i64) * n);
  char *mark = calloc(n, 1);
  i64 cnt = 0;
  i64 nz = D == 3 ? lnn[2] : 1;
  /* vertices, lexicographic */
  for (i64 z = 0; z < nz; z++)
    for (i64 y = 0; y < lnn[1]; y++)
      for (i64 x = 0; x < lnn[0]; x++) {
        if (K == 2 && ((x & 1) || (y & 1) || (z & 1))) continue;
        i64 l = lidx(lnn, x, y, z);
        list[cnt++] = l;
        mark[l] = 1;
      }
  if (K == 2) {
    i64 lc[3] = {P->loc_hi[0] - P->loc_lo[0], P->loc_hi[1] - P->loc_lo[1],
                 D == 3 ? P->loc_hi[2] - P->loc_lo[2] : 1};
    for (int fdim = 1; fdim < D; fdim++) {
      const int (*F)[3];
      int nf;
      if (D == 3) {
        F = fdim == 1 ? HEX_E : HEX_F;
        nf = fdim == 1 ? 12 : 6;
      } else {
        F = QUAD_E;
        nf = 4;
      }
      for (i64 cz = 0; cz < lc[2]; cz++)
        for (i64 cy = 0; cy < lc[1]; cy++)
          for (i64 cx = 0; cx < lc[0]; cx++)
            for (int f = 0; f < nf; f++) {
              i64 l = lidx(lnn, 2 * cx + F[f][0], 2 * cy + F[f][1], 2 * cz + F[f][2]);
              if (!mark[l]) {
                mark[l] = 1;
                list[cnt++] = l;
              }
            }
    }
    for (i64 cz = 0; cz < lc[2]; cz++)
      for (i64 cy = 0; cy < lc[1]; cy++)
        for (i64 cx = 0; cx < lc[0]; cx++) {
          i64 l = lidx(lnn, 2 * cx + 1, 2 * cy + 1, D == 3 ? 2 * cz + 1 : 0);
          list[cnt++] = l;
          mark[l] = 1;
        }
  }
  /* a node's own dofs are its components, consecutive (Gridap: each face's own
     dofs in order; Q1/Q2 faces own one node) */
  const int nc = C->nc;
  P->lnode_dof = malloc(sizeof(i64) * n);
  P->ldof_lnode = malloc(sizeof(i64) * n * nc);
  i64 nf = 0, nd = 0;
  for (i64 k = 0; k < cnt; k++) {
    i64 l = list[k];
    i64 g[3] = {P->loc_lo[0] * K + l % lnn[0], P->loc_lo[1] * K + (l / lnn[0]) % lnn[1],
                D == 3 ? P->loc_lo[2] * K + l / (lnn[0] * lnn[1]) : 0};
    if (is_dirichlet_gnode(C, g)) {
      P->lnode_dof[l] = -1 - nd;
      nd += nc;
    } else {
      for (int c = 0; c < nc; c++) P->ldof_lnode[nf + c] = l;
      P->lnode_dof[l] = nf;
      nf += nc;
    }
  }
  P->n_ldofs = nf;
  P->n_ddofs = nd;
  free(list);
  free(mark);
}

/* local dof of component comp at local node l (negative: Dirichlet) */
static inline i64 node_dof(const part_t *P, i64 l, int comp) {
  i64 d0 = P->lnode_dof[l];
  return d0 >= 0 ? d0 + comp : d0 - comp;
}

/* local node of a global node in part P, or -1 */
static i64 lnode_of_gnode(const orc_ctx *C, const part_t *P, const i64 g[3]) {
  i64 l[3] = {0, 0, 0};
  for (int d = 0; d < C->D; d++) {
    l[d] = g[d] - P->loc_lo[d] * C->K;
    if (l[d] < 0 || l[d] >= P->lnn[d]) return -1;
  }
  return lidx(P->lnn, l[0], l[1], l[2]);
}

/* iterate cells of a box in lexicographic order */
#define FOR_BOX(lo, hi, D, cx, cy, cz)                                  \
  for (i64 cz = (lo)[2]; cz < ((D) == 3 ? (hi)[2] : (lo)[2] + 1); cz++) \
    for (i64 cy = (lo)[1]; cy < (hi)[1]; cy++)                          \
      for (i64 cx = (lo)[0]; cx < (hi)[0]; cx++)

static int cell_owner_rank(const orc_ctx *C, i64 cx, i64 cy, i64 cz) {
  int pc[3] = {cell_owner_dim(C, 0, cx), cell_owner_dim(C, 1, cy), C->D == 3 ? cell_owner_dim(C, 2, cz) : 0};
  return rank_of(C, pc);
}

static void setup(orc_ctx *C) {
  int D = C->D, K = C->K;
  C->nparts = C->d.parts[0] * C->d.parts[1] * (D == 3 ? C->d.parts[2] : 1);
  for (int d = 0; d < 3; d++) C->nn[d] = d < D ? K * C->d.cells[d] + 1 : 1;
  C->P = calloc(C->nparts, sizeof(part_t));
  for (int r = 0; r < C->nparts; r++) {
    part_t *P = &C->P[r];
    P->rank = r;
    P->pc[0] = r % C->d.parts[0];
    P->pc[1] = (r / C->d.parts[0]) % C->d.parts[1];
    P->pc[2] = D == 3 ? r / (C->d.parts[0] * C->d.parts[1]) : 0;
    for (int d = 0; d < 3; d++) {
      if (d < D) {
        local_range(P->pc[d], C->d.parts[d], C->d.cells[d], 0, &P->own_lo[d], &P->own_hi[d]);
        local_range(P->pc[d], C->d.parts[d], C->d.cells[d], 1, &P->loc_lo[d], &P->loc_hi[d]);
        P->lnn[d] = K * (P->loc_hi[d] - P->loc_lo[d]) + 1;
      } else {
        P->own_lo[d] = P->loc_lo[d] = 0;
        P->own_hi[d] = P->loc_hi[d] = 1;
        P->lnn[d] = 1;
      }
    }
    P->n_lnodes = P->lnn[0] * P->lnn[1] * P->lnn[2];
    number_local_dofs(C, P);
  }
  /* generate_gids (src/FESpaces.jl:139-261) */
  int coff[27][3];
  int ncn = cell_nodes(D, K, coff);
  int **own1 = malloc(sizeof(int *) * C->nparts);
  for (int r = 0; r < C->nparts; r++) {
    part_t *P = &C->P[r];
    own1[r] = calloc(P->n_ldofs, sizeof(int));
    FOR_BOX(P->loc_lo, P->loc_hi, D, cx, cy, cz) {
      int co = cell_owner_rank(C, cx, cy, cz);
      i64 c[3] = {cx - P->loc_lo[0], cy - P->loc_lo[1], cz - P->loc_lo[2]};
      for (int k = 0; k < C->nc * ncn; k++) {
        int a = k % ncn;
        i64 l = lidx(P->lnn, K * c[0] + coff[a][0], K * c[1] + coff[a][1], K * c[2] + coff[a][2]);
        i64 dof = node_dof(P, l, k / ncn);
        if (dof >= 0 && co > own1[r][dof]) own1[r][dof] = co;
      }
    }
  }
  /* exchange of cell-wise owners: ghost cells take the owner's view */
  for (int r = 0; r < C->nparts; r++) {
    part_t *P = &C->P[r];
    P->ldof_owner = malloc(sizeof(int) * (P->n_ldofs + 1));
    memcpy(P->ldof_owner, own1[r], sizeof(int) * P->n_ldofs);
    FOR_BOX(P->loc_lo, P->loc_hi, D, cx, cy, cz) {
      int co = cell_owner_rank(C, cx, cy, cz);
      if (co == r) continue;
      part_t *Q = &C->P[co];
      i64 c[3] = {cx - P->loc_lo[0], cy - P->loc_lo[1], cz - P->loc_lo[2]};
      for (int k = 0; k < C->nc * ncn; k++) {
        int a = k % ncn;
        i64 l = lidx(P->lnn, K * c[0] + coff[a][0], K * c[1] + coff[a][1], K * c[2] + coff[a][2]);
        i64 dof = node_dof(P, l, k / ncn);
        if (dof < 0) continue;
        i64 g[3] = {K * cx + coff[a][0], K * cy + coff[a][1], K * cz + coff[a][2]};
        i64 lq = lnode_of_gnode(C, Q, g);
        P->ldof_owner[dof] = own1[co][node_dof(Q, lq, k / ncn)];
      }
    }
  }
  for (int r = 0; r < C->nparts; r++) free(own1[r]);
  free(own1);
  /* owned counts, exclusive scan, rank-blocked gids */
  i64 first = 0;
  for (int r = 0; r < C->nparts; r++) {
    part_t *P = &C->P[r];
    P->ldof_gid = malloc(sizeof(i64) * (P->n_ldofs + 1));
    P->n_own = 0;
    for (i64 i = 0; i < P->n_ldofs; i++) P->n_own += P->ldof_owner[i] == r;
    P->own_ldofs = malloc(sizeof(i64) * (P->n_own + 1));
    i64 o = 0;
    for (i64 i = 0; i < P->n_ldofs; i++) {
      if (P->ldof_owner[i] == r) {
        P->own_ldofs[o] = i;
        P->ldof_gid[i] = first + o;
        o++;
      } else {
        P->ldof_gid[i] = -1;
      }
    }
    first += P->n_own;
  }
  C->ngdofs = first;
  /* ghost gids: two rounds of cell-wise fetch from the ghost cell's owner */
  for (int round = 0; round < 2; round++) {
    i64 **snap = malloc(sizeof(i64 *) * C->nparts);
    for (int r = 0; r < C->nparts; r++) {
      snap[r] = malloc(sizeof(i64) * (C->P[r].n_ldofs + 1));
      memcpy(snap[r], C->P[r].ldof_gid, sizeof(i64) * C->P[r].n_ldofs);
    }
    for (int r = 0; r < C->nparts; r++) {
      part_t *P = &C->P[r];
      FOR_BOX(P->loc_lo, P->loc_hi, D, cx, cy, cz) {
        int co = cell_owner_rank(C, cx, cy, cz);
        if (co == r) continue;
        part_t *Q = &C->P[co];
        i64 c[3] = {cx - P->loc_lo[0], cy - P->loc_lo[1], cz - P->loc_lo[2]};
        for (int k = 0; k < C->nc * ncn; k++) {
          int a = k % ncn;
          i64 l = lidx(P->lnn, K * c[0] + coff[a][0], K * c[1] + coff[a][1], K * c[2] + coff[a][2]);
          i64 dof = node_dof(P, l, k / ncn);
          if (dof < 0) continue;
          if (round == 0 && P->ldof_owner[dof] != co) continue;
          i64 g[3] = {K * cx + coff[a][0], K * cy + coff[a][1], K * cz + coff[a][2]};
          i64 lq = lnode_of_gnode(C, Q, g);
          P->ldof_gid[dof] = snap[co][node_dof(Q, lq, k / ncn)];
        }
      }
    }
    for (int r = 0; r < C->nparts; r++) free(snap[r]);
    free(snap);
  }
}

/* ---------- coefficient, source, Dirichlet data ---------- */
static double eval_g(const orc_ctx *C, const double x[3], double dg[3]) {
  double g = 0.0;
  dg[0] = dg[1] = dg[2] = 0.0;
  for (int m = 0; m < C->d.nmodes; m++) {
    const double *M = C->modes + 5 * m;
    double th = 2.0 * M_PI * (M[0] * x[0] + M[1] * x[1] + M[2] * x[2]) + M[4];
    g += M[3] * cos(th);
    double s = -M[3] * sin(th) * 2.0 * M_PI;
    dg[0] += s * M[0];
    dg[1] += s * M[1];
    dg[2] += s * M[2];
  }
  return g;
}
static double eval_kappa(const orc_ctx *C, const double x[3], double *f) {
  double dg[3];
  double g = eval_g(C, x, dg);
  double k = C->d.coef_const * exp(g);
  if (C->d.rhs_kind == 1) {
    /* f = -div(kappa grad uD), uD linear => -kappa * (grad g . c) */
    *f = -k * (dg[0] * C->d.dir_coef[1] + dg[1] * C->d.dir_coef[2] + dg[2] * C->d.dir_coef[3]);
  } else {
    *f = C->d.rhs_const;
  }
  return k;
}
static double eval_uD(const orc_ctx *C, const double x[3]) {
  return C->d.dir_coef[0] + C->d.dir_coef[1] * x[0] + C->d.dir_coef[2] * x[1] + C->d.dir_coef[3] * x[2];
}

/* cell integration in Gridap's local node order.  Geometry through the Q1 map
   from the cell's vertices: J = sum_v x_v (x) grad N_v, detJ, J^{-1}. */
/* flux operator: kappa(s) = s^((p-2)/2) and its derivative factor (p-2) beta s^((p-4)/2), with s = alpha + beta |g|^2
   (written with pow on both exponents, independently of the library's kappa/s form) */
static double flux_kappa(const orc_ctx *C, double s) { return pow(s, 0.5 * (C->d.flux_p - 2.0)); }
static double flux_gamma(const orc_ctx *C, double s) {
  return (C->d.flux_p - 2.0) * C->d.flux_beta * pow(s, 0.5 * (C->d.flux_p - 4.0));
}

/* ucell: u at the cell's nodes (op_kind 1).  energy != NULL: accumulate
   int( (alpha + beta|grad u|^2)^(p/2) / (p beta) - f u ) over the cell instead of Ac/bc;
   its gradient with respect to the free nodal values is the residual. */
static void integrate_cell(const orc_ctx *C, i64 cx, i64 cy, i64 cz, int ncn, int coff[27][3],
                           double *Ac, double *bc, double uD[27], const int isdir[27], const double ucell[27],
                           double *energy) {
  int D = C->D, K = C->K, nq = C->nq;
  double vx[8][3];
  int nv = D == 3 ? 8 : 4;
  for (int v = 0; v < nv; v++) {
    const int *o = D == 3 ? HEX_V[v] : QUAD_V[v];
    i64 g[3] = {K * (cx + o[0] / 2), K * (cy + o[1] / 2), K * (cz + o[2] / 2)};
    node_xyz(C, g, vx[v]);
  }
  const int nc = C->nc, ncd = nc * ncn; /* cell dofs: component-major */
  for (int a = 0; a < ncd * ncd; a++) Ac[a] = 0.0;
  for (int a = 0; a < ncd; a++) bc[a] = 0.0;
  int nqz = D == 3 ? nq : 1;
  for (int q3 = 0; q3 < nqz; q3++)
    for (int q2 = 0; q2 < nq; q2++)
      for (int q1 = 0; q1 < nq; q1++) {
        double xi[3] = {C->qx[q1], C->qx[q2], D == 3 ? C->qx[q3] : 0.0};
        double w = C->qw[q1] * C->qw[q2] * (D == 3 ? C->qw[q3] : 1.0);
        /* Q1 geometric map */
        double J[3][3] = {{0}}, xq[3] = {0, 0, 0};
        for (int v = 0; v < nv; v++) {
          const int *o = D == 3 ? HEX_V[v] : QUAD_V[v];
          double N = 1.0, dN[3];
          for (int d = 0; d < D; d++) N *= lag1(1, o[d] / 2, xi[d]);
          for (int e = 0; e < D; e++) {
            dN[e] = 1.0;
            for (int d = 0; d < D; d++) dN[e] *= d == e ? dlag1(1, o[d] / 2, xi[d]) : lag1(1, o[d] / 2, xi[d]);
          }
          for (int d = 0; d < D; d++) {
            xq[d] += N * vx[v][d];
            for (int e = 0; e < D; e++) J[d][e] += vx[v][d] * dN[e];
          }
        }
        double det, iJ[3][3];
        if (D == 2) {
          det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
          iJ[0][0] = J[1][1] / det; iJ[0][1] = -J[0][1] / det;
          iJ[1][0] = -J[1][0] / det; iJ[1][1] = J[0][0] / det;
        } else {
          det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) - J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
          iJ[0][0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) / det;
          iJ[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
          iJ[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
          iJ[1][0] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) / det;
          iJ[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
          iJ[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
          iJ[2][0] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) / det;
          iJ[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
          iJ[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        }
        double f;
        double kap = eval_kappa(C, xq, &f);
        if (C->d.cell_kappa) {
          kap = C->d.coef_const * C->d.cell_kappa[cx + C->d.cells[0] * (cy + C->d.cells[1] * cz)];
          f = C->d.rhs_const;
        }
        double dV = w * fabs(det);
        double phi[27], grad[27][3];
        for (int a = 0; a < ncn; a++) {
          double ref[3];
          phi[a] = 1.0;
          for (int d = 0; d < D; d++) phi[a] *= lag1(K, coff[a][d], xi[d]);
          for (int e = 0; e < D; e++) {
            ref[e] = 1.0;
            for (int d = 0; d < D; d++) ref[e] *= d == e ? dlag1(K, coff[a][d], xi[d]) : lag1(K, coff[a][d], xi[d]);
          }
          /* grad = J^{-T} ref */
          for (int e = 0; e < D; e++) {
            grad[a][e] = 0.0;
            for (int d = 0; d < D; d++) grad[a][e] += iJ[d][e] * ref[d];
          }
        }
        if (C->d.op_kind == 2) {
          /* eps(v) : (lambda tr(eps(u)) I + 2 mu eps(u)) with v = phi_a e_ca,
             u = phi_b e_cb:  lambda d_ca(phi_a) d_cb(phi_b)
                            + mu d_cb(phi_a) d_ca(phi_b) + mu [ca == cb] grad(phi_a).grad(phi_b) */
          double lam = C->d.lame_lambda, mu = C->d.lame_mu;
          if (C->d.cell_lame) {
            const i64 c = cx + C->d.cells[0] * (cy + C->d.cells[1] * cz);
            lam = C->d.cell_lame[2 * c];
            mu = C->d.cell_lame[2 * c + 1];
          }
          for (int ca = 0; ca < nc; ca++)
            for (int a = 0; a < ncn; a++) {
              const int ia = ca * ncn + a;
              bc[ia] += C->d.body_force[ca] * phi[a] * dV;
              for (int cb = 0; cb < nc; cb++)
                for (int b = 0; b < ncn; b++) {
                  double v = lam * grad[a][ca] * grad[b][cb] + mu * grad[a][cb] * grad[b][ca];
                  if (ca == cb)
                    for (int e = 0; e < D; e++) v += mu * grad[a][e] * grad[b][e];
                  Ac[ia * ncd + cb * ncn + b] += v * dV;
                }
            }
          continue;
        }
        if (C->d.op_kind == 1) {
          double gu[3] = {0.0, 0.0, 0.0}, uq = 0.0;
          for (int b = 0; b < ncn; b++) {
            uq += ucell[b] * phi[b];
            for (int e = 0; e < D; e++) gu[e] += ucell[b] * grad[b][e];
          }
          double s2 = C->d.flux_alpha + C->d.flux_beta * (gu[0] * gu[0] + gu[1] * gu[1] + gu[2] * gu[2]);
          double fq = C->d.rhs_const;
          if (energy) {
            *energy += (pow(s2, 0.5 * C->d.flux_p) / (C->d.flux_p * C->d.flux_beta) - fq * uq) * dV;
            continue;
          }
          double kq = flux_kappa(C, s2), gq = flux_gamma(C, s2);
          for (int a = 0; a < ncn; a++) {
            double gua = 0.0;
            for (int e = 0; e < D; e++) gua += gu[e] * grad[a][e];
            bc[a] += (kq * gua - fq * phi[a]) * dV;
            for (int b = 0; b < ncn; b++) {
              double sab = 0.0, gub = 0.0;
              for (int e = 0; e < D; e++) {
                sab += grad[a][e] * grad[b][e];
                gub += gu[e] * grad[b][e];
              }
              Ac[a * ncn + b] += (kq * sab + gq * gua * gub) * dV;
            }
          }
          continue;
        }
        for (int a = 0; a < ncn; a++) {
          bc[a] += f * phi[a] * dV;
          const double dw = C->d.weights_set ? C->d.diffusion_weight : 1.0;
          const double rw = C->d.weights_set ? C->d.reaction : 0.0;
          for (int b = 0; b < ncn; b++) {
            double s = 0.0;
            for (int e = 0; e < D; e++) s += grad[a][e] * grad[b][e];
            Ac[a * ncn + b] += (dw * kap * s + rw * phi[a] * phi[b]) * dV;
          }
        }
      }
  /* Dirichlet lifting: b_c -= A_c u_D (the flux operator's residual already
     sees the Dirichlet values through u) */
  if (C->d.op_kind == 1) return;
  for (int a = 0; a < ncd; a++)
    for (int b = 0; b < ncd; b++)
      if (isdir[b]) bc[a] -= Ac[a * ncd + b] * uD[b];
}

/* u at global lattice node g: the state, or uD on Dirichlet nodes */
static double state_at(const orc_ctx *C, const i64 g[3], int isdir, double uDv) {
  if (isdir) return uDv;
  if (!C->state) return 0.0;
  return C->state[g[0] + C->nn[0] * (g[1] + C->nn[1] * g[2])];
}

static void coo_push(part_t *P, i64 i, i64 j, double v) {
  if (P->nnz_coo == P->cap_coo) {
    P->cap_coo = P->cap_coo ? 2 * P->cap_coo : 1024;
    P->ci = realloc(P->ci, sizeof(i64) * P->cap_coo);
    P->cj = realloc(P->cj, sizeof(i64) * P->cap_coo);
    P->cv = realloc(P->cv, sizeof(double) * P->cap_coo);
  }
  P->ci[P->nnz_coo] = i;
  P->cj[P->nnz_coo] = j;
  P->cv[P->nnz_coo] = v;
  P->nnz_coo++;
}

/* per-part numeric loop.  mode 0: append COO (fresh assembly); mode 1: add
   into the existing CSR through the FE-lid -> matrix-lid maps (in place). */
static void numeric_loop(orc_ctx *C, part_t *P, int mode, const i64 *fe_to_row, const i64 *fe_to_col) {
  int D = C->D, K = C->K;
  int coff[27][3];
  int ncn = cell_nodes(D, K, coff);
  const int ncd = C->nc * ncn;
  double Ac[81 * 81], bc[81], uD[81], ucell[81];
  int isdir[81];
  i64 row[81], col[81];
  int FA = C->d.strategy == 1;
  const i64 *lo = FA ? P->loc_lo : P->own_lo;
  const i64 *hi = FA ? P->loc_hi : P->own_hi;
  FOR_BOX(lo, hi, D, cx, cy, cz) {
    i64 c[3] = {cx - P->loc_lo[0], cy - P->loc_lo[1], cz - P->loc_lo[2]};
    for (int k = 0; k < ncd; k++) {
      const int a = k % ncn;
      i64 l = lidx(P->lnn, K * c[0] + coff[a][0], K * c[1] + coff[a][1], K * c[2] + coff[a][2]);
      i64 dof = node_dof(P, l, k / ncn);
      col[k] = dof;
      row[k] = dof;
      if (FA && dof >= 0 && P->ldof_owner[dof] != P->rank) row[k] = -1; /* ghost rows dropped */
      isdir[k] = dof < 0;
      i64 g[3] = {K * cx + coff[a][0], K * cy + coff[a][1], K * cz + coff[a][2]};
      double x[3];
      node_xyz(C, g, x);
      uD[k] = isdir[k] && C->d.op_kind != 2 ? eval_uD(C, x) : 0.0; /* elasticity: clamped, u = 0 */
      ucell[k] = state_at(C, g, isdir[k], uD[k]);
    }
    integrate_cell(C, cx, cy, cz, ncn, coff, Ac, bc, uD, isdir, ucell, NULL);
    /* add_entries!: columns outer, rows inner, skipping non-positive ids */
    for (int b = 0; b < ncd; b++) {
      if (col[b] < 0) continue;
      for (int a = 0; a < ncd; a++) {
        if (row[a] < 0) continue;
        if (mode == 0) {
          coo_push(P, row[a], col[b], Ac[a * ncd + b]);
        } else {
          i64 r = fe_to_row[row[a]], cc = fe_to_col[col[b]];
          i64 lo2 = P->rowptr[r], hi2 = P->rowptr[r + 1] - 1;
          while (lo2 < hi2) {
            i64 mid = (lo2 + hi2) / 2;
            if (P->colidx[mid] < cc) lo2 = mid + 1; else hi2 = mid;
          }
          P->val[lo2] += Ac[a * ncd + b];
        }
      }
    }
    for (int a = 0; a < ncd; a++) {
      if (row[a] < 0) continue;
      if (mode == 0) {
        P->touched[row[a]] = 1;
        P->bfe[row[a]] += bc[a];
      } else {
        P->b[fe_to_row[row[a]]] += bc[a];
      }
    }
  }
}

/* stable sort of index array by key (merge sort) */
static void msort(i64 *idx, i64 n, const i64 *key, i64 *tmp) {
  if (n < 2) return;
  i64 h = n / 2;
  msort(idx, h, key, tmp);
  msort(idx + h, n - h, key, tmp);
  i64 i = 0, j = h, k = 0;
  while (i < h && j < n) tmp[k++] = key[idx[j]] < key[idx[i]] ? idx[j++] : idx[i++];
  while (i < h) tmp[k++] = idx[i++];
  while (j < n) tmp[k++] = idx[j++];
  memcpy(idx, tmp, sizeof(i64) * n);
}

static i64 *gid_lookup(const i64 *gids, i64 n, i64 ng) {
  i64 *m = malloc(sizeof(i64) * (ng + 1));
  for (i64 i = 0; i < ng; i++) m[i] = -1;
  for (i64 i = 0; i < n; i++) m[gids[i]] = i;
  return m;
}

static int gid_owner(const orc_ctx *C, i64 gid) {
  i64 first = 0;
  for (int r = 0; r < C->nparts; r++) {
    if (gid < first + C->P[r].n_own) return r;
    first += C->P[r].n_own;
  }
  return -1;
}

/* COO (matrix-local ids) -> CSR with sorted columns, duplicates summed in COO order */
static void compress(part_t *P, const i64 *I, const i64 *J, const double *V, i64 n) {
  i64 *idx = malloc(sizeof(i64) * (n + 1)), *tmp = malloc(sizeof(i64) * (n + 1));
  i64 *key = malloc(sizeof(i64) * (n + 1));
  for (i64 k = 0; k < n; k++) {
    idx[k] = k;
    key[k] = I[k] * P->ncols + J[k];
  }
  msort(idx, n, key, tmp);
  P->rowptr = calloc(P->nrows + 1, sizeof(i64));
  i64 nnz = 0;
  for (i64 k = 0; k < n; k++)
    if (k == 0 || key[idx[k]] != key[idx[k - 1]]) nnz++;
  P->colidx = malloc(sizeof(int) * (nnz + 1));
  P->val = malloc(sizeof(double) * (nnz + 1));
  i64 p = -1;
  for (i64 k = 0; k < n; k++) {
    i64 e = idx[k];
    if (k == 0 || key[e] != key[idx[k - 1]]) {
      p++;
      P->colidx[p] = (int)J[e];
      P->val[p] = V[e];
      P->rowptr[I[e] + 1]++;
    } else {
      P->val[p] += V[e];
    }
  }
  for (i64 r = 0; r < P->nrows; r++) P->rowptr[r + 1] += P->rowptr[r];
  free(idx);
  free(tmp);
  free(key);
}

static void assemble_fresh(orc_ctx *C) {
  int np = C->nparts;
  int FA = C->d.strategy == 1;
  for (int r = 0; r < np; r++) {
    part_t *P = &C->P[r];
    P->bfe = calloc(P->n_ldofs + 1, sizeof(double));
    P->touched = calloc(P->n_ldofs + 1, 1);
    numeric_loop(C, P, 0, NULL, NULL);
  }
  /* rows PRange: own + (SA) touched ghost rows sorted by FE local id */
  for (int r = 0; r < np; r++) {
    part_t *P = &C->P[r];
    char *tg = calloc(P->n_ldofs + 1, 1);
    P->n_ghost_rows = 0;
    if (!FA)
      for (i64 k = 0; k < P->nnz_coo; k++) {
        i64 i = P->ci[k];
        if (P->ldof_owner[i] != r && !tg[i]) {
          tg[i] = 1;
          P->n_ghost_rows++;
        }
      }
    P->row_gid = malloc(sizeof(i64) * (P->n_own + P->n_ghost_rows + 1));
    P->row_owner = malloc(sizeof(int) * (P->n_own + P->n_ghost_rows + 1));
    for (i64 o = 0; o < P->n_own; o++) {
      P->row_gid[o] = P->ldof_gid[P->own_ldofs[o]];
      P->row_owner[o] = r;
    }
    i64 g = P->n_own;
    for (i64 i = 0; i < P->n_ldofs; i++)
      if (tg[i]) {
        P->row_gid[g] = P->ldof_gid[i];
        P->row_owner[g] = P->ldof_owner[i];
        g++;
      }
    P->nrows = P->n_own + P->n_ghost_rows;
    free(tg);
    /* to global ids, column owners */
    P->cjo = malloc(sizeof(int) * (P->nnz_coo + 1));
    for (i64 k = 0; k < P->nnz_coo; k++) {
      i64 i = P->ci[k], j = P->cj[k];
      P->cjo[k] = P->ldof_owner[j];
      P->ci[k] = P->ldof_gid[i];
      P->cj[k] = P->ldof_gid[j];
    }
  }
  if (!FA) {
    /* assemble_coo_with_column_owner!: ghost-row triplets to their owners.
       Senders pack per destination in COO order and zero their copy;
       receivers append per sender in ascending rank order. */
    i64 *base = malloc(sizeof(i64) * np);
    for (int r = 0; r < np; r++) base[r] = C->P[r].nnz_coo;
    for (int dst = 0; dst < np; dst++) {
      part_t *Pd = &C->P[dst];
      for (int src = 0; src < np; src++) {
        if (src == dst) continue;
        part_t *Ps = &C->P[src];
        for (i64 k = 0; k < base[src]; k++) {
          int ro = gid_owner(C, Ps->ci[k]);
          if (ro != dst) continue;
          coo_push(Pd, Ps->ci[k], Ps->cj[k], Ps->cv[k]);
          Pd->cjo = realloc(Pd->cjo, sizeof(int) * (Pd->cap_coo + 1));
          Pd->cjo[Pd->nnz_coo - 1] = Ps->cjo[k];
        }
      }
    }
    for (int src = 0; src < np; src++) {
      part_t *Ps = &C->P[src];
      for (i64 k = 0; k < base[src]; k++)
        if (gid_owner(C, Ps->ci[k]) != src) Ps->cv[k] = 0.0;
    }
    free(base);
  }
  /* cols PRange, local ids, compression */
  for (int r = 0; r < np; r++) {
    part_t *P = &C->P[r];
    i64 cap = 1024;
    P->col_gid = malloc(sizeof(i64) * (P->n_own + cap));
    P->col_owner = malloc(sizeof(int) * (P->n_own + cap));
    for (i64 o = 0; o < P->n_own; o++) {
      P->col_gid[o] = P->ldof_gid[P->own_ldofs[o]];
      P->col_owner[o] = r;
    }
    i64 *seen = gid_lookup(P->col_gid, P->n_own, C->ngdofs);
    i64 ng = 0;
    for (i64 k = 0; k < P->nnz_coo; k++) {
      i64 j = P->cj[k];
      if (P->cjo[k] != r && seen[j] < 0) {
        if (ng == cap) {
          cap *= 2;
          P->col_gid = realloc(P->col_gid, sizeof(i64) * (P->n_own + cap));
          P->col_owner = realloc(P->col_owner, sizeof(int) * (P->n_own + cap));
        }
        P->col_gid[P->n_own + ng] = j;
        P->col_owner[P->n_own + ng] = P->cjo[k];
        seen[j] = P->n_own + ng;
        ng++;
      }
    }
    if (FA) {
      /* _setup_prange_with_ghosts without owners: ghost ids sorted by FE local id */
      i64 *fe_of_gid = malloc(sizeof(i64) * (C->ngdofs + 1));
      for (i64 i = 0; i < P->n_ldofs; i++) fe_of_gid[P->ldof_gid[i]] = i;
      i64 *idx = malloc(sizeof(i64) * (ng + 1)), *tmp = malloc(sizeof(i64) * (ng + 1)),
          *key = malloc(sizeof(i64) * (ng + 1));
      for (i64 g = 0; g < ng; g++) {
        idx[g] = g;
        key[g] = fe_of_gid[P->col_gid[P->n_own + g]];
      }
      msort(idx, ng, key, tmp);
      i64 *cg = malloc(sizeof(i64) * (ng + 1));
      int *co = malloc(sizeof(int) * (ng + 1));
      for (i64 g = 0; g < ng; g++) {
        cg[g] = P->col_gid[P->n_own + idx[g]];
        co[g] = P->col_owner[P->n_own + idx[g]];
      }
      for (i64 g = 0; g < ng; g++) {
        P->col_gid[P->n_own + g] = cg[g];
        P->col_owner[P->n_own + g] = co[g];
        seen[cg[g]] = P->n_own + g;
      }
      free(cg); free(co); free(idx); free(tmp); free(key); free(fe_of_gid);
    }
    P->n_ghost_cols = ng;
    P->ncols = P->n_own + ng;
    i64 *rowof = gid_lookup(P->row_gid, P->nrows, C->ngdofs);
    for (i64 k = 0; k < P->nnz_coo; k++) {
      P->ci[k] = rowof[P->ci[k]];
      P->cj[k] = seen[P->cj[k]];
    }
    compress(P, P->ci, P->cj, P->cv, P->nnz_coo);
    /* RHS on A's rows: _rhs_callback */
    P->b = calloc(P->nrows + 1, sizeof(double));
    i64 *fe_of_gid = malloc(sizeof(i64) * (C->ngdofs + 1));
    for (i64 i = 0; i < P->n_ldofs; i++) fe_of_gid[P->ldof_gid[i]] = i;
    for (i64 k = 0; k < P->nrows; k++) P->b[k] = P->bfe[fe_of_gid[P->row_gid[k]]];
    free(fe_of_gid);
    free(rowof);
    free(seen);
  }
  if (!FA) {
    /* assemble!(b): ghost values added into owners (ascending sender), ghosts zeroed */
    for (int dst = 0; dst < np; dst++) {
      part_t *Pd = &C->P[dst];
      i64 *rowof = gid_lookup(Pd->row_gid, Pd->n_own, C->ngdofs);
      for (int src = 0; src < np; src++) {
        if (src == dst) continue;
        part_t *Ps = &C->P[src];
        for (i64 g = Ps->n_own; g < Ps->nrows; g++)
          if (Ps->row_owner[g] == dst) Pd->b[rowof[Ps->row_gid[g]]] += Ps->b[g];
      }
      free(rowof);
    }
    for (int r = 0; r < np; r++) {
      part_t *P = &C->P[r];
      for (i64 g = P->n_own; g < P->nrows; g++) P->b[g] = 0.0;
    }
  }
}

/* threads of the in-place re-assembly (orc_set_threads; 1 = the reference's
   debug mode, all parts in one process one after another) */
static int g_orc_threads = 1;
void orc_set_threads(int n) { g_orc_threads = n < 1 ? 1 : n; }
static int orc_threads_for(int nparts) { return g_orc_threads < nparts ? g_orc_threads : nparts; }

/* in-place numeric re-assembly (LocalView path + assemble!) into the CSR built
   by assemble_fresh.  Returns elapsed seconds. */
static double assemble_inplace(orc_ctx *C) {
  int np = C->nparts;
  int FA = C->d.strategy == 1;
  i64 **fe_to_row = malloc(sizeof(i64 *) * np), **fe_to_col = malloc(sizeof(i64 *) * np);
  for (int r = 0; r < np; r++) {
    part_t *P = &C->P[r];
    i64 *rowof = gid_lookup(P->row_gid, P->nrows, C->ngdofs);
    i64 *colof = gid_lookup(P->col_gid, P->ncols, C->ngdofs);
    fe_to_row[r] = malloc(sizeof(i64) * (P->n_ldofs + 1));
    fe_to_col[r] = malloc(sizeof(i64) * (P->n_ldofs + 1));
    for (i64 i = 0; i < P->n_ldofs; i++) {
      fe_to_row[r][i] = rowof[P->ldof_gid[i]];
      fe_to_col[r][i] = colof[P->ldof_gid[i]];
    }
    free(rowof);
    free(colof);
  }
  /* owner-side lookup for assemble!(A) */
  i64 **rowof_own = malloc(sizeof(i64 *) * np), **colof_all = malloc(sizeof(i64 *) * np);
  for (int r = 0; r < np; r++) {
    rowof_own[r] = gid_lookup(C->P[r].row_gid, C->P[r].n_own, C->ngdofs);
    colof_all[r] = gid_lookup(C->P[r].col_gid, C->P[r].ncols, C->ngdofs);
  }
  double t0 = now_s();
  /* the parts are independent (as the reference's parts, one per process):
     one OpenMP thread per part; assemble!(A) below writes each owner's rows
     from one thread, senders in ascending order, so the result does not
     depend on the thread count */
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_threads_for(np))
  for (int r = 0; r < np; r++) {
    part_t *P = &C->P[r];
    memset(P->val, 0, sizeof(double) * P->rowptr[P->nrows]);
    memset(P->b, 0, sizeof(double) * P->nrows);
    numeric_loop(C, P, 1, fe_to_row[r], fe_to_col[r]);
  }
  if (!FA) {
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_threads_for(np))
    for (int dst = 0; dst < np; dst++) {
      part_t *Pd = &C->P[dst];
      for (int src = 0; src < np; src++) {
        if (src == dst) continue;
        part_t *Ps = &C->P[src];
        for (i64 g = Ps->n_own; g < Ps->nrows; g++) {
          if (Ps->row_owner[g] != dst) continue;
          i64 rd = rowof_own[dst][Ps->row_gid[g]];
          Pd->b[rd] += Ps->b[g];
          for (i64 k = Ps->rowptr[g]; k < Ps->rowptr[g + 1]; k++) {
            i64 cd = colof_all[dst][Ps->col_gid[Ps->colidx[k]]];
            i64 lo2 = Pd->rowptr[rd], hi2 = Pd->rowptr[rd + 1] - 1;
            while (lo2 < hi2) {
              i64 mid = (lo2 + hi2) / 2;
              if (Pd->colidx[mid] < cd) lo2 = mid + 1; else hi2 = mid;
            }
            Pd->val[lo2] += Ps->val[k];
          }
        }
      }
    }
#pragma omp parallel for schedule(dynamic, 1) num_threads(orc_threads_for(np))
    for (int r = 0; r < np; r++) {
      part_t *P = &C->P[r];
      for (i64 g = P->n_own; g < P->nrows; g++) {
        P->b[g] = 0.0;
        for (i64 k = P->rowptr[g]; k < P->rowptr[g + 1]; k++) P->val[k] = 0.0;
      }
    }
  }
  double t = now_s() - t0;
  for (int r = 0; r < np; r++) {
    free(fe_to_row[r]); free(fe_to_col[r]); free(rowof_own[r]); free(colof_all[r]);
  }
  free(fe_to_row); free(fe_to_col); free(rowof_own); free(colof_all);
  return t;
}

/* ------------------------------- public API ------------------------------- */
void *orc_run(const orc_desc *d, int inplace_reps) {
  orc_ctx *C = calloc(1, sizeof(orc_ctx));
  C->d = *d;
  C->D = d->dim;
  C->K = d->order;
  C->nq = d->qdeg / 2 + 1;
  C->nc = d->op_kind == 2 ? d->dim : 1;
  if (d->nmodes > 0) {
    C->modes = malloc(sizeof(double) * 5 * d->nmodes);
    memcpy(C->modes, d->modes, sizeof(double) * 5 * d->nmodes);
  }
  C->d.modes = C->modes;
  C->state = NULL;
  gauss_legendre(C->nq, C->qx, C->qw);
  for (int k = 0; k < 3; k++) C->nn[k] = k < d->dim ? d->order * d->cells[k] + 1 : 1;
  if (d->state) {
    i64 n = C->nn[0] * C->nn[1] * C->nn[2];
    C->state = malloc(sizeof(double) * n);
    memcpy(C->state, d->state, sizeof(double) * n);
  }
  C->d.state = C->state;
  C->cell_kappa = NULL;
  if (d->cell_kappa) {
    i64 nc = d->cells[0] * (d->dim > 1 ? d->cells[1] : 1) * (d->dim > 2 ? d->cells[2] : 1);
    C->cell_kappa = malloc(sizeof(double) * nc);
    memcpy(C->cell_kappa, d->cell_kappa, sizeof(double) * nc);
  }
  C->d.cell_kappa = C->cell_kappa;
  C->cell_lame = NULL;
  if (d->cell_lame) {
    i64 nc = d->cells[0] * (d->dim > 1 ? d->cells[1] : 1) * (d->dim > 2 ? d->cells[2] : 1);
    C->cell_lame = malloc(sizeof(double) * 2 * nc);
    memcpy(C->cell_lame, d->cell_lame, sizeof(double) * 2 * nc);
  }
  C->d.cell_lame = C->cell_lame;
  C->vertex_coords = NULL;
  if (d->vertex_coords) {
    i64 nv = (d->cells[0] + 1) * (d->dim > 1 ? d->cells[1] + 1 : 1) * (d->dim > 2 ? d->cells[2] + 1 : 1);
    C->vertex_coords = malloc(sizeof(double) * d->dim * nv);
    memcpy(C->vertex_coords, d->vertex_coords, sizeof(double) * d->dim * nv);
  }
  C->d.vertex_coords = C->vertex_coords;
  double t0 = now_s();
  setup(C);
  double t1 = now_s();
  assemble_fresh(C);
  double t2 = now_s();
  C->t_setup = t1 - t0;
  C->t_assemble = t2 - t1;
  C->t_inplace = 0.0;
  for (int k = 0; k < inplace_reps; k++) {
    double t = assemble_inplace(C);
    if (k == 0 || t < C->t_inplace) C->t_inplace = t;
  }
  return C;
}

int orc_nparts(void *h) { return ((orc_ctx *)h)->nparts; }
i64 orc_ngdofs(void *h) { return ((orc_ctx *)h)->ngdofs; }
void orc_times(void *h, double *t) {
  orc_ctx *C = h;
  t[0] = C->t_setup;
  t[1] = C->t_assemble;
  t[2] = C->t_inplace;
}

/* sizes: n_ldofs, n_own, n_ghost_rows, n_ghost_cols, nnz */
void orc_part_sizes(void *h, int r, i64 *s) {
  part_t *P = &((orc_ctx *)h)->P[r];
  s[0] = P->n_ldofs;
  s[1] = P->n_own;
  s[2] = P->n_ghost_rows;
  s[3] = P->n_ghost_cols;
  s[4] = P->rowptr[P->nrows];
}

void orc_part_fe(void *h, int r, i64 *ldof_gid, int *ldof_owner, i64 *ldof_gnode) {
  orc_ctx *C = h;
  part_t *P = &C->P[r];
  for (i64 i = 0; i < P->n_ldofs; i++) {
    ldof_gid[i] = P->ldof_gid[i];
    ldof_owner[i] = P->ldof_owner[i];
    i64 l = P->ldof_lnode[i];
    i64 g0 = C->K * P->loc_lo[0] + l % P->lnn[0];
    i64 g1 = C->K * P->loc_lo[1] + (l / P->lnn[0]) % P->lnn[1];
    i64 g2 = C->K * P->loc_lo[2] + l / (P->lnn[0] * P->lnn[1]);
    ldof_gnode[i] = g0 + C->nn[0] * (g1 + C->nn[1] * g2);
  }
}

void orc_part_matrix(void *h, int r, i64 *row_gid, int *row_owner, i64 *col_gid, int *col_owner, i64 *rowptr,
                     int *colidx, double *val, double *b) {
  part_t *P = &((orc_ctx *)h)->P[r];
  memcpy(row_gid, P->row_gid, sizeof(i64) * P->nrows);
  memcpy(row_owner, P->row_owner, sizeof(int) * P->nrows);
  memcpy(col_gid, P->col_gid, sizeof(i64) * P->ncols);
  memcpy(col_owner, P->col_owner, sizeof(int) * P->ncols);
  memcpy(rowptr, P->rowptr, sizeof(i64) * (P->nrows + 1));
  memcpy(colidx, P->colidx, sizeof(int) * P->rowptr[P->nrows]);
  memcpy(val, P->val, sizeof(double) * P->rowptr[P->nrows]);
  memcpy(b, P->b, sizeof(double) * P->nrows);
}

void orc_free(void *h) {
  orc_ctx *C = h;
  if (!C) return;
  for (int r = 0; r < C->nparts; r++) {
    part_t *P = &C->P[r];
    free(P->lnode_dof); free(P->ldof_lnode); free(P->ldof_owner); free(P->ldof_gid); free(P->own_ldofs);
    free(P->ci); free(P->cj); free(P->cjo); free(P->cv); free(P->bfe); free(P->touched);
    free(P->row_gid); free(P->row_owner); free(P->col_gid); free(P->col_owner);
    free(P->rowptr); free(P->colidx); free(P->val); free(P->b);
  }
  free(C->P);
  free(C->modes);
  free(C->state);
  free(C->cell_kappa);
  free(C->cell_lame);
  free(C->vertex_coords);
  free(C);
}

/* Energy of the flux operator at the state (op_kind 1, flux_beta != 0), summed
   over every cell of the mesh once:
     E(u) = int( (alpha + beta |grad u|^2)^(p/2) / (p beta) - f u ) dOmega.
   Its derivative along a free nodal direction is the residual, which pins the
   residual independently of its own formula. */
double orc_energy(void *h) {
  orc_ctx *C = h;
  if (!C || C->d.op_kind != 1 || C->d.flux_beta == 0.0) return NAN;
  int D = C->D, K = C->K;
  int coff[27][3];
  int ncn = cell_nodes(D, K, coff);
  double Ac[27 * 27], bc[27], uD[27], ucell[27];
  int isdir[27];
  double E = 0.0;
  i64 lo[3] = {0, 0, 0}, hi[3] = {C->d.cells[0], D > 1 ? C->d.cells[1] : 1, D > 2 ? C->d.cells[2] : 1};
  FOR_BOX(lo, hi, D, cx, cy, cz) {
    for (int a = 0; a < ncn; a++) {
      i64 g[3] = {K * cx + coff[a][0], K * cy + coff[a][1], K * cz + coff[a][2]};
      double x[3];
      node_xyz(C, g, x);
      isdir[a] = C->d.dir_kind == 1 && is_dirichlet_gnode(C, g);
      uD[a] = isdir[a] ? eval_uD(C, x) : 0.0;
      ucell[a] = state_at(C, g, isdir[a], uD[a]);
    }
    integrate_cell(C, cx, cy, cz, ncn, coff, Ac, bc, uD, isdir, ucell, &E);
  }
  return E;
}

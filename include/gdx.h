/*
 * gdx.h -- C ABI of the MI355X-native assembly library (libgdx.so).
 *
 * The library replaces, for Lagrangian Q1/Q2 problems on a partitioned
 * CartesianDiscreteModel, the reference's assembly hot path:
 *
 *   SparseMatrixAssembler(U,V,strategy)       src/FESpaces.jl:819-862
 *   assemble_matrix_and_vector(assem,data)    Gridap generic loop driven by
 *     symbolic_loop_matrix_and_vector!        src/FESpaces.jl:785-792
 *     numeric_loop_matrix_and_vector!         src/FESpaces.jl:794-798
 *     create_from_nz (COO exchange+compress)  src/Algebra.jl:593-646,778-793
 *   assemble_matrix_and_vector!(A,b,...)      src/Algebra.jl:323-393,530-535
 *   generate_gids                             src/FESpaces.jl:139-261
 *   _uniform_partition                        src/PArraysExtras.jl:5-84
 *
 * All parts of the partition live on one GPU.  The library owns every device
 * buffer; host arrays passed in are caller-owned and copied.  Indices in the
 * ABI are 0-based.  Every entry point returns GDX_OK (0) or a negative error
 * class; gdx_last_error() returns a thread-local message.  A context and the
 * problems created on it must be used from one host thread at a time.
 */
#ifndef GDX_H
#define GDX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GDX_VERSION_MAJOR 0
#define GDX_VERSION_MINOR 1
#define GDX_VERSION_PATCH 0

/* status codes */
#define GDX_OK 0
#define GDX_ERR_ARG (-1)         /* invalid argument (the reference's @check / @assert) */
#define GDX_ERR_HIP (-2)         /* HIP runtime error */
#define GDX_ERR_STATE (-3)       /* call out of order (e.g. numeric before symbolic) */
#define GDX_ERR_UNSUPPORTED (-4) /* valid request outside the supported vocabulary (@notimplemented) */
#define GDX_ERR_ALLOC (-5)       /* device or host allocation failed */

/* assembly strategies: src/Algebra.jl:397-398 */
#define GDX_SUB_ASSEMBLED_ROWS 0
#define GDX_FULLY_ASSEMBLED_ROWS 1

/* numeric precision of the cell contractions; the CSR values are FP64 */
#define GDX_PREC_FP64 0
#define GDX_PREC_FP32 1 /* mixed form: sum-factorised contractions in FP32 (3D, z-marching or direct-store kernel);
                           for GDX_OP_FLUX the cell Jacobians in FP32 (fused 3D Q1 kernel), residual FP64 */

/* operator kinds */
#define GDX_OP_DIFFUSION 0 /* affine: a(u,v) and l(v) below */
#define GDX_OP_FLUX 1      /* nonlinear flux operator at a state, see gdx_set_state */
#define GDX_OP_ELASTICITY 2 /* affine, vector-valued: linear elasticity, see below */
#define GDX_DIR_CLAMPED_X0 2 /* dir_kind: u = 0 on the face x = x0 (closure) */

/* right-hand side kinds */
#define GDX_RHS_CONST 0        /* f = rhs_const */
#define GDX_RHS_MANUFACTURED 1 /* f = -div(kappa grad uD) for the linear uD below */

typedef struct gdx_ctx gdx_ctx;
typedef struct gdx_problem gdx_problem;
typedef struct gdx_nn gdx_nn; /* learned preconditioner, declared below */

/*
 * Problem descriptor: the restricted weak-form vocabulary the GPU path lowers.
 *   a(u,v) = int( kappa * grad(v) . grad(u) ) dOmega
 *   l(v)   = int( f * v ) dOmega
 * on CartesianDiscreteModel(ranks, parts, domain, cells) with
 * ReferenceFE(lagrangian, Float64, order), Measure(Omega, qdeg) and
 * TrialFESpace(uD, V) with V = TestFESpace(model, reffe; dirichlet_tags="boundary")
 * (test/FESpacesTests.jl:12-27, test/PoissonTests.jl:14-45).
 *   kappa(x) = coef_const * exp( sum_m a_m cos(2 pi k_m . x + phi_m) )
 *   modes: nmodes rows of (k1, k2, k3, a, phi)
 *   uD(x)  = dir_coef[0] + dir_coef[1] x + dir_coef[2] y + dir_coef[3] z
 *
 * op_kind GDX_OP_FLUX (order 1, f = rhs_const, no kappa modes) replaces the
 * FEOperator(r, j, U, V, assem) of test/PLaplacianTests.jl:24-42:
 *   r(u,v)    = int( grad(v) . sigma(grad(u)) - v f ) dOmega
 *   j(u,du,v) = int( grad(v) . dsigma(grad(du), grad(u)) ) dOmega
 *   sigma(g)  = (flux_alpha + flux_beta |g|^2)^((flux_p - 2)/2) g
 * (the reference test is flux_p = 4, flux_alpha = flux_beta = 1; the
 * p-Laplacian is flux_p = p with flux_alpha = eps^2 > 0, flux_beta = 1).
 *
 * op_kind GDX_OP_ELASTICITY (dim 3, order 1 or 2, no kappa modes) is linear
 * elasticity on ReferenceFE(lagrangian, VectorValue{3,Float64}, order):
 *   a(u,v) = int( eps(v) : (lame_lambda tr(eps(u)) I + 2 lame_mu eps(u)) ) dOmega
 *   l(v)   = int( v . body_force ) dOmega
 * with dir_kind 0 or GDX_DIR_CLAMPED_X0 (u = 0 at every node of the face
 * x = x0, its edges and vertices included; dir_coef must be zero).  Each node's
 * three components are consecutive dofs (node-major); cell dofs are
 * component-major, which fixes the first-encounter order of ghost columns.
 */
typedef struct gdx_problem_desc {
  int32_t dim;      /* 2 or 3 */
  int32_t order;    /* 1 or 2 */
  int32_t qdeg;     /* quadrature degree: qdeg/2+1 Gauss points per direction (<= 4) */
  int32_t strategy; /* GDX_SUB_ASSEMBLED_ROWS or GDX_FULLY_ASSEMBLED_ROWS */
  int64_t cells[3];
  int32_t parts[3];
  double domain[6]; /* x0 x1 y0 y1 z0 z1 */
  double coef_const;
  int32_t nmodes;
  const double *modes;
  int32_t rhs_kind;
  double rhs_const;
  int32_t dir_kind; /* 0: none, 1: whole boundary, GDX_DIR_CLAMPED_X0 (elasticity) */
  double dir_coef[4];
  int32_t op_kind;  /* GDX_OP_DIFFUSION, GDX_OP_FLUX or GDX_OP_ELASTICITY */
  double flux_p, flux_alpha, flux_beta;
  double lame_lambda, lame_mu; /* GDX_OP_ELASTICITY */
  double body_force[3];
} gdx_problem_desc;

/* sizes reported per part by gdx_part_sizes (index into the int64 array) */
#define GDX_SZ_NLDOFS 0      /* free local dofs of the FE space (own + ghost) */
#define GDX_SZ_NOWN 1        /* owned dofs = owned rows = owned cols */
#define GDX_SZ_NGHOST_ROWS 2 /* ghost rows of the matrix row PRange */
#define GDX_SZ_NGHOST_COLS 3 /* ghost cols of the matrix col PRange */
#define GDX_SZ_NNZ 4         /* stored entries of the local CSR */
#define GDX_SZ_COUNT 5

/* library */
int gdx_version(int32_t *major, int32_t *minor, int32_t *patch);
const char *gdx_last_error(void);
/* replaces with_debug()/distribute: one context per GPU device */
int gdx_init(int32_t device, gdx_ctx **out);
int gdx_finalize(gdx_ctx *ctx);

/* FE space + assembler construction.
   Replaces CartesianDiscreteModel(ranks,parts,...) (src/Geometry.jl:378-411),
   FESpace/TrialFESpace (src/FESpaces.jl:570-614,396-401) and
   SparseMatrixAssembler(U,V,strategy) (src/FESpaces.jl:819-862).
   ctx == NULL makes a host-only problem: gdx_symbolic and the structure
   exports work (values and RHS export as zeros, like allocate_matrix_and_vector),
   gdx_numeric returns GDX_ERR_STATE. */
int gdx_problem_create(gdx_ctx *ctx, const gdx_problem_desc *desc, gdx_problem **out);
int gdx_problem_destroy(gdx_problem *prob);

/* Symbolic phase, run once: partition, dof numbering and generate_gids
   (src/FESpaces.jl:139-261), row/col PRanges (src/Algebra.jl:1136-1257),
   local CSR pattern and the cell-dof -> CSR position map
   (src/Algebra.jl:410-418,849-857 and the compression of :636). */
int gdx_symbolic(gdx_problem *prob);

/* Numeric phase: fills the CSR values and the RHS on the device.  Replaces
   numeric_loop_matrix_and_vector! (src/FESpaces.jl:794-798) followed by the
   owner/ghost exchange and compression (src/Algebra.jl:593-646, :1048-1131)
   or, on an allocated matrix, assemble_matrix_and_vector! + assemble!
   (src/Algebra.jl:530-535,755-758).  Ghost rows end with zero values. */
int gdx_numeric(gdx_problem *prob, int32_t precision);

/* 3D matrix kernel of a problem: variant 0 = z-marching columns with rows_z z
   rows per workgroup (default 32), 1 = per-tile kernel, 2 = per-tile kernel
   storing straight from registers, 3 = per-tile kernel whose waves own whole
   rows (Q2 only; rows_z ignored for 1-3).  All give the same matrix to
   rounding; gdx_batch_autotune picks the fastest. */
int gdx_problem_set_kernel(gdx_problem *prob, int32_t variant, int32_t rows_z);
/* variant flag: every tile reads its position patterns from HBM, as the tiles
   with more distinct patterns than the kernel's LDS holds always do (same
   values; lets tests run those kernel instances on small meshes) */
#define GDX_KERNEL_HBM_PATTERNS 0x100
int gdx_problem_get_kernel(gdx_problem *prob, int32_t *variant, int32_t *rows_z);

/* queries / exports (copies device data into caller arrays) */
int gdx_num_parts(gdx_problem *prob, int32_t *nparts);
int gdx_num_global_dofs(gdx_problem *prob, int64_t *ngdofs);
int gdx_part_sizes(gdx_problem *prob, int32_t part, int64_t *sizes /* [GDX_SZ_COUNT] */);
/* FE dof PRange of a part (LocalIndices of src/FESpaces.jl:249-255) plus the
   global lattice node of each free local dof */
int gdx_export_fe(gdx_problem *prob, int32_t part, int64_t *ldof_gid, int32_t *ldof_owner, int64_t *ldof_gnode);
/* row / col PRanges: own then ghost (OwnAndGhostIndices, src/Algebra.jl:1240-1252) */
int gdx_export_rows(gdx_problem *prob, int32_t part, int64_t *gid, int32_t *owner);
int gdx_export_cols(gdx_problem *prob, int32_t part, int64_t *gid, int32_t *owner);
/* local CSR (SparseMatrixCSR{0,Float64,Int} layout, test/PoissonTests.jl:40) */
int gdx_export_csr(gdx_problem *prob, int32_t part, int64_t *rowptr, int32_t *colidx, double *values);
/* RHS on the row PRange of A (src/Algebra.jl:720-753) */
int gdx_export_rhs(gdx_problem *prob, int32_t part, double *b);

/* Solve side, on the assembled owned rows in global-id order (the system the
   reference solves with A\b, README.md:76, and applies with A*x / mul!,
   test/FESpacesTests.jl:24).  x, y, xsol are host arrays of length ngdofs. */
int gdx_spmv(gdx_problem *prob, const double *x, double *y);
/* Device time (ms) of one SpMV of the assembled matrix, averaged over reps
   launches between two hipEvents on the library's stream (the solver's kernel
   on its own vectors; a measurement helper with no counterpart in the reference).
   bytes (may be NULL) receives the SpMV's algorithmic bytes per launch: values
   (8 B), column ids (4 B, or 2 B in blocks read as 16-bit offsets) and 28 B per
   row (row start, length, x and y). */
int gdx_spmv_timed(gdx_problem *prob, int32_t reps, double *ms, double *bytes);
/* Jacobi-preconditioned CG from x0 = 0 until ||r|| <= rtol ||b||; relres is the
   true FP64 residual ||b - A x|| / ||b|| recomputed at the end, ms the wall time
   of the solve on the device (the copy of the solution to xsol is outside it) */
int gdx_pcg(gdx_problem *prob, double rtol, int32_t maxiter, double *xsol, int32_t *iters, double *relres,
            double *ms);

/* Cell-wise coefficient (the lognormal fields of the learned-preconditioner
   workload): kappa is coef_const * kappa_c on every cell c instead of the
   Fourier form; the source must be constant (GDX_RHS_CONST).  Replaces a
   CellField coefficient of the bilinear form a(u,v) = int( kappa grad(v).grad(u) ).
   gdx_set_cell_coefficient: host array of ncells values (x fastest), NULL returns
   to the descriptor's kappa.  gdx_cell_coefficient_grf: kappa = exp(log_scale g),
   g a unit-variance Gaussian random field with correlation exp(-r^2 / (2 corr_len^2))
   drawn on the device from seed.  Both take effect at the next gdx_numeric. */
int gdx_set_cell_coefficient(gdx_problem *prob, const double *kappa);
int gdx_cell_coefficient_grf(gdx_problem *prob, uint64_t seed, double corr_len, double log_scale);
int gdx_export_cell_coefficient(gdx_problem *prob, double *kappa);
/* Piecewise-constant material for GDX_OP_ELASTICITY: the constitutive law
   sigma(eps) = lambda_c tr(eps) I + 2 mu_c eps on cell c (the Lame parameters of
   a(u,v) = int( eps(v) : sigma(eps(u)) ) as CellFields,
   test/MultiFieldTests.jl:25 for the vector-valued space).
   gdx_set_cell_lame: host array of 2 ncells values, (lambda_c, mu_c) per cell,
   cells x fastest; each pair needs mu_c > 0 and lambda_c > -2 mu_c / 3; NULL
   returns to the descriptor's constants.  On an elasticity problem
   gdx_set_cell_coefficient and gdx_cell_coefficient_grf set kappa_c scaling both
   constants (a Young's modulus field at fixed Poisson ratio).  The last call of
   the three decides; it takes effect at the next gdx_numeric. */
int gdx_set_cell_lame(gdx_problem *prob, const double *lame);

/* Mapped Cartesian grid: CartesianDiscreteModel(ranks, parts, domain, cells;
   map) (src/Geometry.jl:378-388, the map forwarded to the CartesianDescriptor).
   xyz holds the mapped coordinates of every mesh vertex, dim values each,
   vertices x fastest over (cells_0 + 1) x (cells_1 + 1) [x (cells_2 + 1)];
   each cell takes the multilinear (Q1) map of its vertices and a Q2 node sits
   at that map's image of its reference position (the mean of the vertices of
   the edge, face or cell whose midpoint it is).  Quadrature uses |det J| at
   every Gauss point; kappa and f are evaluated at the mapped points, uD at the
   mapped nodes.  Diffusion only (GDX_ERR_UNSUPPORTED otherwise), FP64; NULL
   returns to the uniform box.  Call after gdx_symbolic (a new symbolic phase
   returns to the uniform box); takes effect at the next gdx_numeric. */
int gdx_set_vertex_coords(gdx_problem *prob, const double *xyz);
/* Diffusion-reaction form on the mapped-grid path:
   a(u,v) = int( diffusion_weight kappa grad(u).grad(v) + reaction u v ), the
   terms of a multi-field block such as u2 v2 or u1 v2 in
   test/BlockSparseMatrixAssemblersTests.jl:110-114.  Needs
   gdx_set_vertex_coords (the uniform box's own vertices for an unmapped mesh);
   clearing the coordinates or a new symbolic phase returns to (1, 0).  Weights
   finite, non-negative, not both zero. */
int gdx_set_reaction(gdx_problem *prob, double diffusion_weight, double reaction);

/* Nonlinear operators (GDX_OP_FLUX).  The state uh of
   residual_and_jacobian(op, uh) (test/PLaplacianTests.jl:44-48) is given by its
   free values x (host array of length ngdofs, global-id order; NULL = zero);
   Dirichlet values come from the descriptor's uD.  After gdx_set_state,
   gdx_numeric assembles the Jacobian into the CSR values and the residual into
   the RHS (residual_and_jacobian; called again it is residual_and_jacobian!). */
int gdx_set_state(gdx_problem *prob, const double *x);

#define GDX_NEWTON_INIT_GIVEN 0   /* start from x as passed (NULL = zero) */
#define GDX_NEWTON_INIT_POISSON 1 /* first solve the linear problem flux_p = 2, alpha = 1, beta = 0 */
typedef struct gdx_newton_opts {
  double rtol;         /* stop when ||r(x)|| <= max(atol, rtol ||r(x_given)||) */
  double atol;
  int32_t maxiter;     /* Newton iterations */
  double lin_rtol;     /* relative tolerance of each Newton correction's linear solve */
  int32_t lin_maxiter;
  int32_t init;        /* GDX_NEWTON_INIT_* */
  gdx_nn *precond;     /* NULL: Jacobi-PCG corrections.  Else flexible PCG with this
                          network's multilevel preconditioner (gdx_nn_set_options and
                          gdx_nn_set_multigrid apply), re-attached to every new
                          Jacobian; the problem must fit the network (2n x 2n x 2n
                          Q1 cells). */
  int32_t jac_precision; /* GDX_PREC_FP64, or GDX_PREC_FP32: the Jacobians' cell
                            contractions in FP32 (the fused 3D Q1 flux kernel),
                            rows summed and stored in FP64; the residual, the
                            Newton test and the corrections' solves stay FP64 */
} gdx_newton_opts;
typedef struct gdx_newton_report {
  int32_t iters;       /* Newton corrections applied */
  int32_t lin_iters;   /* PCG iterations over all corrections (initial guess included) */
  int32_t converged;
  double res_ref;      /* ||r(x_given)|| (2-norm over free dofs) */
  double res_final;    /* ||r(x)|| at the returned x */
  double ms_total, ms_assembly, ms_solve; /* wall times; ms_total from the initial guess
                                             on the device to the solution there (the
                                             host copies of x are outside it) */
  int32_t lin_fallbacks; /* corrections in which FPCG fell back to Jacobi-PCG */
} gdx_newton_report;
/* Newton's method on the flux operator: NLSolver(method=:newton) +
   FESolver + solve(solver, op) (test/PLaplacianTests.jl:54-57), with the
   Jacobian and residual re-assembled in place every iteration and each
   correction solved on the device (Jacobi-PCG, or FPCG with opts->precond).  x: host array of length
   ngdofs, in = initial guess (NULL with GDX_NEWTON_INIT_GIVEN means zero),
   out = the solution's free values. */
int gdx_newton(gdx_problem *prob, const gdx_newton_opts *opts, double *x, gdx_newton_report *report);

/* Learned preconditioner.  A small BF16 convolutional multilevel network
   (3 levels on n^3, n^3/8, n^3/64 grids, channels 8/16/24, 3x3x3 convolutions,
   about 60k parameters) gives the coarse-grid guess of a multilevel
   preconditioner (damped Jacobi smoothing, full weighting, trilinear
   prolongation, Galerkin coarse operators with a V-cycle below the network's
   grid) for 3D Q1 problems on 2n x 2n x 2n cells.  It is trained on the
   device by the library's own forward/backward kernels from systems assembled
   by gdx_numeric, and used inside flexible PCG (gdx_fpcg).  These entry points
   have no counterpart in the reference, whose solve is a direct LU
   (README.md:76); they extend its A\b. */
int gdx_nn_create(gdx_ctx *ctx, int32_t n, uint64_t seed, gdx_nn **out);
int gdx_nn_destroy(gdx_nn *nn);
int gdx_nn_num_params(gdx_nn *nn, int64_t *nparams);
/* FP32 master parameters, layer by layer: W[out][in][27] then b[out] */
int gdx_nn_save(gdx_nn *nn, float *theta);
int gdx_nn_load(gdx_nn *nn, const float *theta);
/* network alone on host arrays: inp [2][n^3], out [n^3]; backward of sum(dout * out)
   gives grad [nparams] and, when dinp is not NULL, the input gradient [2][n^3] */
int gdx_nn_forward(gdx_nn *nn, const float *inp, float *out);
int gdx_nn_backward(gdx_nn *nn, const float *dout, float *grad, float *dinp);
/* bind to a problem's current matrix (call after every gdx_numeric that changes it) */
int gdx_nn_attach(gdx_nn *nn, gdx_problem *prob);
/* one Adam step on a fresh sample (smooth coarse error of the given seed,
   smoothed by smooth_passes 3x3x3 means) from the attached matrix; loss = relative squared error */
int gdx_nn_train_step(gdx_nn *nn, uint64_t seed, float lr, int32_t smooth_passes, float *loss);
/* coarse damped-Jacobi sweeps on the Galerkin coarse operator after the
   network's coarse solve in the two-grid form (default 8; see
   gdx_nn_set_multigrid); use_network = 0 starts the coarse correction from
   zero instead (the same preconditioner without the network) */
int gdx_nn_set_options(gdx_nn *nn, int32_t coarse_sweeps, int32_t use_network);
/* coarse correction form: enable = 1 (the default) runs a V-cycle from the
   network's guess: nu damped-Jacobi sweeps before and after each coarse-grid
   correction, Galerkin operators P^T A P computed on the device level by
   level, halving the grid while its cell count is even and above 4, damped
   Jacobi sweeps in one workgroup on the coarsest grid (<= 8^3 cells).
   enable = 0 uses the coarse sweeps of gdx_nn_set_options (two-grid form).
   levels (may be NULL): grids of the V-cycle, the network grid included. */
int gdx_nn_set_multigrid(gdx_nn *nn, int32_t enable, int32_t nu, int32_t *levels);
/* level l of the V-cycle (0 = the network grid): its cells per direction m
   and, when stencil is not NULL, its Galerkin operator as [27][m^3] in the
   layout of gdx_nn_coarse_operator */
int gdx_nn_level_operator(gdx_nn *nn, int32_t level, int32_t *m, double *stencil);
/* the Galerkin coarse operator of the attached matrix as a 27-point stencil:
   stencil[(dx+1) + 3(dy+1) + 9(dz+1)][j] on the network grid (j = I - 1 per
   direction for interior coarse node I; the last plane is zero padding) */
int gdx_nn_coarse_operator(gdx_nn *nn, double *stencil);
/* z = M r on host vectors (global-id order) */
int gdx_nn_apply(gdx_nn *nn, const double *r, double *z);
/* gdx_nn_apply without any device allocation.  No counterpart in the
   reference: it is for callers that cannot allocate (serving loops, harnesses
   that forbid hipMalloc).  The device staging of r and z lives in a static
   array of 32 MiB that each GPU running the library holds until the process
   ends, so the attached problem may have at most 2,097,152 free dofs (larger:
   GDX_ERR_ARG).  Every call on a GPU shares that array: calls that overlap on
   one GPU (several contexts or threads) race.  Same kernels and results as
   gdx_nn_apply. */
int gdx_nn_apply_noalloc(gdx_nn *nn, const double *r, double *z);
/* Flexible PCG with the learned multilevel preconditioner (nn attached to prob),
   or Jacobi-PCG when nn is NULL; falls back to Jacobi-PCG from the current
   iterate when the flexible iteration stagnates (fellback = 1).  rhs: host
   array of ngdofs (NULL = the assembled RHS); stops at ||r|| <= rtol ||b||,
   relres = true FP64 residual recomputed at the end; ms = wall time of the
   solve from the right-hand side on the device to the solution there (the
   host copies of rhs and xsol are outside it). */
int gdx_fpcg(gdx_problem *prob, gdx_nn *nn, const double *rhs, double rtol, int32_t maxiter, double *xsol,
             int32_t *iters, double *relres, double *ms, int32_t *fellback);

/* Batched numeric assembly: independent problems of uneven sizes and
   partitions (each after gdx_symbolic, all on one context) assembled by one
   hipGraph launch.  gdx_batch_autotune times each 3D problem's matrix kernels
   (z-marching columns of 16, 32 and 64 z rows per workgroup, the per-tile
   kernel and its direct-store form) and keeps the fastest (chosen, per
   problem, may be NULL: the z rows per workgroup, -1 for the per-tile kernel,
   -2 for the direct-store form, -3 for the wave-slab form); the matrices
   agree to rounding for every choice.  gdx_batch_capture records the numeric launches of every problem
   (no memsets) as independent branches of one graph: the problems are spread over four streams
   (largest first, to the least loaded), forked from and joined back into the context's stream, so
   small problems run beside each other; gdx_batch_launch replays the graph on the context's stream
   and the results equal separate gdx_numeric calls bit for bit.
   The reference has no counterpart: it assembles one problem per call
   (src/FESpaces.jl:785-798). */
typedef struct gdx_batch gdx_batch;
int gdx_batch_create(gdx_ctx *ctx, gdx_batch **out);
int gdx_batch_add(gdx_batch *batch, gdx_problem *prob);
int gdx_batch_autotune(gdx_batch *batch, int32_t reps, int32_t *chosen);
int gdx_batch_capture(gdx_batch *batch);
int gdx_batch_launch(gdx_batch *batch);
int gdx_batch_destroy(gdx_batch *batch);

/* Device PVectors on a problem's row or column PRange and the PArrays operations
   of the solve side.  A PVector holds every part's local vector (own entries,
   then ghost entries, in the order of gdx_export_rows / gdx_export_cols)
   concatenated in part order, in device memory the library owns; on the row
   PRange that is the layout of the assembled b.  PVectors belong to their
   problem: gdx_problem_destroy (or gdx_finalize) releases their memory and
   their later calls return GDX_ERR_STATE; gdx_pvector_destroy frees the handle.
     consistent!    ghost entries take their owner's value
     assemble!      ghost entries are added into their owner's entry (owner
                    first, then the copies in ascending part and local index),
                    then zeroed
     axpy!          y = alpha x + y, then consistent!(y)      src/Algebra.jl:42-49
     change_ghost   own values of src onto dst's PRange, dst's ghosts zero,
                    then consistent!(dst) if make_consistent  src/Algebra.jl:270-290
     mul!(y, A, x)  consistent!(x), then every part's local CSR times its local
                    x; y on A's row PRange, x on its column PRange (the
                    PSparseMatrix product the reference's tests use,
                    test/PLaplacianTests.jl:50-52)
   Calls queue work on the context's stream; set and get synchronize. */
#define GDX_PRANGE_ROWS 0
#define GDX_PRANGE_COLS 1
typedef struct gdx_pvector gdx_pvector;
int gdx_pvector_create(gdx_problem *prob, int32_t which, gdx_pvector **out);
int gdx_pvector_destroy(gdx_pvector *v);
int gdx_pvector_length(gdx_pvector *v, int64_t *n);
int gdx_pvector_device(gdx_pvector *v, double **data);
int gdx_pvector_set(gdx_pvector *v, const double *host);
int gdx_pvector_get(gdx_pvector *v, double *host);
int gdx_pvector_fill(gdx_pvector *v, double value);
int gdx_pvector_copy_rhs(gdx_pvector *v);
int gdx_pvector_consistent(gdx_pvector *v);
int gdx_pvector_assemble(gdx_pvector *v);
int gdx_pvector_axpy(double alpha, gdx_pvector *x, gdx_pvector *y);
int gdx_pvector_change_ghost(gdx_pvector *src, gdx_pvector *dst, int32_t make_consistent);
int gdx_psparse_mul(gdx_problem *prob, gdx_pvector *y, gdx_pvector *x);

/* device views and timing */
int gdx_device_values(gdx_problem *prob, double **values, int64_t *nnz_total);
int gdx_stream(gdx_problem *prob, void **stream);
int gdx_sync(gdx_problem *prob);
/* per-phase device times (ms) of the last gdx_numeric, measured with hipEvents
   on the library's stream: [0] coefficient, [1] rhs, [2] matrix, [3] total */
int gdx_phase_times(gdx_problem *prob, double *ms, int32_t n);
/* number of owned free x free stored entries (the Gnnz/s numerator) */
int gdx_owned_nnz(gdx_problem *prob, int64_t *nnz);

#ifdef __cplusplus
}
#endif
#endif /* GDX_H */

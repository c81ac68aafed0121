// Internal data structures of libgdx (host side).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/gdx.h"

namespace gdx {

constexpr int kMaxQ = 4;     // Gauss points per direction
constexpr int kMaxK = 2;     // polynomial order
constexpr uint8_t kSlotAbsent = 0xFF;
constexpr uint8_t kSlotDirichlet = 0xFE;
// tiles of the fused 3D flux kernel: 8 x kFluxR2 x kFluxR3 owned rows, kFluxNT
// lanes (8 x 4 x 4: the cell records of two workgroups fit in one CU's LDS)
constexpr int kFluxR2 = 4, kFluxR3 = 4, kFluxNT = 256;

// One part (subdomain) of the Cartesian partition; all parts live on one GPU.
struct Part {
  int rank = 0, pc[3] = {0, 0, 0};
  int64_t own_lo[3], own_hi[3];  // owned cells [lo,hi)
  int64_t loc_lo[3], loc_hi[3];  // owned + one ghost layer
  int64_t lnn[3];                // local node counts
  std::vector<int64_t> lnode_dof;  // local node -> free ldof (>=0) or -1
  std::vector<int64_t> ldof_lnode;
  std::vector<int32_t> ldof_owner;
  std::vector<int64_t> ldof_gid;
  int64_t n_ldofs = 0, n_own = 0;
  std::vector<int64_t> own_ldofs;
  std::vector<int64_t> ldof_col;  // ldof -> local col id (or -1)
  // matrix PRanges
  std::vector<int64_t> row_gid, col_gid;
  std::vector<int32_t> row_owner, col_owner;
  int64_t nrows = 0, ncols = 0, n_ghost_rows = 0, n_ghost_cols = 0;
  std::vector<int64_t> row_ldof;  // local row -> ldof
  // local CSR pattern (host copies: made on first host use when the device
  // symbolic phase built them)
  std::vector<int64_t> rowptr;
  std::vector<int32_t> colidx;
  int64_t nnz = 0;  // stored entries of the local CSR
  int64_t val_base = 0, row_base = 0;
  // offsets of the part's arrays among all parts' (device symbolic phase)
  int64_t lnode_base = 0, ldof_base = 0, own_base = 0, col_base = 0;
  // owned node box (global node coords)
  int64_t onb_lo[3], onb_n[3];
  int64_t rowinfo_base = 0;
};

// Per owned node of a part's owned node box (device layout).
struct RowInfo {
  long long rowstart;  // offset of the row's first value in the global values buffer
  long long bidx;      // offset of the row in the global RHS buffer
  int pid;             // position pattern id, -1 when the node is not a free row
  int lenf;            // row length | has-Dirichlet-neighbours << 16 | tile-local pattern index << 17
};

// Pattern rows are padded to whole dwords (stride of pat_fwd / pat_inv), with
// at least one spare byte (pat_invr keeps the Dirichlet count there).
constexpr int pattern_stride(int S) { return (S + 4) & ~3; }

// Reduced stencil numbering: along a dimension a node with i % K == 0 couples
// to offsets -K..K (2K+1 slots), any other node to -1..1 (3 slots).  A row's
// reduced slot is x-fastest over its own box of coupled offsets.
__host__ __device__ constexpr int stencil_width(int K, int i) { return (i % K == 0) ? 2 * K + 1 : 3; }
__host__ __device__ constexpr int stencil_lo(int K, int i) { return (i % K == 0) ? -K : -1; }

// A tile of owned rows processed by one workgroup.
struct Tile {
  int part;
  int org[3];     // global node coords of the tile origin (multiple of order)
  int ext[3];     // rows of the tile inside the owned node box
  int onb_lo[3];  // owned node box of the part
  int onb_n[3];
  int npat;       // distinct position patterns of the tile's rows
  long long rowinfo_base;
  long long pat_off;  // offset of the tile's pattern id list in tile_pats
};

// Problem-wide constants passed to the kernels by value.
struct KParams {
  int dim, K, NQ, S;             // S = (2K+1)^dim stencil slots
  int ncells[3];
  int qn[3];                     // qp lattice extents NQ*ncells
  int nn[3];                     // node lattice extents K*ncells+1
  double x0[3], h[3], hK[3];  // hK = h / K: node spacing
  double phi[kMaxK + 1][kMaxQ], dphi[kMaxK + 1][kMaxQ], qw[kMaxQ], qx[kMaxQ];
  // 1D pair factors per dimension: G = w dphi_a dphi_b / h, M = w h phi_a phi_b
  double fG[3][kMaxK + 1][kMaxK + 1][kMaxQ], fM[3][kMaxK + 1][kMaxK + 1][kMaxQ];
  double coef_const;
  int nmodes;
  int rhs_kind;
  double rhs_const;
  int dir_kind;
  double dir_coef[4];
};

// Numbering, global ids and PRanges made by the device symbolic phase: each
// array holds every part's array back to back in part order (offsets in
// Part::lnode_base, ldof_base, own_base, row_base, col_base).
struct DevNumbering {
  int64_t *lnode_dof = nullptr, *ldof_lnode = nullptr, *ldof_gid = nullptr, *own_ldofs = nullptr,
          *ldof_col = nullptr, *row_ldof = nullptr, *row_gid = nullptr, *col_gid = nullptr;
  int32_t *ldof_owner = nullptr, *row_owner = nullptr, *col_owner = nullptr;
};

struct Device {
  double *W = nullptr;       // kappa at quadrature points (global qp lattice)
  double *F = nullptr;       // f at quadrature points
  double *E = nullptr;       // separable Fourier tables [3][nmodes][qn][2]
  double *modes = nullptr;   // [nmodes][5]
  double *values = nullptr;  // concatenated local CSR values of all parts
  double *b = nullptr;       // concatenated RHS (rows of every part)
  RowInfo *rowinfo = nullptr;
  uint8_t *pat_fwd = nullptr;  // [npat][S]: slot -> offset | absent | Dirichlet
  uint8_t *pat_inv = nullptr;  // [npat][S]: offset -> slot
  uint8_t *pat_invr = nullptr; // [npat][SP]: offset -> reduced slot; Dirichlet slots after len; last byte = their count
  int *pat_len = nullptr;
  int *pat_dir = nullptr;      // pattern has Dirichlet slots
  Tile *tiles = nullptr;
  int *tile_pats = nullptr;    // per tile: its distinct pattern ids
  double *cellk = nullptr;     // kappa per cell (cell-wise coefficient mode)
  double *cell_lame = nullptr; // elasticity: (lambda, mu) per cell (cell-wise Lame mode)
  // local CSR of every part, concatenated (device symbolic phase): global row
  // offsets (rows_total + 1 entries) and local column ids
  long long *csr_rowptr = nullptr;
  int *csr_colidx = nullptr;
};

}  // namespace gdx

struct gdx_nn;
struct gdx_batch;
struct gdx_pvector;

struct gdx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // live objects created on this context: gdx_finalize releases their device
  // resources and detaches them (their later calls return GDX_ERR_STATE)
  std::vector<gdx_problem *> probs;
  std::vector<gdx_nn *> nns;
  std::vector<gdx_batch *> batches;
};

struct gdx_problem {
  gdx_ctx *ctx = nullptr;
  gdx_problem_desc desc{};
  std::vector<double> modes;
  int nparts = 0;
  int K = 1, D = 2, NQ = 2;
  int ncomp = 1;  // dofs per node: 3 for GDX_OP_ELASTICITY
  int64_t nn[3] = {1, 1, 1};
  int64_t ngdofs = 0;
  std::vector<gdx::Part> P;
  bool symbolic_done = false;
  // structured-kernel data (host copies)
  std::vector<gdx::RowInfo> rowinfo;
  std::vector<uint8_t> pat_fwd, pat_inv, pat_invr;
  std::vector<int> pat_len, pat_dir;
  // vector-valued rows (ncomp > 1): position -> stencil slot * ncomp + column component
  std::vector<uint16_t> pat16_inv;
  int pat16_stride = 0;
  int npat = 0;
  std::vector<gdx::Tile> tiles;
  std::vector<int> tile_pats;
  int n_fast_tiles = 0;  // leading tiles whose position patterns fit in LDS
  int kvariant3d = 0;    // 3D matrix kernel: 0 z-marching columns, 1 per-tile, 2 per-tile direct-store,
                         // 3 per-tile wave-slab (Q2); chosen by the autotuner
  bool hbm_patterns = false;  // every tile takes the HBM-pattern kernel instance (GDX_KERNEL_HBM_PATTERNS)
  int precision = 0;     // contraction precision of the current numeric phase (GDX_PREC_*)
  int tileR[3] = {1, 1, 1};
  int64_t nnz_total = 0, rows_total = 0, owned_nnz = 0;
  // the device symbolic phase built the CSR, row table and tiles in HBM; their
  // host copies (P.rowptr/colidx, rowinfo) are made on first host use
  bool dev_structure = false;
  bool host_csr_stale = false, host_rowinfo_stale = false;
  // the numbering, ids and PRanges were made on the device (in dn); their host
  // copies in P are made on first host use
  bool dev_numbering = false, host_fe_stale = false;
  gdx::DevNumbering dn;
  gdx::KParams kp{};
  gdx::Device dev;
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  double phase_ms[4] = {0, 0, 0, 0};
  void *solve = nullptr;  // solver state (gdx_solve.hip), built on the first solve
  void *flux = nullptr;   // nonlinear operator state (gdx_flux.hip)
  void *elast = nullptr;  // elasticity tables (gdx_elast.hip)
  bool coef_cell = false; // kappa from dev.cellk instead of the Fourier modes
                          // (elasticity: kappa scales lambda and mu on each cell)
  bool lame_cell = false; // elasticity: (lambda, mu) from dev.cell_lame
  bool mapped = false;    // diffusion on a mapped grid (vertex coordinates in mapped_data)
  double diff_w = 1.0, reac_w = 0.0;  // mapped grids: diff_w kappa grad(u).grad(v) + reac_w u v
  void *mapped_data = nullptr;  // gdx_mapped.hip
  void *pv = nullptr;     // PRange maps of the device PVectors (gdx_pvector.hip)
  std::vector<gdx_pvector *> pvecs;  // live PVectors on this problem
};

namespace gdx {
void set_error(const std::string &msg);
void release_problem_device(gdx_problem *p);  // device buffers, solver and flux state, events
void release_nn_device(gdx_nn *nn);
void release_batch_device(gdx_batch *b);
void pvector_release(gdx_problem *p);  // PRange maps; the problem's PVectors lose their device memory
void batch_forget_problem(gdx_ctx *ctx, gdx_problem *p);  // a destroyed problem leaves every batch
// every batch holding p drops its graph (its launch returns GDX_ERR_STATE until
// captured again): called whenever p's buffers, kernels or coefficient mode change
void batch_invalidate_problem(gdx_ctx *ctx, gdx_problem *p);
template <class T>
void unregister(std::vector<T *> &v, T *x) {
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] == x) {
      v.erase(v.begin() + i);
      return;
    }
}
int symbolic(gdx_problem *p);  // symbolic phase (gdx_symbolic.cpp)
// The device allocations of the symbolic phase go through dev_malloc: it is
// hipMalloc, except that GDX_FAULT_ALLOC=n, read when gdx_symbolic starts
// (FaultScope), makes the n-th of that call's allocations (from 0) fail as
// exhausted HBM would, once, so that the error paths can be tested.
hipError_t dev_malloc(void **p, size_t bytes);
template <class T>
hipError_t dev_malloc(T **p, size_t bytes) {
  return dev_malloc(reinterpret_cast<void **>(p), bytes);
}
struct FaultScope {
  FaultScope();
  ~FaultScope();
};
// device stage of the symbolic phase (gdx_symbolic_dev.hip): CSR, position
// patterns and row table from the host numbering and PRanges.  Returns
// GDX_OK, a GDX_ERR_* code, or kUseHost when the host stage must build them.
constexpr int kUseHost = 1;
int structure_device(gdx_problem *p);
// GDX_TRACE_SYMBOLIC=1: the wall time of each symbolic stage since the last
// mark (or since `start`) goes to stderr
void symbolic_mark(const char *stage, bool start = false);
int build_tiles_device(gdx_problem *p, const int R[3]);
// device numbering, owners, global ids and row/col PRanges (the part geometry
// in P is set): owner_dim[d][N] = part index along d owning lattice node N;
// off = the reference cell's dofs as order-K lattice offsets (ncd of them);
// E, F = edges and faces of the cell polytope as doubled-lattice offsets
int numbering_device(gdx_problem *p, const std::vector<int> owner_dim[3], const int off[27][3], int ncd,
                     const int (*E)[3], int nE, const int (*F)[3], int nF);
void free_numbering(gdx_problem *p);
int host_fe(gdx_problem *p);  // host copies of the numbering, ids and PRanges (no-op when current)
// free global id -> lattice node (x fastest), on the device, from the device numbering
int gid_node_device(gdx_problem *p, long long *gid_node);
int host_csr(gdx_problem *p);      // host copies of the local CSRs (no-op when current)
int host_rowinfo(gdx_problem *p);  // host copy of the row table (no-op when current)
// tables of one position pattern from its key (slot -> offset, kSlotAbsent or
// kSlotDirichlet) and the node coordinates of a row that has it; returns
// whether the pattern has Dirichlet slots
bool scalar_pattern_tables(int D, int K, const uint8_t *key, int S, int SP, const int64_t g[3], int len,
                           uint8_t *fwd, uint8_t *inv, uint8_t *invr);
// vector-valued rows: key over slot * ncomp + component (16 bit: offset, 0xFFFE
// Dirichlet, 0xFFFF absent); inv = position -> slot * ncomp + component;
// returns whether the pattern has Dirichlet slots
bool vector_pattern_inv(const uint16_t *key, int SN, int SP, uint16_t *inv);
// device address of the no-alloc staging array (gdx_twogrid.hip): resolved once
// per device with that device current (gdx_init), then looked up
int noalloc_setup(int device);
double *noalloc_scratch(int device);
void fill_kparams(gdx_problem *p);
void tile_shape(int dim, int K, int R[3]);
void build_tiles(gdx_problem *p, const int R[3]);
int retile(gdx_problem *p, const int R[3]);  // rebuild, order and upload the tiles (gdx_capi.hip)
int numeric_launches(gdx_problem *p, hipStream_t st);
int pattern_capacity_3d(int K, int NQ);
int pattern_capacity_tile3d(int K, int NQ);
void tile3d_shape(int K, int R[3]);
int launch_matrix_tile3d(gdx_problem *p, hipStream_t st);
int pattern_capacity_tile3d_direct(int K, int NQ);
int launch_matrix_tile3d_direct(gdx_problem *p, hipStream_t st);
bool wslab_supported(int K, int NQ);
int pattern_capacity_tile3d_wslab(int K, int NQ);
int launch_matrix_tile3d_wslab(gdx_problem *p, hipStream_t st);
int set_kernel_variant(gdx_problem *p, int variant, int rows_z);  // retiles for the variant
void solve_free(gdx_problem *p);
int solve_newton(gdx_problem *p, const gdx_newton_opts *o, double *x, gdx_newton_report *rep);
// preconditioner callback of the flexible CG: z = M r (device vectors, global-id order)
struct FpcgPrec {
  int (*apply)(void *ctx, const double *r, double *z, hipStream_t st);
  void *ctx;
};
int solve_fpcg(gdx_problem *p, const FpcgPrec *prec, const double *rhs_h, double rtol, int maxiter, double *xh,
               int *iters, double *relres, double *ms, int *fellback);
int solve_dinv(gdx_problem *p, const double **dinv, int *n);
int solve_csr_view(gdx_problem *p, const long long **rstart, const int **rlen, const int **cgid);
void solve_spmv_dev(gdx_problem *p, const double *x, double *y, hipStream_t st);
bool flux_supported(int dim, int order, int nq);
bool flux_mixed_supported(const gdx_problem *p);  // the fused 3D kernel (FP32 cell Jacobians)
int64_t num_cells(const gdx_problem *p);
int cell_coef_alloc(gdx_problem *p);
int launch_cell_coefficients(gdx_problem *p, hipStream_t st);
int grf_cell_coefficient(gdx_problem *p, uint64_t seed, double corr_len, double log_scale);
int elast_setup(gdx_problem *p);  // factor tables and patterns to HBM
void elast_free(gdx_problem *p);
int elast_assemble(gdx_problem *p, hipStream_t st);
int mapped_set(gdx_problem *p, const double *xyz);  // vertex coordinates; cell records and row list
void mapped_free(gdx_problem *p);
int mapped_assemble(gdx_problem *p, hipStream_t st, int stage);  // stage 0: cell records, 1: rows
int flux_setup(gdx_problem *p);
void flux_free(gdx_problem *p);
double *flux_state(gdx_problem *p);
int flux_set_state(gdx_problem *p, const double *xh);
int flux_assemble(gdx_problem *p, const double *x, bool linear, bool j32, hipStream_t st);
int solve_spmv(gdx_problem *p, const double *x, double *y);
int solve_spmv_timed(gdx_problem *p, int reps, double *ms, double *bytes);
int solve_pcg(gdx_problem *p, double rtol, int maxiter, double *x, int *iters, double *relres, double *ms);
}  // namespace gdx

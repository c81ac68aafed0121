// C ABI of libgdx (declared in include/gdx.h).
#include <hip/hip_runtime.h>

#include <cmath>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>

#include "gdx_internal.h"
#include "gdx_nn.h"

namespace gdx {
int launch_coefficients(gdx_problem *p, hipStream_t st);
int launch_rhs(gdx_problem *p, hipStream_t st);
int launch_matrix(gdx_problem *p, hipStream_t st);
bool matrix_kernel_supported(int dim, int K, int NQ);

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

static thread_local long g_fault_alloc = -1;  // allocations left before the injected failure
FaultScope::FaultScope() {
  const char *e = std::getenv("GDX_FAULT_ALLOC");
  g_fault_alloc = (e && *e) ? std::atol(e) : -1;
}
FaultScope::~FaultScope() { g_fault_alloc = -1; }
hipError_t dev_malloc(void **p, size_t bytes) {
  if (g_fault_alloc == 0) {
    g_fault_alloc = -1;
    *p = nullptr;
    return hipErrorOutOfMemory;
  }
  if (g_fault_alloc > 0) --g_fault_alloc;
  return hipMalloc(p, bytes);
}
}  // namespace gdx

using gdx::set_error;

#define GDX_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t e_ = (call);                                                                    \
    if (e_ != hipSuccess) {                                                                    \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));                            \
      return e_ == hipErrorOutOfMemory ? GDX_ERR_ALLOC : GDX_ERR_HIP;                          \
    }                                                                                          \
  } while (0)

#define GDX_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) {                   \
      set_error(msg);                \
      return code;                   \
    }                                \
  } while (0)

namespace {

template <class T>
int upload(T **dst, const std::vector<T> &src) {
  if (*dst) {
    (void)hipFree(*dst);
    *dst = nullptr;
  }
  size_t n = src.size() ? src.size() : 1;
  GDX_HIP(gdx::dev_malloc(reinterpret_cast<void **>(dst), n * sizeof(T)));
  if (!src.empty()) GDX_HIP(hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return GDX_OK;
}

void free_device(gdx_problem *p) {
  gdx::Device &d = p->dev;
  void *ptrs[] = {d.W, d.F, d.E, d.modes, d.values, d.b, d.rowinfo, d.pat_fwd, d.pat_inv, d.pat_invr, d.pat_len,
                  d.pat_dir, d.tiles, d.tile_pats, d.cellk, d.cell_lame, d.csr_rowptr, d.csr_colidx};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  d = gdx::Device{};
  gdx::free_numbering(p);
}

int check_desc(const gdx_problem_desc *d) {
  GDX_REQUIRE(d, GDX_ERR_ARG, "problem descriptor is NULL");
  GDX_REQUIRE(d->dim == 2 || d->dim == 3, GDX_ERR_ARG, "dim must be 2 or 3");
  GDX_REQUIRE(d->order == 1 || d->order == 2, GDX_ERR_UNSUPPORTED, "only Lagrangian order 1 and 2 are supported");
  GDX_REQUIRE(d->qdeg >= 0, GDX_ERR_ARG, "quadrature degree must be >= 0");
  GDX_REQUIRE(d->qdeg / 2 + 1 <= gdx::kMaxQ, GDX_ERR_UNSUPPORTED, "quadrature degree above 7 is not supported");
  GDX_REQUIRE(d->strategy == GDX_SUB_ASSEMBLED_ROWS || d->strategy == GDX_FULLY_ASSEMBLED_ROWS, GDX_ERR_ARG,
              "unknown assembly strategy");
  for (int k = 0; k < d->dim; ++k) {
    GDX_REQUIRE(d->parts[k] >= 1, GDX_ERR_ARG, "parts must be >= 1 in every direction");
    GDX_REQUIRE(d->cells[k] >= d->parts[k], GDX_ERR_ARG, "each part needs at least one cell per direction");
    GDX_REQUIRE(d->cells[k] * d->order < (int64_t)1 << 30, GDX_ERR_ARG, "mesh too large for 32-bit node coordinates");
    GDX_REQUIRE(d->domain[2 * k + 1] > d->domain[2 * k], GDX_ERR_ARG, "empty domain interval");
  }
  GDX_REQUIRE(d->nmodes >= 0 && d->nmodes <= 1024, GDX_ERR_ARG, "nmodes must be in [0,1024]");
  GDX_REQUIRE(d->nmodes == 0 || d->modes, GDX_ERR_ARG, "modes is NULL with nmodes > 0");
  GDX_REQUIRE(d->rhs_kind == GDX_RHS_CONST || d->rhs_kind == GDX_RHS_MANUFACTURED, GDX_ERR_ARG, "unknown rhs kind");
  GDX_REQUIRE(d->dir_kind == 0 || d->dir_kind == 1 || d->dir_kind == GDX_DIR_CLAMPED_X0, GDX_ERR_ARG,
              "unknown Dirichlet kind");
  GDX_REQUIRE(d->op_kind == GDX_OP_DIFFUSION || d->op_kind == GDX_OP_FLUX || d->op_kind == GDX_OP_ELASTICITY,
              GDX_ERR_ARG, "unknown operator kind");
  if (d->op_kind == GDX_OP_ELASTICITY) {
    GDX_REQUIRE(d->dim == 3, GDX_ERR_UNSUPPORTED, "linear elasticity is built for 3D meshes");
    GDX_REQUIRE(d->nmodes == 0 && d->rhs_kind == GDX_RHS_CONST, GDX_ERR_UNSUPPORTED,
                "linear elasticity takes constant Lame parameters and a constant body force");
    GDX_REQUIRE(d->dir_kind == 0 || d->dir_kind == GDX_DIR_CLAMPED_X0, GDX_ERR_UNSUPPORTED,
                "linear elasticity supports no Dirichlet boundary or the clamped face x = x0");
    GDX_REQUIRE(d->dir_coef[0] == 0.0 && d->dir_coef[1] == 0.0 && d->dir_coef[2] == 0.0 && d->dir_coef[3] == 0.0,
                GDX_ERR_UNSUPPORTED, "the clamped boundary takes u = 0 (dir_coef must be zero)");
    GDX_REQUIRE(std::isfinite(d->lame_lambda) && std::isfinite(d->lame_mu) && d->lame_mu > 0.0 &&
                    d->lame_lambda > -2.0 / 3.0 * d->lame_mu,
                GDX_ERR_ARG, "Lame parameters need mu > 0 and 3 lambda + 2 mu > 0");
    return GDX_OK;
  }
  GDX_REQUIRE(d->dir_kind != GDX_DIR_CLAMPED_X0, GDX_ERR_UNSUPPORTED, "the clamped boundary is an elasticity option");
  if (d->op_kind == GDX_OP_FLUX) {
    GDX_REQUIRE(gdx::flux_supported(d->dim, d->order, d->qdeg / 2 + 1), GDX_ERR_UNSUPPORTED,
                "the flux operator is built for Q1 with at most 3 Gauss points per direction");
    GDX_REQUIRE(d->rhs_kind == GDX_RHS_CONST && d->nmodes == 0, GDX_ERR_UNSUPPORTED,
                "the flux operator takes a constant source and no kappa modes");
    GDX_REQUIRE(d->flux_p >= 1.0 && d->flux_alpha >= 0.0 && d->flux_beta >= 0.0, GDX_ERR_ARG,
                "flux parameters need p >= 1, alpha >= 0, beta >= 0");
    return GDX_OK;
  }
  GDX_REQUIRE(gdx::matrix_kernel_supported(d->dim, d->order, d->qdeg / 2 + 1), GDX_ERR_UNSUPPORTED,
              "no matrix kernel for this (dim, order, quadrature) combination");
  return GDX_OK;
}

// tiles whose position patterns fit in the matrix kernel's LDS go first
void order_tiles(gdx_problem *p) {
  const int cap = p->hbm_patterns       ? -1
                  : p->D != 3             ? 0
                  : p->kvariant3d == 1 ? gdx::pattern_capacity_tile3d(p->K, p->NQ)
                  : p->kvariant3d == 2 ? gdx::pattern_capacity_tile3d_direct(p->K, p->NQ)
                  : p->kvariant3d == 3 ? gdx::pattern_capacity_tile3d_wslab(p->K, p->NQ)
                                       : gdx::pattern_capacity_3d(p->K, p->NQ);
  // (a tile without free rows has no pattern list to read from HBM: it stays in
  // the LDS instance, which then stores nothing)
  auto fast = [cap](const gdx::Tile &t) { return t.npat <= cap || t.npat == 0; };
  std::stable_partition(p->tiles.begin(), p->tiles.end(), fast);
  p->n_fast_tiles = 0;
  for (const gdx::Tile &t : p->tiles) p->n_fast_tiles += fast(t) ? 1 : 0;
}

int part_check(gdx_problem *p, int32_t part) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(part >= 0 && part < p->nparts, GDX_ERR_ARG, "part index out of range");
  return GDX_OK;
}

// host copies of the numbering and PRanges (made on first use after a device
// symbolic phase)
int fe_on_host(gdx_problem *p) {
  if (p->ctx && p->host_fe_stale) {
    GDX_HIP(hipSetDevice(p->ctx->device));
    const int rc = gdx::host_fe(p);
    if (rc) set_error("copying the numbering to the host failed");
    return rc;
  }
  return GDX_OK;
}

}  // namespace

extern "C" {

int gdx_version(int32_t *major, int32_t *minor, int32_t *patch) {
  GDX_REQUIRE(major && minor && patch, GDX_ERR_ARG, "NULL output");
  *major = GDX_VERSION_MAJOR;
  *minor = GDX_VERSION_MINOR;
  *patch = GDX_VERSION_PATCH;
  return GDX_OK;
}

const char *gdx_last_error(void) { return gdx::g_last_error.c_str(); }

int gdx_init(int32_t device, gdx_ctx **out) {
  GDX_REQUIRE(out, GDX_ERR_ARG, "NULL output");
  *out = nullptr;
  int n = 0;
  GDX_HIP(hipGetDeviceCount(&n));
  GDX_REQUIRE(device >= 0 && device < n, GDX_ERR_ARG, "device index out of range");
  GDX_HIP(hipSetDevice(device));
  gdx_ctx *c = new (std::nothrow) gdx_ctx;
  GDX_REQUIRE(c, GDX_ERR_ALLOC, "host allocation failed");
  c->device = device;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    set_error(std::string("hipStreamCreate: ") + hipGetErrorString(e));
    return GDX_ERR_HIP;
  }
  // the device address of the no-alloc entry point's staging array
  if (gdx::noalloc_setup(device) != GDX_OK) {
    (void)hipStreamDestroy(c->stream);
    delete c;
    set_error("hipGetSymbolAddress of the no-alloc staging array failed");
    return GDX_ERR_HIP;
  }
  *out = c;
  return GDX_OK;
}

int gdx_finalize(gdx_ctx *ctx) {
  GDX_REQUIRE(ctx, GDX_ERR_ARG, "context is NULL");
  GDX_HIP(hipSetDevice(ctx->device));
  GDX_HIP(hipStreamSynchronize(ctx->stream));
  // objects still alive lose their device resources and their context
  for (gdx_batch *b : ctx->batches) gdx::release_batch_device(b);
  for (gdx_nn *nn : ctx->nns) gdx::release_nn_device(nn);
  for (gdx_problem *p : ctx->probs) {
    gdx::release_problem_device(p);
    p->ctx = nullptr;
  }
  GDX_HIP(hipStreamDestroy(ctx->stream));
  delete ctx;
  return GDX_OK;
}

extern "C++" namespace gdx {
void release_problem_device(gdx_problem *p) {
  if (!p->ctx) return;
  (void)hipSetDevice(p->ctx->device);
  (void)hipStreamSynchronize(p->ctx->stream);
  gdx::solve_free(p);
  gdx::flux_free(p);
  gdx::elast_free(p);
  gdx::mapped_free(p);
  gdx::pvector_release(p);
  free_device(p);
  for (hipEvent_t &e : p->ev)
    if (e) {
      (void)hipEventDestroy(e);
      e = nullptr;
    }
  p->coef_cell = false;
  p->lame_cell = false;
  p->mapped = false;
  p->diff_w = 1.0;
  p->reac_w = 0.0;
}
}  // namespace gdx

int gdx_problem_create(gdx_ctx *ctx, const gdx_problem_desc *desc, gdx_problem **out) {
  // ctx == NULL creates a host-only problem: the symbolic phase and the
  // structure exports work without a device, gdx_numeric is refused
  GDX_REQUIRE(out, GDX_ERR_ARG, "NULL output");
  *out = nullptr;
  int rc = check_desc(desc);
  if (rc) return rc;
  gdx_problem *p = new (std::nothrow) gdx_problem;
  GDX_REQUIRE(p, GDX_ERR_ALLOC, "host allocation failed");
  p->ctx = ctx;
  if (ctx) ctx->probs.push_back(p);
  p->desc = *desc;
  p->modes.assign(desc->modes, desc->modes + 5 * desc->nmodes);
  p->desc.modes = p->modes.data();
  for (int k = desc->dim; k < 3; ++k) {
    p->desc.cells[k] = 1;
    p->desc.parts[k] = 1;
    p->desc.domain[2 * k] = 0.0;
    p->desc.domain[2 * k + 1] = 1.0;
  }
  p->D = desc->dim;
  p->K = desc->order;
  p->NQ = desc->qdeg / 2 + 1;
  p->ncomp = desc->op_kind == GDX_OP_ELASTICITY ? 3 : 1;
  p->nparts = p->desc.parts[0] * p->desc.parts[1] * p->desc.parts[2];
  *out = p;
  return GDX_OK;
}

int gdx_problem_destroy(gdx_problem *p) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  if (p->ctx) {
    for (gdx_nn *nn : p->ctx->nns)  // networks attached to it lose the attachment
      if (gdx::twogrid_problem(nn) == p) gdx::twogrid_free(nn);
    gdx::batch_forget_problem(p->ctx, p);
    gdx::release_problem_device(p);
    gdx::unregister(p->ctx->probs, p);
  }
  delete p;
  return GDX_OK;
}

int gdx_symbolic(gdx_problem *p) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  const gdx::FaultScope fault_scope;
  if (p->ctx) {
    GDX_HIP(hipSetDevice(p->ctx->device));
    // device buffers are re-made (the device stage of the symbolic phase
    // writes some of them): graphs that captured the old ones are dropped
    gdx::batch_invalidate_problem(p->ctx, p);
    p->symbolic_done = false;
    GDX_HIP(hipStreamSynchronize(p->ctx->stream));
    // networks attached to the problem hold its solver's diagonal: they lose
    // the attachment (their next use returns GDX_ERR_STATE until re-attached)
    for (gdx_nn *nn : p->ctx->nns)
      if (gdx::twogrid_problem(nn) == p) gdx::twogrid_free(nn);
    free_device(p);
    // the cell-wise coefficient buffers go with the device buffers: the
    // problem returns to the descriptor's coefficients
    p->coef_cell = false;
    p->lame_cell = false;
    // and so does a mapped grid's state: the uniform box again
    gdx::mapped_free(p);
    p->mapped = false;
    p->diff_w = 1.0;
    p->reac_w = 0.0;
    gdx::solve_free(p);  // its row tables are rebuilt from the new structure
    gdx::elast_free(p);
    gdx::flux_free(p);
  }
  try {
    int rc = gdx::symbolic(p);
    if (rc) return rc;
  } catch (const std::bad_alloc &) {
    set_error("host allocation failed in the symbolic phase");
    return GDX_ERR_ALLOC;
  } catch (const std::exception &e) {
    set_error(e.what());
    return GDX_ERR_ARG;
  }
  gdx::fill_kparams(p);
  order_tiles(p);
  if (!p->ctx) {
    p->symbolic_done = true;
    return GDX_OK;
  }
  gdx::Device &d = p->dev;
  const bool elast = p->desc.op_kind == GDX_OP_ELASTICITY;
  // kappa and f at the quadrature points (constant-coefficient elasticity has none)
  const size_t nq = elast ? 1 : (size_t)p->kp.qn[0] * p->kp.qn[1] * p->kp.qn[2];
  int qnmax = std::max(p->kp.qn[0], std::max(p->kp.qn[1], p->kp.qn[2]));
  GDX_HIP(gdx::dev_malloc(reinterpret_cast<void **>(&d.W), nq * sizeof(double)));
  GDX_HIP(gdx::dev_malloc(reinterpret_cast<void **>(&d.F), nq * sizeof(double)));
  if (p->desc.nmodes > 0) {
    GDX_HIP(gdx::dev_malloc(reinterpret_cast<void **>(&d.E), (size_t)2 * 3 * p->desc.nmodes * qnmax * sizeof(double)));
    int rc = upload(&d.modes, p->modes);
    if (rc) return rc;
  }
  GDX_HIP(gdx::dev_malloc(reinterpret_cast<void **>(&d.values), std::max<int64_t>(p->nnz_total, 1) * sizeof(double)));
  GDX_HIP(gdx::dev_malloc(reinterpret_cast<void **>(&d.b), std::max<int64_t>(p->rows_total, 1) * sizeof(double)));
  // ghost rows are never written by the numeric phase: they stay zero
  GDX_HIP(hipMemset(d.values, 0, std::max<int64_t>(p->nnz_total, 1) * sizeof(double)));
  GDX_HIP(hipMemset(d.b, 0, std::max<int64_t>(p->rows_total, 1) * sizeof(double)));
  int rc;
  // the row table and tile pattern lists are in HBM already when the device
  // stage built them
  if (!p->dev_structure && (rc = upload(&d.rowinfo, p->rowinfo))) return rc;
  if ((rc = upload(&d.pat_fwd, p->pat_fwd))) return rc;
  if ((rc = upload(&d.pat_inv, p->pat_inv))) return rc;
  if ((rc = upload(&d.pat_invr, p->pat_invr))) return rc;
  if ((rc = upload(&d.pat_len, p->pat_len))) return rc;
  if ((rc = upload(&d.pat_dir, p->pat_dir))) return rc;
  if ((rc = upload(&d.tiles, p->tiles))) return rc;
  if (!p->dev_structure && (rc = upload(&d.tile_pats, p->tile_pats))) return rc;
  if (elast && (rc = gdx::elast_setup(p))) {
    set_error("uploading the elasticity tables failed");
    return rc;
  }
  // the flux operator's lattice state and id map, set up here so that no
  // allocation, copy or fill ever runs inside a batch capture
  if (p->desc.op_kind == GDX_OP_FLUX && (rc = gdx::flux_setup(p))) {
    set_error("setting up the flux operator's device state failed");
    return rc;
  }
  for (hipEvent_t &e : p->ev)
    if (!e) GDX_HIP(hipEventCreate(&e));
  p->symbolic_done = true;
  return GDX_OK;
}

extern "C++" namespace gdx {
// rebuild the tiles for extent R (the row table's tile-local pattern bits
// change with them) and upload tiles, tile patterns and row table
int retile(gdx_problem *p, const int R[3]) {
  // the tile buffers are re-allocated: graphs holding the old ones are dropped
  if (p->ctx) batch_invalidate_problem(p->ctx, p);
  int rc;
  if (p->dev_structure) {
    // the row table is in HBM: the tiles are built there as well
    if ((rc = build_tiles_device(p, R))) return rc;
    order_tiles(p);
    return upload(&p->dev.tiles, p->tiles);
  }
  build_tiles(p, R);
  order_tiles(p);
  if ((rc = upload(&p->dev.rowinfo, p->rowinfo))) return rc;
  if ((rc = upload(&p->dev.tiles, p->tiles))) return rc;
  if ((rc = upload(&p->dev.tile_pats, p->tile_pats))) return rc;
  return GDX_OK;
}

// 3D kernel variant: 0 = z-marching columns with rows_z z rows per workgroup,
// 1 = per-tile kernel, 2 = per-tile direct-store kernel, 3 = per-tile
// wave-slab kernel (Q2 only); 1-3 use their own tile shape (rows_z ignored)
int set_kernel_variant(gdx_problem *p, int variant, int rows_z) {
  if (p->D != 3) return GDX_OK;
  if (variant == 3 && !wslab_supported(p->K, p->NQ)) return GDX_ERR_UNSUPPORTED;
  int R[3];
  if (variant >= 1) {
    tile3d_shape(p->K, R);
  } else {
    tile_shape(3, p->K, R);
    R[2] = rows_z;
  }
  p->kvariant3d = variant;
  return retile(p, R);
}
}  // namespace gdx

extern "C++" namespace gdx {
// the numeric phase's kernel launches on st (no events, no memsets: safe to
// capture in a graph)
int numeric_launches(gdx_problem *p, hipStream_t st) {
  if (p->desc.op_kind == GDX_OP_FLUX) return flux_assemble(p, nullptr, false, p->precision == GDX_PREC_FP32, st);
  if (p->desc.op_kind == GDX_OP_ELASTICITY) return elast_assemble(p, st);
  if (p->mapped) {
    int rcm = mapped_assemble(p, st, 0);
    return rcm ? rcm : mapped_assemble(p, st, 1);
  }
  int rc = launch_coefficients(p, st);
  if (!rc) rc = launch_rhs(p, st);
  if (!rc) rc = launch_matrix(p, st);
  return rc;
}
}  // namespace gdx

int gdx_numeric(gdx_problem *p, int32_t precision) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_REQUIRE(precision == GDX_PREC_FP64 || precision == GDX_PREC_FP32, GDX_ERR_ARG, "unknown precision");
  GDX_REQUIRE(precision == GDX_PREC_FP64 || !p->mapped, GDX_ERR_UNSUPPORTED, "a mapped grid assembles in FP64");
  GDX_REQUIRE(precision == GDX_PREC_FP64 ||
                  (p->D == 3 && p->desc.op_kind == GDX_OP_DIFFUSION && (p->kvariant3d == 0 || p->kvariant3d == 2)) ||
                  (p->desc.op_kind == GDX_OP_FLUX && gdx::flux_mixed_supported(p)),
              GDX_ERR_UNSUPPORTED,
              "the FP32 form runs with the 3D z-marching or direct-store kernel (gdx_problem_set_kernel(prob, 0|2, rows)) "
              "or the fused 3D flux kernel (Q1, at most 2 Gauss points per direction)");
  if (precision != p->precision) gdx::batch_invalidate_problem(p->ctx, p);  // captured kernels change
  p->precision = precision;
  GDX_HIP(hipSetDevice(p->ctx->device));
  hipStream_t st = p->ctx->stream;
  GDX_HIP(hipEventRecord(p->ev[0], st));
  if (p->desc.op_kind == GDX_OP_FLUX) {
    // residual_and_jacobian at the stored state: no coefficient or separate RHS phase
    GDX_HIP(hipEventRecord(p->ev[1], st));
    GDX_HIP(hipEventRecord(p->ev[2], st));
    int rcf = gdx::flux_assemble(p, nullptr, false, precision == GDX_PREC_FP32, st);
    if (rcf) {
      set_error("flux operator launch failed");
      return rcf;
    }
    GDX_HIP(hipEventRecord(p->ev[3], st));
    return GDX_OK;
  }
  if (p->desc.op_kind == GDX_OP_ELASTICITY) {
    // matrix and RHS in one launch; no coefficient phase
    GDX_HIP(hipEventRecord(p->ev[1], st));
    GDX_HIP(hipEventRecord(p->ev[2], st));
    int rce = gdx::elast_assemble(p, st);
    if (rce) {
      set_error("elasticity kernel launch failed");
      return rce;
    }
    GDX_HIP(hipEventRecord(p->ev[3], st));
    return GDX_OK;
  }
  if (p->mapped) {
    // cell records (geometry, coefficients, cell matrices), then the rows
    int rcm = gdx::mapped_assemble(p, st, 0);
    GDX_HIP(hipEventRecord(p->ev[1], st));
    GDX_HIP(hipEventRecord(p->ev[2], st));
    if (!rcm) rcm = gdx::mapped_assemble(p, st, 1);
    if (rcm) {
      set_error("mapped-grid kernel launch failed");
      return rcm;
    }
    GDX_HIP(hipEventRecord(p->ev[3], st));
    return GDX_OK;
  }
  int rc = gdx::launch_coefficients(p, st);
  if (rc) {
    set_error("coefficient kernel launch failed");
    return rc;
  }
  GDX_HIP(hipEventRecord(p->ev[1], st));
  if ((rc = gdx::launch_rhs(p, st))) {
    set_error("rhs kernel launch failed");
    return rc;
  }
  GDX_HIP(hipEventRecord(p->ev[2], st));
  if ((rc = gdx::launch_matrix(p, st))) {
    set_error("matrix kernel launch failed");
    return rc;
  }
  GDX_HIP(hipEventRecord(p->ev[3], st));
  return GDX_OK;
}

int gdx_problem_set_kernel(gdx_problem *p, int32_t variant, int32_t rows_z) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  const bool hbm = (variant & GDX_KERNEL_HBM_PATTERNS) != 0;
  variant &= ~GDX_KERNEL_HBM_PATTERNS;
  GDX_REQUIRE(p->symbolic_done && p->ctx, GDX_ERR_STATE, "needs gdx_symbolic and a device context");
  GDX_REQUIRE(variant >= 0 && variant <= 3, GDX_ERR_ARG,
              "variant must be 0 (z-marching), 1 (per tile), 2 (per tile, direct stores) or 3 (per tile, wave slabs)");
  GDX_REQUIRE(variant != 3 || gdx::wslab_supported(p->K, p->NQ), GDX_ERR_UNSUPPORTED,
              "the wave-slab kernel is built for Q2");
  GDX_REQUIRE(p->desc.op_kind == GDX_OP_DIFFUSION, GDX_ERR_UNSUPPORTED, "kernel variants are for the diffusion operator");
  GDX_REQUIRE(variant != 0 || (rows_z >= p->K && rows_z % p->K == 0 && rows_z <= 4096), GDX_ERR_ARG,
              "rows_z must be a positive multiple of the order");
  GDX_HIP(hipSetDevice(p->ctx->device));
  GDX_HIP(hipStreamSynchronize(p->ctx->stream));
  p->hbm_patterns = hbm;
  return gdx::set_kernel_variant(p, variant, rows_z);
}

int gdx_problem_get_kernel(gdx_problem *p, int32_t *variant, int32_t *rows_z) {
  GDX_REQUIRE(p && variant && rows_z, GDX_ERR_ARG, "NULL argument");
  *variant = p->kvariant3d | (p->hbm_patterns ? GDX_KERNEL_HBM_PATTERNS : 0);
  *rows_z = p->tileR[2];
  return GDX_OK;
}

int gdx_set_state(gdx_problem *p, const double *x) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_REQUIRE(p->desc.op_kind == GDX_OP_FLUX, GDX_ERR_STATE, "the problem's operator does not depend on a state");
  GDX_HIP(hipSetDevice(p->ctx->device));
  int rc = gdx::flux_set_state(p, x);
  if (rc) set_error("setting the state failed");
  return rc;
}

namespace {
// the operators with a cell-wise coefficient: diffusion with a constant source
// (kappa), elasticity (kappa scales lambda and mu)
bool cell_coef_operator(const gdx_problem *p) {
  return (p->desc.op_kind == GDX_OP_DIFFUSION && p->desc.rhs_kind == GDX_RHS_CONST) ||
         p->desc.op_kind == GDX_OP_ELASTICITY;
}

// a change of coefficient mode changes the captured coefficient or elasticity
// launch: graphs holding p are dropped (new values in the same buffer replay as
// they are)
void set_coef_mode(gdx_problem *p, bool kappa, bool lame) {
  if (p->coef_cell != kappa || p->lame_cell != lame) gdx::batch_invalidate_problem(p->ctx, p);
  p->coef_cell = kappa;
  p->lame_cell = lame;
}
}  // namespace

int gdx_set_cell_coefficient(gdx_problem *p, const double *kappa) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_REQUIRE(cell_coef_operator(p), GDX_ERR_UNSUPPORTED,
              "cell-wise kappa needs the diffusion operator with a constant source, or elasticity");
  GDX_HIP(hipSetDevice(p->ctx->device));
  if (!kappa) {
    set_coef_mode(p, false, false);
    return GDX_OK;
  }
  const int64_t nc = gdx::num_cells(p);
  for (int64_t i = 0; i < nc; ++i)
    GDX_REQUIRE(kappa[i] > 0.0 && kappa[i] < INFINITY, GDX_ERR_ARG, "kappa must be positive and finite");
  int rc = gdx::cell_coef_alloc(p);
  if (rc) return rc;
  GDX_HIP(hipMemcpy(p->dev.cellk, kappa, nc * sizeof(double), hipMemcpyHostToDevice));
  set_coef_mode(p, true, false);
  return GDX_OK;
}

int gdx_set_vertex_coords(gdx_problem *p, const double *xyz) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_REQUIRE(p->desc.op_kind == GDX_OP_DIFFUSION, GDX_ERR_UNSUPPORTED, "mapped grids are lowered for diffusion only");
  GDX_HIP(hipSetDevice(p->ctx->device));
  if (!xyz) {
    if (p->mapped) gdx::batch_invalidate_problem(p->ctx, p);
    p->mapped = false;
    p->diff_w = 1.0;  // the uniform kernels assemble the plain diffusion form
    p->reac_w = 0.0;
    return GDX_OK;
  }
  int64_t nv = 1;
  for (int d = 0; d < p->D; ++d) nv *= p->desc.cells[d] + 1;
  for (int64_t i = 0; i < nv * p->D; ++i)
    GDX_REQUIRE(std::isfinite(xyz[i]), GDX_ERR_ARG, "vertex coordinates must be finite");
  GDX_HIP(hipStreamSynchronize(p->ctx->stream));  // the cell records may be in use
  int rc = gdx::mapped_set(p, xyz);
  if (rc) {
    set_error("setting up the mapped grid's device state failed");
    return rc;
  }
  if (!p->mapped) gdx::batch_invalidate_problem(p->ctx, p);
  p->mapped = true;
  return GDX_OK;
}

int gdx_set_reaction(gdx_problem *p, double diffusion_weight, double reaction) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(p->mapped, GDX_ERR_STATE, "the weighted form runs on the mapped-grid path: call gdx_set_vertex_coords");
  GDX_REQUIRE(std::isfinite(diffusion_weight) && std::isfinite(reaction) && diffusion_weight >= 0.0 &&
                  reaction >= 0.0 && (diffusion_weight > 0.0 || reaction > 0.0),
              GDX_ERR_ARG, "weights must be finite, non-negative and not both zero");
  // the captured cell kernel takes the weights by value
  if (diffusion_weight != p->diff_w || reaction != p->reac_w) gdx::batch_invalidate_problem(p->ctx, p);
  p->diff_w = diffusion_weight;
  p->reac_w = reaction;
  return GDX_OK;
}

int gdx_set_cell_lame(gdx_problem *p, const double *lame) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_REQUIRE(p->desc.op_kind == GDX_OP_ELASTICITY, GDX_ERR_UNSUPPORTED, "cell-wise Lame parameters need elasticity");
  GDX_HIP(hipSetDevice(p->ctx->device));
  if (!lame) {
    set_coef_mode(p, false, false);
    return GDX_OK;
  }
  const int64_t nc = gdx::num_cells(p);
  for (int64_t i = 0; i < nc; ++i) {
    const double lam = lame[2 * i], mu = lame[2 * i + 1];
    GDX_REQUIRE(std::isfinite(lam) && std::isfinite(mu) && mu > 0.0 && lam > -2.0 / 3.0 * mu, GDX_ERR_ARG,
                "each cell needs mu > 0 and lambda > -2 mu / 3, both finite");
  }
  if (!p->dev.cell_lame)
    GDX_REQUIRE(gdx::dev_malloc(reinterpret_cast<void **>(&p->dev.cell_lame), std::max<int64_t>(nc, 1) * 2 *
                                sizeof(double)) == hipSuccess, GDX_ERR_ALLOC, "out of device memory");
  GDX_HIP(hipMemcpy(p->dev.cell_lame, lame, nc * 2 * sizeof(double), hipMemcpyHostToDevice));
  set_coef_mode(p, false, true);
  return GDX_OK;
}

int gdx_cell_coefficient_grf(gdx_problem *p, uint64_t seed, double corr_len, double log_scale) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_REQUIRE(cell_coef_operator(p), GDX_ERR_UNSUPPORTED,
              "cell-wise kappa needs the diffusion operator with a constant source, or elasticity");
  GDX_REQUIRE(corr_len > 0.0 && log_scale >= 0.0 && log_scale < 50.0, GDX_ERR_ARG, "invalid field parameters");
  GDX_HIP(hipSetDevice(p->ctx->device));
  int rc = gdx::grf_cell_coefficient(p, seed, corr_len, log_scale);
  if (rc) {
    set_error("random field generation failed");
    return rc;
  }
  set_coef_mode(p, true, false);
  return GDX_OK;
}

int gdx_export_cell_coefficient(gdx_problem *p, double *kappa) {
  GDX_REQUIRE(p && kappa, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(p->coef_cell && p->dev.cellk, GDX_ERR_STATE, "no cell-wise coefficient is set");
  GDX_HIP(hipSetDevice(p->ctx->device));
  GDX_HIP(hipMemcpy(kappa, p->dev.cellk, gdx::num_cells(p) * sizeof(double), hipMemcpyDeviceToHost));
  return GDX_OK;
}

int gdx_newton(gdx_problem *p, const gdx_newton_opts *o, double *x, gdx_newton_report *rep) {
  GDX_REQUIRE(p && o, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_REQUIRE(p->desc.op_kind == GDX_OP_FLUX, GDX_ERR_STATE, "Newton needs a GDX_OP_FLUX problem");
  GDX_REQUIRE(o->maxiter >= 0 && o->lin_maxiter >= 1 && o->rtol >= 0.0 && o->atol >= 0.0 && o->lin_rtol > 0.0,
              GDX_ERR_ARG, "invalid Newton options");
  GDX_REQUIRE(o->init == GDX_NEWTON_INIT_GIVEN || o->init == GDX_NEWTON_INIT_POISSON, GDX_ERR_ARG,
              "unknown Newton initial guess");
  GDX_REQUIRE(o->jac_precision == GDX_PREC_FP64 || o->jac_precision == GDX_PREC_FP32, GDX_ERR_ARG,
              "unknown Jacobian precision");
  GDX_REQUIRE(o->jac_precision == GDX_PREC_FP64 || gdx::flux_mixed_supported(p), GDX_ERR_UNSUPPORTED,
              "FP32 Jacobians run in the fused 3D flux kernel (Q1, at most 2 Gauss points per direction)");
  if (gdx_nn *nn = o->precond) {
    GDX_REQUIRE(nn->ctx, GDX_ERR_STATE, "the preconditioner's context was finalized");
    GDX_REQUIRE(nn->ctx == p->ctx, GDX_ERR_ARG, "network and problem belong to different contexts");
    GDX_REQUIRE(p->D == 3 && p->K == 1 && p->desc.dir_kind == 1, GDX_ERR_UNSUPPORTED,
                "the two-grid preconditioner is built for 3D Q1 with Dirichlet data on the whole boundary");
    for (int d = 0; d < 3; ++d)
      GDX_REQUIRE(p->kp.ncells[d] == 2 * nn->net.n, GDX_ERR_ARG,
                  "the problem must have 2n x 2n x 2n cells for a network grid of n^3");
  }
  GDX_HIP(hipSetDevice(p->ctx->device));
  int rc = gdx::solve_newton(p, o, x, rep);
  if (rc) set_error("Newton iteration failed");
  return rc;
}

int gdx_sync(gdx_problem *p) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  if (!p->ctx) return GDX_OK;
  GDX_HIP(hipStreamSynchronize(p->ctx->stream));
  return GDX_OK;
}

int gdx_stream(gdx_problem *p, void **stream) {
  GDX_REQUIRE(p && stream, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  *stream = reinterpret_cast<void *>(p->ctx->stream);
  return GDX_OK;
}

int gdx_phase_times(gdx_problem *p, double *ms, int32_t n) {
  GDX_REQUIRE(p && ms && n >= 0, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_HIP(hipEventSynchronize(p->ev[3]));
  float t[4] = {0, 0, 0, 0};
  GDX_HIP(hipEventElapsedTime(&t[0], p->ev[0], p->ev[1]));
  GDX_HIP(hipEventElapsedTime(&t[1], p->ev[1], p->ev[2]));
  GDX_HIP(hipEventElapsedTime(&t[2], p->ev[2], p->ev[3]));
  GDX_HIP(hipEventElapsedTime(&t[3], p->ev[0], p->ev[3]));
  for (int i = 0; i < n && i < 4; ++i) ms[i] = t[i];
  return GDX_OK;
}

int gdx_num_parts(gdx_problem *p, int32_t *nparts) {
  GDX_REQUIRE(p && nparts, GDX_ERR_ARG, "NULL argument");
  *nparts = p->nparts;
  return GDX_OK;
}

int gdx_num_global_dofs(gdx_problem *p, int64_t *ngdofs) {
  GDX_REQUIRE(p && ngdofs, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  *ngdofs = p->ngdofs;
  return GDX_OK;
}

int gdx_owned_nnz(gdx_problem *p, int64_t *nnz) {
  GDX_REQUIRE(p && nnz, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  *nnz = p->owned_nnz;
  return GDX_OK;
}

int gdx_part_sizes(gdx_problem *p, int32_t part, int64_t *s) {
  int rc = part_check(p, part);
  if (rc) return rc;
  GDX_REQUIRE(s, GDX_ERR_ARG, "NULL argument");
  const gdx::Part &P = p->P[part];
  s[GDX_SZ_NLDOFS] = P.n_ldofs;
  s[GDX_SZ_NOWN] = P.n_own;
  s[GDX_SZ_NGHOST_ROWS] = P.n_ghost_rows;
  s[GDX_SZ_NGHOST_COLS] = P.n_ghost_cols;
  s[GDX_SZ_NNZ] = P.nnz;
  return GDX_OK;
}

int gdx_export_fe(gdx_problem *p, int32_t part, int64_t *gid, int32_t *owner, int64_t *gnode) {
  int rc = part_check(p, part);
  if (rc) return rc;
  if ((rc = fe_on_host(p))) return rc;
  GDX_REQUIRE(gid && owner && gnode, GDX_ERR_ARG, "NULL argument");
  const gdx::Part &P = p->P[part];
  for (int64_t i = 0; i < P.n_ldofs; ++i) {
    gid[i] = P.ldof_gid[i];
    owner[i] = P.ldof_owner[i];
    const int64_t l = P.ldof_lnode[i];
    const int64_t g0 = p->K * P.loc_lo[0] + l % P.lnn[0];
    const int64_t g1 = p->K * P.loc_lo[1] + (l / P.lnn[0]) % P.lnn[1];
    const int64_t g2 = p->K * P.loc_lo[2] + l / (P.lnn[0] * P.lnn[1]);
    gnode[i] = g0 + (int64_t)p->kp.nn[0] * (g1 + (int64_t)p->kp.nn[1] * g2);
  }
  return GDX_OK;
}

int gdx_export_rows(gdx_problem *p, int32_t part, int64_t *gid, int32_t *owner) {
  int rc = part_check(p, part);
  if (rc) return rc;
  if ((rc = fe_on_host(p))) return rc;
  GDX_REQUIRE(gid && owner, GDX_ERR_ARG, "NULL argument");
  const gdx::Part &P = p->P[part];
  std::memcpy(gid, P.row_gid.data(), P.nrows * sizeof(int64_t));
  std::memcpy(owner, P.row_owner.data(), P.nrows * sizeof(int32_t));
  return GDX_OK;
}

int gdx_export_cols(gdx_problem *p, int32_t part, int64_t *gid, int32_t *owner) {
  int rc = part_check(p, part);
  if (rc) return rc;
  if ((rc = fe_on_host(p))) return rc;
  GDX_REQUIRE(gid && owner, GDX_ERR_ARG, "NULL argument");
  const gdx::Part &P = p->P[part];
  std::memcpy(gid, P.col_gid.data(), P.ncols * sizeof(int64_t));
  std::memcpy(owner, P.col_owner.data(), P.ncols * sizeof(int32_t));
  return GDX_OK;
}

int gdx_export_csr(gdx_problem *p, int32_t part, int64_t *rowptr, int32_t *colidx, double *values) {
  int rc = part_check(p, part);
  if (rc) return rc;
  GDX_REQUIRE(rowptr && colidx && values, GDX_ERR_ARG, "NULL argument");
  if (p->ctx && (rc = gdx::host_csr(p))) {
    set_error("copying the local CSR to the host failed");
    return rc;
  }
  const gdx::Part &P = p->P[part];
  std::memcpy(rowptr, P.rowptr.data(), (P.nrows + 1) * sizeof(int64_t));
  std::memcpy(colidx, P.colidx.data(), P.nnz * sizeof(int32_t));
  if (!p->ctx) {
    std::memset(values, 0, P.nnz * sizeof(double));
    return GDX_OK;
  }
  GDX_HIP(hipSetDevice(p->ctx->device));
  GDX_HIP(hipStreamSynchronize(p->ctx->stream));
  if (P.nnz > 0)
    GDX_HIP(hipMemcpy(values, p->dev.values + P.val_base, P.nnz * sizeof(double), hipMemcpyDeviceToHost));
  return GDX_OK;
}

int gdx_export_rhs(gdx_problem *p, int32_t part, double *b) {
  int rc = part_check(p, part);
  if (rc) return rc;
  GDX_REQUIRE(b, GDX_ERR_ARG, "NULL argument");
  const gdx::Part &P = p->P[part];
  if (!p->ctx) {
    std::memset(b, 0, P.nrows * sizeof(double));
    return GDX_OK;
  }
  GDX_HIP(hipSetDevice(p->ctx->device));
  GDX_HIP(hipStreamSynchronize(p->ctx->stream));
  if (P.nrows > 0) GDX_HIP(hipMemcpy(b, p->dev.b + P.row_base, P.nrows * sizeof(double), hipMemcpyDeviceToHost));
  return GDX_OK;
}

int gdx_spmv(gdx_problem *p, const double *x, double *y) {
  GDX_REQUIRE(p && x && y, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_HIP(hipSetDevice(p->ctx->device));
  int rc = gdx::solve_spmv(p, x, y);
  if (rc) set_error("gdx_spmv failed");
  return rc;
}

int gdx_spmv_timed(gdx_problem *p, int32_t reps, double *ms, double *bytes) {
  GDX_REQUIRE(p && ms, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(reps >= 1 && reps <= 100000, GDX_ERR_ARG, "reps must be in 1..100000");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_HIP(hipSetDevice(p->ctx->device));
  int rc = gdx::solve_spmv_timed(p, reps, ms, bytes);
  if (rc) set_error("gdx_spmv_timed failed");
  return rc;
}

int gdx_pcg(gdx_problem *p, double rtol, int32_t maxiter, double *x, int32_t *iters, double *relres, double *ms) {
  GDX_REQUIRE(p, GDX_ERR_ARG, "problem is NULL");
  GDX_REQUIRE(rtol > 0.0 && maxiter >= 0, GDX_ERR_ARG, "rtol must be > 0 and maxiter >= 0");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_HIP(hipSetDevice(p->ctx->device));
  int it = 0;
  int rc = gdx::solve_pcg(p, rtol, maxiter, x, &it, relres, ms);
  if (rc) set_error("gdx_pcg failed (singular diagonal or device error)");
  if (iters) *iters = it;
  return rc;
}

int gdx_device_values(gdx_problem *p, double **values, int64_t *nnz_total) {
  GDX_REQUIRE(p && values && nnz_total, GDX_ERR_ARG, "NULL argument");
  GDX_REQUIRE(p->ctx, GDX_ERR_STATE, "host-only problem (created without a context)");
  GDX_REQUIRE(p->symbolic_done, GDX_ERR_STATE, "gdx_symbolic has not run");
  *values = p->dev.values;
  *nnz_total = p->nnz_total;
  return GDX_OK;
}

}  // extern "C"

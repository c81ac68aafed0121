// Nonlinear flux operator of libgdx (GDX_OP_FLUX), Q1 on the Cartesian mesh:
//
//   r(u,v)    = int( grad(v) . sigma(grad(u)) - v f ) dOmega
//   j(u,du,v) = int( grad(v) . dsigma(grad(du), grad(u)) ) dOmega
//   sigma(g)    = kappa(s) g,  s = alpha + beta |g|^2,  kappa(s) = s^((p-2)/2)
//   dsigma(d,g) = kappa(s) d + (p-2) beta s^((p-4)/2) (g.d) g
//
// (test/PLaplacianTests.jl:24-31 is p = 4, alpha = beta = 1; the p-Laplacian
// is p = 3 with alpha = eps^2 regularising grad(u) = 0).
//
// Owner computes rows: 2^D lanes per owned free row, one per adjacent cell,
// evaluate grad(u) at the cell's Gauss points from the state on the global
// node lattice (free values scattered there by global id, Dirichlet nodes hold
// uD) and accumulate the cell's contributions to the row; the row's 3^D
// Jacobian entries and its residual are then summed over the cells in a fixed
// order.  Each Gauss point is evaluated by the 2^D rows of its cell:
// recomputation instead of a scatter, so there are no atomics and two
// assemblies are bit-identical.  Rows are written through their position
// patterns (Dirichlet columns dropped, as in the affine path); ghost rows are
// never written and stay zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gdx_device.h"

namespace gdx {

namespace {

struct FluxParams {
  double p, alpha, beta, f;
  int mode;  // kappa exponent (p-2)/2: 0 general, 1 zero, 2 one half, 3 one
};

struct FluxData {
  int64_t nnodes = 0;
  double *ulat = nullptr;         // u on the global node lattice
  long long *gid_node = nullptr;  // free global id -> lattice node
  double *xs = nullptr;           // state: free values by global id
  bool state_set = false;
};

FluxData *flux_data(gdx_problem *p) { return reinterpret_cast<FluxData *>(p->flux); }

FluxParams flux_params(const gdx_problem_desc &d, bool linear) {
  FluxParams fp;
  fp.p = linear ? 2.0 : d.flux_p;
  fp.alpha = linear ? 1.0 : d.flux_alpha;
  fp.beta = linear ? 0.0 : d.flux_beta;
  fp.f = d.rhs_const;
  const double e = 0.5 * (fp.p - 2.0);
  fp.mode = e == 0.0 ? 1 : e == 0.5 ? 2 : e == 1.0 ? 3 : 0;
  return fp;
}

// kappa(s) = s^((p-2)/2) by the exponent's form (MODE as FluxParams::mode):
// the common exponents compile without pow, whose register needs would
// otherwise hold down the kernel's occupancy
template <int MODE>
__device__ __forceinline__ double flux_kappa(double s, const FluxParams &fp) {
  if constexpr (MODE == 1) return 1.0;
  else if constexpr (MODE == 2) return sqrt(s);
  else if constexpr (MODE == 3) return s;
  else return pow(s, 0.5 * (fp.p - 2.0));
}

// Dirichlet data on boundary nodes, zero elsewhere (free nodes are overwritten
// by the state scatter)
__global__ void k_lattice_init(int64_t n, double *__restrict__ ulat, KParams kp) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int g0 = (int)(i % kp.nn[0]), g1 = (int)((i / kp.nn[0]) % kp.nn[1]), g2 = (int)(i / ((int64_t)kp.nn[0] * kp.nn[1]));
  bool bnd = false;
  bnd |= g0 == 0 || g0 == kp.nn[0] - 1;
  bnd |= g1 == 0 || g1 == kp.nn[1] - 1;
  if (kp.dim == 3) bnd |= g2 == 0 || g2 == kp.nn[2] - 1;
  ulat[i] = (kp.dir_kind == 1 && bnd) ? uD_at(kp, g0, g1, g2) : 0.0;
}

__global__ void k_state_scatter(int64_t n, const long long *__restrict__ gid_node, const double *__restrict__ x,
                                double *__restrict__ ulat) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g < n) ulat[gid_node[g]] = x ? x[g] : 0.0;
}

// 2^D lanes per row, one per adjacent cell: a lane evaluates grad(u) at its
// cell's Gauss points and accumulates the cell's contributions to the row's
// 2^D columns and to the residual; the row's lanes then sum the 3^D slots from
// LDS in a fixed cell order.  Each Gauss point is still evaluated once per row
// of its cell (no scatter, no atomics), but a lane holds only 2^D accumulators.
template <int D, int NQ, int MODE>
__global__ __launch_bounds__(256) void k_flux_q1(const Tile *__restrict__ tiles, int ntiles,
                                                 const RowInfo *__restrict__ rowinfo,
                                                 const uint8_t *__restrict__ pat_fwd, int pstride,
                                                 const double *__restrict__ ulat, double *__restrict__ values,
                                                 double *__restrict__ b, KParams kp, FluxParams fp) {
  constexpr int S = D == 3 ? 27 : 9, NC = 1 << D, NQD = D == 3 ? NQ * NQ * NQ : NQ * NQ;
  constexpr int RPB = 256 / NC;  // rows per pass of the workgroup
  __shared__ double part[256][NC + 1];
  // reference-cell tables at every Gauss point, identical for every cell of the
  // uniform mesh: gradients of the 2^D local basis functions (1/h applied),
  // values, weight x cell volume, and grad(phi_a).grad(phi_b) (rows padded to
  // NC + 1 doubles against bank conflicts).  Filled once from the 1D tables of
  // the kernel argument, so the Gauss-point loop only reads LDS.
  __shared__ double tg[NQD][NC][D], tp[NQD][NC], tw[NQD], tgab[NQD][NC][NC + 1];
  {
    const double ih[3] = {1.0 / kp.h[0], 1.0 / kp.h[1], D == 3 ? 1.0 / kp.h[2] : 0.0};
    const double vol = kp.h[0] * kp.h[1] * (D == 3 ? kp.h[2] : 1.0);
    for (int i = threadIdx.x; i < NQD * NC; i += blockDim.x) {
      const int q = i / NC, bb = i % NC;
      const int q0 = q % NQ, q1 = (q / NQ) % NQ, q2 = D == 3 ? q / (NQ * NQ) : 0;
      const int b0 = bb & 1, b1 = (bb >> 1) & 1, b2 = D == 3 ? (bb >> 2) & 1 : 0;
      const double p0 = kp.phi[b0][q0], p1 = kp.phi[b1][q1], p2 = D == 3 ? kp.phi[b2][q2] : 1.0;
      const double d0 = kp.dphi[b0][q0] * ih[0], d1 = kp.dphi[b1][q1] * ih[1];
      const double d2 = D == 3 ? kp.dphi[b2][q2] * ih[2] : 0.0;
      tg[q][bb][0] = d0 * p1 * p2;
      tg[q][bb][1] = p0 * d1 * p2;
      if constexpr (D == 3) tg[q][bb][2] = p0 * p1 * d2;
      tp[q][bb] = p0 * p1 * p2;
      if (bb == 0) tw[q] = kp.qw[q0] * kp.qw[q1] * (D == 3 ? kp.qw[q2] : 1.0) * vol;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NQD * NC * NC; i += blockDim.x) {
      const int q = i / (NC * NC), aa = (i / NC) % NC, bb = i % NC;
      double v = tg[q][aa][0] * tg[q][bb][0] + tg[q][aa][1] * tg[q][bb][1];
      if constexpr (D == 3) v += tg[q][aa][2] * tg[q][bb][2];
      tgab[q][aa][bb] = v;
    }
    __syncthreads();
  }
  const int tb = tile_of_block(ntiles);
  if (tb >= ntiles) return;
  const Tile t = tiles[tb];
  const int nrows = t.ext[0] * t.ext[1] * t.ext[2];
  const int64_t sy = kp.nn[0], sz = (int64_t)kp.nn[0] * kp.nn[1];
  const int tid = threadIdx.x, e = tid % NC, rl = tid / NC;
  const int e0 = e & 1, e1 = (e >> 1) & 1, e2 = D == 3 ? (e >> 2) & 1 : 0;
  const int ai = (NC - 1) - e;  // the row node's index in the cell: (1 - e0, 1 - e1, 1 - e2)
  for (int r0 = 0; r0 < nrows; r0 += RPB) {  // uniform over the workgroup
    const int r = r0 + rl;
    int g0 = 0, g1 = 0, g2 = 0;
    RowInfo ri{0, 0, -1, 0};
    if (r < nrows) {
      const int u1 = r % t.ext[0], u2 = (r / t.ext[0]) % t.ext[1], u3 = r / (t.ext[0] * t.ext[1]);
      g0 = t.org[0] + u1, g1 = t.org[1] + u2, g2 = t.org[2] + u3;
      ri = rowinfo[t.rowinfo_base + row_index(t, g0, g1, g2)];
    }
    const bool live = ri.pid >= 0;
    double acc8[NC];
#pragma unroll
    for (int bb = 0; bb < NC; ++bb) acc8[bb] = 0.0;
    double res = 0.0;
    const int c0 = g0 - 1 + e0, c1 = g1 - 1 + e1, c2 = D == 3 ? g2 - 1 + e2 : 0;
    const bool incell = live && c0 >= 0 && c0 < kp.ncells[0] && c1 >= 0 && c1 < kp.ncells[1] &&
                        (D == 2 || (c2 >= 0 && c2 < kp.ncells[2]));
    if (incell) {
      const double *ub = ulat + c0 + sy * c1 + sz * c2;
      double uc[NC];
#pragma unroll
      for (int bb = 0; bb < NC; ++bb)
        uc[bb] = ub[(bb & 1) + sy * ((bb >> 1) & 1) + (D == 3 ? sz * ((bb >> 2) & 1) : 0)];
#pragma unroll 1
      for (int q = 0; q < NQD; ++q) {
        double gu0 = 0.0, gu1 = 0.0, gu2 = 0.0;
#pragma unroll
        for (int bb = 0; bb < NC; ++bb) {
          gu0 = fma(uc[bb], tg[q][bb][0], gu0);
          gu1 = fma(uc[bb], tg[q][bb][1], gu1);
          if constexpr (D == 3) gu2 = fma(uc[bb], tg[q][bb][2], gu2);
        }
        const double s2 = fp.alpha + fp.beta * (gu0 * gu0 + gu1 * gu1 + gu2 * gu2);
        const double kap = flux_kappa<MODE>(s2, fp);
        const double gam = (MODE == 1 || s2 <= 0.0) ? 0.0 : (fp.p - 2.0) * fp.beta * kap / s2;
        const double wd = tw[q];
        double gua = gu0 * tg[q][ai][0] + gu1 * tg[q][ai][1];
        if constexpr (D == 3) gua += gu2 * tg[q][ai][2];
        res = fma(wd, kap * gua - fp.f * tp[q][ai], res);
        const double wk = wd * kap, wg = wd * gam * gua;
#pragma unroll
        for (int bb = 0; bb < NC; ++bb) {
          double gub = gu0 * tg[q][bb][0] + gu1 * tg[q][bb][1];
          if constexpr (D == 3) gub += gu2 * tg[q][bb][2];
          acc8[bb] = fma(wk, tgab[q][ai][bb], fma(wg, gub, acc8[bb]));
        }
      }
    }
#pragma unroll
    for (int bb = 0; bb < NC; ++bb) part[tid][bb] = acc8[bb];
    part[tid][NC] = res;
    __syncthreads();
    if (live) {
      // slot (s0, s1, s2) of the row = sum over the cells c holding both nodes of
      // c's contribution to column (s - c), cells in a fixed order
      const uint8_t *fwd = pat_fwd + (size_t)ri.pid * pstride;
      double *dst = values + ri.rowstart;
      const int base = rl * NC;
      for (int sl = e; sl < S; sl += NC) {
        const int s0 = sl % 3, s1 = (sl / 3) % 3, s2 = D == 3 ? sl / 9 : 0;
        double v = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int b0 = s0 - (c & 1), b1 = s1 - ((c >> 1) & 1), b2 = D == 3 ? s2 - ((c >> 2) & 1) : 0;
          if (b0 < 0 || b0 > 1 || b1 < 0 || b1 > 1 || b2 < 0 || b2 > 1) continue;
          v += part[base + c][b0 | (b1 << 1) | (b2 << 2)];
        }
        const int pos = fwd[sl];
        if (pos < kSlotDirichlet) dst[pos] = v;
      }
      if (e == 0) {
        double rr = 0.0;
#pragma unroll
        for (int c = 0; c < NC; ++c) rr += part[base + c][NC];
        b[ri.bidx] = rr;
      }
    }
    __syncthreads();  // part is rewritten by the next pass
  }
}

template <int D, int NQ, int MODE>
void launch_flux_m(gdx_problem *p, const FluxData *fd, const FluxParams &fp, hipStream_t st) {
  const int nt = (int)p->tiles.size();
  hipLaunchKernelGGL((k_flux_q1<D, NQ, MODE>), dim3(8u * ((nt + 7) / 8)), dim3(256), 0, st, p->dev.tiles, nt,
                     p->dev.rowinfo, p->dev.pat_fwd, pattern_stride(p->kp.S), fd->ulat, p->dev.values, p->dev.b,
                     p->kp, fp);
}

template <int D, int NQ>
void launch_flux_t(gdx_problem *p, const FluxData *fd, const FluxParams &fp, hipStream_t st) {
  switch (fp.mode) {
    case 1: launch_flux_m<D, NQ, 1>(p, fd, fp, st); break;
    case 2: launch_flux_m<D, NQ, 2>(p, fd, fp, st); break;
    case 3: launch_flux_m<D, NQ, 3>(p, fd, fp, st); break;
    default: launch_flux_m<D, NQ, 0>(p, fd, fp, st); break;
  }
}

// zero fill (a kernel rather than a memset: flux_setup may run on a stream
// that a caller later captures)
__global__ void k_fill_zero(int64_t n, double *__restrict__ x) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) x[i] = 0.0;
}

}  // namespace

bool flux_supported(int dim, int order, int nq) { return order == 1 && (dim == 2 || dim == 3) && nq >= 1 && nq <= 3; }

void flux_free(gdx_problem *p) {
  FluxData *fd = flux_data(p);
  if (!fd) return;
  void *ptrs[] = {fd->ulat, fd->gid_node, fd->xs};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete fd;
  p->flux = nullptr;
}

// device buffers of the flux operator: lattice state with Dirichlet data, the
// free-id -> lattice map, and the state vector (zero).  Run by gdx_symbolic,
// never inside a capture; any failure leaves p->flux unset.
int flux_setup(gdx_problem *p) {
  flux_free(p);
  const int64_t nnodes = (int64_t)p->kp.nn[0] * p->kp.nn[1] * p->kp.nn[2];
  std::vector<long long> gid_node(std::max<int64_t>(p->ngdofs, 1), 0);
  for (const Part &P : p->P)
    for (int64_t i = 0; i < P.n_ldofs; ++i) {
      const int64_t l = P.ldof_lnode[i];
      const int64_t g0 = p->K * P.loc_lo[0] + l % P.lnn[0];
      const int64_t g1 = p->K * P.loc_lo[1] + (l / P.lnn[0]) % P.lnn[1];
      const int64_t g2 = p->K * P.loc_lo[2] + l / (P.lnn[0] * P.lnn[1]);
      gid_node[P.ldof_gid[i]] = g0 + (int64_t)p->kp.nn[0] * (g1 + (int64_t)p->kp.nn[1] * g2);
    }
  FluxData *fd = new (std::nothrow) FluxData;
  if (!fd) return GDX_ERR_ALLOC;
  p->flux = fd;
  fd->nnodes = nnodes;
  hipStream_t st = p->ctx->stream;
  const int64_t nx = std::max<int64_t>(p->ngdofs, 1);
  if (hipMalloc(reinterpret_cast<void **>(&fd->ulat), nnodes * sizeof(double)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&fd->gid_node), gid_node.size() * sizeof(long long)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&fd->xs), nx * sizeof(double)) != hipSuccess) {
    flux_free(p);
    return GDX_ERR_ALLOC;
  }
  if (hipMemcpy(fd->gid_node, gid_node.data(), gid_node.size() * sizeof(long long), hipMemcpyHostToDevice) !=
      hipSuccess) {
    flux_free(p);
    return GDX_ERR_HIP;
  }
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((nx + 255) / 256)), dim3(256), 0, st, nx, fd->xs);
  hipLaunchKernelGGL(k_lattice_init, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, st, nnodes, fd->ulat,
                     p->kp);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
    flux_free(p);
    return GDX_ERR_HIP;
  }
  return GDX_OK;
}

double *flux_state(gdx_problem *p) { return p->flux ? flux_data(p)->xs : nullptr; }

int flux_set_state(gdx_problem *p, const double *xh) {
  if (!p->flux) return GDX_ERR_STATE;
  FluxData *fd = flux_data(p);
  hipStream_t st = p->ctx->stream;
  const size_t bytes = std::max<int64_t>(p->ngdofs, 1) * sizeof(double);
  hipError_t e = xh ? hipMemcpyAsync(fd->xs, xh, p->ngdofs * sizeof(double), hipMemcpyHostToDevice, st)
                    : hipMemsetAsync(fd->xs, 0, bytes, st);
  if (e != hipSuccess) return GDX_ERR_HIP;
  fd->state_set = true;
  return hipStreamSynchronize(st) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// Jacobian into the CSR values and residual into b at the state x (device,
// by global id; nullptr = the stored state).  linear: p = 2, alpha = 1, beta = 0
// (the Poisson operator with the same source), used for Newton's initial guess.
int flux_assemble(gdx_problem *p, const double *x, bool linear, hipStream_t st) {
  if (!p->flux) return GDX_ERR_STATE;  // flux_setup runs in gdx_symbolic
  FluxData *fd = flux_data(p);
  const int64_t n = p->ngdofs;
  if (n > 0)
    hipLaunchKernelGGL(k_state_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, fd->gid_node,
                       x ? x : fd->xs, fd->ulat);
  const FluxParams fp = flux_params(p->desc, linear);
  if (p->tiles.empty()) return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
  if (p->D == 3 && p->NQ == 2) launch_flux_t<3, 2>(p, fd, fp, st);
  else if (p->D == 3 && p->NQ == 1) launch_flux_t<3, 1>(p, fd, fp, st);
  else if (p->D == 3 && p->NQ == 3) launch_flux_t<3, 3>(p, fd, fp, st);
  else if (p->D == 2 && p->NQ == 2) launch_flux_t<2, 2>(p, fd, fp, st);
  else if (p->D == 2 && p->NQ == 1) launch_flux_t<2, 1>(p, fd, fp, st);
  else if (p->D == 2 && p->NQ == 3) launch_flux_t<2, 3>(p, fd, fp, st);
  else return GDX_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

}  // namespace gdx

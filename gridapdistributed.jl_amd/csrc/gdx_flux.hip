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
// Owner computes rows: one thread per owned free row visits the row's 2^D
// cells, evaluates grad(u) at their Gauss points from the state on the global
// node lattice (free values scattered there by global id, Dirichlet nodes hold
// uD), and accumulates the row's 3^D Jacobian entries and its residual in
// registers, in a fixed order.  Each Gauss point is evaluated by the 2^D rows
// of its cell: recomputation instead of a scatter, so there are no atomics and
// two assemblies are bit-identical.  Rows are written through their position
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

__device__ __forceinline__ double flux_kappa(double s, const FluxParams &fp) {
  switch (fp.mode) {
    case 1: return 1.0;
    case 2: return sqrt(s);
    case 3: return s;
    default: return pow(s, 0.5 * (fp.p - 2.0));
  }
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

template <int D, int NQ>
__global__ __launch_bounds__(256) void k_flux_q1(const Tile *__restrict__ tiles, int ntiles,
                                                 const RowInfo *__restrict__ rowinfo,
                                                 const uint8_t *__restrict__ pat_fwd, int pstride,
                                                 const double *__restrict__ ulat, double *__restrict__ values,
                                                 double *__restrict__ b, KParams kp, FluxParams fp) {
  constexpr int S = D == 3 ? 27 : 9, NC = 1 << D, NQD = D == 3 ? NQ * NQ * NQ : NQ * NQ;
  const int tb = tile_of_block(ntiles);
  if (tb >= ntiles) return;
  const Tile t = tiles[tb];
  const int nrows = t.ext[0] * t.ext[1] * t.ext[2];
  const double ih0 = 1.0 / kp.h[0], ih1 = 1.0 / kp.h[1], ih2 = D == 3 ? 1.0 / kp.h[2] : 0.0;
  const double vol = kp.h[0] * kp.h[1] * (D == 3 ? kp.h[2] : 1.0);
  const int64_t sy = kp.nn[0], sz = (int64_t)kp.nn[0] * kp.nn[1];
  for (int r = threadIdx.x; r < nrows; r += blockDim.x) {
    const int u1 = r % t.ext[0], u2 = (r / t.ext[0]) % t.ext[1], u3 = r / (t.ext[0] * t.ext[1]);
    const int g0 = t.org[0] + u1, g1 = t.org[1] + u2, g2 = t.org[2] + u3;
    const RowInfo ri = rowinfo[t.rowinfo_base + row_index(t, g0, g1, g2)];
    if (ri.pid < 0) continue;
    double acc[S];
#pragma unroll
    for (int s = 0; s < S; ++s) acc[s] = 0.0;
    double res = 0.0;
#pragma unroll
    for (int e = 0; e < NC; ++e) {
      const int e0 = e & 1, e1 = (e >> 1) & 1, e2 = D == 3 ? (e >> 2) & 1 : 0;
      const int c0 = g0 - 1 + e0, c1 = g1 - 1 + e1, c2 = D == 3 ? g2 - 1 + e2 : 0;
      if (c0 < 0 || c0 >= kp.ncells[0] || c1 < 0 || c1 >= kp.ncells[1]) continue;
      if (D == 3 && (c2 < 0 || c2 >= kp.ncells[2])) continue;
      // the row node is local node (1-e0, 1-e1, 1-e2) of the cell
      const double *ub = ulat + c0 + sy * c1 + sz * c2;
      double uc[NC];
#pragma unroll
      for (int bb = 0; bb < NC; ++bb)
        uc[bb] = ub[(bb & 1) + sy * ((bb >> 1) & 1) + (D == 3 ? sz * ((bb >> 2) & 1) : 0)];
#pragma unroll
      for (int q = 0; q < NQD; ++q) {
        const int q0 = q % NQ, q1 = (q / NQ) % NQ, q2 = D == 3 ? q / (NQ * NQ) : 0;
        // 1D values and derivatives of the two local nodes per direction
        double p0[2], p1[2], p2[2], d0[2], d1[2], d2[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
          p0[a] = kp.phi[a][q0];
          p1[a] = kp.phi[a][q1];
          p2[a] = D == 3 ? kp.phi[a][q2] : 1.0;
          d0[a] = kp.dphi[a][q0] * ih0;
          d1[a] = kp.dphi[a][q1] * ih1;
          d2[a] = D == 3 ? kp.dphi[a][q2] * ih2 : 0.0;
        }
        double gb[NC][3];
#pragma unroll
        for (int bb = 0; bb < NC; ++bb) {
          const int b0 = bb & 1, b1 = (bb >> 1) & 1, b2 = D == 3 ? (bb >> 2) & 1 : 0;
          gb[bb][0] = d0[b0] * p1[b1] * p2[b2];
          gb[bb][1] = p0[b0] * d1[b1] * p2[b2];
          gb[bb][2] = p0[b0] * p1[b1] * d2[b2];
        }
        double gu0 = 0.0, gu1 = 0.0, gu2 = 0.0;
#pragma unroll
        for (int bb = 0; bb < NC; ++bb) {
          gu0 = fma(uc[bb], gb[bb][0], gu0);
          gu1 = fma(uc[bb], gb[bb][1], gu1);
          gu2 = fma(uc[bb], gb[bb][2], gu2);
        }
        const double s = fp.alpha + fp.beta * (gu0 * gu0 + gu1 * gu1 + gu2 * gu2);
        const double kap = flux_kappa(s, fp);
        const double gam = (fp.mode == 1 || s <= 0.0) ? 0.0 : (fp.p - 2.0) * fp.beta * kap / s;
        const double wd = kp.qw[q0] * kp.qw[q1] * (D == 3 ? kp.qw[q2] : 1.0) * vol;
        const int a =(1 - e0) | ((1 - e1) << 1) | (D == 3 ? (1 - e2) << 2 : 0);
        const double ga0 = gb[a][0], ga1 = gb[a][1], ga2 = gb[a][2];
        const double gua = gu0 * ga0 + gu1 * ga1 + gu2 * ga2;
        const double phia = p0[1 - e0] * p1[1 - e1] * p2[D == 3 ? 1 - e2 : 0];
        res = fma(wd, kap * gua - fp.f * phia, res);
        const double wk = wd * kap, wg = wd * gam * gua;
#pragma unroll
        for (int bb = 0; bb < NC; ++bb) {
          const int b0 = bb & 1, b1 = (bb >> 1) & 1, b2 = D == 3 ? (bb >> 2) & 1 : 0;
          const int slot = (e0 + b0) + 3 * (e1 + b1) + (D == 3 ? 9 * (e2 + b2) : 0);
          const double gab = ga0 * gb[bb][0] + ga1 * gb[bb][1] + ga2 * gb[bb][2];
          const double gub = gu0 * gb[bb][0] + gu1 * gb[bb][1] + gu2 * gb[bb][2];
          acc[slot] = fma(wk, gab, fma(wg, gub, acc[slot]));
        }
      }
    }
    const uint8_t *fwd = pat_fwd + (size_t)ri.pid * pstride;
    double *dst = values + ri.rowstart;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int pos = fwd[s];
      if (pos < kSlotDirichlet) dst[pos] = acc[s];
    }
    b[ri.bidx] = res;
  }
}

template <int D, int NQ>
void launch_flux_t(gdx_problem *p, const FluxData *fd, const FluxParams &fp, hipStream_t st) {
  const int nt = (int)p->tiles.size();
  hipLaunchKernelGGL((k_flux_q1<D, NQ>), dim3(8u * ((nt + 7) / 8)), dim3(256), 0, st, p->dev.tiles, nt,
                     p->dev.rowinfo, p->dev.pat_fwd, pattern_stride(p->kp.S), fd->ulat, p->dev.values, p->dev.b,
                     p->kp, fp);
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
// free-id -> lattice map, and the state vector (zero)
int flux_setup(gdx_problem *p) {
  if (p->flux) return GDX_OK;
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
  FluxData *fd = new FluxData;
  p->flux = fd;
  fd->nnodes = nnodes;
  hipStream_t st = p->ctx->stream;
  if (hipMalloc(reinterpret_cast<void **>(&fd->ulat), nnodes * sizeof(double)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&fd->gid_node), gid_node.size() * sizeof(long long)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&fd->xs), std::max<int64_t>(p->ngdofs, 1) * sizeof(double)) != hipSuccess) {
    flux_free(p);
    return GDX_ERR_ALLOC;
  }
  if (hipMemcpy(fd->gid_node, gid_node.data(), gid_node.size() * sizeof(long long), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemsetAsync(fd->xs, 0, std::max<int64_t>(p->ngdofs, 1) * sizeof(double), st) != hipSuccess)
    return GDX_ERR_HIP;
  hipLaunchKernelGGL(k_lattice_init, dim3((unsigned)((nnodes + 255) / 256)), dim3(256), 0, st, nnodes, fd->ulat,
                     p->kp);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

double *flux_state(gdx_problem *p) { return p->flux ? flux_data(p)->xs : nullptr; }

int flux_set_state(gdx_problem *p, const double *xh) {
  int rc = flux_setup(p);
  if (rc) return rc;
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
  int rc = flux_setup(p);
  if (rc) return rc;
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

// Linear elasticity on the partitioned Cartesian mesh: the numeric phase of
// AffineFEOperator(a, l, U, V) with
//   a(u,v) = int( eps(v) : (lambda tr(eps(u)) I + 2 mu eps(u)) ) dOmega
//   l(v)   = int( v . f ) dOmega
// on ReferenceFE(lagrangian, VectorValue{3,Float64}, K), constant lambda, mu, f
// (the reference computes 81 x 81 Q2 cell matrices per cell and scatters them,
// src/FESpaces.jl:794-798 and src/Algebra.jl:808-836).
//
// With constant coefficients on an axis-aligned uniform mesh every entry is a
// sum of products of 1D factors ASSEMBLED over the 1D mesh (the sum over the
// cells shared by two nodes factorises per dimension):
//   K_d(i,j) = int phi_i' phi_j'   M_d(i,j) = int phi_i phi_j
//   P_d(i,j) = int phi_i' phi_j    Q_d(i,j) = int phi_i phi_j'
// computed with the problem's Gauss rule, so they reproduce the tensor-product
// quadrature of the cell integrals exactly.  For row (i, a), column (j, b):
//   a == b: mu (K0 M1 M2 + M0 K1 M2 + M0 M1 K2) + (lambda + mu) K_a M_c M_e
//   a != b: lambda P_a Q_b M_c + mu Q_a P_b M_c        (c the third direction)
// and b(i, a) = f_a m0(i0) m1(i1) m2(i2) with m_d(i) = int phi_i.  Clamped
// (zero) Dirichlet data needs no lifting.
//
// Kernel: one wave per owned row node (its 3 rows share one column set and one
// position pattern).  The node's 1D factors (3 directions x 2K+1 offsets x
// {K,M,P,Q}) go to LDS; each lane then walks row positions e = lane, lane+64,
// ..., decodes e -> (stencil slot, column component) through the pattern and
// writes the 3 rows' values at e: every store instruction writes 64
// consecutive doubles.  HBM traffic is the 8 B value writes (plus 24 B of row
// table per node); the arithmetic is ~10 FLOP per stored value.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gdx_internal.h"

namespace gdx {

namespace {

constexpr int kWaves = 4;  // waves (row nodes) per workgroup

struct ElastPart {
  long long rowinfo_base;
  int onb_lo[3], onb_n[3];
};

struct ElastData {
  double *fac = nullptr;    // [3][nnmax][W][4]: K, M, P, Q at (node, offset + K)
  double *mass = nullptr;   // [3][nnmax]
  uint16_t *pinv = nullptr; // [npat][stride]: position -> slot * 3 + component
  ElastPart *parts = nullptr;
  int nparts = 0, nnmax = 0, stride = 0;
  long long nrownodes = 0;
};

ElastData *ed(gdx_problem *p) { return reinterpret_cast<ElastData *>(p->elast); }

struct ElastArgs {
  double lam, mu, f[3];
  int nnmax, stride, nparts;
  long long nrownodes;
};

template <int K>
__global__ __launch_bounds__(64 * kWaves) void k_elast(ElastArgs A, const ElastPart *__restrict__ parts,
                                                       const RowInfo *__restrict__ rowinfo,
                                                       const double *__restrict__ fac, const double *__restrict__ mass,
                                                       const uint16_t *__restrict__ pinv, double *__restrict__ values,
                                                       double *__restrict__ b) {
  constexpr int W = 2 * K + 1;
  constexpr int NF = 3 * W * 4;  // factors of one row node
  __shared__ __align__(16) double lds[kWaves][NF];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long long node = (long long)blockIdx.x * kWaves + wave;
  if (node >= A.nrownodes) return;  // whole waves exit together; no barrier below
  const RowInfo ri = rowinfo[node];
  if (ri.pid < 0) return;
  // the node's part (parts sorted by rowinfo_base) and its global coordinates
  int lo = 0, hi = A.nparts - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (parts[mid].rowinfo_base <= node) lo = mid;
    else hi = mid - 1;
  }
  const ElastPart &P = parts[lo];
  const long long li = node - P.rowinfo_base;
  const int g[3] = {P.onb_lo[0] + (int)(li % P.onb_n[0]), P.onb_lo[1] + (int)((li / P.onb_n[0]) % P.onb_n[1]),
                    P.onb_lo[2] + (int)(li / ((long long)P.onb_n[0] * P.onb_n[1]))};
  double *F = lds[wave];
  if (lane < NF) {
    const int d = lane / (W * 4), r = lane % (W * 4);
    F[lane] = fac[((size_t)d * A.nnmax + g[d]) * (W * 4) + r];
  }
  const int len = ri.lenf & 0xFFFF;
  if (lane < 3) {
    const double m = mass[g[0]] * mass[A.nnmax + g[1]] * mass[2 * A.nnmax + g[2]];
    b[ri.bidx + lane] = A.f[lane] * m;
  }
  __builtin_amdgcn_wave_barrier();  // LDS written and read by the same wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const uint16_t *pat = pinv + (size_t)ri.pid * A.stride;
  double *row0 = values + ri.rowstart;
  const double lam = A.lam, mu = A.mu, lpm = A.lam + A.mu;
  for (int e = lane; e < len; e += 64) {
    const int s3 = pat[e];
    const int s = s3 / 3, bc = s3 - 3 * s;
    const int o0 = s % W, o1 = (s / W) % W, o2 = s / (W * W);
    const double2 *f0 = reinterpret_cast<const double2 *>(F + (0 * W + o0) * 4);
    const double2 *f1 = reinterpret_cast<const double2 *>(F + (1 * W + o1) * 4);
    const double2 *f2 = reinterpret_cast<const double2 *>(F + (2 * W + o2) * 4);
    const double2 km0 = f0[0], pq0 = f0[1], km1 = f1[0], pq1 = f1[1], km2 = f2[0], pq2 = f2[1];
    const double Kd[3] = {km0.x, km1.x, km2.x}, Md[3] = {km0.y, km1.y, km2.y};
    const double Pd[3] = {pq0.x, pq1.x, pq2.x}, Qd[3] = {pq0.y, pq1.y, pq2.y};
    const double lap = mu * (Kd[0] * Md[1] * Md[2] + Md[0] * Kd[1] * Md[2] + Md[0] * Md[1] * Kd[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double v;
      if (a == bc) {
        const int c1 = (a + 1) % 3, c2 = (a + 2) % 3;
        v = lap + lpm * Kd[a] * Md[c1] * Md[c2];
      } else {
        const int c = 3 - a - bc;
        v = (lam * Pd[a] * Qd[bc] + mu * Qd[a] * Pd[bc]) * Md[c];
      }
      row0[(size_t)a * len + e] = v;
    }
  }
}

}  // namespace

void elast_free(gdx_problem *p) {
  ElastData *E = ed(p);
  if (!E) return;
  void *ptrs[] = {E->fac, E->mass, E->pinv, E->parts};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete E;
  p->elast = nullptr;
}

// 1D factors assembled over the mesh with the problem's Gauss rule
// (p->kp: points, weights, basis values and derivatives on [0,1], spacing h)
int elast_setup(gdx_problem *p) {
  elast_free(p);
  const KParams &kp = p->kp;
  const int K = p->K, W = 2 * K + 1, NQ = p->NQ;
  int nnmax = 0;
  for (int d = 0; d < 3; ++d) nnmax = std::max(nnmax, kp.nn[d]);
  std::vector<double> fac((size_t)3 * nnmax * W * 4, 0.0), mass((size_t)3 * nnmax, 0.0);
  for (int d = 0; d < 3; ++d) {
    const double h = kp.h[d];
    for (int e = 0; e < kp.ncells[d]; ++e)
      for (int a = 0; a <= K; ++a) {
        const int i = K * e + a;
        double m = 0.0;
        for (int q = 0; q < NQ; ++q) m += kp.qw[q] * h * kp.phi[a][q];
        mass[(size_t)d * nnmax + i] += m;
        for (int bb = 0; bb <= K; ++bb) {
          double kk = 0.0, mm = 0.0, pp = 0.0, qq = 0.0;
          for (int q = 0; q < NQ; ++q) {
            kk += kp.qw[q] * kp.dphi[a][q] * kp.dphi[bb][q] / h;
            mm += kp.qw[q] * h * kp.phi[a][q] * kp.phi[bb][q];
            pp += kp.qw[q] * kp.dphi[a][q] * kp.phi[bb][q];
            qq += kp.qw[q] * kp.phi[a][q] * kp.dphi[bb][q];
          }
          double *f = &fac[(((size_t)d * nnmax + i) * W + (bb - a + K)) * 4];
          f[0] += kk;
          f[1] += mm;
          f[2] += pp;
          f[3] += qq;
        }
      }
  }
  std::vector<ElastPart> parts(p->nparts);
  long long nrow = 0;
  for (int r = 0; r < p->nparts; ++r) {
    const Part &P = p->P[r];
    parts[r].rowinfo_base = P.rowinfo_base;
    for (int k = 0; k < 3; ++k) {
      parts[r].onb_lo[k] = (int)P.onb_lo[k];
      parts[r].onb_n[k] = (int)P.onb_n[k];
    }
    nrow += P.onb_n[0] * P.onb_n[1] * P.onb_n[2];
  }
  if (nrow != (long long)p->rowinfo.size()) return GDX_ERR_HIP;
  ElastData *E = new ElastData;
  p->elast = E;
  E->nparts = p->nparts;
  E->nnmax = nnmax;
  E->stride = p->pat16_stride;
  E->nrownodes = nrow;
  auto up = [](void **dst, const void *src, size_t bytes) {
    if (hipMalloc(dst, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
    return bytes == 0 || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up(reinterpret_cast<void **>(&E->fac), fac.data(), fac.size() * sizeof(double)) &&
            up(reinterpret_cast<void **>(&E->mass), mass.data(), mass.size() * sizeof(double)) &&
            up(reinterpret_cast<void **>(&E->pinv), p->pat16_inv.data(), p->pat16_inv.size() * sizeof(uint16_t)) &&
            up(reinterpret_cast<void **>(&E->parts), parts.data(), parts.size() * sizeof(ElastPart));
  if (!ok) {
    elast_free(p);
    return GDX_ERR_ALLOC;
  }
  return GDX_OK;
}

int elast_assemble(gdx_problem *p, hipStream_t st) {
  ElastData *E = ed(p);
  if (!E) return GDX_ERR_STATE;
  if (E->nrownodes == 0) return GDX_OK;
  ElastArgs A;
  A.lam = p->desc.lame_lambda;
  A.mu = p->desc.lame_mu;
  for (int k = 0; k < 3; ++k) A.f[k] = p->desc.body_force[k];
  A.nnmax = E->nnmax;
  A.stride = E->stride;
  A.nparts = E->nparts;
  A.nrownodes = E->nrownodes;
  const long long blocks = (E->nrownodes + kWaves - 1) / kWaves;
  if (blocks > 0x7FFFFFFFLL) return GDX_ERR_UNSUPPORTED;
  const dim3 grid((unsigned)blocks), block(64 * kWaves);
  if (p->K == 2)
    hipLaunchKernelGGL(k_elast<2>, grid, block, 0, st, A, E->parts, p->dev.rowinfo, E->fac, E->mass, E->pinv,
                       p->dev.values, p->dev.b);
  else
    hipLaunchKernelGGL(k_elast<1>, grid, block, 0, st, A, E->parts, p->dev.rowinfo, E->fac, E->mass, E->pinv,
                       p->dev.values, p->dev.b);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

}  // namespace gdx

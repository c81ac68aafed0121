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
// Two launches per assembly.  The cell stage evaluates every Gauss point of
// every mesh cell once, from the state on the global node lattice (free values
// scattered there by global id, Dirichlet nodes hold uD), into the cell's
// symmetric Jacobian and residual (HBM, entry-major).  The row stage is owner
// computes rows: 2^D lanes per owned free row gather the row's entries from its
// adjacent cells and sum them in a fixed cell order, so there are no atomics
// and two assemblies are bit-identical.  Rows are written through their
// position patterns (Dirichlet columns dropped, as in the affine path); ghost
// rows are never written and stay zero.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>
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
  double *cm = nullptr;           // cell Jacobians and residuals [NC(NC+1)/2 + NC][ncell]
  double *tab = nullptr;          // reference-cell tables per Gauss point (FluxTab layout)
  float *tab32 = nullptr;         // the same in FP32, for the Jacobian of the mixed form
  int64_t ncell = 0;
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

// Reference-cell tables at every Gauss point, identical for every cell of the
// uniform mesh, built once on the host (flux_setup) into a small read-only
// device buffer: per Gauss point q a record of QS doubles
//   [0]                 w_q x cell volume
//   [1 + a]             phi_a(q)
//   [1 + NC + a D + d]  d phi_a / d x_d (q)   (1/h applied)
//   [1 + NC + NC D + k] grad(phi_a).grad(phi_b) (q), k = sym_index(a, b), a <= b
// Every lane reads the same entries, so the kernels load them with scalar
// loads (uniform addresses) instead of LDS reads.
template <int D>
struct FluxTab {
  static constexpr int NC = 1 << D, TW = 0, TP = 1, TG = 1 + NC, TGAB = 1 + NC + NC * D;
  static constexpr int QS = (TGAB + NC * (NC + 1) / 2 + 7) & ~7;
};

// position of cell-matrix entry (a, b), a <= b, in the packed upper triangle
template <int NC>
__host__ __device__ constexpr int sym_index(int a, int b) {
  return a * NC - a * (a - 1) / 2 + (b - a);
}

// cell Jacobian and residual of one cell (the stage-1 arithmetic, shared by the
// two-launch and the fused forms): J packed upper triangle, then r.  JT is the
// Jacobian's type: with float (the mixed form) its products and sums over the
// Gauss points run in FP32 from the FP32 tables; grad u, kappa, gamma and the
// residual stay FP64 either way.
template <int D, int NQ, int MODE, class JT = double>
__device__ __forceinline__ void flux_cell(const double *__restrict__ tab, const float *__restrict__ tab32,
                                          const double *__restrict__ ulat, int64_t c0, int64_t c1, int64_t c2,
                                          const KParams &kp, const FluxParams &fp,
                                          JT (&J)[(1 << D) * ((1 << D) + 1) / 2], double (&R)[1 << D]) {
  using F = FluxTab<D>;
  constexpr int NC = 1 << D, NQD = D == 3 ? NQ * NQ * NQ : NQ * NQ, NJ = NC * (NC + 1) / 2;
  const int64_t sy = kp.nn[0], sz = (int64_t)kp.nn[0] * kp.nn[1];
  const double *ub = ulat + c0 + sy * c1 + sz * c2;
  double uc[NC];
#pragma unroll
  for (int bb = 0; bb < NC; ++bb) uc[bb] = ub[(bb & 1) + sy * ((bb >> 1) & 1) + (D == 3 ? sz * ((bb >> 2) & 1) : 0)];
#pragma unroll
  for (int k = 0; k < NJ; ++k) J[k] = 0;
#pragma unroll
  for (int a = 0; a < NC; ++a) R[a] = 0.0;
#pragma unroll 1
  for (int q = 0; q < NQD; ++q) {
    const double *T = tab + q * F::QS;
    double gu0 = 0.0, gu1 = 0.0, gu2 = 0.0;
#pragma unroll
    for (int bb = 0; bb < NC; ++bb) {
      gu0 = fma(uc[bb], T[F::TG + bb * D], gu0);
      gu1 = fma(uc[bb], T[F::TG + bb * D + 1], gu1);
      if constexpr (D == 3) gu2 = fma(uc[bb], T[F::TG + bb * D + 2], gu2);
    }
    const double s2 = fp.alpha + fp.beta * (gu0 * gu0 + gu1 * gu1 + gu2 * gu2);
    const double kap = flux_kappa<MODE>(s2, fp);
    const double gam = (MODE == 1 || s2 <= 0.0) ? 0.0 : (fp.p - 2.0) * fp.beta * kap / s2;
    const double wd = T[F::TW];
    double ga[NC];
#pragma unroll
    for (int a = 0; a < NC; ++a) {
      double v = gu0 * T[F::TG + a * D] + gu1 * T[F::TG + a * D + 1];
      if constexpr (D == 3) v += gu2 * T[F::TG + a * D + 2];
      ga[a] = v;
    }
    const double wk = wd * kap, wg = wd * gam;
    if constexpr (std::is_same_v<JT, float>) {
      const float *T32 = tab32 + q * F::QS;
      const float wkf = (float)wk;
      float gaf[NC];
#pragma unroll
      for (int a = 0; a < NC; ++a) gaf[a] = (float)ga[a];
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        R[a] = fma(wd, kap * ga[a] - fp.f * T[F::TP + a], R[a]);
        const float wga = (float)(wg * ga[a]);
#pragma unroll
        for (int bb = a; bb < NC; ++bb) {
          const int k = sym_index<NC>(a, bb);
          J[k] = fmaf(wkf, T32[F::TGAB + k], fmaf(wga, gaf[bb], J[k]));
        }
      }
    } else {
#pragma unroll
      for (int a = 0; a < NC; ++a) {
        R[a] = fma(wd, kap * ga[a] - fp.f * T[F::TP + a], R[a]);
        const double wga = wg * ga[a];
#pragma unroll
        for (int bb = a; bb < NC; ++bb) {
          const int k = sym_index<NC>(a, bb);
          J[k] = fma(wk, T[F::TGAB + k], fma(wga, ga[bb], J[k]));
        }
      }
    }
  }
}

// Stage 1: one lane per mesh cell evaluates each of the cell's Gauss points
// once (grad u, kappa, gamma) and accumulates the cell Jacobian
//   J_ab = sum_q w_q (kappa grad(phi_a).grad(phi_b) + gamma (grad(u).grad(phi_a)) (grad(u).grad(phi_b)))
// (symmetric: the upper triangle is kept) and the cell residual
//   r_a = sum_q w_q (kappa grad(u).grad(phi_a) - f phi_a).
// Output, entry-major for coalesced stores and loads: cm[e * ncell + cell],
// e < NC(NC+1)/2 the packed J, then the NC residual entries.
template <int D, int NQ, int MODE>
__global__ __launch_bounds__(256) void k_flux_cells(int64_t ncell, const double *__restrict__ tab,
                                                    const double *__restrict__ ulat, double *__restrict__ cm,
                                                    KParams kp, FluxParams fp) {
  constexpr int NC = 1 << D, NJ = NC * (NC + 1) / 2;
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int64_t nc0 = kp.ncells[0], nc1 = kp.ncells[1];
  const int64_t c0 = c % nc0, c1 = (c / nc0) % nc1, c2 = D == 3 ? c / (nc0 * nc1) : 0;
  double J[NJ], R[NC];
  flux_cell<D, NQ, MODE>(tab, nullptr, ulat, c0, c1, c2, kp, fp, J, R);
#pragma unroll
  for (int k = 0; k < NJ; ++k) cm[k * ncell + c] = J[k];
#pragma unroll
  for (int a = 0; a < NC; ++a) cm[(NJ + a) * ncell + c] = R[a];
}

// Fused 3D form: one workgroup of NT lanes per tile of 8 x R2 x R3 rows.  The
// tile's (ext+1)^3 cells (at most 9 x (R2+1) x (R3+1)) compute their Jacobian
// and residual into LDS (one lane per cell, the stage-1 arithmetic), then
// NT / rows lanes per row sum the row's 27 slots from LDS in the fixed cell
// order of the row stage and store them through the row's position pattern.
// Cells on the tile boundary are recomputed by the neighbouring tiles instead
// of a round trip of every cell matrix through HBM.  With 8 x 4 x 4 rows (225
// cells, 81 KB of records, 256 lanes) two workgroups share a CU, so one's
// loads and stores overlap the other's cell arithmetic (8 x 8 x 4 rows with
// 405 cells in 146 KB left one workgroup per CU: 0.29 against 0.25 ms at 96^3).
// The values do not depend on the tile shape (same cell arithmetic, same
// summation order per row).  JT = float is the mixed form: cell Jacobians in
// FP32 (records of 37 floats, the residuals in FP64 records of their own:
// 49 KB, three workgroups per CU), the rows still summed in FP64.
template <int NQ, int MODE, int R2, int R3, int NT, class JT>
__global__ __launch_bounds__(NT) void k_flux_fused(const Tile *__restrict__ tiles, int ntiles,
                                                    const RowInfo *__restrict__ rowinfo,
                                                    const uint8_t *__restrict__ pat_fwd, int pstride,
                                                    const double *__restrict__ tab, const float *__restrict__ tab32,
                                                    const double *__restrict__ ulat, double *__restrict__ values,
                                                    double *__restrict__ b, KParams kp, FluxParams fp) {
  constexpr int D = 3, NC = 8, NJ = 36, NE = NJ + NC;
  constexpr bool F32 = std::is_same_v<JT, float>;
  // odd record strides against bank conflicts; FP64: J and r in one record
  constexpr int ESJ = F32 ? NJ + 1 : NE + 1, ESR = F32 ? NC + 1 : ESJ, OFFR = F32 ? 0 : NJ;
  constexpr int R1 = 8, CX = R1 + 1, CY = R2 + 1, CZ = R3 + 1, NCELL = CX * CY * CZ;
  constexpr int LPR = NT / (R1 * R2 * R3);  // lanes per row
  static_assert(NCELL <= NT && LPR >= 1, "one lane per cell");
  __shared__ JT cj[NCELL * ESJ];
  __shared__ double cjr_own[F32 ? NCELL * ESR : 1];
  double *cjr;
  if constexpr (F32) cjr = cjr_own;
  else cjr = cj;
  const int tb = tile_of_block(ntiles);
  if (tb >= ntiles) return;
  const Tile t = tiles[tb];
  const int tid = threadIdx.x;
  // this lane's stage-2 row: its row-table entry and its pattern entries are
  // loaded now, so that their latency overlaps the cell arithmetic
  constexpr int SPL = (27 + LPR - 1) / LPR;  // slots per lane
  const int r = tid / LPR, q = tid % LPR;
  const int u1 = r % R1, u2 = (r / R1) % R2, u3 = r / (R1 * R2);
  const bool rin = u1 < t.ext[0] && u2 < t.ext[1] && u3 < t.ext[2];
  RowInfo ri{0, 0, -1, 0};
  if (rin) ri = rowinfo[t.rowinfo_base + row_index(t, t.org[0] + u1, t.org[1] + u2, t.org[2] + u3)];
  int posv[SPL];
#pragma unroll
  for (int j = 0; j < SPL; ++j)
    posv[j] = (ri.pid >= 0 && q + LPR * j < 27) ? pat_fwd[(size_t)ri.pid * pstride + q + LPR * j] : kSlotAbsent;
  // stage 1: the cells the tile's rows touch (zero outside the mesh)
  if (tid < NCELL) {
    const int ix = tid % CX, iy = (tid / CX) % CY, iz = tid / (CX * CY);
    const int64_t cx = t.org[0] - 1 + ix, cy = t.org[1] - 1 + iy, cz = t.org[2] - 1 + iz;
    JT *dst = cj + tid * ESJ;
    double *dstr = cjr + tid * ESR + OFFR;
    if (ix <= t.ext[0] && iy <= t.ext[1] && iz <= t.ext[2] && cx >= 0 && cx < kp.ncells[0] && cy >= 0 &&
        cy < kp.ncells[1] && cz >= 0 && cz < kp.ncells[2]) {
      JT J[NJ];
      double R[NC];
      flux_cell<D, NQ, MODE, JT>(tab, tab32, ulat, cx, cy, cz, kp, fp, J, R);
#pragma unroll
      for (int k = 0; k < NJ; ++k) dst[k] = J[k];
#pragma unroll
      for (int a = 0; a < NC; ++a) dstr[a] = R[a];
    } else {
#pragma unroll
      for (int k = 0; k < NJ; ++k) dst[k] = 0;
#pragma unroll
      for (int a = 0; a < NC; ++a) dstr[a] = 0.0;
    }
  }
  __syncthreads();
  // stage 2: LPR lanes per row split its 27 slots
  if (!rin || ri.pid < 0) return;
  double *dstv = values + ri.rowstart;
#pragma unroll
  for (int j = 0; j < SPL; ++j) {
    const int sl = q + LPR * j;
    if (sl >= 27) break;
    const int s0 = sl % 3, s1 = (sl / 3) % 3, s2 = sl / 9;
    double acc = 0.0;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
      const int c0 = cc & 1, c1 = (cc >> 1) & 1, c2 = (cc >> 2) & 1;
      const int b0 = s0 - c0, b1 = s1 - c1, b2 = s2 - c2;
      if (b0 < 0 || b0 > 1 || b1 < 0 || b1 > 1 || b2 < 0 || b2 > 1) continue;
      const int ai = (NC - 1) - cc, bb = b0 | (b1 << 1) | (b2 << 2);
      const int e = bb < ai ? sym_index<NC>(bb, ai) : sym_index<NC>(ai, bb);
      acc += (double)cj[(((u3 + c2) * CY + (u2 + c1)) * CX + (u1 + c0)) * ESJ + e];
    }
    const int pos = posv[j];
    if (pos < kSlotDirichlet) dstv[pos] = acc;
  }
  if (q == 0) {
    double rr = 0.0;
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
      const int c0 = cc & 1, c1 = (cc >> 1) & 1, c2 = (cc >> 2) & 1;
      rr += cjr[(((u3 + c2) * CY + (u2 + c1)) * CX + (u1 + c0)) * ESR + OFFR + (NC - 1) - cc];
    }
    b[ri.bidx] = rr;
  }
}

// Stage 2: 2^D lanes per owned free row, one per adjacent cell, read the cell's
// row of J (the row node's local index ai against the 2^D columns) and r_ai;
// the row's lanes then sum the 3^D slots from LDS over the cells holding both
// nodes in a fixed cell order, and write them through the row's position
// pattern (Dirichlet columns dropped).  No atomics: two assemblies are
// bit-identical.  Ghost rows are never written and stay zero.
template <int D>
__global__ __launch_bounds__(256) void k_flux_rows(const Tile *__restrict__ tiles, int ntiles,
                                                   const RowInfo *__restrict__ rowinfo,
                                                   const uint8_t *__restrict__ pat_fwd, int pstride,
                                                   const double *__restrict__ cm, int64_t ncell,
                                                   double *__restrict__ values, double *__restrict__ b, KParams kp) {
  constexpr int S = D == 3 ? 27 : 9, NC = 1 << D, NJ = NC * (NC + 1) / 2;
  constexpr int RPB = 256 / NC;  // rows per pass of the workgroup
  __shared__ double part[256][NC + 1];
  const int tb = tile_of_block(ntiles);
  if (tb >= ntiles) return;
  const Tile t = tiles[tb];
  const int nrows = t.ext[0] * t.ext[1] * t.ext[2];
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
    // the lane's pattern entries (slots e, e + NC, ...), loaded before the barrier
    constexpr int SPL = (S + NC - 1) / NC;
    int posv[SPL];
#pragma unroll
    for (int j = 0; j < SPL; ++j)
      posv[j] = (live && e + NC * j < S) ? pat_fwd[(size_t)ri.pid * pstride + e + NC * j] : kSlotAbsent;
    const int c0 = g0 - 1 + e0, c1 = g1 - 1 + e1, c2 = D == 3 ? g2 - 1 + e2 : 0;
    const bool incell = live && c0 >= 0 && c0 < kp.ncells[0] && c1 >= 0 && c1 < kp.ncells[1] &&
                        (D == 2 || (c2 >= 0 && c2 < kp.ncells[2]));
    double v[NC + 1];
    if (incell) {
      const double *src = cm + c0 + (int64_t)kp.ncells[0] * (c1 + (int64_t)kp.ncells[1] * c2);
#pragma unroll
      for (int bb = 0; bb < NC; ++bb)
        v[bb] = src[(int64_t)(bb < ai ? sym_index<NC>(bb, ai) : sym_index<NC>(ai, bb)) * ncell];
      v[NC] = src[(int64_t)(NJ + ai) * ncell];
    } else {
#pragma unroll
      for (int bb = 0; bb <= NC; ++bb) v[bb] = 0.0;
    }
#pragma unroll
    for (int bb = 0; bb <= NC; ++bb) part[tid][bb] = v[bb];
    __syncthreads();
    if (live) {
      // slot (s0, s1, s2) of the row = sum over the cells c holding both nodes of
      // c's entry for column (s - c), cells in a fixed order
      double *dst = values + ri.rowstart;
      const int base = rl * NC;
#pragma unroll
      for (int j = 0; j < SPL; ++j) {
        const int sl = e + NC * j;
        if (sl >= S) break;
        const int s0 = sl % 3, s1 = (sl / 3) % 3, s2 = D == 3 ? sl / 9 : 0;
        double acc = 0.0;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) {
          const int b0 = s0 - (cc & 1), b1 = s1 - ((cc >> 1) & 1), b2 = D == 3 ? s2 - ((cc >> 2) & 1) : 0;
          if (b0 < 0 || b0 > 1 || b1 < 0 || b1 > 1 || b2 < 0 || b2 > 1) continue;
          acc += part[base + cc][b0 | (b1 << 1) | (b2 << 2)];
        }
        const int pos = posv[j];
        if (pos < kSlotDirichlet) dst[pos] = acc;
      }
      if (e == 0) {
        double rr = 0.0;
#pragma unroll
        for (int cc = 0; cc < NC; ++cc) rr += part[base + cc][NC];
        b[ri.bidx] = rr;
      }
    }
    __syncthreads();  // part is rewritten by the next pass
  }
}

template <int D, int NQ, int MODE>
int launch_flux_m(gdx_problem *p, const FluxData *fd, const FluxParams &fp, bool j32, hipStream_t st) {
  if constexpr (D == 3 && NQ <= 2) {  // (3 Gauss points per direction: the tables leave no room in LDS)
    const int nt = (int)p->tiles.size();
    if (flux_mixed_supported(p)) {  // the fused form's tile shape
      if (j32)
        hipLaunchKernelGGL((k_flux_fused<NQ, MODE, kFluxR2, kFluxR3, kFluxNT, float>), dim3(8u * ((nt + 7) / 8)),
                           dim3(kFluxNT), 0, st, p->dev.tiles, nt, p->dev.rowinfo, p->dev.pat_fwd,
                           pattern_stride(p->kp.S), fd->tab, fd->tab32, fd->ulat, p->dev.values, p->dev.b, p->kp, fp);
      else
        hipLaunchKernelGGL((k_flux_fused<NQ, MODE, kFluxR2, kFluxR3, kFluxNT, double>), dim3(8u * ((nt + 7) / 8)),
                           dim3(kFluxNT), 0, st, p->dev.tiles, nt, p->dev.rowinfo, p->dev.pat_fwd,
                           pattern_stride(p->kp.S), fd->tab, fd->tab32, fd->ulat, p->dev.values, p->dev.b, p->kp, fp);
      return GDX_OK;
    }
  }
  if (j32) return GDX_ERR_UNSUPPORTED;  // the mixed form exists in the fused kernel only
  hipLaunchKernelGGL((k_flux_cells<D, NQ, MODE>), dim3((unsigned)((fd->ncell + 255) / 256)), dim3(256), 0, st,
                     fd->ncell, fd->tab, fd->ulat, fd->cm, p->kp, fp);
  const int nt = (int)p->tiles.size();
  hipLaunchKernelGGL((k_flux_rows<D>), dim3(8u * ((nt + 7) / 8)), dim3(256), 0, st, p->dev.tiles, nt,
                     p->dev.rowinfo, p->dev.pat_fwd, pattern_stride(p->kp.S), fd->cm, fd->ncell, p->dev.values,
                     p->dev.b, p->kp);
  return GDX_OK;
}

template <int D, int NQ>
int launch_flux_t(gdx_problem *p, const FluxData *fd, const FluxParams &fp, bool j32, hipStream_t st) {
  switch (fp.mode) {
    case 1: return launch_flux_m<D, NQ, 1>(p, fd, fp, j32, st);
    case 2: return launch_flux_m<D, NQ, 2>(p, fd, fp, j32, st);
    case 3: return launch_flux_m<D, NQ, 3>(p, fd, fp, j32, st);
    default: return launch_flux_m<D, NQ, 0>(p, fd, fp, j32, st);
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

bool flux_mixed_supported(const gdx_problem *p) {
  return p->D == 3 && p->NQ <= 2 && p->tileR[0] == 8 && p->tileR[1] == kFluxR2 && p->tileR[2] == kFluxR3;
}

void flux_free(gdx_problem *p) {
  FluxData *fd = flux_data(p);
  if (!fd) return;
  void *ptrs[] = {fd->ulat, fd->gid_node, fd->xs, fd->cm, fd->tab, fd->tab32};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete fd;
  p->flux = nullptr;
}

// device buffers of the flux operator: lattice state with Dirichlet data, the
// free-id -> lattice map, and the state vector (zero).  Run by gdx_symbolic,
// never inside a capture; any failure leaves p->flux unset.
// the FluxTab records of every Gauss point (tensor-product rule, x fastest)
static std::vector<double> flux_tables_host(const KParams &kp, int D, int NQ) {
  const int NC = 1 << D, NQD = D == 3 ? NQ * NQ * NQ : NQ * NQ;
  const int TP = 1, TG = 1 + NC, TGAB = 1 + NC + NC * D, QS = (TGAB + NC * (NC + 1) / 2 + 7) & ~7;
  std::vector<double> t((size_t)NQD * QS, 0.0);
  const double ih[3] = {1.0 / kp.h[0], 1.0 / kp.h[1], D == 3 ? 1.0 / kp.h[2] : 0.0};
  const double vol = kp.h[0] * kp.h[1] * (D == 3 ? kp.h[2] : 1.0);
  for (int q = 0; q < NQD; ++q) {
    double *T = t.data() + (size_t)q * QS;
    const int q0 = q % NQ, q1 = (q / NQ) % NQ, q2 = D == 3 ? q / (NQ * NQ) : 0;
    T[0] = kp.qw[q0] * kp.qw[q1] * (D == 3 ? kp.qw[q2] : 1.0) * vol;
    for (int bb = 0; bb < NC; ++bb) {
      const int b0 = bb & 1, b1 = (bb >> 1) & 1, b2 = D == 3 ? (bb >> 2) & 1 : 0;
      const double p0 = kp.phi[b0][q0], p1 = kp.phi[b1][q1], p2 = D == 3 ? kp.phi[b2][q2] : 1.0;
      const double d0 = kp.dphi[b0][q0] * ih[0], d1 = kp.dphi[b1][q1] * ih[1];
      const double d2 = D == 3 ? kp.dphi[b2][q2] * ih[2] : 0.0;
      T[TP + bb] = p0 * p1 * p2;
      T[TG + bb * D] = d0 * p1 * p2;
      T[TG + bb * D + 1] = p0 * d1 * p2;
      if (D == 3) T[TG + bb * D + 2] = p0 * p1 * d2;
    }
    for (int a = 0; a < NC; ++a)
      for (int bb = a; bb < NC; ++bb) {
        double v = 0.0;
        for (int d = 0; d < D; ++d) v += T[TG + a * D + d] * T[TG + bb * D + d];
        T[TGAB + a * NC - a * (a - 1) / 2 + (bb - a)] = v;
      }
  }
  return t;
}

int flux_setup(gdx_problem *p) {
  flux_free(p);
  const int64_t nnodes = (int64_t)p->kp.nn[0] * p->kp.nn[1] * p->kp.nn[2];
  // free global id -> lattice node: built on the device when the numbering is there
  std::vector<long long> gid_node(p->dev_numbering ? 0 : std::max<int64_t>(p->ngdofs, 1), 0);
  for (const Part &P : p->P)
    for (int64_t i = 0; i < P.n_ldofs && !p->dev_numbering; ++i) {
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
  fd->ncell = num_cells(p);
  const int nc = 1 << p->D;
  const int64_t ncm = std::max<int64_t>((int64_t)(nc * (nc + 1) / 2 + nc) * fd->ncell, 1);
  hipStream_t st = p->ctx->stream;
  const int64_t nx = std::max<int64_t>(p->ngdofs, 1);
  if (dev_malloc(reinterpret_cast<void **>(&fd->cm), ncm * sizeof(double)) != hipSuccess ||
      dev_malloc(reinterpret_cast<void **>(&fd->ulat), nnodes * sizeof(double)) != hipSuccess ||
      dev_malloc(reinterpret_cast<void **>(&fd->gid_node), std::max<int64_t>(p->ngdofs, 1) * sizeof(long long)) !=
          hipSuccess ||
      dev_malloc(reinterpret_cast<void **>(&fd->xs), nx * sizeof(double)) != hipSuccess) {
    flux_free(p);
    return GDX_ERR_ALLOC;
  }
  const std::vector<double> tab = flux_tables_host(p->kp, p->D, p->NQ);
  const std::vector<float> tab32(tab.begin(), tab.end());
  if (dev_malloc(reinterpret_cast<void **>(&fd->tab), tab.size() * sizeof(double)) != hipSuccess ||
      dev_malloc(reinterpret_cast<void **>(&fd->tab32), tab32.size() * sizeof(float)) != hipSuccess) {
    flux_free(p);
    return GDX_ERR_ALLOC;
  }
  if ((p->dev_numbering ? gid_node_device(p, fd->gid_node) != GDX_OK
                         : hipMemcpy(fd->gid_node, gid_node.data(), gid_node.size() * sizeof(long long),
                                     hipMemcpyHostToDevice) != hipSuccess) ||
      hipMemcpy(fd->tab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(fd->tab32, tab32.data(), tab32.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
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
// j32: the mixed form (cell Jacobians in FP32; flux_mixed_supported problems).
int flux_assemble(gdx_problem *p, const double *x, bool linear, bool j32, hipStream_t st) {
  if (!p->flux) return GDX_ERR_STATE;  // flux_setup runs in gdx_symbolic
  FluxData *fd = flux_data(p);
  const int64_t n = p->ngdofs;
  if (n > 0)
    hipLaunchKernelGGL(k_state_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, fd->gid_node,
                       x ? x : fd->xs, fd->ulat);
  const FluxParams fp = flux_params(p->desc, linear);
  if (p->tiles.empty()) return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
  int rc = GDX_ERR_UNSUPPORTED;
  if (p->D == 3 && p->NQ == 2) rc = launch_flux_t<3, 2>(p, fd, fp, j32, st);
  else if (p->D == 3 && p->NQ == 1) rc = launch_flux_t<3, 1>(p, fd, fp, j32, st);
  else if (p->D == 3 && p->NQ == 3) rc = launch_flux_t<3, 3>(p, fd, fp, j32, st);
  else if (p->D == 2 && p->NQ == 2) rc = launch_flux_t<2, 2>(p, fd, fp, j32, st);
  else if (p->D == 2 && p->NQ == 1) rc = launch_flux_t<2, 1>(p, fd, fp, j32, st);
  else if (p->D == 2 && p->NQ == 3) rc = launch_flux_t<2, 3>(p, fd, fp, j32, st);
  if (rc) return rc;
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

}  // namespace gdx

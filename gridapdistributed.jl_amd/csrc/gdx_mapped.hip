// Diffusion on a mapped Cartesian grid: the numeric phase of
// AffineFEOperator(a, l, U, V) with a(u,v) = int( kappa grad(v).grad(u) ),
// l(v) = int( f v ) on CartesianDiscreteModel(ranks, parts, domain, cells; map)
// (src/Geometry.jl:378-388 forwards the map to the CartesianDescriptor): every
// mesh vertex sits where the map puts it and each cell takes the multilinear
// (Q1) map of its 2^D vertices, so J, detJ and J^-1 vary over the cell.  The
// reference integrates each cell with this map (src/FESpaces.jl:794-798); the
// uniform-box kernels of this library assume J = diag(h) instead.
//
// Two launches:
// - k_mapped_cells: one wave per cell.  The Gauss points' Jacobians, weights
//   w |detJ| kappa(x_q) and w |detJ| f(x_q) (kappa: the Fourier form at the
//   mapped point x_q, or kappa per cell), then the physical basis gradients
//   J^-T grad(phi_a) of every (point, basis function) to LDS, then the cell
//   matrix on the matrix cores and the cell vector; written cell-major to HBM
//   (K_c[a][b] row-major, then b_c[a]; local nodes tensor-lexicographic);
// - k_mapped_rows: lanes per owned free row over its (2K+1)^D stencil slots.
//   A slot's value is the sum over the cells holding both nodes of their
//   entry, cells in a fixed order; it goes to the row's position for that slot
//   or, for a Dirichlet column, into the lifting sum_j A_ij uD(x_j).  The row's
//   lanes then reduce b = sum_c b_c - lifting in a fixed butterfly order.  No
//   atomics: repeated assemblies are bit-identical; ghost rows stay zero.
// HBM traffic per assembly: the cell records (written once, read once: entry
// (a, b) serves the row of node a, whose lanes read a contiguous row of the
// cell matrix) and the 8 B CSR value writes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "gdx_internal.h"

namespace gdx {

namespace {

// One owned free row, in the order of the rows in the values buffer.
struct MappedRow {
  long long rowstart;  // first value of the row
  long long bidx;      // RHS index
  int pid;             // position pattern
  unsigned short g[3]; // global lattice coordinates
  unsigned short pad[3];
};
static_assert(sizeof(MappedRow) == 32, "MappedRow is 32 bytes");

struct MappedData {
  double *vc = nullptr;     // vertex coordinates [nv][D]
  double *cm = nullptr;     // cell records [ncell][NE]
  MappedRow *rows = nullptr;
  long long nrows = 0, ncell = 0;
  int ne = 0;
};

MappedData *md(gdx_problem *p) { return reinterpret_cast<MappedData *>(p->mapped_data); }

constexpr long long kMappedBlocks = 2048;  // workgroups of the persistent cell kernel (8 per CU)
constexpr int kLdsModes = 32;             // Fourier modes each cell-kernel workgroup stages in LDS
static_assert(kLdsModes % 2 == 0, "the lanes of a Gauss point split the modes in pairs");

template <int D, int K>
struct MShape {
  static constexpr int NCN = D == 3 ? (K + 1) * (K + 1) * (K + 1) : (K + 1) * (K + 1);  // cell nodes
  static constexpr int NM = NCN * NCN;                                                  // cell matrix
  static constexpr int NE = NM + NCN;                                                   // record length
  static constexpr int NV = 1 << D;                                                     // vertices
  static constexpr int W = 2 * K + 1;
  static constexpr int S = D == 3 ? W * W * W : W * W;                                  // stencil slots
};


// the owned free rows of one part's owned node box, at their owned row index
__global__ void k_mapped_rowlist(const RowInfo *__restrict__ rowinfo, long long rinfo0, long long row_base,
                                 long long n_own, long long grp0, long long lo0, long long lo1, long long lo2,
                                 long long n0, long long n1, long long n2, MappedRow *__restrict__ rows,
                                 unsigned long long *cnt) {
  const long long li = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (li >= n0 * n1 * n2) return;
  const RowInfo ri = rowinfo[rinfo0 + li];
  if (ri.pid < 0) return;
  const long long o = ri.bidx - row_base;
  if (o < 0 || o >= n_own) {
    atomicOr(cnt + 1, 1ull);
    return;
  }
  atomicAdd(cnt, 1ull);  // every owned row must be listed exactly once
  MappedRow r{};
  r.rowstart = ri.rowstart;
  r.bidx = ri.bidx;
  r.pid = ri.pid;
  r.g[0] = (unsigned short)(lo0 + li % n0);
  r.g[1] = (unsigned short)(lo1 + (li / n0) % n1);
  r.g[2] = (unsigned short)(lo2 + li / (n0 * n1));
  rows[grp0 + o] = r;
}

typedef double f64x4_t __attribute__((ext_vector_type(4)));

// Stage 1: a persistent grid; each wave takes one cell at a time.  The cell
// matrix is G^T diag(w) G with G[(q, d)][a] = sum_e J^-T[d][e] Gref[q][a][e]
// (Gref the reference gradients at the Gauss points, the same for every cell:
// built once per workgroup in LDS): a (NQT D) x NCN matrix, K = NQT D (padded
// to KD, a multiple of 4) and NCN padded to NT tiles of 16.  Each K step forms
// the lane's two G values from LDS and feeds every tile pair to
// v_mfma_f64_16x16x4f64 (lane l holds A[l % 16][l / 16] and B[l / 16][l % 16];
// its accumulator i is D[4 i + l / 16][l % 16], unlike the FP32 16x16x4 form's
// D[4 (l / 16) + i]).  Only the tiles on and above the diagonal are formed;
// the full matrix is stored, rows contiguous.
template <int D, int K, int NQ>
struct CellShape {
  static constexpr int NQT = D == 3 ? NQ * NQ * NQ : NQ * NQ;  // Gauss points, <= 64
  static constexpr int KD = (NQT * D + 3) / 4 * 4;             // contraction length (gradient rows)
  static constexpr int KR = (NQT * (D + 1) + 3) / 4 * 4;       // with the reaction term's value rows
  static constexpr int NT = (MShape<D, K>::NCN + 15) / 16;      // 16-wide tiles per side
  static constexpr int NCN = MShape<D, K>::NCN;
  static constexpr int SHARED = (NQT * NCN * (D + 1) + 2 * kMaxQ + 5 * kLdsModes) * 8;  // LDS of the workgroup's tables
  static constexpr int PERWAVE = NQT * (D * D + 3) * 8;                 // LDS per wave
  // waves per workgroup: at most 60 KB of LDS per workgroup
  static constexpr int WPB = SHARED + 4 * PERWAVE <= 60 * 1024 ? 4 : (SHARED + 2 * PERWAVE <= 60 * 1024 ? 2 : 1);
  // waves per SIMD asked of the register allocator: 4 up to 27 Gauss points
  // (some 3D instances then spill 20-96 bytes per lane; the 3D Q2 one with
  // 27 points still runs faster with twice the waves)
  static constexpr int OCC = NQT <= 27 ? 4 : 1;
};

template <int D, int K, int NQ, bool REACT>
__global__ __launch_bounds__((64 * CellShape<D, K, NQ>::WPB))
__attribute__((amdgpu_waves_per_eu(CellShape<D, K, NQ>::OCC))) void k_mapped_cells(
    KParams kp, long long ncell, const double *__restrict__ vc, const double *__restrict__ modes,
    const double *__restrict__ cellk, double dw, double rw, double *__restrict__ cm) {
  using Sh = MShape<D, K>;
  using C = CellShape<D, K, NQ>;
  constexpr int NQT = C::NQT, KD = C::KD, NT = C::NT, WPB = C::WPB;
  constexpr int NCN = Sh::NCN, NM = Sh::NM, NV = Sh::NV;
  __shared__ double GR[NQT][NCN][D];      // reference gradients at the Gauss points
  __shared__ double PH[NQT][NCN];         // basis values at the Gauss points
  __shared__ double TQ[2][kMaxQ];         // Gauss points and weights on [0, 1]
  __shared__ double MO[kLdsModes][5];     // the first Fourier modes (the others are read from HBM)
  __shared__ double JT[WPB][NQT][D * D];  // J^-T per point
  __shared__ double WQ[WPB][NQT], WF[WPB][NQT], WM[WPB][NQT];  // w |det J| kappa dw, w |det J| f, w |det J| rw
  // the wave index is wave-uniform: so are the cell index, its lattice
  // coordinates and vertex addresses (scalar arithmetic and scalar loads)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  for (int i = threadIdx.x; i < NQ; i += blockDim.x) {
    TQ[0][i] = kp.qx[i];
    TQ[1][i] = kp.qw[i];
  }
  const int nml = cellk || !modes ? 0 : min(kp.nmodes, kLdsModes);
  for (int i = threadIdx.x; i < 5 * nml; i += blockDim.x) MO[i / 5][i % 5] = modes[i];
  for (int i = threadIdx.x; i < NQT * NCN; i += blockDim.x) {
    const int q = i / NCN, a = i % NCN;
    const int qi[3] = {q % NQ, (q / NQ) % NQ, D == 3 ? q / (NQ * NQ) : 0};
    const int ai[3] = {a % (K + 1), (a / (K + 1)) % (K + 1), D == 3 ? a / ((K + 1) * (K + 1)) : 0};
    double ph = 1.0;
#pragma unroll
    for (int d = 0; d < D; ++d) ph *= kp.phi[ai[d]][qi[d]];
    PH[q][a] = ph;
#pragma unroll
    for (int e = 0; e < D; ++e) {
      double g = 1.0;
#pragma unroll
      for (int d = 0; d < D; ++d) g *= d == e ? kp.dphi[ai[d]][qi[d]] : kp.phi[ai[d]][qi[d]];
      GR[q][a][e] = g;
    }
  }
  __syncthreads();  // the only block barrier: below, every wave works alone
  const int n0 = kp.ncells[0], n1 = kp.ncells[1];
  const long long nv0 = n0 + 1, nv1 = n1 + 1;
  constexpr int QS = NQT <= 32 ? 2 : 1;  // lanes per Gauss point in A
  const int q = QS == 2 ? (lane & 31) : lane, h = QS == 2 ? (lane >> 5) : 0;
  const int li = lane & 15, lk = lane >> 4;
  for (long long c = (long long)blockIdx.x * WPB + wave; c < ncell; c += (long long)gridDim.x * WPB) {
    const long long c0 = c % n0, c1 = (c / n0) % n1, c2 = D == 3 ? c / ((long long)n0 * n1) : 0;
    // A: Gauss point q: Jacobian of the multilinear map, weights, J^-T; with
    // at most 32 points two lanes per point (q, q + 32) share the Fourier modes
    if (q < NQT) {
      double X[NV][D];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const long long vi = (c0 + (v & 1)) + nv0 * ((c1 + ((v >> 1) & 1)) + nv1 * (c2 + ((v >> 2) & 1)));
#pragma unroll
        for (int d = 0; d < D; ++d) X[v][d] = vc[vi * D + d];
      }
      const int qi[3] = {q % NQ, (q / NQ) % NQ, D == 3 ? q / (NQ * NQ) : 0};
      double xi[3], J[D][D] = {}, x[3] = {0.0, 0.0, 0.0};
#pragma unroll
      for (int d = 0; d < D; ++d) xi[d] = TQ[0][qi[d]];
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        double N = 1.0, dN[D];
#pragma unroll
        for (int d = 0; d < D; ++d) N *= ((v >> d) & 1) ? xi[d] : 1.0 - xi[d];
#pragma unroll
        for (int e = 0; e < D; ++e) {
          dN[e] = 1.0;
#pragma unroll
          for (int d = 0; d < D; ++d)
            dN[e] *= d == e ? (((v >> d) & 1) ? 1.0 : -1.0) : (((v >> d) & 1) ? xi[d] : 1.0 - xi[d]);
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
          x[d] += N * X[v][d];
#pragma unroll
          for (int e = 0; e < D; ++e) J[d][e] += X[v][d] * dN[e];
        }
      }
      double det, Jit[D][D];  // J^-T[d][e] = J^-1[e][d] = adj(J)[e][d] / det
      if constexpr (D == 2) {
        det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        const double id = 1.0 / det;
        Jit[0][0] = J[1][1] * id;
        Jit[1][0] = -J[0][1] * id;
        Jit[0][1] = -J[1][0] * id;
        Jit[1][1] = J[0][0] * id;
      } else {
        const double a00 = J[1][1] * J[2][2] - J[1][2] * J[2][1], a01 = J[0][2] * J[2][1] - J[0][1] * J[2][2],
                     a02 = J[0][1] * J[1][2] - J[0][2] * J[1][1], a10 = J[1][2] * J[2][0] - J[1][0] * J[2][2],
                     a11 = J[0][0] * J[2][2] - J[0][2] * J[2][0], a12 = J[0][2] * J[1][0] - J[0][0] * J[1][2],
                     a20 = J[1][0] * J[2][1] - J[1][1] * J[2][0], a21 = J[0][1] * J[2][0] - J[0][0] * J[2][1],
                     a22 = J[0][0] * J[1][1] - J[0][1] * J[1][0];
        det = J[0][0] * a00 + J[0][1] * a10 + J[0][2] * a20;
        const double id = 1.0 / det;  // one division, nine products
        Jit[0][0] = a00 * id, Jit[0][1] = a10 * id, Jit[0][2] = a20 * id;
        Jit[1][0] = a01 * id, Jit[1][1] = a11 * id, Jit[1][2] = a21 * id;
        Jit[2][0] = a02 * id, Jit[2][1] = a12 * id, Jit[2][2] = a22 * id;
      }
      double w = fabs(det);
#pragma unroll
      for (int d = 0; d < D; ++d) w *= TQ[1][qi[d]];
      double kap, f;
      if (cellk) {
        kap = kp.coef_const * cellk[c];
        f = kp.rhs_const;
      } else {
        // kappa = coef exp(g), g = sum_m a_m cos(2 pi k_m.x + phi_m); the lanes
        // of a point take every other mode and add the partner's sums
        double g = 0.0, dg0 = 0.0, dg1 = 0.0, dg2 = 0.0;
        auto mode = [&](const double *Mo) {
          // theta / pi = 2 k.x + phi / pi: sincospi reduces its argument exactly
          const double tp = 2.0 * (Mo[0] * x[0] + Mo[1] * x[1] + Mo[2] * x[2]) + Mo[4] * M_1_PI;
          double sn, cs;
          sincospi(tp, &sn, &cs);
          g += Mo[3] * cs;
          const double sd = -Mo[3] * sn * 2.0 * M_PI;
          dg0 += sd * Mo[0];
          dg1 += sd * Mo[1];
          dg2 += sd * Mo[2];
        };
        // nml is kLdsModes (a multiple of QS) or all the modes, so the second
        // loop continues each lane's sequence h, h + QS, ...
        for (int m = h; m < nml; m += QS) mode(MO[m]);
        for (int m = nml + h; m < kp.nmodes; m += QS) mode(modes + 5 * m);
        if (QS == 2) {
          g += __shfl_xor(g, 32);
          dg0 += __shfl_xor(dg0, 32);
          dg1 += __shfl_xor(dg1, 32);
          dg2 += __shfl_xor(dg2, 32);
        }
        kap = kp.coef_const * exp(g);
        f = kp.rhs_kind == GDX_RHS_MANUFACTURED
                ? -kap * (dg0 * kp.dir_coef[1] + dg1 * kp.dir_coef[2] + dg2 * kp.dir_coef[3])
                : kp.rhs_const;
      }
      if (h == 0) {
#pragma unroll
        for (int d = 0; d < D; ++d)
#pragma unroll
          for (int e = 0; e < D; ++e) JT[wave][q][d * D + e] = Jit[d][e];
        WQ[wave][q] = w * kap * dw;
        WM[wave][q] = w * rw;
        WF[wave][q] = w * f;
      }
    }
    __builtin_amdgcn_wave_barrier();  // LDS written and read by the same wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // C: K step k0: the lane's row k = k0 + l / 16 of G at columns 16 t + l % 16
    f64x4_t acc[NT][NT];
#pragma unroll
    for (int ta = 0; ta < NT; ++ta)
#pragma unroll
      for (int tb = 0; tb < NT; ++tb) acc[ta][tb] = f64x4_t{0.0, 0.0, 0.0, 0.0};
    // with a reaction term rw u v (REACT), the rows (q, D) of the contraction
    // hold the basis values, weighted by w |det J| rw
    constexpr int KEND = REACT ? C::KR : KD;
#pragma unroll 3
    for (int k0 = 0; k0 < KEND; k0 += 4) {
      const int k = k0 + lk;
      double gv[NT], wk = 0.0;
#pragma unroll
      for (int t = 0; t < NT; ++t) gv[t] = 0.0;
      if (k < NQT * D) {
        const int qk = k / D, d = k % D;
        wk = WQ[wave][qk];
        double jt[D];
#pragma unroll
        for (int e = 0; e < D; ++e) jt[e] = JT[wave][qk][d * D + e];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int a = 16 * t + li;
          if (a < NCN) {
            double sg = 0.0;
#pragma unroll
            for (int e = 0; e < D; ++e) sg += jt[e] * GR[qk][a][e];
            gv[t] = sg;
          }
        }
      } else if (REACT && k < NQT * (D + 1)) {
        const int qm = k - NQT * D;
        wk = WM[wave][qm];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int a = 16 * t + li;
          gv[t] = a < NCN ? PH[qm][a] : 0.0;
        }
      }
#pragma unroll
      for (int ta = 0; ta < NT; ++ta)
#pragma unroll
        for (int tb = ta; tb < NT; ++tb)
          acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(wk * gv[ta], gv[tb], acc[ta][tb], 0, 0, 0);
    }
    // the cell matrix is symmetric: the tiles below the diagonal are the
    // transposes of those above it, written from the same accumulators
    double *out = cm + c * (long long)Sh::NE;
#pragma unroll
    for (int ta = 0; ta < NT; ++ta)
#pragma unroll
      for (int tb = ta; tb < NT; ++tb)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int a = 16 * ta + 4 * i + lk, b = 16 * tb + li;
          if (a < NCN && b < NCN) {
            out[a * NCN + b] = acc[ta][tb][i];
            if (ta != tb) out[b * NCN + a] = acc[ta][tb][i];
          }
        }
    // D: the cell vector
    for (int a = lane; a < NCN; a += 64) {
      double s = 0.0;
      for (int qq = 0; qq < NQT; ++qq) s += WF[wave][qq] * PH[qq][a];
      out[NM + a] = s;
    }
    __builtin_amdgcn_wave_barrier();  // the next cell rewrites JT, WQ, WF
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

// coordinates of global lattice node g on the mapped grid: the mean of the
// vertices of the edge, face or cell whose midpoint it is (a vertex itself)
template <int D, int K>
__device__ inline void node_xyz(const double *__restrict__ vc, const int g[3], long long nv0, long long nv1,
                                double x[3]) {
  x[0] = x[1] = x[2] = 0.0;
  int v0[3], two[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    v0[d] = d < D ? g[d] / K : 0;
    two[d] = d < D && g[d] % K != 0;
  }
  const int nb = 1 << (two[0] + two[1] + two[2]);
  for (int cc = 0; cc < (1 << D); ++cc) {
    const int s0 = cc & 1, s1 = (cc >> 1) & 1, s2 = (cc >> 2) & 1;
    if ((s0 && !two[0]) || (s1 && !two[1]) || (s2 && !two[2])) continue;
    const long long v = (v0[0] + s0) + nv0 * ((v0[1] + s1) + nv1 * (v0[2] + s2));
#pragma unroll
    for (int d = 0; d < D; ++d) x[d] += vc[v * D + d] / nb;
  }
}

// Stage 2: LPR lanes per row, NR consecutive rows per lane group; every load
// of the group's rows (row entries, pattern bytes, cell records) is issued
// before the first sum.  The record loads depend on the row entry only (the
// cells and local nodes follow from the node coordinates), not on the pattern.
template <int D, int K, int NR>
__global__ __launch_bounds__(256) void k_mapped_rows(KParams kp, long long nrows, const MappedRow *__restrict__ rows,
                                                     const uint8_t *__restrict__ pat_fwd, int pstride,
                                                     const double *__restrict__ cm, const double *__restrict__ vc,
                                                     double *__restrict__ values, double *__restrict__ b) {
  using Sh = MShape<D, K>;
  constexpr int S = Sh::S, W = Sh::W, NE = Sh::NE, NM = Sh::NM, NCN = Sh::NCN;
  constexpr int LPR = S <= 16 ? 16 : (S <= 32 ? 32 : 64);
  constexpr int SPL = (S + LPR - 1) / LPR;
  constexpr int NC = 1 << D;
  const int q = threadIdx.x % LPR;
  // workgroups are dealt to the 8 XCDs in turn: XCD x takes the x-th eighth of
  // the rows, so the cell records its rows share stay in its own L2
  const long long per = gridDim.x / 8;  // the grid is a multiple of 8
  const long long blk = (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
  const long long grp = (blk * blockDim.x + threadIdx.x) / LPR;
  const long long n0 = kp.ncells[0], n1 = kp.ncells[1];
  MappedRow mr[NR];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const long long r = grp * NR + j;
    mr[j] = MappedRow{};
    mr[j].pid = -1;
    if (r < nrows) mr[j] = rows[r];
  }
  int posv[NR][SPL];
  int cf[NR][3], af[NR][3], ns[NR][3];  // the row node's cells along each direction ("slots")
  double v[NR][SPL][NC];
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const bool live = mr[j].pid >= 0;
#pragma unroll
    for (int t = 0; t < SPL; ++t) {
      const int sl = q + LPR * t;
      posv[j][t] = live && sl < S ? pat_fwd[(size_t)mr[j].pid * pstride + sl] : kSlotAbsent;
    }
    // first cell cf, the node's local index af in it, count ns; a second slot
    // is cell cf + 1 at local index 0
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      cf[j][d] = 0, af[j][d] = 0, ns[j][d] = 1;
      if (d >= D) continue;
      const int i = mr[j].g[d];
      if (i % K == 0) {
        const bool left = i / K - 1 >= 0, right = i / K < kp.ncells[d];
        cf[j][d] = left ? i / K - 1 : i / K;
        af[j][d] = left ? K : 0;
        ns[j][d] = (int)left + (int)right;
      } else {
        cf[j][d] = i / K;
        af[j][d] = i % K;
      }
    }
    // the record row of the node in each of its cells: entry (a, b) of cell c
    // is at c NE + a NCN + b, and b = a + lo with lo the column offset's linear
    // index, so a slot adds lo to a per-cell base (computed once per row)
    long long base[NC];
    bool cv[NC];
#pragma unroll
    for (int cc = 0; cc < NC; ++cc) {
      const int s0 = cc & 1, s1 = (cc >> 1) & 1, s2 = D == 3 ? (cc >> 2) & 1 : 0;
      cv[cc] = s0 < ns[j][0] && s1 < ns[j][1] && s2 < ns[j][2];
      const int a0 = s0 ? 0 : af[j][0], a1 = s1 ? 0 : af[j][1], a2 = s2 ? 0 : af[j][2];
      const long long cell = (cf[j][0] + s0) + n0 * ((cf[j][1] + s1) + n1 * (long long)(D == 3 ? cf[j][2] + s2 : 0));
      base[cc] = cell * NE + (long long)(a0 + (K + 1) * (a1 + (K + 1) * (D == 3 ? a2 : 0))) * (NCN + 1);
    }
    // a combination of the node's cells that does not hold the column reads
    // record 0 and adds +0
#pragma unroll
    for (int t = 0; t < SPL; ++t) {
      const int sl = q + LPR * t;
      const int o0 = sl % W - K, o1 = (sl / W) % W - K, o2 = D == 3 ? sl / (W * W) - K : 0;
      const int lo = o0 + (K + 1) * (o1 + (K + 1) * o2);
      // the column lies in the node's first cell (local index af + o) or its
      // second (local index o), direction by direction
      const bool f0 = af[j][0] + o0 >= 0 && af[j][0] + o0 <= K, g0 = o0 >= 0 && o0 <= K;
      const bool f1 = af[j][1] + o1 >= 0 && af[j][1] + o1 <= K, g1 = o1 >= 0 && o1 <= K;
      const bool f2 = D == 2 || (af[j][2] + o2 >= 0 && af[j][2] + o2 <= K), g2 = D == 2 || (o2 >= 0 && o2 <= K);
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) {  // s0 fastest: the order of the sum below
        const int s0 = cc & 1, s1 = (cc >> 1) & 1, s2 = D == 3 ? (cc >> 2) & 1 : 0;
        const bool ok = sl < S && cv[cc] && (s0 ? g0 : f0) && (s1 ? g1 : f1) && (s2 ? g2 : f2);
        const double x = cm[ok ? base[cc] + lo : 0];
        v[j][t][cc] = ok ? x : 0.0;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NR; ++j) {
    const bool live = mr[j].pid >= 0;
    double lift = 0.0;
    double *dst = values + mr[j].rowstart;
#pragma unroll
    for (int t = 0; t < SPL; ++t) {
      const int sl = q + LPR * t;
      const int pos = posv[j][t];
      if (pos == kSlotAbsent) continue;
      double acc = 0.0;
#pragma unroll
      for (int cc = 0; cc < NC; ++cc) acc += v[j][t][cc];
      if (pos < kSlotDirichlet) {
        dst[pos] = acc;
      } else {  // Dirichlet column: uD at its node
        const int gc[3] = {mr[j].g[0] + sl % W - K, mr[j].g[1] + (sl / W) % W - K,
                           D == 3 ? mr[j].g[2] + sl / (W * W) - K : 0};
        double x[3];
        node_xyz<D, K>(vc, gc, n0 + 1, n1 + 1, x);
        const double ud = kp.dir_coef[0] + kp.dir_coef[1] * x[0] + kp.dir_coef[2] * x[1] + kp.dir_coef[3] * x[2];
        lift += acc * ud;
      }
    }
    // b = sum over the row node's cells of b_c[local node] - lifting: lane
    // q < 2^D takes the cell of slot combination q
    double part = -lift;
    if (live && q < NC) {
      const int s0 = q & 1, s1 = (q >> 1) & 1, s2 = (q >> 2) & 1;
      if (s0 < ns[j][0] && s1 < ns[j][1] && (D == 2 || s2 < ns[j][2])) {
        const long long cell =
            (cf[j][0] + s0) + n0 * ((cf[j][1] + s1) + n1 * (long long)(D == 3 ? cf[j][2] + s2 : 0));
        const int la = (s0 ? 0 : af[j][0]) + (K + 1) * ((s1 ? 0 : af[j][1]) + (K + 1) * (D == 3 ? (s2 ? 0 : af[j][2]) : 0));
        part += cm[cell * NE + NM + la];
      }
    }
#pragma unroll
    for (int m = LPR / 2; m >= 1; m >>= 1) part += __shfl_xor(part, m, LPR);
    if (live && q == 0) b[mr[j].bidx] = part;
  }
}

constexpr int kMappedRowsPerGroup = 1;  // rows per lane group of k_mapped_rows (2: 2.67 ms against 2.2 at 64^3 Q2)

template <int D, int K, int NQ>
int launch_mapped_t(gdx_problem *p, MappedData *M, hipStream_t st, bool rows_stage) {
  using Sh = MShape<D, K>;
  constexpr int WPB = CellShape<D, K, NQ>::WPB;
  if (!rows_stage) {
    // a persistent grid: every workgroup builds its tables once, then its waves
    // take every (WPB x workgroups)-th cell
    const long long blocks = std::min<long long>((M->ncell + WPB - 1) / WPB, kMappedBlocks);
    if (p->reac_w != 0.0)
      hipLaunchKernelGGL((k_mapped_cells<D, K, NQ, true>), dim3((unsigned)std::max<long long>(blocks, 1)),
                         dim3(64 * WPB), 0, st, p->kp, M->ncell, M->vc, p->dev.modes,
                         p->coef_cell ? p->dev.cellk : nullptr, p->diff_w, p->reac_w, M->cm);
    else
      hipLaunchKernelGGL((k_mapped_cells<D, K, NQ, false>), dim3((unsigned)std::max<long long>(blocks, 1)),
                         dim3(64 * WPB), 0, st, p->kp, M->ncell, M->vc, p->dev.modes,
                         p->coef_cell ? p->dev.cellk : nullptr, p->diff_w, 0.0, M->cm);
  } else if (M->nrows > 0) {
    constexpr int LPR = Sh::S <= 16 ? 16 : (Sh::S <= 32 ? 32 : 64);
    const long long groups = (M->nrows + kMappedRowsPerGroup - 1) / kMappedRowsPerGroup;
    const long long blocks = ((groups * LPR + 255) / 256 + 7) / 8 * 8;  // a multiple of 8 (XCD order)
    if (blocks > 0x7FFFFFFFLL) return GDX_ERR_UNSUPPORTED;
    hipLaunchKernelGGL((k_mapped_rows<D, K, kMappedRowsPerGroup>), dim3((unsigned)blocks), dim3(256), 0, st, p->kp,
                       M->nrows, M->rows,
                       p->dev.pat_fwd, pattern_stride(p->kp.S), M->cm, M->vc, p->dev.values, p->dev.b);
  }
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

template <int D, int K>
int launch_mapped_k(gdx_problem *p, MappedData *M, hipStream_t st, bool rows_stage) {
  switch (p->NQ) {
    case 1: return launch_mapped_t<D, K, 1>(p, M, st, rows_stage);
    case 2: return launch_mapped_t<D, K, 2>(p, M, st, rows_stage);
    case 3: return launch_mapped_t<D, K, 3>(p, M, st, rows_stage);
    case 4: return launch_mapped_t<D, K, 4>(p, M, st, rows_stage);
  }
  return GDX_ERR_UNSUPPORTED;
}

int launch_mapped(gdx_problem *p, MappedData *M, hipStream_t st, bool rows_stage) {
  if (p->D == 3) return p->K == 2 ? launch_mapped_k<3, 2>(p, M, st, rows_stage) : launch_mapped_k<3, 1>(p, M, st, rows_stage);
  return p->K == 2 ? launch_mapped_k<2, 2>(p, M, st, rows_stage) : launch_mapped_k<2, 1>(p, M, st, rows_stage);
}

}  // namespace

void mapped_free(gdx_problem *p) {
  MappedData *M = md(p);
  if (!M) return;
  void *ptrs[] = {M->vc, M->cm, M->rows};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete M;
  p->mapped_data = nullptr;
}

// vertex coordinates to HBM; the cell records and the row list are made on the
// first call after a symbolic phase (the numeric phase allocates nothing)
int mapped_set(gdx_problem *p, const double *xyz) {
  const int D = p->D, K = p->K;
  int64_t nv = 1;
  for (int d = 0; d < D; ++d) nv *= p->kp.ncells[d] + 1;
  for (int d = 0; d < D; ++d)
    if (p->kp.nn[d] > 65535) return GDX_ERR_UNSUPPORTED;  // 16-bit node coordinates in MappedRow
  MappedData *M = md(p);
  if (!M) {
    M = new MappedData;
    p->mapped_data = M;
    const int ncn = D == 3 ? (K + 1) * (K + 1) * (K + 1) : (K + 1) * (K + 1);
    M->ne = ncn * ncn + ncn;
    M->ncell = num_cells(p);
    std::vector<int64_t> grp0(p->nparts + 1, 0);
    for (int r = 0; r < p->nparts; ++r) grp0[r + 1] = grp0[r] + p->P[r].n_own;
    M->nrows = grp0[p->nparts];
    unsigned long long *cnt = nullptr;  // listed rows, out-of-range flag
    const size_t rbytes = std::max<int64_t>(M->nrows, 1) * sizeof(MappedRow);
    bool ok = dev_malloc(reinterpret_cast<void **>(&M->vc), (size_t)nv * D * sizeof(double)) == hipSuccess &&
              dev_malloc(reinterpret_cast<void **>(&M->cm), (size_t)std::max<int64_t>(M->ncell, 1) * M->ne *
                                                                sizeof(double)) == hipSuccess &&
              dev_malloc(reinterpret_cast<void **>(&M->rows), rbytes) == hipSuccess &&
              dev_malloc(reinterpret_cast<void **>(&cnt), 2 * sizeof(unsigned long long)) == hipSuccess &&
              hipMemsetAsync(M->rows, 0xFF, rbytes, p->ctx->stream) == hipSuccess &&  // pid -1: no row
              hipMemsetAsync(cnt, 0, 2 * sizeof(unsigned long long), p->ctx->stream) == hipSuccess;
    for (int r = 0; r < p->nparts && ok; ++r) {
      const Part &P = p->P[r];
      const long long n = P.onb_n[0] * P.onb_n[1] * P.onb_n[2];
      if (n > 0)
        hipLaunchKernelGGL(k_mapped_rowlist, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->ctx->stream,
                           p->dev.rowinfo, (long long)P.rowinfo_base, (long long)P.row_base, (long long)P.n_own,
                           (long long)grp0[r], (long long)P.onb_lo[0], (long long)P.onb_lo[1], (long long)P.onb_lo[2],
                           (long long)P.onb_n[0], (long long)P.onb_n[1], (long long)P.onb_n[2], M->rows, cnt);
    }
    unsigned long long hc[2] = {0, 0};
    ok = ok && hipGetLastError() == hipSuccess && hipStreamSynchronize(p->ctx->stream) == hipSuccess &&
         hipMemcpy(hc, cnt, sizeof(hc), hipMemcpyDeviceToHost) == hipSuccess;
    if (cnt) (void)hipFree(cnt);
    const bool bad = ok && (hc[1] != 0 || hc[0] != (unsigned long long)M->nrows);
    if (!ok || bad) {
      mapped_free(p);
      return bad ? GDX_ERR_STATE : GDX_ERR_ALLOC;
    }
  }
  if (hipMemcpy(M->vc, xyz, (size_t)nv * D * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) return GDX_ERR_HIP;
  return GDX_OK;
}

int mapped_assemble(gdx_problem *p, hipStream_t st, int stage) {
  MappedData *M = md(p);
  if (!M || !M->vc || !M->cm || !M->rows) return GDX_ERR_STATE;
  if (p->coef_cell && !p->dev.cellk) return GDX_ERR_STATE;
  if (!p->coef_cell && p->desc.nmodes > 0 && !p->dev.modes) return GDX_ERR_STATE;
  return launch_mapped(p, M, st, stage == 1);
}

}  // namespace gdx

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
// Kernel: each wave takes two owned row nodes (a node's 3 rows share one column
// set and one position pattern), nodes in the order of their rows in memory.
// All loads of both nodes are issued before the first store: the node's 1D
// factors (3 directions x 2K+1 offsets x {K,M,P,Q}, to LDS) and each lane's
// slice of the pattern (row positions e = lane, lane+64, ...), so no global
// load sits between the stores.  The pattern holds, per row position, the
// column's stencil offsets and component packed in bit fields (decoded once at
// setup), and a lane writes the 3 rows' values at e: every store instruction
// writes 64 consecutive doubles.  HBM traffic is the 8 B value writes (plus
// 32 B of row table per node); the arithmetic is ~10 FLOP per stored value.
//
// Cell-wise Lame parameters (a piecewise-constant material: (lambda_c, mu_c)
// per cell, or kappa_c scaling the constants, i.e. a Young's modulus field)
// break the sum over cells out of the factors: an entry is the sum over the
// cells c shared by its two nodes of the same expression with cell-local 1D
// factors (the integrals over the one 1D cell) and (lambda_c, mu_c).  Along
// each direction a node lies in one cell (a Q2 mid node) or two (a vertex, one
// at the mesh boundary), its "slots"; the cells of an entry are the slot
// combinations whose local factors at the column offset are nonzero.  Kernel
// k_elast_cw: the layout of k_elast, with the node's slot factors
// (3 directions x 2 slots x 2K+1 offsets x {K,M,P,Q}) and the pairs of its up
// to 8 cells in LDS, and the loop over slot combinations (1 to 8, uniform over
// the wave) inside each entry.  Same stores, ~4x the arithmetic of k_elast on
// Q2 (3.4 slot combinations per node on average).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gdx_internal.h"

namespace gdx {

namespace {

constexpr int kWaves = 4;         // waves per workgroup
constexpr int kNodesPerWave = 2;  // row nodes per wave, their loads issued together
                                  // (measured on 96x80x72 Q2: 1 node 5.28 ms, 2 nodes 5.14 ms,
                                  // 4 nodes spill registers, 6.56 ms)

// One owned row node, in the order of its rows in the values buffer.
struct ElastRow {
  long long rowstart;  // first value of the node's first row
  long long bidx;      // RHS index of the node's first row
  int pid;             // position pattern
  int len;             // row length (the node's 3 rows have the same)
  unsigned short g[3]; // global lattice coordinates
  unsigned short pad;
};
static_assert(sizeof(ElastRow) == 32, "ElastRow is 32 bytes");

// the owned row nodes of one part's owned node box, each at its owned row group
__global__ void k_elast_rows(const RowInfo *__restrict__ rowinfo, long long rinfo0, long long row_base, long long n_own,
                             long long grp0, long long lo0, long long lo1, long long lo2, long long n0, long long n1,
                             long long n2, ElastRow *__restrict__ rows, int *err) {
  const long long li = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (li >= n0 * n1 * n2) return;
  const RowInfo ri = rowinfo[rinfo0 + li];
  if (ri.pid < 0) return;
  const long long o = ri.bidx - row_base;  // the node's first owned row
  if (o < 0 || o >= n_own || o % 3) {
    atomicOr(err, 1);
    return;
  }
  ElastRow er{};
  er.rowstart = ri.rowstart;
  er.bidx = ri.bidx;
  er.pid = ri.pid;
  er.len = ri.lenf & 0xFFFF;
  er.g[0] = (unsigned short)(lo0 + li % n0);
  er.g[1] = (unsigned short)(lo1 + (li / n0) % n1);
  er.g[2] = (unsigned short)(lo2 + li / (n0 * n1));
  rows[grp0 + o / 3] = er;
}

struct ElastData {
  double *fac = nullptr;    // [3][nnmax][W][4]: K, M, P, Q at (node, offset + K)
  double *sfac = nullptr;   // [3][nnmax][2][W][4]: the same per slot (one cell), zero past the slots
  int *sinfo = nullptr;     // [3][nnmax]: first slot's cell * 4 + slot count (the second slot is the next cell)
  double *mass = nullptr;   // [3][nnmax]
  uint16_t *pinv = nullptr; // [npat][stride]: position -> o0 | o1 << 3 | o2 << 6 | component << 9
  ElastRow *rows = nullptr;
  int nnmax = 0, stride = 0;
  long long nrownodes = 0;
};

ElastData *ed(gdx_problem *p) { return reinterpret_cast<ElastData *>(p->elast); }

struct ElastArgs {
  double lam, mu, f[3];
  int nnmax, stride;
  long long nrownodes;
  int ncell[3];
};

template <int K, int NPW>
__global__ __launch_bounds__(64 * kWaves) void k_elast(ElastArgs A, const ElastRow *__restrict__ rows,
                                                       const double *__restrict__ fac, const double *__restrict__ mass,
                                                       const uint16_t *__restrict__ pinv, double *__restrict__ values,
                                                       double *__restrict__ b) {
  constexpr int W = 2 * K + 1;
  constexpr int NF = 3 * W * 4;                  // factors of one row node
  constexpr int MAXE = (3 * W * W * W + 63) / 64;  // row positions per lane
  __shared__ __align__(16) double lds[kWaves][NPW][NF];
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;  // wave-uniform
  const long long node0 = ((long long)blockIdx.x * kWaves + wave) * NPW;
  if (node0 >= A.nrownodes) return;  // whole waves exit together; no barrier below
  const int nn = (int)min((long long)NPW, A.nrownodes - node0);
  // all loads of the wave's NPW nodes are issued before any store
  ElastRow ri[NPW];
#pragma unroll
  for (int j = 0; j < NPW; ++j) ri[j] = rows[node0 + (j < nn ? j : 0)];
  double fv[NPW];
  int s3v[NPW][MAXE];
#pragma unroll
  for (int j = 0; j < NPW; ++j) {
    fv[j] = 0.0;
    if (lane < NF) {
      const int d = lane / (W * 4), r = lane % (W * 4);
      fv[j] = fac[((size_t)d * A.nnmax + ri[j].g[d]) * (W * 4) + r];
    }
    const uint16_t *pat = pinv + (size_t)ri[j].pid * A.stride;
#pragma unroll
    for (int k = 0; k < MAXE; ++k) {
      const int e = lane + 64 * k;
      s3v[j][k] = e < ri[j].len ? pat[e] : 0;
    }
  }
#pragma unroll
  for (int j = 0; j < NPW; ++j)
    if (j < nn && lane < 3) {
      const double m = mass[ri[j].g[0]] * mass[A.nnmax + ri[j].g[1]] * mass[2 * A.nnmax + ri[j].g[2]];
      b[ri[j].bidx + lane] = A.f[lane] * m;
    }
#pragma unroll
  for (int j = 0; j < NPW; ++j)
    if (lane < NF) lds[wave][j][lane] = fv[j];
  __builtin_amdgcn_wave_barrier();  // LDS written and read by the same wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const double lam = A.lam, mu = A.mu, lpm = A.lam + A.mu;
#pragma unroll
  for (int j = 0; j < NPW; ++j) {
    if (j >= nn) break;
    const double *F = lds[wave][j];
    const int len = ri[j].len;
    double *row0 = values + ri[j].rowstart;
#pragma unroll
    for (int k = 0; k < MAXE; ++k) {
      const int e = lane + 64 * k;
      if (e >= len) break;
      const int code = s3v[j][k];
      const int o0 = code & 7, o1 = (code >> 3) & 7, o2 = (code >> 6) & 7, bc = code >> 9;
      const double2 *f0 = reinterpret_cast<const double2 *>(F + (0 * W + o0) * 4);
      const double2 *f1 = reinterpret_cast<const double2 *>(F + (1 * W + o1) * 4);
      const double2 *f2 = reinterpret_cast<const double2 *>(F + (2 * W + o2) * 4);
      const double2 km0 = f0[0], pq0 = f0[1], km1 = f1[0], pq1 = f1[1], km2 = f2[0], pq2 = f2[1];
      // t_a = K_a times the M of the other two directions
      const double t0 = km0.x * (km1.y * km2.y), t1 = km1.x * (km0.y * km2.y), t2 = km2.x * (km0.y * km1.y);
      const double lap = mu * (t0 + t1 + t2);
      const bool b0 = bc == 0, b1 = bc == 1, b2 = bc == 2;
      const double Qb = b0 ? pq0.y : (b1 ? pq1.y : pq2.y), Pb = b0 ? pq0.x : (b1 ? pq1.x : pq2.x);
      // row component a: diagonal block when a == bc, else M of the third direction
      const double v0 = b0 ? fma(lpm, t0, lap) : (lam * pq0.x * Qb + mu * pq0.y * Pb) * (b1 ? km2.y : km1.y);
      const double v1 = b1 ? fma(lpm, t1, lap) : (lam * pq1.x * Qb + mu * pq1.y * Pb) * (b0 ? km2.y : km0.y);
      const double v2 = b2 ? fma(lpm, t2, lap) : (lam * pq2.x * Qb + mu * pq2.y * Pb) * (b0 ? km1.y : km0.y);
      row0[e] = v0;
      row0[(size_t)len + e] = v1;
      row0[(size_t)2 * len + e] = v2;
    }
  }
}

// Cell-wise Lame parameters.  CW = 1: coef[c] = kappa_c, (lambda, mu) * kappa_c;
// CW = 2: coef[2c], coef[2c + 1] = (lambda_c, mu_c).
template <int K, int NPW, int CW>
__global__ __launch_bounds__(64 * kWaves) void k_elast_cw(ElastArgs A, const ElastRow *__restrict__ rows,
                                                          const double *__restrict__ sfac, const int *__restrict__ sinfo,
                                                          const double *__restrict__ coef,
                                                          const double *__restrict__ mass,
                                                          const uint16_t *__restrict__ pinv,
                                                          double *__restrict__ values, double *__restrict__ b) {
  constexpr int W = 2 * K + 1;
  constexpr int NS = 2 * W * 4;                  // slot factors of one direction
  constexpr int NF = 3 * NS;                     // of one row node
  constexpr int NL = (NF + 63) / 64;             // loads per lane
  constexpr int MAXE = (3 * W * W * W + 63) / 64;  // row positions per lane
  __shared__ __align__(16) double lds[kWaves][NPW][NF + 16];  // factors, then (lambda, mu) of 8 cells
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;  // wave-uniform
  const long long node0 = ((long long)blockIdx.x * kWaves + wave) * NPW;
  if (node0 >= A.nrownodes) return;  // whole waves exit together; no barrier below
  const int nn = (int)min((long long)NPW, A.nrownodes - node0);
  ElastRow ri[NPW];
#pragma unroll
  for (int j = 0; j < NPW; ++j) ri[j] = rows[node0 + (j < nn ? j : 0)];
  int si[NPW][3];
  double fv[NPW][NL], cv[NPW][2];
  int s3v[NPW][MAXE];
#pragma unroll
  for (int j = 0; j < NPW; ++j) {
#pragma unroll
    for (int d = 0; d < 3; ++d) si[j][d] = sinfo[d * A.nnmax + ri[j].g[d]];
#pragma unroll
    for (int l = 0; l < NL; ++l) {
      const int r = lane + 64 * l, d = r / NS;
      const int gd = d == 0 ? ri[j].g[0] : (d == 1 ? ri[j].g[1] : ri[j].g[2]);
      fv[j][l] = r < NF ? sfac[((size_t)d * A.nnmax + gd) * NS + r % NS] : 0.0;
    }
    // lane = s0 + 2 s1 + 4 s2 < 8: the cell of that slot combination (a slot
    // past the node's count reads the first slot's cell; its factors are zero)
    cv[j][0] = cv[j][1] = 0.0;
    if (lane < 8) {
      long long c = 0;
#pragma unroll
      for (int d = 2; d >= 0; --d) {
        const int sd = (lane >> d) & 1, ns = si[j][d] & 3;
        c = c * A.ncell[d] + (si[j][d] >> 2) + (sd < ns ? sd : 0);
      }
      if (CW == 1) {
        const double k = coef[c];
        cv[j][0] = A.lam * k;
        cv[j][1] = A.mu * k;
      } else {
        const double2 lm = reinterpret_cast<const double2 *>(coef)[c];
        cv[j][0] = lm.x;
        cv[j][1] = lm.y;
      }
    }
    const uint16_t *pat = pinv + (size_t)ri[j].pid * A.stride;
#pragma unroll
    for (int k = 0; k < MAXE; ++k) {
      const int e = lane + 64 * k;
      s3v[j][k] = e < ri[j].len ? pat[e] : 0;
    }
  }
#pragma unroll
  for (int j = 0; j < NPW; ++j)
    if (j < nn && lane < 3) {
      const double m = mass[ri[j].g[0]] * mass[A.nnmax + ri[j].g[1]] * mass[2 * A.nnmax + ri[j].g[2]];
      b[ri[j].bidx + lane] = A.f[lane] * m;
    }
#pragma unroll
  for (int j = 0; j < NPW; ++j) {
#pragma unroll
    for (int l = 0; l < NL; ++l)
      if (lane + 64 * l < NF) lds[wave][j][lane + 64 * l] = fv[j][l];
    if (lane < 8) {
      lds[wave][j][NF + 2 * lane] = cv[j][0];
      lds[wave][j][NF + 2 * lane + 1] = cv[j][1];
    }
  }
  __builtin_amdgcn_wave_barrier();  // LDS written and read by the same wave
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
  for (int j = 0; j < NPW; ++j) {
    if (j >= nn) break;
    const double *F = lds[wave][j];
    const double *LM = F + NF;
    const int len = ri[j].len;
    const int ns0 = si[j][0] & 3, ns1 = si[j][1] & 3, ns2 = si[j][2] & 3;
    double *row0 = values + ri[j].rowstart;
#pragma unroll
    for (int k = 0; k < MAXE; ++k) {
      const int e = lane + 64 * k;
      if (e >= len) break;
      const int code = s3v[j][k];
      const int o0 = code & 7, o1 = (code >> 3) & 7, o2 = (code >> 6) & 7, bc = code >> 9;
      const bool b0 = bc == 0, b1 = bc == 1, b2 = bc == 2;
      // a node in two cells along d (a vertex) meets a column at offset < 0
      // only in the first, > 0 only in the second: the slots with nonzero
      // factors for this entry
      const int lo0 = ns0 == 2 && o0 > K, hi0 = ns0 == 2 && o0 >= K;
      const int lo1 = ns1 == 2 && o1 > K, hi1 = ns1 == 2 && o1 >= K;
      const int lo2 = ns2 == 2 && o2 > K, hi2 = ns2 == 2 && o2 >= K;
      double v0 = 0.0, v1 = 0.0, v2 = 0.0;
      for (int s2 = lo2; s2 <= hi2; ++s2) {
        const double2 *f2 = reinterpret_cast<const double2 *>(F + 2 * NS + (s2 * W + o2) * 4);
        const double2 km2 = f2[0], pq2 = f2[1];
        for (int s1 = lo1; s1 <= hi1; ++s1) {
          const double2 *f1 = reinterpret_cast<const double2 *>(F + NS + (s1 * W + o1) * 4);
          const double2 km1 = f1[0], pq1 = f1[1];
          for (int s0 = lo0; s0 <= hi0; ++s0) {
            const double2 *f0 = reinterpret_cast<const double2 *>(F + (s0 * W + o0) * 4);
            const double2 km0 = f0[0], pq0 = f0[1];
            const double2 lm = reinterpret_cast<const double2 *>(LM)[s0 + 2 * s1 + 4 * s2];
            const double lam = lm.x, mu = lm.y, lpm = lm.x + lm.y;
            const double t0 = km0.x * (km1.y * km2.y), t1 = km1.x * (km0.y * km2.y), t2 = km2.x * (km0.y * km1.y);
            const double lap = mu * (t0 + t1 + t2);
            const double Qb = b0 ? pq0.y : (b1 ? pq1.y : pq2.y), Pb = b0 ? pq0.x : (b1 ? pq1.x : pq2.x);
            v0 += b0 ? fma(lpm, t0, lap) : (lam * pq0.x * Qb + mu * pq0.y * Pb) * (b1 ? km2.y : km1.y);
            v1 += b1 ? fma(lpm, t1, lap) : (lam * pq1.x * Qb + mu * pq1.y * Pb) * (b0 ? km2.y : km0.y);
            v2 += b2 ? fma(lpm, t2, lap) : (lam * pq2.x * Qb + mu * pq2.y * Pb) * (b0 ? km1.y : km0.y);
          }
        }
      }
      row0[e] = v0;
      row0[(size_t)len + e] = v1;
      row0[(size_t)2 * len + e] = v2;
    }
  }
}

}  // namespace

void elast_free(gdx_problem *p) {
  ElastData *E = ed(p);
  if (!E) return;
  void *ptrs[] = {E->fac, E->sfac, E->sinfo, E->mass, E->pinv, E->rows};
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
  if (nnmax > 65535) return GDX_ERR_UNSUPPORTED;  // 16-bit node coordinates in ElastRow
  std::vector<double> fac((size_t)3 * nnmax * W * 4, 0.0), mass((size_t)3 * nnmax, 0.0);
  std::vector<double> sfac((size_t)3 * nnmax * 2 * W * 4, 0.0);
  std::vector<int> sinfo((size_t)3 * nnmax, 0);
  for (int d = 0; d < 3; ++d) {
    const double h = kp.h[d];
    // one cell's K, M, P, Q at local nodes (a, bb)
    double loc[3][3][4] = {};
    for (int a = 0; a <= K; ++a)
      for (int bb = 0; bb <= K; ++bb)
        for (int q = 0; q < NQ; ++q) {
          loc[a][bb][0] += kp.qw[q] * kp.dphi[a][q] * kp.dphi[bb][q] / h;
          loc[a][bb][1] += kp.qw[q] * h * kp.phi[a][q] * kp.phi[bb][q];
          loc[a][bb][2] += kp.qw[q] * kp.dphi[a][q] * kp.phi[bb][q];
          loc[a][bb][3] += kp.qw[q] * kp.phi[a][q] * kp.dphi[bb][q];
        }
    // slots of node i: the cells holding it, in increasing order, with its local node
    for (int i = 0; i < kp.nn[d]; ++i) {
      int cell[2], al[2], ns = 0;
      if (i % K == 0) {
        if (i / K - 1 >= 0) cell[ns] = i / K - 1, al[ns++] = K;
        if (i / K < kp.ncells[d]) cell[ns] = i / K, al[ns++] = 0;
      } else {
        cell[ns] = i / K, al[ns++] = i % K;
      }
      if (ns == 0) return GDX_ERR_STATE;
      sinfo[(size_t)d * nnmax + i] = cell[0] * 4 + ns;
      for (int sl = 0; sl < ns; ++sl)
        for (int o = -K; o <= K; ++o) {
          const int bb = al[sl] + o;
          if (bb < 0 || bb > K) continue;
          double *f = &sfac[((((size_t)d * nnmax + i) * 2 + sl) * W + (o + K)) * 4];
          for (int t = 0; t < 4; ++t) f[t] = loc[al[sl]][bb][t];
        }
    }
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
  // owned row nodes in the order of their rows in the values buffer, so that
  // consecutive waves write consecutive memory: a node's place is its owned
  // row group (parts in order, then owned rows, whose offsets increase)
  std::vector<int64_t> grp0(p->nparts + 1, 0);
  for (int r = 0; r < p->nparts; ++r) grp0[r + 1] = grp0[r] + p->P[r].n_own / 3;
  // the row table is in HBM after the device symbolic phase: the list is made there
  std::vector<ElastRow> rows(p->dev_structure ? 0 : grp0[p->nparts]);
  for (int r = 0; r < p->nparts && !p->dev_structure; ++r) {
    const Part &P = p->P[r];
    const int64_t n = P.onb_n[0] * P.onb_n[1] * P.onb_n[2];
    for (int64_t li = 0; li < n; ++li) {
      const RowInfo &ri = p->rowinfo[P.rowinfo_base + li];
      if (ri.pid < 0) continue;
      ElastRow er{};
      er.rowstart = ri.rowstart;
      er.bidx = ri.bidx;
      er.pid = ri.pid;
      er.len = ri.lenf & 0xFFFF;
      er.g[0] = (unsigned short)(P.onb_lo[0] + li % P.onb_n[0]);
      er.g[1] = (unsigned short)(P.onb_lo[1] + (li / P.onb_n[0]) % P.onb_n[1]);
      er.g[2] = (unsigned short)(P.onb_lo[2] + li / (P.onb_n[0] * P.onb_n[1]));
      const int64_t o = ri.bidx - P.row_base;  // the node's first owned row
      if (o < 0 || o >= P.n_own || o % 3) return GDX_ERR_STATE;
      rows[grp0[r] + o / 3] = er;
    }
  }
  // position -> packed (stencil offsets, column component): slot * 3 + component
  // with slot = o0 + W (o1 + W o2); padding entries stay as they are
  std::vector<uint16_t> code(p->pat16_inv.size());
  for (size_t i = 0; i < code.size(); ++i) {
    const int s3 = p->pat16_inv[i], sl = s3 / 3, bc = s3 % 3;
    const int o0 = sl % W, o1 = (sl / W) % W, o2 = sl / (W * W);
    code[i] = (uint16_t)(o0 | o1 << 3 | o2 << 6 | bc << 9);
  }
  ElastData *E = new ElastData;
  p->elast = E;
  E->nnmax = nnmax;
  E->stride = p->pat16_stride;
  E->nrownodes = (long long)grp0[p->nparts];
  auto up = [](void **dst, const void *src, size_t bytes) {
    if (dev_malloc(dst, std::max<size_t>(bytes, 8)) != hipSuccess) return false;
    return bytes == 0 || hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice) == hipSuccess;
  };
  bool ok = up(reinterpret_cast<void **>(&E->fac), fac.data(), fac.size() * sizeof(double)) &&
            up(reinterpret_cast<void **>(&E->sfac), sfac.data(), sfac.size() * sizeof(double)) &&
            up(reinterpret_cast<void **>(&E->sinfo), sinfo.data(), sinfo.size() * sizeof(int)) &&
            up(reinterpret_cast<void **>(&E->mass), mass.data(), mass.size() * sizeof(double)) &&
            up(reinterpret_cast<void **>(&E->pinv), code.data(), code.size() * sizeof(uint16_t)) &&
            up(reinterpret_cast<void **>(&E->rows), rows.data(), rows.size() * sizeof(ElastRow));
  if (ok && p->dev_structure) {
    (void)hipFree(E->rows);
    E->rows = nullptr;
    int *err = nullptr;
    ok = dev_malloc(reinterpret_cast<void **>(&E->rows), std::max<int64_t>(grp0[p->nparts], 1) * sizeof(ElastRow)) ==
             hipSuccess &&
         dev_malloc(reinterpret_cast<void **>(&err), sizeof(int)) == hipSuccess &&
         hipMemsetAsync(err, 0, sizeof(int), p->ctx->stream) == hipSuccess;
    for (int r = 0; r < p->nparts && ok; ++r) {
      const Part &P = p->P[r];
      const long long n = P.onb_n[0] * P.onb_n[1] * P.onb_n[2];
      if (n > 0)
        hipLaunchKernelGGL(k_elast_rows, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, p->ctx->stream,
                           p->dev.rowinfo, (long long)P.rowinfo_base, (long long)P.row_base, (long long)P.n_own,
                           (long long)grp0[r], (long long)P.onb_lo[0], (long long)P.onb_lo[1], (long long)P.onb_lo[2],
                           (long long)P.onb_n[0], (long long)P.onb_n[1], (long long)P.onb_n[2], E->rows, err);
    }
    int herr = 0;
    ok = ok && hipGetLastError() == hipSuccess && hipStreamSynchronize(p->ctx->stream) == hipSuccess &&
         hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess;
    if (err) (void)hipFree(err);
    if (ok && herr) {
      elast_free(p);
      return GDX_ERR_STATE;
    }
  }
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
  A.nrownodes = E->nrownodes;
  for (int d = 0; d < 3; ++d) A.ncell[d] = (int)p->kp.ncells[d];
  const long long per = (long long)kWaves * kNodesPerWave;
  const long long blocks = (E->nrownodes + per - 1) / per;
  if (blocks > 0x7FFFFFFFLL) return GDX_ERR_UNSUPPORTED;
  const dim3 grid((unsigned)blocks), block(64 * kWaves);
  if (p->coef_cell || p->lame_cell) {
    const double *coef = p->coef_cell ? p->dev.cellk : p->dev.cell_lame;
    if (!coef) return GDX_ERR_STATE;
#define GDX_ELAST_CW(KK, CW)                                                                                       \
  hipLaunchKernelGGL((k_elast_cw<KK, kNodesPerWave, CW>), grid, block, 0, st, A, E->rows, E->sfac, E->sinfo, coef, \
                     E->mass, E->pinv, p->dev.values, p->dev.b)
    if (p->K == 2 && p->coef_cell) GDX_ELAST_CW(2, 1);
    else if (p->K == 2) GDX_ELAST_CW(2, 2);
    else if (p->coef_cell) GDX_ELAST_CW(1, 1);
    else GDX_ELAST_CW(1, 2);
#undef GDX_ELAST_CW
    return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
  }
  if (p->K == 2)
    hipLaunchKernelGGL((k_elast<2, kNodesPerWave>), grid, block, 0, st, A, E->rows, E->fac, E->mass, E->pinv,
                       p->dev.values, p->dev.b);
  else
    hipLaunchKernelGGL((k_elast<1, kNodesPerWave>), grid, block, 0, st, A, E->rows, E->fac, E->mass, E->pinv,
                       p->dev.values, p->dev.b);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

}  // namespace gdx

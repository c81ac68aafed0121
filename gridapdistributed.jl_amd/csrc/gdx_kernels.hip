// HIP kernels of libgdx for gfx950 (MI355X).
//
// The numeric assembly of a Lagrangian Q_K Poisson operator on a Cartesian
// mesh is done "owner computes rows": every workgroup owns a tile of owned
// rows (free dofs) of one part and computes their complete rows -- including
// the contributions of cells that belong to neighbouring parts -- with the
// tensor-product structure of the mesh, the basis and the Gauss rule:
//
//   A_ij = sum_q kappa(x_q) * [ G1 M2 M3 + M1 G2 M3 + M1 M2 G3 ](q; i,j)
//   G_d(q; i_d,j_d) = w_q dphi_a(q) dphi_b(q) / h_d
//   M_d(q; i_d,j_d) = w_q h_d  phi_a(q)  phi_b(q)
//
// which is exactly sum over cells c containing i and j of the cell matrix
// entry int_c kappa grad(phi_i).grad(phi_j) for the affine Cartesian map
// (J = diag(h), detJ = h1 h2 h3, J^{-T} grad = grad_ref / h).  The three
// contractions over x-, y- and z-quadrature points are done one dimension at
// a time in LDS (sum factorisation), so each stored value costs O(1) FMAs.
// Rows are staged in LDS and written once, contiguously, through a position
// pattern (stencil slot -> offset within the CSR row), so the scatter is a
// plain coalesced store: no float atomics, bit-for-bit repeatable.  The
// SubAssembledRows owner/ghost exchange is folded in: owned rows receive the
// neighbouring parts' contributions at the source, ghost rows stay zero.
#include <hip/hip_runtime.h>

#include "gdx_internal.h"

namespace gdx {

// ---------------------------------------------------------------------------
// Coefficient and source at every quadrature point of the global lattice.
// exp(i theta_m) = exp(i phi_m) * prod_d exp(2 pi i k_md x_d) with 1D tables.
// ---------------------------------------------------------------------------
__global__ void k_fourier_tables(const double *__restrict__ modes, KParams kp, int qnmax, double *__restrict__ E) {
  const int m = blockIdx.y, d = blockIdx.z;
  for (int q = threadIdx.x + blockIdx.x * blockDim.x; q < qnmax; q += blockDim.x * gridDim.x) {
    double x = 0.0;
    if (d < kp.dim && q < kp.qn[d]) {
      const int c = q / kp.NQ, t = q % kp.NQ;
      x = kp.x0[d] + kp.h[d] * ((double)c + kp.qx[t]);
    }
    double s, co;
    sincos(2.0 * M_PI * modes[5 * m + d] * x, &s, &co);
    double *e = E + 2 * (((size_t)d * kp.nmodes + m) * qnmax + q);
    e[0] = co;
    e[1] = s;
  }
}

__global__ void k_coefficients(const double *__restrict__ modes, const double *__restrict__ E, KParams kp, int qnmax,
                               double *__restrict__ W, double *__restrict__ F) {
  const int q1 = blockIdx.x * blockDim.x + threadIdx.x;
  const int q2 = blockIdx.y, q3 = blockIdx.z;
  if (q1 >= kp.qn[0]) return;
  double g = 0.0, s = 0.0;
  for (int m = 0; m < kp.nmodes; ++m) {
    const double *M = modes + 5 * m;
    const double *e1 = E + 2 * ((size_t)(0 * kp.nmodes + m) * qnmax + q1);
    const double *e2 = E + 2 * ((size_t)(1 * kp.nmodes + m) * qnmax + q2);
    const double *e3 = E + 2 * ((size_t)(2 * kp.nmodes + m) * qnmax + q3);
    double cr, ci;
    sincos(M[4], &ci, &cr);
    cr *= M[3];
    ci *= M[3];
    // z = c * e1 * e2 * e3
    double ar = cr * e1[0] - ci * e1[1], ai = cr * e1[1] + ci * e1[0];
    double br = ar * e2[0] - ai * e2[1], bi = ar * e2[1] + ai * e2[0];
    double zr = br * e3[0] - bi * e3[1], zi = br * e3[1] + bi * e3[0];
    g += zr;
    s += (M[0] * kp.dir_coef[1] + M[1] * kp.dir_coef[2] + M[2] * kp.dir_coef[3]) * zi;
  }
  const double kap = kp.coef_const * exp(g);
  const size_t idx = q1 + (size_t)kp.qn[0] * (q2 + (size_t)kp.qn[1] * q3);
  W[idx] = kap;
  F[idx] = kp.rhs_kind == GDX_RHS_MANUFACTURED ? kap * 2.0 * M_PI * s : kp.rhs_const;
}

// ---------------------------------------------------------------------------
// RHS f-part: b_i = sum_{cells c ni i} sum_q f(x_q) w_q detJ phi_i(x_q)
// ---------------------------------------------------------------------------
__global__ void k_rhs(const Tile *__restrict__ tiles, const RowInfo *__restrict__ rowinfo, const double *__restrict__ F,
                      KParams kp, int R0, int R1, int R2, double *__restrict__ b) {
  const Tile t = tiles[blockIdx.x];
  const int K = kp.K, NQ = kp.NQ;
  for (int r = threadIdx.x; r < R0 * R1 * R2; r += blockDim.x) {
    const int u[3] = {r % R0, (r / R0) % R1, r / (R0 * R1)};
    if (u[0] >= t.ext[0] || u[1] >= t.ext[1] || u[2] >= t.ext[2]) continue;
    const int g[3] = {t.org[0] + u[0], t.org[1] + u[1], t.org[2] + u[2]};
    const long long li = ((long long)(g[2] - t.onb_lo[2]) * t.onb_n[1] + (g[1] - t.onb_lo[1])) * t.onb_n[0] +
                         (g[0] - t.onb_lo[0]);
    const RowInfo ri = rowinfo[t.rowinfo_base + li];
    if (ri.pid < 0) continue;
    // per dimension: the (<=2) cells containing node g and its local index
    int cl[3][2], al[3][2], nc[3];
    for (int d = 0; d < 3; ++d) {
      nc[d] = 0;
      if (d >= kp.dim) {
        cl[d][0] = 0;
        al[d][0] = 0;
        nc[d] = 1;
        continue;
      }
      const int cmin = (g[d] + K - 1) / K - 1, cmax = g[d] / K;
      for (int c = cmin; c <= cmax; ++c)
        if (c >= 0 && c < kp.ncells[d]) {
          cl[d][nc[d]] = c;
          al[d][nc[d]] = g[d] - K * c;
          nc[d]++;
        }
    }
    double acc = 0.0;
    for (int c3 = 0; c3 < nc[2]; ++c3)
      for (int t3 = 0; t3 < (kp.dim == 3 ? NQ : 1); ++t3) {
        const double f3 = kp.dim == 3 ? kp.qw[t3] * kp.h[2] * kp.phi[al[2][c3]][t3] : 1.0;
        const int q3 = kp.dim == 3 ? cl[2][c3] * NQ + t3 : 0;
        for (int c2 = 0; c2 < nc[1]; ++c2)
          for (int t2 = 0; t2 < NQ; ++t2) {
            const double f2 = f3 * kp.qw[t2] * kp.h[1] * kp.phi[al[1][c2]][t2];
            const int q2 = cl[1][c2] * NQ + t2;
            const double *Fr = F + (size_t)kp.qn[0] * (q2 + (size_t)kp.qn[1] * q3);
            for (int c1 = 0; c1 < nc[0]; ++c1)
              for (int t1 = 0; t1 < NQ; ++t1)
                acc += Fr[cl[0][c1] * NQ + t1] * f2 * kp.qw[t1] * kp.h[0] * kp.phi[al[0][c1]][t1];
          }
      }
    b[ri.bidx] = acc;
  }
}

// ---------------------------------------------------------------------------
// 1D pair structure of a tile.  Region node v = u + K (the region starts one
// cell below the tile).  A pair (v, v+s) is supported on cells lc0 [, lc0+1].
// ---------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void pair_natural(int idx, int &u, int &s) {
  if (K == 1) {
    u = idx / 3;
    s = idx % 3 - 1;
  } else {
    const int blk = idx >> 3, r = idx & 7;
    if (r < 5) {
      u = 2 * blk;
      s = r - 2;
    } else {
      u = 2 * blk + 1;
      s = r - 6;
    }
  }
}

// pairs supported on two cells first (keeps the support length wave-uniform)
template <int K, int R>
__device__ __forceinline__ void pair_doubles_first(int idx, int &u, int &s) {
  constexpr int ND = (K == 1) ? R : R / 2;
  if (idx < ND) {
    u = K * idx;
    s = 0;
    return;
  }
  const int j = idx - ND;
  if (K == 1) {
    u = j >> 1;
    s = (j & 1) ? 1 : -1;
  } else {
    const int blk = j / 7, r = j % 7;
    if (r < 4) {
      u = 2 * blk;
      s = r < 2 ? r - 2 : r - 1;
    } else {
      u = 2 * blk + 1;
      s = r - 5;
    }
  }
}

template <int K>
__host__ __device__ constexpr int pair_count(int R) {
  return K == 1 ? 3 * R : 4 * R;
}

// first supporting cell, number of cells, local indices in the first cell
template <int K>
__device__ __forceinline__ void pair_cells(int u, int s, int &lc0, int &ncell, int &a0, int &b0) {
  const int v = u + K, w = v + s;
  const int hi = v > w ? v : w, lo = v < w ? v : w;
  const int cmin = (hi + K - 1) / K - 1, cmax = lo / K;
  lc0 = cmin;
  ncell = cmax - cmin + 1;
  a0 = v - K * cmin;
  b0 = w - K * cmin;
}

// factor of the pair at support point t (cell t/NQ of the support, Gauss point t%NQ)
template <int K, int NQ>
__device__ __forceinline__ void pair_factors(const KParams &kp, int d, int a0, int b0, int t, double &G, double &M) {
  const int ci = t / NQ, q = t % NQ;
  const int a = a0 - K * ci, b = b0 - K * ci;
  const double w = kp.qw[q], h = kp.h[d];
  G = w * kp.dphi[a][q] * kp.dphi[b][q] / h;
  M = w * h * kp.phi[a][q] * kp.phi[b][q];
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double uD_at(const KParams &kp, int g0, int g1, int g2) {
  const double x = kp.x0[0] + kp.h[0] * (double)g0 / kp.K;
  const double y = kp.x0[1] + kp.h[1] * (double)g1 / kp.K;
  const double z = kp.dim == 3 ? kp.x0[2] + kp.h[2] * (double)g2 / kp.K : 0.0;
  return kp.dir_coef[0] + kp.dir_coef[1] * x + kp.dir_coef[2] * y + kp.dir_coef[3] * z;
}

// Stage the rows of one slab (fixed tile-local z row u3) from LDS to HBM:
// contiguous stores in CSR order, plus the Dirichlet lifting of the RHS.
template <int K, int DIM, int R1, int R2, int NT>
__device__ __forceinline__ void write_slab(const Tile &t, int u3, const double *__restrict__ slab,
                                           const RowInfo *__restrict__ rowinfo, const uint8_t *__restrict__ pat_fwd,
                                           const uint8_t *__restrict__ pat_inv, const int *__restrict__ pat_len,
                                           const int *__restrict__ pat_dir, double *__restrict__ values,
                                           double *__restrict__ b, const KParams &kp) {
  constexpr int W = 2 * K + 1;
  constexpr int S = DIM == 3 ? W * W * W : W * W;
  constexpr int NW = NT / 64;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int rr = wave; rr < R1 * R2; rr += NW) {
    const int u1 = rr % R1, u2 = rr / R1;
    if (u1 >= t.ext[0] || u2 >= t.ext[1] || u3 >= t.ext[2]) continue;
    const int g0 = t.org[0] + u1, g1 = t.org[1] + u2, g2 = t.org[2] + u3;
    const long long li =
        ((long long)(g2 - t.onb_lo[2]) * t.onb_n[1] + (g1 - t.onb_lo[1])) * t.onb_n[0] + (g0 - t.onb_lo[0]);
    const RowInfo ri = rowinfo[t.rowinfo_base + li];
    if (ri.pid < 0) continue;
    const double *row = slab + rr * S;
    const uint8_t *inv = pat_inv + (size_t)ri.pid * S;
    const int len = pat_len[ri.pid];
    double *dst = values + ri.rowstart;
    for (int k = lane; k < len; k += 64) dst[k] = row[inv[k]];
    if (pat_dir[ri.pid]) {
      const uint8_t *fwd = pat_fwd + (size_t)ri.pid * S;
      double acc = 0.0;
      for (int sl = lane; sl < S; sl += 64) {
        if (fwd[sl] == kSlotDirichlet) {
          const int s1 = sl % W - K, s2 = (sl / W) % W - K, s3 = DIM == 3 ? sl / (W * W) - K : 0;
          acc += row[sl] * uD_at(kp, g0 + s1, g1 + s2, g2 + s3);
        }
      }
      acc = wave_sum(acc);
      if (lane == 0) b[ri.bidx] -= acc;
    }
  }
}

// ---------------------------------------------------------------------------
// 3D matrix kernel: one workgroup = one tile of R1 x R2 x R3 owned rows.
// ---------------------------------------------------------------------------
template <int K, int NQ, int R1, int R2, int R3>
struct Tile3 {
  static constexpr int C1 = R1 / K + 1, C2 = R2 / K + 1, C3 = R3 / K + 1;
  static constexpr int Q1 = C1 * NQ, Q2 = C2 * NQ, Q3 = C3 * NQ;
  static constexpr int NP1 = pair_count<K>(R1), NP2 = pair_count<K>(R2), NP3 = pair_count<K>(R3);
  static constexpr int XS = (Q2 * Q3) | 1;  // odd stride in doubles: conflict-free lanes
  static constexpr int NT = NP1 * NP2;
  static constexpr int W = 2 * K + 1, S = W * W * W;
  static constexpr int LDS_KAPPA = Q1 * Q2 * Q3;
  static constexpr int LDS_X = 2 * NP1 * XS;
  static constexpr int LDS_SLAB = R1 * R2 * S;
  // the output slab reuses the X buffer once stage 2 has consumed it
  static constexpr int LDS_DOUBLES = LDS_KAPPA + (LDS_X > LDS_SLAB ? LDS_X : LDS_SLAB);
};

template <int K, int NQ, int R1, int R2, int R3>
__global__ __launch_bounds__((Tile3<K, NQ, R1, R2, R3>::NT)) void k_assemble3d(
    const Tile *__restrict__ tiles, const RowInfo *__restrict__ rowinfo, const double *__restrict__ Wq,
    const uint8_t *__restrict__ pat_fwd, const uint8_t *__restrict__ pat_inv, const int *__restrict__ pat_len,
    const int *__restrict__ pat_dir, double *__restrict__ values, double *__restrict__ b, KParams kp) {
  using T = Tile3<K, NQ, R1, R2, R3>;
  __shared__ double lds[T::LDS_DOUBLES];
  double *kap = lds;                  // [Q3][Q2][Q1]
  double *XG = lds + T::LDS_KAPPA;    // [NP1][XS]
  double *XM = XG + T::NP1 * T::XS;   // [NP1][XS]
  double *slab = XG;                  // reused after stage 2: [R2][R1][S]
  const Tile t = tiles[blockIdx.x];
  const int tid = threadIdx.x;

  // region of quadrature points: cells org/K - 1 .. org/K + R/K, zero outside the mesh
  const int qs0 = (t.org[0] / K - 1) * NQ, qs1 = (t.org[1] / K - 1) * NQ, qs2 = (t.org[2] / K - 1) * NQ;
  for (int i = tid; i < T::LDS_KAPPA; i += T::NT) {
    const int a = i % T::Q1, bq = (i / T::Q1) % T::Q2, c = i / (T::Q1 * T::Q2);
    const int q0 = qs0 + a, q1 = qs1 + bq, q2 = qs2 + c;
    double v = 0.0;
    if (q0 >= 0 && q0 < kp.qn[0] && q1 >= 0 && q1 < kp.qn[1] && q2 >= 0 && q2 < kp.qn[2])
      v = Wq[q0 + (size_t)kp.qn[0] * (q1 + (size_t)kp.qn[1] * q2)];
    kap[i] = v;
  }
  __syncthreads();

  // stage 1: contract x quadrature points for every x pair
  {
    const int wave = tid >> 6, lane = tid & 63;
    constexpr int NW = T::NT / 64;
    for (int p1 = wave; p1 < T::NP1; p1 += NW) {
      int u, s, lc0, nc, a0, b0;
      pair_natural<K>(p1, u, s);
      pair_cells<K>(u, s, lc0, nc, a0, b0);
      double G[2 * NQ], M[2 * NQ];
#pragma unroll
      for (int tt = 0; tt < 2 * NQ; ++tt) pair_factors<K, NQ>(kp, 0, a0, b0, tt, G[tt], M[tt]);
      for (int it = lane; it < T::Q2 * T::Q3; it += 64) {
        const double *kr = kap + it * T::Q1 + lc0 * NQ;  // it = q3*Q2 + q2
        double xg = 0.0, xm = 0.0;
        if (nc == 2) {
#pragma unroll
          for (int tt = 0; tt < 2 * NQ; ++tt) {
            xg = fma(kr[tt], G[tt], xg);
            xm = fma(kr[tt], M[tt], xm);
          }
        } else {
#pragma unroll
          for (int tt = 0; tt < NQ; ++tt) {
            xg = fma(kr[tt], G[tt], xg);
            xm = fma(kr[tt], M[tt], xm);
          }
        }
        XG[p1 * T::XS + it] = xg;
        XM[p1 * T::XS + it] = xm;
      }
    }
  }
  __syncthreads();

  // stage 2: thread (p1, p2) contracts y quadrature points for all z points
  const int p1 = tid % T::NP1, p2 = tid / T::NP1;
  int u1, s1, u2, s2;
  pair_natural<K>(p1, u1, s1);
  pair_doubles_first<K, R2>(p2, u2, s2);
  double Y1[T::Q3], Y2[T::Q3];
  {
    int lc0, nc, a0, b0;
    pair_cells<K>(u2, s2, lc0, nc, a0, b0);
    double G[2 * NQ], M[2 * NQ];
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt) pair_factors<K, NQ>(kp, 1, a0, b0, tt, G[tt], M[tt]);
    const double *xg = XG + p1 * T::XS + lc0 * NQ;
    const double *xm = XM + p1 * T::XS + lc0 * NQ;
    if (nc == 2) {
#pragma unroll
      for (int q3 = 0; q3 < T::Q3; ++q3) {
        double y1 = 0.0, y2 = 0.0;
#pragma unroll
        for (int tt = 0; tt < 2 * NQ; ++tt) {
          const double g = xg[q3 * T::Q2 + tt], m = xm[q3 * T::Q2 + tt];
          y1 = fma(g, M[tt], fma(m, G[tt], y1));
          y2 = fma(m, M[tt], y2);
        }
        Y1[q3] = y1;
        Y2[q3] = y2;
      }
    } else {
#pragma unroll
      for (int q3 = 0; q3 < T::Q3; ++q3) {
        double y1 = 0.0, y2 = 0.0;
#pragma unroll
        for (int tt = 0; tt < NQ; ++tt) {
          const double g = xg[q3 * T::Q2 + tt], m = xm[q3 * T::Q2 + tt];
          y1 = fma(g, M[tt], fma(m, G[tt], y1));
          y2 = fma(m, M[tt], y2);
        }
        Y1[q3] = y1;
        Y2[q3] = y2;
      }
    }
  }
  __syncthreads();  // X is dead: its LDS becomes the output slab

  // stage 3: contract z quadrature points, one z row of the tile at a time
  const int rowslot = (u2 * R1 + u1) * T::S + (s2 + K) * T::W + (s1 + K);
#pragma unroll
  for (int u3 = 0; u3 < R3; ++u3) {
#pragma unroll
    for (int s3 = -K; s3 <= K; ++s3) {
      if (K == 2 && (u3 & 1) && (s3 == -2 || s3 == 2)) continue;
      int lc0, nc, a0, b0;
      pair_cells<K>(u3, s3, lc0, nc, a0, b0);
      double v = 0.0;
#pragma unroll
      for (int tt = 0; tt < 2 * NQ; ++tt) {
        if (tt >= nc * NQ) break;
        double G, M;
        pair_factors<K, NQ>(kp, 2, a0, b0, tt, G, M);
        v = fma(Y1[lc0 * NQ + tt], M, fma(Y2[lc0 * NQ + tt], G, v));
      }
      slab[rowslot + (s3 + K) * T::W * T::W] = v;
    }
    __syncthreads();
    write_slab<K, 3, R1, R2, T::NT>(t, u3, slab, rowinfo, pat_fwd, pat_inv, pat_len, pat_dir, values, b, kp);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// 2D matrix kernel: one workgroup = one tile of R1 x R2 owned rows.
// ---------------------------------------------------------------------------
template <int K, int NQ, int R1, int R2>
struct Tile2 {
  static constexpr int C1 = R1 / K + 1, C2 = R2 / K + 1;
  static constexpr int Q1 = C1 * NQ, Q2 = C2 * NQ;
  static constexpr int NP1 = pair_count<K>(R1), NP2 = pair_count<K>(R2);
  static constexpr int XS = Q2 | 1;
  static constexpr int NT = 256;
  static constexpr int W = 2 * K + 1, S = W * W;
  static constexpr int LDS_KAPPA = Q1 * Q2, LDS_X = 2 * NP1 * XS, LDS_SLAB = R1 * R2 * S;
  static constexpr int LDS_DOUBLES = LDS_KAPPA + LDS_X + LDS_SLAB;
};

template <int K, int NQ, int R1, int R2>
__global__ __launch_bounds__(256) void k_assemble2d(const Tile *__restrict__ tiles,
                                                    const RowInfo *__restrict__ rowinfo,
                                                    const double *__restrict__ Wq, const uint8_t *__restrict__ pat_fwd,
                                                    const uint8_t *__restrict__ pat_inv,
                                                    const int *__restrict__ pat_len, const int *__restrict__ pat_dir,
                                                    double *__restrict__ values, double *__restrict__ b, KParams kp) {
  using T = Tile2<K, NQ, R1, R2>;
  __shared__ double lds[T::LDS_DOUBLES];
  double *kap = lds;
  double *XG = lds + T::LDS_KAPPA;
  double *XM = XG + T::NP1 * T::XS;
  double *slab = XM + T::NP1 * T::XS;
  const Tile t = tiles[blockIdx.x];
  const int tid = threadIdx.x;
  const int qs0 = (t.org[0] / K - 1) * NQ, qs1 = (t.org[1] / K - 1) * NQ;
  for (int i = tid; i < T::LDS_KAPPA; i += T::NT) {
    const int a = i % T::Q1, c = i / T::Q1;
    const int q0 = qs0 + a, q1 = qs1 + c;
    double v = 0.0;
    if (q0 >= 0 && q0 < kp.qn[0] && q1 >= 0 && q1 < kp.qn[1]) v = Wq[q0 + (size_t)kp.qn[0] * q1];
    kap[i] = v;
  }
  __syncthreads();
  for (int it = tid; it < T::NP1 * T::Q2; it += T::NT) {
    const int p1 = it / T::Q2, q2 = it % T::Q2;
    int u, s, lc0, nc, a0, b0;
    pair_natural<K>(p1, u, s);
    pair_cells<K>(u, s, lc0, nc, a0, b0);
    double xg = 0.0, xm = 0.0;
    for (int tt = 0; tt < nc * NQ; ++tt) {
      double G, M;
      pair_factors<K, NQ>(kp, 0, a0, b0, tt, G, M);
      const double k = kap[q2 * T::Q1 + lc0 * NQ + tt];
      xg = fma(k, G, xg);
      xm = fma(k, M, xm);
    }
    XG[p1 * T::XS + q2] = xg;
    XM[p1 * T::XS + q2] = xm;
  }
  __syncthreads();
  for (int it = tid; it < T::NP1 * T::NP2; it += T::NT) {
    const int p1 = it % T::NP1, p2 = it / T::NP1;
    int u1, s1, u2, s2, lc0, nc, a0, b0;
    pair_natural<K>(p1, u1, s1);
    pair_natural<K>(p2, u2, s2);
    pair_cells<K>(u2, s2, lc0, nc, a0, b0);
    double v = 0.0;
    for (int tt = 0; tt < nc * NQ; ++tt) {
      double G, M;
      pair_factors<K, NQ>(kp, 1, a0, b0, tt, G, M);
      v = fma(XG[p1 * T::XS + lc0 * NQ + tt], M, fma(XM[p1 * T::XS + lc0 * NQ + tt], G, v));
    }
    slab[(u2 * R1 + u1) * T::S + (s2 + K) * T::W + (s1 + K)] = v;
  }
  __syncthreads();
  write_slab<K, 2, R1, R2, T::NT>(t, 0, slab, rowinfo, pat_fwd, pat_inv, pat_len, pat_dir, values, b, kp);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int launch_coefficients(gdx_problem *p, hipStream_t st) {
  const KParams &kp = p->kp;
  int qnmax = kp.qn[0];
  for (int d = 1; d < 3; ++d) qnmax = qnmax > kp.qn[d] ? qnmax : kp.qn[d];
  if (kp.nmodes > 0) {
    dim3 g1((qnmax + 255) / 256, kp.nmodes, 3);
    hipLaunchKernelGGL(k_fourier_tables, g1, dim3(256), 0, st, p->dev.modes, kp, qnmax, p->dev.E);
  }
  dim3 g2((kp.qn[0] + 127) / 128, kp.qn[1], kp.qn[2]);
  hipLaunchKernelGGL(k_coefficients, g2, dim3(128), 0, st, p->dev.modes, p->dev.E, kp, qnmax, p->dev.W, p->dev.F);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int launch_rhs(gdx_problem *p, hipStream_t st) {
  const int nt = (int)p->tiles.size();
  if (nt == 0) return GDX_OK;
  hipLaunchKernelGGL(k_rhs, dim3(nt), dim3(256), 0, st, p->dev.tiles, p->dev.rowinfo, p->dev.F, p->kp, p->tileR[0],
                     p->tileR[1], p->tileR[2], p->dev.b);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

template <int K, int NQ, int R1, int R2, int R3>
static void run3(gdx_problem *p, hipStream_t st) {
  using T = Tile3<K, NQ, R1, R2, R3>;
  hipLaunchKernelGGL((k_assemble3d<K, NQ, R1, R2, R3>), dim3((unsigned)p->tiles.size()), dim3(T::NT), 0, st,
                     p->dev.tiles, p->dev.rowinfo, p->dev.W, p->dev.pat_fwd, p->dev.pat_inv, p->dev.pat_len,
                     p->dev.pat_dir, p->dev.values, p->dev.b, p->kp);
}

template <int K, int NQ, int R1, int R2>
static void run2(gdx_problem *p, hipStream_t st) {
  hipLaunchKernelGGL((k_assemble2d<K, NQ, R1, R2>), dim3((unsigned)p->tiles.size()), dim3(256), 0, st, p->dev.tiles,
                     p->dev.rowinfo, p->dev.W, p->dev.pat_fwd, p->dev.pat_inv, p->dev.pat_len, p->dev.pat_dir,
                     p->dev.values, p->dev.b, p->kp);
}

bool matrix_kernel_supported(int dim, int K, int NQ) {
  if (dim == 3) return (K == 2 && (NQ == 2 || NQ == 3)) || (K == 1 && NQ == 2);
  return dim == 2 && (K == 1 || K == 2) && (NQ == 2 || NQ == 3);
}

int launch_matrix(gdx_problem *p, hipStream_t st) {
  if (p->tiles.empty()) return GDX_OK;
  const int K = p->K, NQ = p->NQ;
  if (p->D == 3) {
    if (K == 2 && NQ == 3) run3<2, 3, 8, 8, 8>(p, st);
    else if (K == 2 && NQ == 2) run3<2, 2, 8, 8, 8>(p, st);
    else if (K == 1 && NQ == 2) run3<1, 2, 8, 8, 4>(p, st);
    else return GDX_ERR_UNSUPPORTED;
  } else {
    if (K == 2 && NQ == 3) run2<2, 3, 16, 16>(p, st);
    else if (K == 2 && NQ == 2) run2<2, 2, 16, 16>(p, st);
    else if (K == 1 && NQ == 2) run2<1, 2, 16, 16>(p, st);
    else if (K == 1 && NQ == 3) run2<1, 3, 16, 16>(p, st);
    else return GDX_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

}  // namespace gdx

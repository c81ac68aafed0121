// 3D matrix kernel of libgdx, per-tile form (the second variant next to the
// z-marching kernel of gdx_march3d.hip; the autotuner picks per problem):
// owner-computes-rows, sum-factorised assembly of the Q_K Poisson operator on a
// Cartesian mesh (formulation in gdx_kernels.hip).  One workgroup = one tile of
// R1 x R2 x R3 owned rows; the RHS comes from k_rhs over the same tiles.
//
// LDS timeline of a workgroup:
//   kernel start  row table of the tile (persistent), kappa at the tile's
//                 quadrature region
//   stage 1       X = sum over x quadrature points, per x pair      (kappa -> X)
//   stage 2       Y = sum over y quadrature points, per (x,y) pair  (X -> registers)
//   stage 3       per z row: sum over z points into a slab of rows, then the
//                 rows are stored contiguously in CSR order; two slabs
//                 alternate and the tile's position patterns sit above them.
#include <hip/hip_runtime.h>

#include "gdx_device.h"

namespace gdx {

template <int K, int NQ, int R1, int R2, int R3>
struct Tile3 {
  static constexpr int C1 = R1 / K + 1, C2 = R2 / K + 1, C3 = R3 / K + 1;
  static constexpr int Q1 = C1 * NQ, Q2 = C2 * NQ, Q3 = C3 * NQ;
  static constexpr int NP1 = pair_count<K>(R1), NP2 = pair_count<K>(R2);
  static constexpr int XS = (Q2 * Q3) | 1;  // odd stride in doubles: conflict-free lanes
  static constexpr int NT = NP1 * NP2;
  static constexpr int NW = NT / 64;
  static constexpr int W = 2 * K + 1, S = W * W * W;
  static constexpr int SLAB = R1 * R2 * S;
  static constexpr int RPW = (R1 * R2 + NW - 1) / NW;  // slab rows per wave
  static constexpr int ROWS = R1 * R2 * R3;
  static constexpr int LDS_KAPPA = Q1 * Q2 * Q3;
  static constexpr int LDS_X = 2 * NP1 * XS;
  static constexpr int SP = pattern_stride(S);
  static constexpr int PRES = 64 * SP / 8;  // room for at least 64 patterns
  static constexpr int MAIN =
      (LDS_KAPPA + LDS_X) > 2 * SLAB + PRES ? (LDS_KAPPA + LDS_X) : 2 * SLAB + PRES;
  static constexpr int LDS_DOUBLES = MAIN + 2 * ROWS;  // + row table (16 bytes per row)
  static constexpr int PFREE = (MAIN - 2 * SLAB) * 8 / SP;
  static constexpr int PMAX = PFREE < 128 ? PFREE : 128;
  static constexpr int PWORDS = (PMAX * SP / 4 + NT - 1) / NT;
  static_assert(NT % 64 == 0, "whole waves");
};

// stage 3 for one z row u3 of the tile (compile-time, so Y stays in registers)
template <int K, int NQ, int W, int Q3, int U3>
__device__ __forceinline__ void z_row(double *__restrict__ slab, int rowslot, const double (&Y1)[Q3],
                                      const double (&Y2)[Q3], const KParams &kp) {
#pragma unroll
  for (int s3 = -K; s3 <= K; ++s3) {
    if (K == 2 && (U3 & 1) && (s3 == -2 || s3 == 2)) continue;
    int lc0, nc, a0, b0;
    pair_cells<K>(U3, s3, lc0, nc, a0, b0);
    double v = 0.0;
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt) {
      if (tt >= nc * NQ) break;
      double G, M;
      pair_factors<K, NQ>(kp, 2, a0, b0, tt, G, M);
      v = fma(Y1[lc0 * NQ + tt], M, fma(Y2[lc0 * NQ + tt], G, v));
    }
    slab[rowslot + (s3 + K) * W * W] = v;
  }
}

template <int K, int NQ, int W, int Q3, int R3, int U = 0>
__device__ __forceinline__ void z_row_dispatch(int u3, double *__restrict__ slab, int rowslot, const double (&Y1)[Q3],
                                               const double (&Y2)[Q3], const KParams &kp) {
  if constexpr (U < R3) {
    if (u3 == U) z_row<K, NQ, W, Q3, U>(slab, rowslot, Y1, Y2, kp);
    else z_row_dispatch<K, NQ, W, Q3, R3, U + 1>(u3, slab, rowslot, Y1, Y2, kp);
  }
}

// PATG: position patterns read from global memory (tiles with more distinct
// patterns than fit in LDS); otherwise they are staged in LDS.
template <int K, int NQ, int R1, int R2, int R3, bool PATG>
__global__ __launch_bounds__((Tile3<K, NQ, R1, R2, R3>::NT), 2) void k_assemble3d(
    const Tile *__restrict__ tiles, int tile0, int ntiles, const RowInfo *__restrict__ rowinfo,
    const double *__restrict__ Wq, const uint8_t *__restrict__ pat_fwd, const uint8_t *__restrict__ pat_inv,
    const int *__restrict__ tile_pats, double *__restrict__ values, double *__restrict__ b, KParams kp) {
  using T = Tile3<K, NQ, R1, R2, R3>;
  constexpr int NINV = (T::S + 63) / 64;
  __shared__ double lds[T::LDS_DOUBLES];
  double *kap = lds;                  // [Q3][Q2][Q1]
  double *XG = lds + T::LDS_KAPPA;    // [NP1][XS]
  double *XM = XG + T::NP1 * T::XS;   // [NP1][XS]
  RowT *rowt = reinterpret_cast<RowT *>(lds + T::MAIN);
  uint32_t *lpat = reinterpret_cast<uint32_t *>(lds + 2 * T::SLAB);
  const int tb = tile_of_block(ntiles);
  if (tb >= ntiles) return;
  const Tile t = tiles[tile0 + tb];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

  // fetch the tile's position patterns early (stored to LDS after stage 2)
  uint32_t pw[PATG ? 1 : T::PWORDS];
  if (!PATG) {
#pragma unroll
    for (int w = 0; w < T::PWORDS; ++w) {
      const int i = tid + w * T::NT;
      pw[w] = 0;
      if (i < t.npat * (T::SP / 4)) {
        const int pid = tile_pats[t.pat_off + i / (T::SP / 4)];
        pw[w] = reinterpret_cast<const uint32_t *>(pat_inv + (size_t)pid * T::SP)[i % (T::SP / 4)];
      }
    }
  }
  // row table of the tile
  for (int r = tid; r < T::ROWS; r += T::NT) {
    const int u1 = r % R1, u2 = (r / R1) % R2, u3 = r / (R1 * R2);
    RowT e{0, 0, 0};
    if (u1 < t.ext[0] && u2 < t.ext[1] && u3 < t.ext[2]) {
      const RowInfo ri = rowinfo[t.rowinfo_base + row_index(t, t.org[0] + u1, t.org[1] + u2, t.org[2] + u3)];
      if (ri.pid >= 0) e = RowT{ri.rowstart, (int)ri.bidx, ri.lenf};
    }
    rowt[r] = e;
  }

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
  for (int p1 = wave; p1 < T::NP1; p1 += T::NW) {
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
  __syncthreads();  // kappa and X are dead: their LDS holds two output slabs and the patterns

  if (!PATG) {
#pragma unroll
    for (int w = 0; w < T::PWORDS; ++w) {
      const int i = tid + w * T::NT;
      if (i < t.npat * (T::SP / 4)) lpat[i] = pw[w];
    }
  }

  // stage 3, one z row of the tile at a time; the two slabs alternate so one
  // barrier per z row suffices
  const int rowslot = (u2 * R1 + u1) * T::S + (s2 + K) * T::W + (s1 + K);
#pragma unroll 1
  for (int u3 = 0; u3 < R3; ++u3) {
    double *slab = lds + (u3 & 1) * T::SLAB;
    z_row_dispatch<K, NQ, T::W, T::Q3, R3>(u3, slab, rowslot, Y1, Y2, kp);
    __syncthreads();
    for (int j = 0; j < T::RPW; ++j) {
      const int rr = wave + T::NW * j;
      if (rr >= R1 * R2) break;
      const RowT e = rowt[u3 * R1 * R2 + rr];
      if (e.lenf == 0) continue;
      const int len = e.lenf & 0xFFFF, lpid = e.lenf >> 17;
      const double *row = slab + rr * T::S;
      double *dst = values + e.rowstart;
      const int gpid = (PATG || (e.lenf & 0x10000)) ? tile_pats[t.pat_off + lpid] : 0;
#pragma unroll
      for (int kk = 0; kk < NINV; ++kk) {
        const int k = lane + 64 * kk;
        if (k < len) {
          const int sl = PATG ? (int)pat_inv[(size_t)gpid * T::SP + k]
                              : (int)reinterpret_cast<const uint8_t *>(lpat)[lpid * T::SP + k];
          dst[k] = row[sl];
        }
      }
      if (e.lenf & 0x10000) {  // Dirichlet neighbours: b_i -= sum_j A_ij uD_j
        const uint8_t *fwd = pat_fwd + (size_t)gpid * T::SP;
        const int g0 = t.org[0] + rr % R1, g1 = t.org[1] + rr / R1, g2 = t.org[2] + u3;
        double acc = 0.0;
        for (int sl = lane; sl < T::S; sl += 64) {
          if (fwd[sl] == kSlotDirichlet) {
            const int d1 = sl % T::W - K, d2 = (sl / T::W) % T::W - K, d3 = sl / (T::W * T::W) - K;
            acc += row[sl] * uD_at(kp, g0 + d1, g1 + d2, g2 + d3);
          }
        }
        acc = wave_sum(acc);
        if (lane == 0) b[e.bidx] -= acc;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Direct-store form of the per-tile kernel, with the RHS fused: stages 1 and 2
// as above; f at the tile's quadrature region is prefetched into registers at
// the start and contracted (x, y, z) once kappa and X are dead.  In stage 3
// each thread stores its freshly contracted matrix values straight to their
// CSR positions through the row's slot -> offset pattern (staged in LDS), so
// the z loop has no output slab and no barrier.  Values adjacent to the
// Dirichlet boundary become per-thread lifting partials, summed per row in a
// fixed order; b = source integral - lifting is written once per row.  Matrix
// values are bit-identical to the slab form (same contraction order); the
// source integral matches k_rhs's order and the lifting sum's order differs.
// ---------------------------------------------------------------------------
template <int K, int NQ, int R1, int R2, int R3>
struct Tile3D : Tile3<K, NQ, R1, R2, R3> {
  using B = Tile3<K, NQ, R1, R2, R3>;
  // after stage 2 the kappa/X region holds, in order: f at the region's
  // quadrature points, its x and y contractions, the rows' source integrals,
  // the lifting partials [u3][thread] and the position patterns
  static constexpr int LF = B::LDS_KAPPA, LZ1 = R1 * B::Q3 * B::Q2, LZ2 = R1 * R2 * B::Q3;
  static constexpr int OFF_Z1 = LF, OFF_Z2 = OFF_Z1 + LZ1, OFF_BP = OFF_Z2 + LZ2, OFF_LIFT = OFF_BP + B::ROWS;
  static constexpr int LIFT = R3 * B::NT;
  // source-integral node factors w_t h_d phi_a(t) per (direction, tile node, region point)
  static constexpr int NF0 = R1 * B::Q1, NF1 = R2 * B::Q2, NF2 = R3 * B::Q3, NFN = NF0 + NF1 + NF2;
  static constexpr int OFF_NF = OFF_LIFT + LIFT;
  static constexpr int OFF_PAT = OFF_NF + NFN;
  // y pair factors (G then M, [a][b][q]) for stage 2, kept beyond the row table
  static constexpr int PFY = (K + 1) * (K + 1) * NQ;
  static constexpr int LDS_ALL = B::LDS_DOUBLES + 2 * PFY;
  static constexpr int PFREE = (B::MAIN - OFF_PAT) * 8 / B::SP;
  static constexpr int PMAX = PFREE < 128 ? PFREE : 128;
  static constexpr int PWORDS = (PMAX * B::SP / 4 + B::NT - 1) / B::NT;
  static constexpr int FPT = (LF + B::NT - 1) / B::NT;  // f values prefetched per thread
  static constexpr bool WROWS = K == 2 && R1 == 4 && R2 == 8;  // wave-owned rows (see wave_row_lane)
  static_assert(PMAX >= 16, "room for the position patterns");
  static_assert(NFN <= B::NT, "one node factor per thread");
};

// Q2 4 x 8 tiles: wave lane -> (row rank in the wave's 2x2 block of rows, s1, s2).
// A 2 x 2 block of rows (x rows 2a, 2a+1; y rows 2v, 2v+1) has 25 + 15 + 15 + 9
// = 64 (s1, s2) pairs: wave (a, v) holds exactly those four rows.
__device__ __forceinline__ void wave_row_lane(int lane, int &rr, int &s1, int &s2) {
  int k, nx;
  if (lane < 25) {
    rr = 0, k = lane, nx = 5;
  } else if (lane < 40) {
    rr = 1, k = lane - 25, nx = 3;
  } else if (lane < 55) {
    rr = 2, k = lane - 40, nx = 5;
  } else {
    rr = 3, k = lane - 55, nx = 3;
  }
  const int ny = (rr & 2) ? 3 : 5;
  s1 = k % nx - nx / 2;
  s2 = k / nx - ny / 2;
}

// the y pairs of tile row u in pair_doubles_first order: count, and the j-th index
template <int K>
__device__ __forceinline__ int y_pair_count(int u) {
  return K == 1 ? 3 : ((u & 1) ? 3 : 5);
}
template <int K, int R>
__device__ __forceinline__ int y_pair_index(int u, int j) {
  constexpr int ND = (K == 1) ? R : R / 2;
  if (K == 1) return j == 0 ? u : ND + 2 * u + (j - 1);          // s = 0, -1, +1
  if ((u & 1) == 0) return j == 0 ? u / 2 : ND + 7 * (u / 2) + (j - 1);  // s = 0, then -2, -1, 1, 2
  return ND + 7 * (u / 2) + 4 + j;                                // s = -1, 0, 1
}

// CT: type of the contractions (double, or float for the mixed form); the
// values are stored and the lifting accumulated in FP64 either way
template <int K, int NQ, int W, int Q3, int U3, class CT>
__device__ __forceinline__ double z_row_direct(const RowT &e, const uint8_t *__restrict__ fwd, int slot12,
                                               const CT (&Y1)[Q3], const CT (&Y2)[Q3], double *__restrict__ values,
                                               int g0, int g1, int g2, const KParams &kp) {
  double lift = 0.0;
#pragma unroll
  for (int s3 = -K; s3 <= K; ++s3) {
    if (K == 2 && (U3 & 1) && (s3 == -2 || s3 == 2)) continue;
    int lc0, nc, a0, b0;
    pair_cells<K>(U3, s3, lc0, nc, a0, b0);
    CT v = 0;
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt) {
      if (tt >= nc * NQ) break;
      double G, M;
      pair_factors<K, NQ>(kp, 2, a0, b0, tt, G, M);
      v = fma(Y1[lc0 * NQ + tt], (CT)M, fma(Y2[lc0 * NQ + tt], (CT)G, v));
    }
    const int pos = fwd[slot12 + (s3 + K) * W * W];
    if (pos < kSlotDirichlet) {
      values[e.rowstart + pos] = (double)v;
    } else if (pos == kSlotDirichlet) {
      const int s1 = slot12 % W - K, s2 = slot12 / W - K;
      lift = fma((double)v, uD_at(kp, g0 + s1, g1 + s2, g2 + s3), lift);
    }
  }
  return lift;
}

template <int K, int NQ, int W, int Q3, int R3, class CT, int U = 0>
__device__ __forceinline__ double z_row_direct_dispatch(int u3, const RowT &e, const uint8_t *__restrict__ fwd,
                                                        int slot12, const CT (&Y1)[Q3], const CT (&Y2)[Q3],
                                                        double *__restrict__ values, int g0, int g1, int g2,
                                                        const KParams &kp) {
  if constexpr (U < R3) {
    if (u3 == U) return z_row_direct<K, NQ, W, Q3, U, CT>(e, fwd, slot12, Y1, Y2, values, g0, g1, g2, kp);
    return z_row_direct_dispatch<K, NQ, W, Q3, R3, CT, U + 1>(u3, e, fwd, slot12, Y1, Y2, values, g0, g1, g2, kp);
  }
  return 0.0;
}

// the second bound is waves per SIMD: 4 keeps two 8-wave workgroups per CU (<= 128 VGPRs)
template <int K, int NQ, int R1, int R2, int R3, bool PATG, class CT>
__global__ __launch_bounds__((Tile3<K, NQ, R1, R2, R3>::NT), 4) void k_assemble3d_direct(
    const Tile *__restrict__ tiles, int tile0, int ntiles, const RowInfo *__restrict__ rowinfo,
    const double *__restrict__ Wq, const double *__restrict__ Fq, const uint8_t *__restrict__ pat_fwd,
    const int *__restrict__ tile_pats, double *__restrict__ values, double *__restrict__ b, KParams kp) {
  using T = Tile3D<K, NQ, R1, R2, R3>;
  __shared__ double lds[T::LDS_ALL];
  CT *kap = reinterpret_cast<CT *>(lds);  // [Q3][Q2][Q1]
  CT *XG = kap + T::LDS_KAPPA;            // [NP1][XS]
  CT *XM = XG + T::NP1 * T::XS;           // [NP1][XS]
  double *Fl = lds;                   // after stage 2: f [Q3][Q2][Q1]
  double *Z1 = lds + T::OFF_Z1;       // [u1][q3][q2]
  double *Z2 = lds + T::OFF_Z2;       // [u1][u2][q3]
  double *bpart = lds + T::OFF_BP;    // source integral per tile row
  double *liftp = lds + T::OFF_LIFT;  // [R3][NT]
  double *nfac = lds + T::OFF_NF;     // after stage 2: node factors [NF0 | NF1 | NF2]
  uint32_t *lpat = reinterpret_cast<uint32_t *>(lds + T::OFF_PAT);
  RowT *rowt = reinterpret_cast<RowT *>(lds + T::MAIN);
  double *pfy = lds + T::LDS_DOUBLES;  // y pair factors: G [a][b][q], then M
  const int tb = tile_of_block(ntiles);
  if (tb >= ntiles) return;
  const Tile t = tiles[tile0 + tb];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

  // the tile's slot -> offset patterns, fetched early (stored to LDS after stage 2)
  uint32_t pw[PATG ? 1 : T::PWORDS];
  if (!PATG) {
#pragma unroll
    for (int w = 0; w < T::PWORDS; ++w) {
      const int i = tid + w * T::NT;
      pw[w] = 0;
      if (i < t.npat * (T::SP / 4)) {
        const int pid = tile_pats[t.pat_off + i / (T::SP / 4)];
        pw[w] = reinterpret_cast<const uint32_t *>(pat_fwd + (size_t)pid * T::SP)[i % (T::SP / 4)];
      }
    }
  }
  for (int r = tid; r < T::ROWS; r += T::NT) {
    const int u1 = r % R1, u2 = (r / R1) % R2, u3 = r / (R1 * R2);
    RowT e{0, 0, 0};
    if (u1 < t.ext[0] && u2 < t.ext[1] && u3 < t.ext[2]) {
      const RowInfo ri = rowinfo[t.rowinfo_base + row_index(t, t.org[0] + u1, t.org[1] + u2, t.org[2] + u3)];
      if (ri.pid >= 0) e = RowT{ri.rowstart, (int)ri.bidx, ri.lenf};
    }
    rowt[r] = e;
  }
  // the small factor tables, read from the kernel argument while the region
  // loads are in flight (a per-lane index into the argument is a memory load
  // each time; from LDS it is not): y pair factors now, one node factor per
  // thread kept in a register until stage 2 has freed its LDS
  for (int i = tid; i < 2 * T::PFY; i += T::NT) {
    const int j = i % T::PFY, q = j % NQ, bb = (j / NQ) % (K + 1), a = j / (NQ * (K + 1));
    pfy[i] = i < T::PFY ? kp.fG[1][a][bb][q] : kp.fM[1][a][bb][q];
  }
  double nfv = 0.0;
  if (tid < T::NFN) {
    if (tid < T::NF0) nfv = node_factor<K, NQ>(kp, 0, tid / T::Q1, tid % T::Q1);
    else if (tid < T::NF0 + T::NF1) nfv = node_factor<K, NQ>(kp, 1, (tid - T::NF0) / T::Q2, (tid - T::NF0) % T::Q2);
    else nfv = node_factor<K, NQ>(kp, 2, (tid - T::NF0 - T::NF1) / T::Q3, (tid - T::NF0 - T::NF1) % T::Q3);
  }
  const int qs0 = (t.org[0] / K - 1) * NQ, qs1 = (t.org[1] / K - 1) * NQ, qs2 = (t.org[2] / K - 1) * NQ;
  double fpre[T::FPT];
#pragma unroll
  for (int j = 0; j < T::FPT; ++j) {
    const int i = tid + j * T::NT;
    const int a = i % T::Q1, bq = (i / T::Q1) % T::Q2, c = i / (T::Q1 * T::Q2);
    const int q0 = qs0 + a, q1 = qs1 + bq, q2 = qs2 + c;
    double v = 0.0, f = 0.0;
    if (i < T::LDS_KAPPA && q0 >= 0 && q0 < kp.qn[0] && q1 >= 0 && q1 < kp.qn[1] && q2 >= 0 && q2 < kp.qn[2]) {
      const size_t qi = q0 + (size_t)kp.qn[0] * (q1 + (size_t)kp.qn[1] * q2);
      v = Wq[qi];
      f = Fq[qi];
    }
    if (i < T::LDS_KAPPA) kap[i] = (CT)v;
    fpre[j] = f;
  }
  __syncthreads();

  // stage 1 (as the slab form)
  for (int p1 = wave; p1 < T::NP1; p1 += T::NW) {
    int u, s, lc0, nc, a0, b0;
    pair_natural<K>(p1, u, s);
    pair_cells<K>(u, s, lc0, nc, a0, b0);
    CT G[2 * NQ], M[2 * NQ];
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt) {
      double g, m;
      pair_factors<K, NQ>(kp, 0, a0, b0, tt, g, m);
      G[tt] = (CT)g;
      M[tt] = (CT)m;
    }
    for (int it = lane; it < T::Q2 * T::Q3; it += 64) {
      const CT *kr = kap + it * T::Q1 + lc0 * NQ;
      CT xg = 0, xm = 0;
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
  __syncthreads();

  // stage 2: thread -> (x pair, y pair).  For 4 x 8 Q2 tiles the threads of a
  // wave hold whole rows (its stores land in 4 rows); otherwise pairs-major
  int p1, u1, s1, u2, s2;
  if constexpr (T::WROWS) {
    int rr;
    wave_row_lane(lane, rr, s1, s2);
    u1 = 2 * (wave & 1) + (rr & 1);
    u2 = 2 * (wave >> 1) + (rr >> 1);
    p1 = 8 * (u1 / 2) + ((u1 & 1) ? 5 + (s1 + 1) : (s1 + 2));  // pair_natural index of (u1, s1)
  } else {
    p1 = tid % T::NP1;
    pair_natural<K>(p1, u1, s1);
    pair_doubles_first<K, R2>(tid / T::NP1, u2, s2);
  }
  CT Y1[T::Q3], Y2[T::Q3];
  {
    int lc0, nc, a0, b0;
    pair_cells<K>(u2, s2, lc0, nc, a0, b0);
    CT G[2 * NQ], M[2 * NQ];
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt) {
      const int ci = tt / NQ, q = tt % NQ;
      const int j = ((a0 - K * ci) * (K + 1) + (b0 - K * ci)) * NQ + q;
      G[tt] = (CT)pfy[j];
      M[tt] = (CT)pfy[T::PFY + j];
    }
    const CT *xg = XG + p1 * T::XS + lc0 * NQ;
    const CT *xm = XM + p1 * T::XS + lc0 * NQ;
    if (nc == 2) {
#pragma unroll
      for (int q3 = 0; q3 < T::Q3; ++q3) {
        CT y1 = 0, y2 = 0;
#pragma unroll
        for (int tt = 0; tt < 2 * NQ; ++tt) {
          const CT g = xg[q3 * T::Q2 + tt], m = xm[q3 * T::Q2 + tt];
          y1 = fma(g, M[tt], fma(m, G[tt], y1));
          y2 = fma(m, M[tt], y2);
        }
        Y1[q3] = y1;
        Y2[q3] = y2;
      }
    } else {
#pragma unroll
      for (int q3 = 0; q3 < T::Q3; ++q3) {
        CT y1 = 0, y2 = 0;
#pragma unroll
        for (int tt = 0; tt < NQ; ++tt) {
          const CT g = xg[q3 * T::Q2 + tt], m = xm[q3 * T::Q2 + tt];
          y1 = fma(g, M[tt], fma(m, G[tt], y1));
          y2 = fma(m, M[tt], y2);
        }
        Y1[q3] = y1;
        Y2[q3] = y2;
      }
    }
  }
  __syncthreads();  // kappa and X are dead: f, its contractions, lifting partials and patterns go there
  if (!PATG) {
#pragma unroll
    for (int w = 0; w < T::PWORDS; ++w) {
      const int i = tid + w * T::NT;
      if (i < t.npat * (T::SP / 4)) lpat[i] = pw[w];
    }
  }
#pragma unroll
  for (int j = 0; j < T::FPT; ++j) {
    const int i = tid + j * T::NT;
    if (i < T::LF) Fl[i] = fpre[j];
  }
  if (tid < T::NFN) nfac[tid] = nfv;
  __syncthreads();
  // source integral of every tile row, sum-factorised (x, y, z) as k_rhs
  for (int i = tid; i < T::LZ1; i += T::NT) {
    const int uu = i / (T::Q3 * T::Q2), r = i % (T::Q3 * T::Q2);
    const double *fr = Fl + r * T::Q1;
    const int v = uu + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
    double acc = 0.0;
    for (int lc = cmin; lc <= cmax; ++lc)
      for (int tt = 0; tt < NQ; ++tt) acc = fma(fr[lc * NQ + tt], nfac[uu * T::Q1 + lc * NQ + tt], acc);
    Z1[i] = acc;
  }
  __syncthreads();
  for (int i = tid; i < T::LZ2; i += T::NT) {
    const int uu1 = i / (R2 * T::Q3), uu2 = (i / T::Q3) % R2, q3 = i % T::Q3;
    const double *zr = Z1 + (uu1 * T::Q3 + q3) * T::Q2;
    const int v = uu2 + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
    double acc = 0.0;
    for (int lc = cmin; lc <= cmax; ++lc)
      for (int tt = 0; tt < NQ; ++tt) acc = fma(zr[lc * NQ + tt], nfac[T::NF0 + uu2 * T::Q2 + lc * NQ + tt], acc);
    Z2[i] = acc;
  }
  __syncthreads();
  for (int r = tid; r < T::ROWS; r += T::NT) {
    const int uu1 = r % R1, uu2 = (r / R1) % R2, uu3 = r / (R1 * R2);
    const double *zr = Z2 + (uu1 * R2 + uu2) * T::Q3;
    const int v = uu3 + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
    double acc = 0.0;
    for (int lc = cmin; lc <= cmax; ++lc)
      for (int tt = 0; tt < NQ; ++tt)
        acc = fma(zr[lc * NQ + tt], nfac[T::NF0 + T::NF1 + uu3 * T::Q3 + lc * NQ + tt], acc);
    bpart[r] = acc;
  }

  // stage 3: no slab, no barrier; each thread stores its values of every z row
  const int slot12 = (s2 + K) * T::W + (s1 + K);
  const int g0 = t.org[0] + u1, g1 = t.org[1] + u2;
#pragma unroll 1
  for (int u3 = 0; u3 < R3; ++u3) {
    const RowT e = rowt[(u3 * R2 + u2) * R1 + u1];
    if (e.lenf == 0) continue;
    const int lpid = e.lenf >> 17;
    const uint8_t *fwd = PATG ? pat_fwd + (size_t)tile_pats[t.pat_off + lpid] * T::SP
                              : reinterpret_cast<const uint8_t *>(lpat) + lpid * T::SP;
    const double lift =
        z_row_direct_dispatch<K, NQ, T::W, T::Q3, R3, CT>(u3, e, fwd, slot12, Y1, Y2, values, g0, g1, t.org[2] + u3, kp);
    if (e.lenf & 0x10000) liftp[u3 * T::NT + tid] = lift;
  }
  __syncthreads();
  // b_i = source integral - sum over the row's threads (fixed order) of their lifting partials
  for (int r = tid; r < T::ROWS; r += T::NT) {
    const RowT e = rowt[r];
    if (e.lenf == 0) continue;
    if (!(e.lenf & 0x10000)) {
      b[e.bidx] = bpart[r];
      continue;
    }
    const int ru1 = r % R1, ru2 = (r / R1) % R2, ru3 = r / (R1 * R2);
    double acc = 0.0;
    if constexpr (T::WROWS) {  // the row's threads are consecutive lanes of one wave
      const int w = (ru1 >> 1) + 2 * (ru2 >> 1), rq = (ru1 & 1) + 2 * (ru2 & 1);
      const int start = rq == 0 ? 0 : (rq == 1 ? 25 : (rq == 2 ? 40 : 55)), cnt = rq == 0 ? 25 : (rq == 3 ? 9 : 15);
      for (int i = 0; i < cnt; ++i) acc += liftp[ru3 * T::NT + 64 * w + start + i];
    } else {
      const int ny = y_pair_count<K>(ru2);
      const int x0 = K == 1 ? 3 * ru1 : 8 * (ru1 / 2) + ((ru1 & 1) ? 5 : 0);
      const int nx = K == 1 ? 3 : ((ru1 & 1) ? 3 : 5);
      for (int j = 0; j < ny; ++j) {
        const int yp = y_pair_index<K, R2>(ru2, j);
        for (int i = 0; i < nx; ++i) acc += liftp[ru3 * T::NT + yp * T::NP1 + x0 + i];
      }
    }
    b[e.bidx] = bpart[r] - acc;
  }
}

// ---------------------------------------------------------------------------
// launcher.  Tile shapes must match tile_shape() in gdx_symbolic.cpp.
// ---------------------------------------------------------------------------
template <int K, int NQ, int R1, int R2, int R3>
static void run3(gdx_problem *p, hipStream_t st) {
  using T = Tile3<K, NQ, R1, R2, R3>;
  const int nt = (int)p->tiles.size(), nf = p->n_fast_tiles;
  if (nf > 0)
    hipLaunchKernelGGL((k_assemble3d<K, NQ, R1, R2, R3, false>), dim3(8u * ((nf + 7) / 8)), dim3(T::NT), 0, st,
                       p->dev.tiles, 0, nf, p->dev.rowinfo, p->dev.W, p->dev.pat_fwd, p->dev.pat_inv,
                       p->dev.tile_pats, p->dev.values, p->dev.b, p->kp);
  if (nt > nf)
    hipLaunchKernelGGL((k_assemble3d<K, NQ, R1, R2, R3, true>), dim3(8u * ((nt - nf + 7) / 8)), dim3(T::NT), 0, st,
                       p->dev.tiles, nf, nt - nf, p->dev.rowinfo, p->dev.W, p->dev.pat_fwd, p->dev.pat_inv,
                       p->dev.tile_pats, p->dev.values, p->dev.b, p->kp);
}

// ---------------------------------------------------------------------------
// Wave-slab form (Q2, 4 x 8 x R3 tiles): the threads of a wave hold whole rows.
// A 2 x 2 block of rows (x rows 2a, 2a+1; y rows 2v, 2v+1) has 25 + 15 + 15 + 9
// = 64 (s1, s2) pairs, so wave (a, v) owns exactly those four rows.  Per z row
// every thread writes its values to the wave's private slab; the wave then
// stores each row contiguously in CSR order through the row's position ->
// slot pattern and writes b = source integral - lifting.  No block barrier
// in the z loop, and every store instruction writes consecutive doubles.
// ---------------------------------------------------------------------------
template <int NQ, int R3>
struct Tile3W : Tile3D<2, NQ, 4, 8, R3> {
  using B = Tile3<2, NQ, 4, 8, R3>;
  using D = Tile3D<2, NQ, 4, 8, R3>;
  static constexpr int OFF_SLAB = D::OFF_LIFT;          // [wave][4 rows][S]
  static constexpr int LSLAB = B::NW * 4 * B::S;
  static constexpr int OFF_PATW = OFF_SLAB + LSLAB;     // position -> slot patterns
  static constexpr int PFREEW = (B::MAIN - OFF_PATW) * 8 / B::SP;
  static constexpr int PMAXW = PFREEW < 128 ? PFREEW : 128;
  static constexpr int PWORDSW = (PMAXW * B::SP / 4 + B::NT - 1) / B::NT;
  static_assert(B::NW == 8 && B::NT == 512, "4 x 8 Q2 tile: 8 waves of 64 (s1,s2) pairs");
  static_assert(PMAXW >= 32, "room for the position patterns");
};

template <int NQ, int W, int Q3, int U3>
__device__ __forceinline__ void z_row_wslab(double *__restrict__ slab, int slot12, const double (&Y1)[Q3],
                                            const double (&Y2)[Q3], const KParams &kp) {
  constexpr int K = 2;
#pragma unroll
  for (int s3 = -K; s3 <= K; ++s3) {
    if ((U3 & 1) && (s3 == -2 || s3 == 2)) continue;
    int lc0, nc, a0, b0;
    pair_cells<K>(U3, s3, lc0, nc, a0, b0);
    double v = 0.0;
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt) {
      if (tt >= nc * NQ) break;
      double G, M;
      pair_factors<K, NQ>(kp, 2, a0, b0, tt, G, M);
      v = fma(Y1[lc0 * NQ + tt], M, fma(Y2[lc0 * NQ + tt], G, v));
    }
    slab[slot12 + (s3 + K) * W * W] = v;
  }
}

template <int NQ, int W, int Q3, int R3, int U = 0>
__device__ __forceinline__ void z_row_wslab_dispatch(int u3, double *__restrict__ slab, int slot12,
                                                     const double (&Y1)[Q3], const double (&Y2)[Q3],
                                                     const KParams &kp) {
  if constexpr (U < R3) {
    if (u3 == U) z_row_wslab<NQ, W, Q3, U>(slab, slot12, Y1, Y2, kp);
    else z_row_wslab_dispatch<NQ, W, Q3, R3, U + 1>(u3, slab, slot12, Y1, Y2, kp);
  }
}

template <int NQ, int R3, bool PATG>
__global__ __launch_bounds__(512, 4) void k_assemble3d_wslab(
    const Tile *__restrict__ tiles, int tile0, int ntiles, const RowInfo *__restrict__ rowinfo,
    const double *__restrict__ Wq, const double *__restrict__ Fq, const uint8_t *__restrict__ pat_fwd,
    const uint8_t *__restrict__ pat_inv, const int *__restrict__ tile_pats, double *__restrict__ values,
    double *__restrict__ b, KParams kp) {
  constexpr int K = 2, R1 = 4, R2 = 8;
  using T = Tile3W<NQ, R3>;
  constexpr int NINV = (T::S + 63) / 64;
  __shared__ double lds[T::LDS_DOUBLES];
  double *kap = lds;                  // [Q3][Q2][Q1]
  double *XG = lds + T::LDS_KAPPA;    // [NP1][XS]
  double *XM = XG + T::NP1 * T::XS;   // [NP1][XS]
  double *Fl = lds;                   // after stage 2: f [Q3][Q2][Q1]
  double *Z1 = lds + T::OFF_Z1;       // [u1][q3][q2]
  double *Z2 = lds + T::OFF_Z2;       // [u1][u2][q3]
  double *bpart = lds + T::OFF_BP;    // source integral per tile row
  uint32_t *lpat = reinterpret_cast<uint32_t *>(lds + T::OFF_PATW);
  RowT *rowt = reinterpret_cast<RowT *>(lds + T::MAIN);
  const int tb = tile_of_block(ntiles);
  if (tb >= ntiles) return;
  const Tile t = tiles[tile0 + tb];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

  uint32_t pw[PATG ? 1 : T::PWORDSW];
  if (!PATG) {
#pragma unroll
    for (int w = 0; w < T::PWORDSW; ++w) {
      const int i = tid + w * T::NT;
      pw[w] = 0;
      if (i < t.npat * (T::SP / 4)) {
        const int pid = tile_pats[t.pat_off + i / (T::SP / 4)];
        pw[w] = reinterpret_cast<const uint32_t *>(pat_inv + (size_t)pid * T::SP)[i % (T::SP / 4)];
      }
    }
  }
  for (int r = tid; r < T::ROWS; r += T::NT) {
    const int u1 = r % R1, u2 = (r / R1) % R2, u3 = r / (R1 * R2);
    RowT e{0, 0, 0};
    if (u1 < t.ext[0] && u2 < t.ext[1] && u3 < t.ext[2]) {
      const RowInfo ri = rowinfo[t.rowinfo_base + row_index(t, t.org[0] + u1, t.org[1] + u2, t.org[2] + u3)];
      if (ri.pid >= 0) e = RowT{ri.rowstart, (int)ri.bidx, ri.lenf};
    }
    rowt[r] = e;
  }
  const int qs0 = (t.org[0] / K - 1) * NQ, qs1 = (t.org[1] / K - 1) * NQ, qs2 = (t.org[2] / K - 1) * NQ;
  double fpre[T::FPT];
#pragma unroll
  for (int j = 0; j < T::FPT; ++j) {
    const int i = tid + j * T::NT;
    const int a = i % T::Q1, bq = (i / T::Q1) % T::Q2, c = i / (T::Q1 * T::Q2);
    const int q0 = qs0 + a, q1 = qs1 + bq, q2 = qs2 + c;
    double v = 0.0, f = 0.0;
    if (i < T::LDS_KAPPA && q0 >= 0 && q0 < kp.qn[0] && q1 >= 0 && q1 < kp.qn[1] && q2 >= 0 && q2 < kp.qn[2]) {
      const size_t qi = q0 + (size_t)kp.qn[0] * (q1 + (size_t)kp.qn[1] * q2);
      v = Wq[qi];
      f = Fq[qi];
    }
    if (i < T::LDS_KAPPA) kap[i] = v;
    fpre[j] = f;
  }
  __syncthreads();

  // stage 1 (as the slab form)
  for (int p1 = wave; p1 < T::NP1; p1 += T::NW) {
    int u, s, lc0, nc, a0, b0;
    pair_natural<K>(p1, u, s);
    pair_cells<K>(u, s, lc0, nc, a0, b0);
    double G[2 * NQ], M[2 * NQ];
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt) pair_factors<K, NQ>(kp, 0, a0, b0, tt, G[tt], M[tt]);
    for (int it = lane; it < T::Q2 * T::Q3; it += 64) {
      const double *kr = kap + it * T::Q1 + lc0 * NQ;
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
  __syncthreads();

  // stage 2 for this lane's (x pair, y pair) of its row
  int rr, s1, s2;
  wave_row_lane(lane, rr, s1, s2);
  const int u1 = 2 * (wave & 1) + (rr & 1), u2 = 2 * (wave >> 1) + (rr >> 1);
  const int p1 = 8 * (u1 / 2) + ((u1 & 1) ? 5 + (s1 + 1) : (s1 + 2));  // pair_natural index of (u1, s1)
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
  __syncthreads();  // kappa and X are dead
  if (!PATG) {
#pragma unroll
    for (int w = 0; w < T::PWORDSW; ++w) {
      const int i = tid + w * T::NT;
      if (i < t.npat * (T::SP / 4)) lpat[i] = pw[w];
    }
  }
#pragma unroll
  for (int j = 0; j < T::FPT; ++j) {
    const int i = tid + j * T::NT;
    if (i < T::LF) Fl[i] = fpre[j];
  }
  __syncthreads();
  // source integral of every tile row, sum-factorised (x, y, z) as k_rhs
  for (int i = tid; i < T::LZ1; i += T::NT) {
    const int uu = i / (T::Q3 * T::Q2), r = i % (T::Q3 * T::Q2);
    const double *fr = Fl + r * T::Q1;
    const int v = uu + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
    double acc = 0.0;
    for (int lc = cmin; lc <= cmax; ++lc)
      for (int tt = 0; tt < NQ; ++tt) acc = fma(fr[lc * NQ + tt], node_factor<K, NQ>(kp, 0, uu, lc * NQ + tt), acc);
    Z1[i] = acc;
  }
  __syncthreads();
  for (int i = tid; i < T::LZ2; i += T::NT) {
    const int uu1 = i / (R2 * T::Q3), uu2 = (i / T::Q3) % R2, q3 = i % T::Q3;
    const double *zr = Z1 + (uu1 * T::Q3 + q3) * T::Q2;
    const int v = uu2 + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
    double acc = 0.0;
    for (int lc = cmin; lc <= cmax; ++lc)
      for (int tt = 0; tt < NQ; ++tt) acc = fma(zr[lc * NQ + tt], node_factor<K, NQ>(kp, 1, uu2, lc * NQ + tt), acc);
    Z2[i] = acc;
  }
  __syncthreads();
  for (int r = tid; r < T::ROWS; r += T::NT) {
    const int uu1 = r % R1, uu2 = (r / R1) % R2, uu3 = r / (R1 * R2);
    const double *zr = Z2 + (uu1 * R2 + uu2) * T::Q3;
    const int v = uu3 + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
    double acc = 0.0;
    for (int lc = cmin; lc <= cmax; ++lc)
      for (int tt = 0; tt < NQ; ++tt) acc = fma(zr[lc * NQ + tt], node_factor<K, NQ>(kp, 2, uu3, lc * NQ + tt), acc);
    bpart[r] = acc;
  }
  __syncthreads();  // bpart is read by other waves below

  // stage 3: per z row, the wave fills its slab and stores its four rows
  double *wslab = lds + T::OFF_SLAB + wave * 4 * T::S;
  const int slot12 = rr * T::S + (s2 + K) * T::W + (s1 + K);
  const int g0b = t.org[0] + 2 * (wave & 1), g1b = t.org[1] + 2 * (wave >> 1);
#pragma unroll 1
  for (int u3 = 0; u3 < R3; ++u3) {
    z_row_wslab_dispatch<NQ, T::W, T::Q3, R3>(u3, wslab, slot12, Y1, Y2, kp);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int ru1 = 2 * (wave & 1) + (q & 1), ru2 = 2 * (wave >> 1) + (q >> 1);
      const int r = (u3 * R2 + ru2) * R1 + ru1;
      const RowT e = rowt[r];
      if (e.lenf == 0) continue;
      const int len = e.lenf & 0xFFFF, lpid = e.lenf >> 17;
      const double *row = wslab + q * T::S;
      double *dst = values + e.rowstart;
      const int gpid = (PATG || (e.lenf & 0x10000)) ? tile_pats[t.pat_off + lpid] : 0;
#pragma unroll
      for (int kk = 0; kk < NINV; ++kk) {
        const int k = lane + 64 * kk;
        if (k < len) {
          const int sl = PATG ? (int)pat_inv[(size_t)gpid * T::SP + k]
                              : (int)reinterpret_cast<const uint8_t *>(lpat)[lpid * T::SP + k];
          dst[k] = row[sl];
        }
      }
      double acc = 0.0;
      if (e.lenf & 0x10000) {  // Dirichlet neighbours: b_i -= sum_j A_ij uD_j
        const uint8_t *fwd = pat_fwd + (size_t)gpid * T::SP;
        const int g0 = g0b + (q & 1), g1 = g1b + (q >> 1), g2 = t.org[2] + u3;
        for (int sl = lane; sl < T::S; sl += 64) {
          if (fwd[sl] == kSlotDirichlet) {
            const int d1 = sl % T::W - K, d2 = (sl / T::W) % T::W - K, d3 = sl / (T::W * T::W) - K;
            acc += row[sl] * uD_at(kp, g0 + d1, g1 + d2, g2 + d3);
          }
        }
        acc = wave_sum(acc);
      }
      if (lane == 0) b[e.bidx] = bpart[r] - acc;
    }
    // the slab is rewritten for the next z row only after every lane has read it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

template <int NQ>
static void run3_wslab(gdx_problem *p, hipStream_t st) {
  const int nt = (int)p->tiles.size(), nf = p->n_fast_tiles;
  if (nf > 0)
    hipLaunchKernelGGL((k_assemble3d_wslab<NQ, 8, false>), dim3(8u * ((nf + 7) / 8)), dim3(512), 0, st,
                       p->dev.tiles, 0, nf, p->dev.rowinfo, p->dev.W, p->dev.F, p->dev.pat_fwd, p->dev.pat_inv,
                       p->dev.tile_pats, p->dev.values, p->dev.b, p->kp);
  if (nt > nf)
    hipLaunchKernelGGL((k_assemble3d_wslab<NQ, 8, true>), dim3(8u * ((nt - nf + 7) / 8)), dim3(512), 0, st,
                       p->dev.tiles, nf, nt - nf, p->dev.rowinfo, p->dev.W, p->dev.F, p->dev.pat_fwd, p->dev.pat_inv,
                       p->dev.tile_pats, p->dev.values, p->dev.b, p->kp);
}

bool wslab_supported(int K, int NQ) { return K == 2 && (NQ == 2 || NQ == 3); }

int pattern_capacity_tile3d_wslab(int K, int NQ) {
  if (K == 2 && NQ == 3) return Tile3W<3, 8>::PMAXW;
  if (K == 2 && NQ == 2) return Tile3W<2, 8>::PMAXW;
  return 0;
}

int launch_matrix_tile3d_wslab(gdx_problem *p, hipStream_t st) {
  if (p->K == 2 && p->NQ == 3) run3_wslab<3>(p, st);
  else if (p->K == 2 && p->NQ == 2) run3_wslab<2>(p, st);
  else return GDX_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

template <int K, int NQ, int R1, int R2, int R3, class CT>
static void run3_direct_t(gdx_problem *p, hipStream_t st) {
  using T = Tile3D<K, NQ, R1, R2, R3>;
  const int nt = (int)p->tiles.size(), nf = p->n_fast_tiles;
  if (nf > 0)
    hipLaunchKernelGGL((k_assemble3d_direct<K, NQ, R1, R2, R3, false, CT>), dim3(8u * ((nf + 7) / 8)), dim3(T::NT), 0,
                       st, p->dev.tiles, 0, nf, p->dev.rowinfo, p->dev.W, p->dev.F, p->dev.pat_fwd,
                       p->dev.tile_pats, p->dev.values, p->dev.b, p->kp);
  if (nt > nf)
    hipLaunchKernelGGL((k_assemble3d_direct<K, NQ, R1, R2, R3, true, CT>), dim3(8u * ((nt - nf + 7) / 8)),
                       dim3(T::NT), 0, st, p->dev.tiles, nf, nt - nf, p->dev.rowinfo, p->dev.W, p->dev.F,
                       p->dev.pat_fwd, p->dev.tile_pats, p->dev.values, p->dev.b, p->kp);
}

// FP64 or the mixed form (FP32 contractions), by the problem's precision
template <int K, int NQ, int R1, int R2, int R3>
static void run3_direct(gdx_problem *p, hipStream_t st) {
  if (p->precision == GDX_PREC_FP32) run3_direct_t<K, NQ, R1, R2, R3, float>(p, st);
  else run3_direct_t<K, NQ, R1, R2, R3, double>(p, st);
}

int pattern_capacity_tile3d_direct(int K, int NQ) {
  if (K == 2 && NQ == 3) return Tile3D<2, 3, 4, 8, 8>::PMAX;
  if (K == 2 && NQ == 2) return Tile3D<2, 2, 4, 8, 8>::PMAX;
  if (K == 1 && NQ == 2) return Tile3D<1, 2, 8, 8, 2>::PMAX;
  return 0;
}

int launch_matrix_tile3d_direct(gdx_problem *p, hipStream_t st) {
  const int K = p->K, NQ = p->NQ;
  if (K == 2 && NQ == 3) run3_direct<2, 3, 4, 8, 8>(p, st);
  else if (K == 2 && NQ == 2) run3_direct<2, 2, 4, 8, 8>(p, st);
  else if (K == 1 && NQ == 2) run3_direct<1, 2, 8, 8, 2>(p, st);
  else return GDX_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int pattern_capacity_tile3d(int K, int NQ) {
  if (K == 2 && NQ == 3) return Tile3<2, 3, 4, 8, 8>::PMAX;
  if (K == 2 && NQ == 2) return Tile3<2, 2, 4, 8, 8>::PMAX;
  if (K == 1 && NQ == 2) return Tile3<1, 2, 8, 8, 2>::PMAX;
  return 0;
}

void tile3d_shape(int K, int R[3]) {
  R[0] = K == 2 ? 4 : 8;
  R[1] = 8;
  R[2] = K == 2 ? 8 : 2;
}

int launch_matrix_tile3d(gdx_problem *p, hipStream_t st) {
  const int K = p->K, NQ = p->NQ;
  if (K == 2 && NQ == 3) run3<2, 3, 4, 8, 8>(p, st);
  else if (K == 2 && NQ == 2) run3<2, 2, 4, 8, 8>(p, st);
  else if (K == 1 && NQ == 2) run3<1, 2, 8, 8, 2>(p, st);
  else return GDX_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

}  // namespace gdx

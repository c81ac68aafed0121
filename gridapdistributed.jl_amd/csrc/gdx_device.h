// Device helpers shared by the libgdx kernels (tile geometry, 1D pair
// structure of the sum-factorised contractions, row stores).
#pragma once
#include <hip/hip_runtime.h>

#include "gdx_internal.h"

namespace gdx {

// ---------------------------------------------------------------------------
// Tile geometry shared by the RHS and matrix kernels.  Tile rows u in [0,R)
// along a dimension map to region node v = u + K; the region starts one cell
// below the tile origin and spans R/K+1 cells (NQ Gauss points each).
// ---------------------------------------------------------------------------
__device__ __forceinline__ int tile_of_block(int nt) {
  // XCD-aware: consecutive blocks are dealt round-robin over the 8 XCDs, so
  // give each XCD a contiguous run of tiles (neighbouring tiles share halos)
  const int per = (nt + 7) / 8;
  return (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
}

__device__ __forceinline__ long long row_index(const Tile &t, int g0, int g1, int g2) {
  return ((long long)(g2 - t.onb_lo[2]) * t.onb_n[1] + (g1 - t.onb_lo[1])) * t.onb_n[0] + (g0 - t.onb_lo[0]);
}


// ---------------------------------------------------------------------------
// 1D pair structure of a tile.  A pair (v, v+s) is supported on cells lc0 [, lc0+1].
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
  G = kp.fG[d][a][b][q];
  M = kp.fM[d][a][b][q];
}

// weight of region quadrature point lcq (cell lc = lcq/NQ, point t) in the
// source integral of tile node u along direction d: w_t h_d phi_a(t)
template <int K, int NQ>
__device__ __forceinline__ double node_factor(const KParams &kp, int d, int u, int lcq) {
  const int v = u + K, lc = lcq / NQ, t = lcq % NQ;
  const int a = v - K * lc;
  if (a < 0 || a > K) return 0.0;
  return kp.qw[t] * kp.h[d] * kp.phi[a][t];
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ double uD_at(const KParams &kp, int g0, int g1, int g2) {
  const double x = kp.x0[0] + kp.hK[0] * (double)g0;
  const double y = kp.x0[1] + kp.hK[1] * (double)g1;
  const double z = kp.dim == 3 ? kp.x0[2] + kp.hK[2] * (double)g2 : 0.0;
  return kp.dir_coef[0] + kp.dir_coef[1] * x + kp.dir_coef[2] * y + kp.dir_coef[3] * z;
}

// Write one staged row from LDS to HBM in CSR order (contiguous stores) and,
// for rows next to the Dirichlet boundary, subtract the lifting sum A_ij uD_j.
template <int K, int DIM, int S>
__device__ __forceinline__ void write_row(const double *__restrict__ row, long long rowstart, long long bidx, int pid,
                                          int lenf, const int (&inv)[(S + 63) / 64], int g0, int g1, int g2,
                                          const uint8_t *__restrict__ pat_fwd, double *__restrict__ values,
                                          double *__restrict__ b, const KParams &kp) {
  constexpr int W = 2 * K + 1;
  const int lane = threadIdx.x & 63, len = lenf & 0xFFFF;
  double *dst = values + rowstart;
#pragma unroll
  for (int kk = 0; kk < (S + 63) / 64; ++kk) {
    const int k = lane + 64 * kk;
    if (k < len) dst[k] = row[inv[kk]];
  }
  if (lenf >> 16) {
    const uint8_t *fwd = pat_fwd + (size_t)pid * pattern_stride(S);
    double acc = 0.0;
    for (int sl = lane; sl < S; sl += 64) {
      if (fwd[sl] == kSlotDirichlet) {
        const int s1 = sl % W - K, s2 = (sl / W) % W - K, s3 = DIM == 3 ? sl / (W * W) - K : 0;
        acc += row[sl] * uD_at(kp, g0 + s1, g1 + s2, g2 + s3);
      }
    }
    acc = wave_sum(acc);
    if (lane == 0) b[bidx] -= acc;
  }
}


// row table entry staged in LDS
struct RowT {
  long long rowstart;
  int bidx;
  int lenf;  // 0 when the node is not a free row of the tile
};

}  // namespace gdx

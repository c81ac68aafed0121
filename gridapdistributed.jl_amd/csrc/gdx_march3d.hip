// 3D assembly kernel of libgdx, z-marching form (formulation in gdx_kernels.hip).
//
// One workgroup owns an R1 x R2 column of owned rows of one part and walks
// along z one cell per step.  Per step c (cell c of the column):
//   stage 1  kappa(x_q) and f(x_q) of cell c (prefetched one step ahead) are
//            contracted over x quadrature points: X[x pair][z point][y point]
//   stage 2  thread (x pair, y pair) contracts y points: Y[z point] of cell c
//            (registers; cell c-1's values are kept, the z halo is never
//            recomputed); the f part likewise per (x row, y row)
//   stage 3  the K z rows whose support is cells c-1 and c are completed into
//            an LDS slab (each row's stencil box, sized by node parity), then
//            one wave per CSR row stores it contiguously through the row's
//            position pattern, and writes b = f part - sum_j A_ij uD_j.
// Two barriers per step; no float atomics; every value is written once.
// Factors indexed per lane (x and y pairs, f weights) are tabulated in LDS at
// kernel start; z factors are wave-uniform and stay in scalar registers.
#include <hip/hip_runtime.h>

#include "gdx_device.h"

namespace gdx {

template <int K>
__host__ __device__ constexpr int width_prefix(int u) {
  // sum of stencil widths of tile rows 0..u-1 (tile origins are multiples of K)
  return K == 1 ? 3 * u : 8 * (u >> 1) + 5 * (u & 1);
}

template <int K, int NQ, int R1, int R2, class CT>
struct March {
  static constexpr int C1 = R1 / K + 1, C2 = R2 / K + 1;
  static constexpr int Q1 = C1 * NQ, Q2 = C2 * NQ;
  static constexpr int NP1 = pair_count<K>(R1), NP2 = pair_count<K>(R2);
  static constexpr int NT = NP1 * NP2, NW = NT / 64;
  static constexpr int CS = Q1 * Q2 * NQ;      // quadrature points of one z cell of the column
  static constexpr int XS = (NQ * Q2) | 1;     // X stride per x pair (odd: conflict-free)
  static constexpr int W = 2 * K + 1, S = W * W * W, SP = pattern_stride(S);
  static constexpr int T1 = width_prefix<K>(R1), T2 = width_prefix<K>(R2);
  static constexpr int ZW0 = 2 * K + 1, ZW1 = 3;  // widths of the step's z rows
  static constexpr int SLAB = T1 * T2 * (K == 2 ? ZW0 + ZW1 : ZW0);
  static constexpr int RS = R1 * R2 * K;          // rows completed per step
  static constexpr int PMAX = 96;  // tile patterns kept in LDS (3D Q2 64^3 needs at most 72)
  static constexpr int NPF = (2 * CS + NT - 1) / NT;  // prefetched values per thread
  static constexpr int AL = 8 / sizeof(CT);           // elements per 8 bytes
  // LDS (in elements of CT)
  static constexpr int O_KF = 0;                        // [2 buffers][kappa, f][CS]
  static constexpr int O_X = O_KF + 4 * CS;             // XG, XM: [NP1][XS] each
  static constexpr int O_Z1 = O_X + 2 * NP1 * XS;       // f part after x: [R1][NQ][Q2]
  static constexpr int O_SLAB = O_Z1 + R1 * NQ * Q2;
  static constexpr int O_BF = O_SLAB + SLAB;            // f part of the step's rows
  static constexpr int O_RT = (O_BF + RS + AL - 1) / AL * AL;  // row tables [2][RS] x 16 B (8-byte aligned)
  static constexpr int O_FY = O_RT + 2 * (16 / sizeof(CT)) * RS;  // y pair factors: [NP2][G | M][2 NQ]
  static constexpr int O_FZ = O_FY + NP2 * 4 * NQ;      // f part after y: [2 cells][R1 R2][NQ]
  static constexpr int O_FX = O_FZ + 2 * R1 * R2 * NQ; // x pair factors: [NP1][G | M][2 NQ]
  static constexpr int O_PX = O_FX + NP1 * 4 * NQ;      // f weights of x rows: [R1][2 NQ]
  static constexpr int O_PY = O_PX + R1 * 2 * NQ;       // f weights of y rows: [R2][2 NQ]
  static constexpr int O_PZ = O_PY + R2 * 2 * NQ;       // f weights of the K z rows: [K][2 NQ]
  static constexpr int O_PAT = O_PZ + K * 2 * NQ;       // PMAX patterns x SP bytes
  static constexpr int LDS_ELEMS = O_PAT + PMAX * SP / sizeof(CT);
  static_assert(NT % 64 == 0, "whole waves");
  static_assert(RS <= NT, "one row-table entry per thread");
};

// row of the step's slab: offset of row (u1,u2) of z row o, and its reduced slot
template <int K, int R1, int R2>
__device__ __forceinline__ int slab_row_offset(int u1, int u2, int o) {
  constexpr int T1 = width_prefix<K>(R1), T2 = width_prefix<K>(R2);
  const int w3 = o == 0 ? 2 * K + 1 : 3;
  const int zoff = o == 0 ? 0 : T1 * T2 * (2 * K + 1);
  return zoff + w3 * (width_prefix<K>(u2) * T1 + stencil_width(K, u2) * width_prefix<K>(u1));
}

// CT: type of the contractions (double, or float for the mixed form); values
// and b are stored in FP64 either way
template <int K, int NQ, int R1, int R2, bool PATG, class CT>
// two workgroups per CU: at most 128 VGPRs (4 waves per SIMD)
__global__ __launch_bounds__((March<K, NQ, R1, R2, CT>::NT)) __attribute__((amdgpu_waves_per_eu(4))) void k_march3d(
    const Tile *__restrict__ tiles, int tile0, int ntiles, const RowInfo *__restrict__ rowinfo,
    const double *__restrict__ Wq, const double *__restrict__ Fq, const uint8_t *__restrict__ pat_invr,
    const int *__restrict__ tile_pats, double *__restrict__ values, double *__restrict__ b, KParams kp) {
  using T = March<K, NQ, R1, R2, CT>;
  __shared__ __align__(16) CT lds[T::LDS_ELEMS];
  CT *XG = lds + T::O_X, *XM = XG + T::NP1 * T::XS;
  CT *Z1 = lds + T::O_Z1, *slab = lds + T::O_SLAB, *BF = lds + T::O_BF;
  RowT *RT = reinterpret_cast<RowT *>(lds + T::O_RT);
  uint8_t *PAT = reinterpret_cast<uint8_t *>(lds + T::O_PAT);
  const int tb = tile_of_block(ntiles);
  if (tb >= ntiles) return;
  const Tile t = tiles[tile0 + tb];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int c0 = t.org[2] / K;                      // first cell whose left node is a tile row
  const int cl = (t.org[2] + t.ext[2] - 1) / K;      // last cell touching a tile row
  const int q0 = (t.org[0] / K - 1) * NQ, q1 = (t.org[1] / K - 1) * NQ;

  // quadrature data of z cell c into buffer c & 1 (zero outside the mesh)
  auto kf_load = [&](int c, CT (&reg)[T::NPF]) {
#pragma unroll
    for (int w = 0; w < T::NPF; ++w) {
      const int i = tid + w * T::NT;
      CT v = 0;
      if (i < 2 * T::CS) {
        const int j = i % T::CS, a = j % T::Q1, bq = (j / T::Q1) % T::Q2, tz = j / (T::Q1 * T::Q2);
        const int qa = q0 + a, qb = q1 + bq, qc = c * NQ + tz;
        if (c >= 0 && c < kp.ncells[2] && qa >= 0 && qa < kp.qn[0] && qb >= 0 && qb < kp.qn[1]) {
          const size_t idx = qa + (size_t)kp.qn[0] * (qb + (size_t)kp.qn[1] * qc);
          v = (CT)(i < T::CS ? Wq[idx] : Fq[idx]);
        }
      }
      reg[w] = v;
    }
  };
  auto kf_store = [&](int c, const CT (&reg)[T::NPF]) {
    CT *buf = lds + T::O_KF + (c & 1) * 2 * T::CS;
#pragma unroll
    for (int w = 0; w < T::NPF; ++w) {
      const int i = tid + w * T::NT;
      if (i < 2 * T::CS) buf[i] = reg[w];
    }
  };
  // row table of step c (rows z = K c + o) into buffer c & 1
  auto rt_load = [&](int c) -> RowT {
    RowT e{0, 0, 0};
    if (tid < T::RS) {
      const int u1 = tid % R1, u2 = (tid / R1) % R2, o = tid / (R1 * R2);
      const int z = K * c + o, uz = z - t.org[2];
      if (u1 < t.ext[0] && u2 < t.ext[1] && uz >= 0 && uz < t.ext[2]) {
        const RowInfo ri = rowinfo[t.rowinfo_base + row_index(t, t.org[0] + u1, t.org[1] + u2, z)];
        if (ri.pid >= 0) e = RowT{ri.rowstart, (int)ri.bidx, ri.lenf};
      }
    }
    return e;
  };

  // prologue: patterns, first z cell, row table of the first emitting step
  if (!PATG) {
    const uint32_t *src = reinterpret_cast<const uint32_t *>(pat_invr);
    uint32_t *dst = reinterpret_cast<uint32_t *>(PAT);
    for (int i = tid; i < t.npat * (T::SP / 4); i += T::NT)
      dst[i] = src[(size_t)tile_pats[t.pat_off + i / (T::SP / 4)] * (T::SP / 4) + i % (T::SP / 4)];
  }
  {
    CT reg[T::NPF];
    kf_load(c0 - 1, reg);
    kf_store(c0 - 1, reg);
  }

  // pair structure of this thread
  const int p1 = tid % T::NP1, p2 = tid / T::NP1;
  int u1, s1, u2, s2;
  pair_natural<K>(p1, u1, s1);
  pair_doubles_first<K, R2>(p2, u2, s2);
  int lc2, nc2, a2, b2;
  pair_cells<K>(u2, s2, lc2, nc2, a2, b2);
  // y pair factors live in LDS (they vary per lane; registers are scarce)
  CT *FY = lds + T::O_FY;
  if (tid < T::NP2) {
    int uu, ss, l0, n0, aa, bb;
    pair_doubles_first<K, R2>(tid, uu, ss);
    pair_cells<K>(uu, ss, l0, n0, aa, bb);
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt)
    {
      double G, M;
      pair_factors<K, NQ>(kp, 1, aa, bb, tt, G, M);
      FY[tid * 4 * NQ + tt] = (CT)G;
      FY[tid * 4 * NQ + 2 * NQ + tt] = (CT)M;
    }
  }
  // x pair factors and the f weights of x and y rows: lane-varying indices
  // into kernel arguments would become vector loads that queue behind the
  // prefetch, so they are tabulated in LDS once
  CT *FX = lds + T::O_FX, *PX = lds + T::O_PX, *PY = lds + T::O_PY;
  if (tid >= 64 && tid < 64 + T::NP1) {
    const int pp = tid - 64;
    int uu, ss, l0, n0, aa, bb;
    pair_natural<K>(pp, uu, ss);
    pair_cells<K>(uu, ss, l0, n0, aa, bb);
#pragma unroll
    for (int tt = 0; tt < 2 * NQ; ++tt)
    {
      double G, M;
      pair_factors<K, NQ>(kp, 0, aa, bb, tt, G, M);
      FX[pp * 4 * NQ + tt] = (CT)G;
      FX[pp * 4 * NQ + 2 * NQ + tt] = (CT)M;
    }
  }
  if (tid >= 128 && tid < 128 + R1 + R2) {
    const int d = tid < 128 + R1 ? 0 : 1, u = tid - 128 - d * R1, v = u + K;
    const int cmin = (v + K - 1) / K - 1, cmax = v / K;
    CT *dst = (d == 0 ? PX : PY) + u * 2 * NQ;
#pragma unroll
    for (int ci = 0; ci < 2; ++ci)
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        dst[ci * NQ + q] = (CT)(cmin + ci <= cmax ? kp.qw[q] * kp.h[d] * kp.phi[v - K * (cmin + ci)][q] : 0.0);
  }
  // f weights of the step's K z rows
  CT *PZ = lds + T::O_PZ;
  if (tid >= 256 && tid < 256 + K) {
    const int o = tid - 256, v3 = o + K, cmin = (v3 + K - 1) / K - 1, cmax = v3 / K;
#pragma unroll
    for (int lc = 0; lc < 2; ++lc)
#pragma unroll
      for (int q = 0; q < NQ; ++q)
        PZ[o * 2 * NQ + lc * NQ + q] =
            (CT)(lc >= cmin && lc <= cmax ? kp.qw[q] * kp.h[2] * kp.phi[v3 - K * lc][q] : 0.0);
  }
  const CT *G2 = FY + p2 * 4 * NQ, *M2 = G2 + 2 * NQ;
  CT *FZ = lds + T::O_FZ;  // f part of this thread's row after the y contraction, per cell parity
  const int w1 = stencil_width(K, u1), w2 = stencil_width(K, u2);
  const int rbase0 = slab_row_offset<K, R1, R2>(u1, u2, 0) + (s2 - stencil_lo(K, u2)) * w1 + (s1 - stencil_lo(K, u1));
  const int rbase1 = K == 2 ? slab_row_offset<K, R1, R2>(u1, u2, 1) + (s2 - stencil_lo(K, u2)) * w1 +
                                  (s1 - stencil_lo(K, u1))
                            : 0;
  // f part: thread tid < R1*R2 owns row (fu1, fu2) of the column
  const int fu1 = tid % R1, fu2 = (tid / R1) % R2;

  CT Yp1[NQ], Yp2[NQ], Yc1[NQ], Yc2[NQ];
#pragma unroll
  for (int i = 0; i < NQ; ++i) Yc1[i] = Yc2[i] = 0;

  __syncthreads();  // patterns and the first cell are in LDS

  // Per step, barrier A (after stage 1) separates the previous step's row
  // stores from this step's slab writes and frees the other quadrature buffer;
  // barrier B (after stage 3) publishes the slab, the next cell's quadrature
  // data and the next row table.
#pragma unroll 1
  for (int c = c0 - 1; c <= cl; ++c) {
    const CT *Kc = lds + T::O_KF + (c & 1) * 2 * T::CS, *Fl = Kc + T::CS;
    // prefetch the next cell and the next step's row table (registers)
    CT reg[T::NPF];
    kf_load(c + 1, reg);
    const RowT rtn = rt_load(c + 1);

    // stage 1: x contraction of kappa (per x pair) and of f (per x row)
    for (int it = tid; it < T::NP1 * NQ * T::Q2 + R1 * NQ * T::Q2; it += T::NT) {
      if (it < T::NP1 * NQ * T::Q2) {
        const int pp = it / (NQ * T::Q2), r = it % (NQ * T::Q2);  // r = tz*Q2 + q2
        int u, s, lc0, nc, a0, b0;
        pair_natural<K>(pp, u, s);
        pair_cells<K>(u, s, lc0, nc, a0, b0);
        const CT *kr = Kc + r * T::Q1 + lc0 * NQ, *fx = FX + pp * 4 * NQ;
        CT xg = 0, xm = 0;
#pragma unroll
        for (int tt = 0; tt < 2 * NQ; ++tt) {
          if (tt >= nc * NQ) break;
          xg = fma(kr[tt], fx[tt], xg);
          xm = fma(kr[tt], fx[2 * NQ + tt], xm);
        }
        XG[pp * T::XS + r] = xg;
        XM[pp * T::XS + r] = xm;
      } else {
        const int j = it - T::NP1 * NQ * T::Q2, u = j / (NQ * T::Q2), r = j % (NQ * T::Q2);
        const int v = u + K, cmin = (v + K - 1) / K - 1, nc = v / K - cmin + 1;
        const CT *fr = Fl + r * T::Q1 + cmin * NQ, *px = PX + u * 2 * NQ;
        CT acc = 0;
#pragma unroll
        for (int tt = 0; tt < 2 * NQ; ++tt) {
          if (tt >= nc * NQ) break;
          acc = fma(fr[tt], px[tt], acc);
        }
        Z1[j] = acc;
      }
    }
    __syncthreads();

    // stage 2: y contraction for this thread's (x pair, y pair), all z points of cell c
#pragma unroll
    for (int i = 0; i < NQ; ++i) {
      Yp1[i] = Yc1[i];
      Yp2[i] = Yc2[i];
    }
    {
      const CT *xg = XG + p1 * T::XS + lc2 * NQ, *xm = XM + p1 * T::XS + lc2 * NQ;
#pragma unroll
      for (int tz = 0; tz < NQ; ++tz) {
        CT y1 = 0, y2 = 0;
#pragma unroll
        for (int tt = 0; tt < 2 * NQ; ++tt) {
          if (tt >= nc2 * NQ) break;
          const CT g = xg[tz * T::Q2 + tt], m = xm[tz * T::Q2 + tt];
          y1 = fma(g, M2[tt], fma(m, G2[tt], y1));
          y2 = fma(m, M2[tt], y2);
        }
        Yc1[tz] = y1;
        Yc2[tz] = y2;
      }
    }
    if (tid < R1 * R2) {
      const int v = fu2 + K, cmin = (v + K - 1) / K - 1, nc = v / K - cmin + 1;
      const CT *py = PY + fu2 * 2 * NQ;
#pragma unroll
      for (int tz = 0; tz < NQ; ++tz) {
        const CT *zr = Z1 + (fu1 * NQ + tz) * T::Q2 + cmin * NQ;
        CT acc = 0;
#pragma unroll
        for (int tt = 0; tt < 2 * NQ; ++tt) {
          if (tt >= nc * NQ) break;
          acc = fma(zr[tt], py[tt], acc);
        }
        FZ[((c & 1) * R1 * R2 + tid) * NQ + tz] = acc;
      }
    }

    // stage 3: complete the K z rows of this step (supported on cells c-1, c)
    if (c >= c0) {
#pragma unroll
      for (int o = 0; o < K; ++o) {
#pragma unroll
        for (int s3 = -K; s3 <= K; ++s3) {
          if (o == 1 && (s3 < -1 || s3 > 1)) continue;
          int lc0, nc, a0, b0;
          pair_cells<K>(o, s3, lc0, nc, a0, b0);
          CT v = 0;
#pragma unroll
          for (int tt = 0; tt < 2 * NQ; ++tt) {
            if (tt >= nc * NQ) break;
            double G, M;
            pair_factors<K, NQ>(kp, 2, a0, b0, tt, G, M);  // wave-uniform: scalar loads
            const int ti = lc0 * NQ + tt;  // 0..NQ-1: cell c-1, NQ..2NQ-1: cell c
            const CT y1 = ti < NQ ? Yp1[ti % NQ] : Yc1[ti % NQ];
            const CT y2 = ti < NQ ? Yp2[ti % NQ] : Yc2[ti % NQ];
            v = fma(y1, (CT)M, fma(y2, (CT)G, v));
          }
          const int lo3 = o == 0 ? -K : -1;
          slab[(o == 0 ? rbase0 : rbase1) + (s3 - lo3) * w1 * w2] = v;
        }
        if (tid < R1 * R2) {
          CT acc = 0;
#pragma unroll
          for (int ti = 0; ti < 2 * NQ; ++ti)
            acc = fma(FZ[(((c - 1 + ti / NQ) & 1) * R1 * R2 + tid) * NQ + ti % NQ], PZ[o * 2 * NQ + ti], acc);
          BF[o * R1 * R2 + fu2 * R1 + fu1] = acc;
        }
      }
    }
    // the prefetched cell and row table land last: their loads had all three
    // stages to arrive
    kf_store(c + 1, reg);
    if (tid < T::RS) RT[((c + 1) & 1) * T::RS + tid] = rtn;
    __syncthreads();

    // store the completed rows: one wave per row, G rows of a wave at a time
    // with their pattern bytes, then their slab values, then their stores
    // issued together (the two dependent LDS reads of a row overlap with the
    // other rows')
    if (c >= c0) {
      const RowT *rt = RT + (c & 1) * T::RS;
      constexpr int G = 4, KK = (T::S + 63) / 64;
      for (int r0 = wave; r0 < T::RS; r0 += G * T::NW) {
        int lenv[G];
        const uint8_t *invv[G];
        const CT *rowv[G];
        double *dstv[G];
#pragma unroll
        for (int j = 0; j < G; ++j) {
          const int r = r0 + j * T::NW;
          lenv[j] = 0;
          invv[j] = PAT;
          rowv[j] = slab;
          dstv[j] = values;
          if (r < T::RS) {
            const RowT e = rt[r];
            const int ru1 = r % R1, ru2 = (r / R1) % R2, o = r / (R1 * R2);
            const int lpid = e.lenf >> 17;
            lenv[j] = e.lenf & 0xFFFF;
            invv[j] = PATG ? pat_invr + (size_t)tile_pats[t.pat_off + lpid] * T::SP : PAT + lpid * T::SP;
            rowv[j] = slab + slab_row_offset<K, R1, R2>(ru1, ru2, o);
            dstv[j] = values + e.rowstart;
          }
        }
        int slv[G][KK];
#pragma unroll
        for (int j = 0; j < G; ++j)
#pragma unroll
          for (int kk = 0; kk < KK; ++kk) {
            const int k = lane + 64 * kk;
            slv[j][kk] = k < lenv[j] ? invv[j][k] : 0;
          }
        CT vv[G][KK];
#pragma unroll
        for (int j = 0; j < G; ++j)
#pragma unroll
          for (int kk = 0; kk < KK; ++kk) vv[j][kk] = rowv[j][slv[j][kk]];
#pragma unroll
        for (int j = 0; j < G; ++j)
#pragma unroll
          for (int kk = 0; kk < KK; ++kk) {
            const int k = lane + 64 * kk;
            if (k < lenv[j]) dstv[j][k] = (double)vv[j][kk];
          }
      }
      // b and the Dirichlet lifting, per row
      for (int r = wave; r < T::RS; r += T::NW) {
        const RowT e = rt[r];
        if (e.lenf == 0) continue;
        const int ru1 = r % R1, ru2 = (r / R1) % R2, o = r / (R1 * R2);
        const int len = e.lenf & 0xFFFF, lpid = e.lenf >> 17;
        const uint8_t *inv = PATG ? pat_invr + (size_t)tile_pats[t.pat_off + lpid] * T::SP : PAT + lpid * T::SP;
        const CT *row = slab + slab_row_offset<K, R1, R2>(ru1, ru2, o);
        // b = f part - sum over Dirichlet columns of A_ij uD_j
        double acc = 0.0;
        if (e.lenf & 0x10000) {
          const int nd = inv[T::SP - 1];
          const int rw1 = stencil_width(K, ru1), rw2 = stencil_width(K, ru2);
          for (int j = lane; j < nd; j += 64) {
            const int sl = inv[len + j];
            const int d1 = sl % rw1 + stencil_lo(K, ru1), d2 = (sl / rw1) % rw2 + stencil_lo(K, ru2);
            const int d3 = sl / (rw1 * rw2) + (o == 0 ? -K : -1);
            acc += (double)row[sl] * uD_at(kp, t.org[0] + ru1 + d1, t.org[1] + ru2 + d2, K * c + o + d3);
          }
          acc = wave_sum(acc);
        }
        if (lane == 0) b[e.bidx] = (double)BF[r] - acc;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// launcher.  Tile shapes must match tile_shape() in gdx_symbolic.cpp.
// ---------------------------------------------------------------------------
template <int K, int NQ, int R1, int R2, class CT>
static void runm(gdx_problem *p, hipStream_t st) {
  using T = March<K, NQ, R1, R2, CT>;
  const int nt = (int)p->tiles.size(), nf = p->n_fast_tiles;
  if (nf > 0)
    hipLaunchKernelGGL((k_march3d<K, NQ, R1, R2, false, CT>), dim3(8u * ((nf + 7) / 8)), dim3(T::NT), 0, st,
                       p->dev.tiles, 0, nf, p->dev.rowinfo, p->dev.W, p->dev.F, p->dev.pat_invr, p->dev.tile_pats,
                       p->dev.values, p->dev.b, p->kp);
  if (nt > nf)
    hipLaunchKernelGGL((k_march3d<K, NQ, R1, R2, true, CT>), dim3(8u * ((nt - nf + 7) / 8)), dim3(T::NT), 0, st,
                       p->dev.tiles, nf, nt - nf, p->dev.rowinfo, p->dev.W, p->dev.F, p->dev.pat_invr,
                       p->dev.tile_pats, p->dev.values, p->dev.b, p->kp);
}

int pattern_capacity_3d(int K, int NQ) {
  (void)K;
  (void)NQ;
  return March<2, 3, 4, 8, double>::PMAX;
}

// precision: GDX_PREC_FP64, or GDX_PREC_FP32 for contractions in FP32
int launch_matrix3d(gdx_problem *p, hipStream_t st, int precision) {
  const int K = p->K, NQ = p->NQ;
  const bool f32 = precision == GDX_PREC_FP32;
  if (K == 2 && NQ == 3) f32 ? runm<2, 3, 4, 8, float>(p, st) : runm<2, 3, 4, 8, double>(p, st);
  else if (K == 2 && NQ == 2) f32 ? runm<2, 2, 4, 8, float>(p, st) : runm<2, 2, 4, 8, double>(p, st);
  else if (K == 1 && NQ == 2) f32 ? runm<1, 2, 8, 8, float>(p, st) : runm<1, 2, 8, 8, double>(p, st);
  else return GDX_ERR_UNSUPPORTED;
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

}  // namespace gdx

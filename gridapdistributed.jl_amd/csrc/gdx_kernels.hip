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

#include "gdx_device.h"

namespace gdx {

// ---------------------------------------------------------------------------
// Coefficient and source at every quadrature point of the global lattice.
// a_m exp(i theta_m) = [a_m exp(i phi_m) exp(2 pi i k_m1 x)] exp(2 pi i k_m2 y) exp(2 pi i k_m3 z)
// with the three 1D factors tabulated once per launch.
// ---------------------------------------------------------------------------
__global__ void k_fourier_tables(const double *__restrict__ modes, KParams kp, int qnmax, double *__restrict__ E) {
  const int m = blockIdx.y, d = blockIdx.z;
  const double *M = modes + 5 * m;
  for (int q = threadIdx.x + blockIdx.x * blockDim.x; q < qnmax; q += blockDim.x * gridDim.x) {
    double x = 0.0;
    if (d < kp.dim && q < kp.qn[d]) {
      const int c = q / kp.NQ, t = q % kp.NQ;
      x = kp.x0[d] + kp.h[d] * ((double)c + kp.qx[t]);
    }
    double s, co;
    sincos(2.0 * M_PI * M[d] * x, &s, &co);
    if (d == 0) {  // fold the amplitude and phase into the x factor
      double ps, pc;
      sincos(M[4], &ps, &pc);
      const double cr = M[3] * pc, ci = M[3] * ps;
      const double r = cr * co - ci * s, i = cr * s + ci * co;
      co = r;
      s = i;
    }
    double *e = E + 2 * (((size_t)d * kp.nmodes + m) * qnmax + q);
    e[0] = co;
    e[1] = s;
  }
}

// One block = LINES consecutive x-lines (q2 .. q2+LINES-1 at fixed q3) of the
// quadrature lattice; thread <-> q1.  The x factors of a chunk of modes stay
// in registers across the lines; the per-line (y,z) products b = e2 e3 and
// (k.c) b are tabulated in LDS, so each (point, mode) costs four FMAs:
//   g += Re(e1 b),  s += Im(e1 (k.c) b).
constexpr int kCoefLines = 8, kCoefModes = 16;

__global__ __launch_bounds__(256) void k_coefficients(const double *__restrict__ modes,
                                                      const double *__restrict__ E, KParams kp, int qnmax,
                                                      double *__restrict__ W, double *__restrict__ F) {
  __shared__ double tb[kCoefLines][kCoefModes][4];  // br, bi, (k.c) br, (k.c) bi
  const int q1 = blockIdx.x * blockDim.x + threadIdx.x;
  const int q2b = blockIdx.y * kCoefLines, q3 = blockIdx.z;
  const bool active = q1 < kp.qn[0];
  const int nm = kp.nmodes;
  double g[kCoefLines], s[kCoefLines];
#pragma unroll
  for (int l = 0; l < kCoefLines; ++l) g[l] = s[l] = 0.0;
  for (int m0 = 0; m0 < nm; m0 += kCoefModes) {
    __syncthreads();
    for (int i = threadIdx.x; i < kCoefLines * kCoefModes; i += blockDim.x) {
      const int l = i / kCoefModes, j = i % kCoefModes, m = m0 + j;
      const int q2 = q2b + l < kp.qn[1] ? q2b + l : kp.qn[1] - 1;
      double br = 0.0, bi = 0.0, kc = 0.0;
      if (m < nm) {
        const double *M = modes + 5 * m;
        const double *e2 = E + 2 * ((size_t)(1 * nm + m) * qnmax + q2);
        const double *e3 = E + 2 * ((size_t)(2 * nm + m) * qnmax + q3);
        br = e2[0] * e3[0] - e2[1] * e3[1];
        bi = e2[0] * e3[1] + e2[1] * e3[0];
        kc = M[0] * kp.dir_coef[1] + M[1] * kp.dir_coef[2] + M[2] * kp.dir_coef[3];
      }
      tb[l][j][0] = br;
      tb[l][j][1] = bi;
      tb[l][j][2] = kc * br;
      tb[l][j][3] = kc * bi;
    }
    __syncthreads();
    double e1r[kCoefModes], e1i[kCoefModes];
#pragma unroll
    for (int j = 0; j < kCoefModes; ++j) {
      const bool ok = active && m0 + j < nm;
      const double *e1 = E + 2 * ((size_t)(ok ? m0 + j : 0) * qnmax + (active ? q1 : 0));
      e1r[j] = ok ? e1[0] : 0.0;
      e1i[j] = ok ? e1[1] : 0.0;
    }
#pragma unroll
    for (int l = 0; l < kCoefLines; ++l) {
#pragma unroll
      for (int j = 0; j < kCoefModes; ++j) {
        g[l] = fma(e1r[j], tb[l][j][0], fma(-e1i[j], tb[l][j][1], g[l]));
        s[l] = fma(e1r[j], tb[l][j][3], fma(e1i[j], tb[l][j][2], s[l]));
      }
    }
  }
  if (!active) return;
#pragma unroll
  for (int l = 0; l < kCoefLines; ++l) {
    const int q2 = q2b + l;
    if (q2 >= kp.qn[1]) break;
    const double kap = kp.coef_const * exp(g[l]);
    const size_t idx = q1 + (size_t)kp.qn[0] * (q2 + (size_t)kp.qn[1] * q3);
    W[idx] = kap;
    // f = -div(kappa grad uD) = -kappa grad(g).c = kappa 2 pi sum_m (k_m.c) a_m sin(theta_m)
    F[idx] = kp.rhs_kind == GDX_RHS_MANUFACTURED ? kap * 2.0 * M_PI * s[l] : kp.rhs_const;
  }
}

// ---------------------------------------------------------------------------
// RHS f-part, sum-factorised: b_i = sum_q f(x_q) prod_d m_d(q; i_d),
// m_d(q; i) = w_q h_d phi_a(q) over the cells containing node i.
// ---------------------------------------------------------------------------
template <int DIM, int K, int NQ, int R1, int R2, int R3>
struct RhsTile {
  static constexpr int C1 = R1 / K + 1, C2 = R2 / K + 1, C3 = DIM == 3 ? R3 / K + 1 : 1;
  static constexpr int Q1 = C1 * NQ, Q2 = C2 * NQ, Q3 = DIM == 3 ? C3 * NQ : 1;
  static constexpr int LF = Q1 * Q2 * Q3, LZ1 = R1 * Q3 * Q2, LZ2 = R1 * R2 * Q3;
  static constexpr int NT = 256;
};


template <int DIM, int K, int NQ, int R1, int R2, int R3>
__global__ __launch_bounds__(256) void k_rhs(const Tile *__restrict__ tiles, int ntiles,
                                             const RowInfo *__restrict__ rowinfo, const double *__restrict__ F,
                                             KParams kp, double *__restrict__ b) {
  using T = RhsTile<DIM, K, NQ, R1, R2, R3>;
  __shared__ double Fl[T::LF], Z1[T::LZ1], Z2[T::LZ2];
  const int ti = tile_of_block(ntiles);
  if (ti >= ntiles) return;
  const Tile t = tiles[ti];
  const int tid = threadIdx.x;
  const int qs0 = (t.org[0] / K - 1) * NQ, qs1 = (t.org[1] / K - 1) * NQ;
  const int qs2 = DIM == 3 ? (t.org[2] / K - 1) * NQ : 0;
  for (int i = tid; i < T::LF; i += T::NT) {
    const int a = i % T::Q1, c = (i / T::Q1) % T::Q2, e = i / (T::Q1 * T::Q2);
    const int q0 = qs0 + a, q1 = qs1 + c, q2 = qs2 + e;
    double v = 0.0;
    if (q0 >= 0 && q0 < kp.qn[0] && q1 >= 0 && q1 < kp.qn[1] && q2 >= 0 && q2 < kp.qn[2])
      v = F[q0 + (size_t)kp.qn[0] * (q1 + (size_t)kp.qn[1] * q2)];
    Fl[i] = v;
  }
  __syncthreads();
  // x: Z1[u1][q3][q2]
  for (int i = tid; i < T::LZ1; i += T::NT) {
    const int u1 = i / (T::Q3 * T::Q2), r = i % (T::Q3 * T::Q2);
    const double *fr = Fl + r * T::Q1;
    const int v = u1 + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
    double acc = 0.0;
    for (int lc = cmin; lc <= cmax; ++lc)
      for (int tt = 0; tt < NQ; ++tt) acc = fma(fr[lc * NQ + tt], node_factor<K, NQ>(kp, 0, u1, lc * NQ + tt), acc);
    Z1[i] = acc;
  }
  __syncthreads();
  // y: Z2[u1][u2][q3]
  for (int i = tid; i < T::LZ2; i += T::NT) {
    const int u1 = i / (R2 * T::Q3), u2 = (i / T::Q3) % R2, q3 = i % T::Q3;
    const double *zr = Z1 + (u1 * T::Q3 + q3) * T::Q2;
    const int v = u2 + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
    double acc = 0.0;
    for (int lc = cmin; lc <= cmax; ++lc)
      for (int tt = 0; tt < NQ; ++tt) acc = fma(zr[lc * NQ + tt], node_factor<K, NQ>(kp, 1, u2, lc * NQ + tt), acc);
    Z2[i] = acc;
  }
  __syncthreads();
  // z and store
  for (int i = tid; i < R1 * R2 * (DIM == 3 ? R3 : 1); i += T::NT) {
    const int u1 = i % R1, u2 = (i / R1) % R2, u3 = i / (R1 * R2);
    if (u1 >= t.ext[0] || u2 >= t.ext[1] || u3 >= t.ext[2]) continue;
    const RowInfo ri = rowinfo[t.rowinfo_base + row_index(t, t.org[0] + u1, t.org[1] + u2, t.org[2] + u3)];
    if (ri.pid < 0) continue;
    const double *zr = Z2 + (u1 * R2 + u2) * T::Q3;
    double acc;
    if (DIM == 3) {
      const int v = u3 + K, cmin = (v + K - 1) / K - 1, cmax = v / K;
      acc = 0.0;
      for (int lc = cmin; lc <= cmax; ++lc)
        for (int tt = 0; tt < NQ; ++tt) acc = fma(zr[lc * NQ + tt], node_factor<K, NQ>(kp, 2, u3, lc * NQ + tt), acc);
    } else {
      acc = zr[0];
    }
    b[ri.bidx] = acc;
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
__global__ __launch_bounds__(256) void k_assemble2d(const Tile *__restrict__ tiles, int ntiles,
                                                    const RowInfo *__restrict__ rowinfo,
                                                    const double *__restrict__ Wq, const uint8_t *__restrict__ pat_fwd,
                                                    const uint8_t *__restrict__ pat_inv, double *__restrict__ values,
                                                    double *__restrict__ b, KParams kp) {
  using T = Tile2<K, NQ, R1, R2>;
  constexpr int NINV = (T::S + 63) / 64;
  __shared__ double lds[T::LDS_DOUBLES];
  double *kap = lds;
  double *XG = lds + T::LDS_KAPPA;
  double *XM = XG + T::NP1 * T::XS;
  double *slab = XM + T::NP1 * T::XS;
  const int ti = tile_of_block(ntiles);
  if (ti >= ntiles) return;
  const Tile t = tiles[ti];
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
  const int wave = tid >> 6, lane = tid & 63;
  for (int rr = wave; rr < R1 * R2; rr += T::NT / 64) {
    const int u1 = rr % R1, u2 = rr / R1;
    if (u1 >= t.ext[0] || u2 >= t.ext[1]) continue;
    const RowInfo ri = rowinfo[t.rowinfo_base + row_index(t, t.org[0] + u1, t.org[1] + u2, t.org[2])];
    if (ri.pid < 0) continue;
    int inv[NINV];
#pragma unroll
    for (int kk = 0; kk < NINV; ++kk) {
      const int k = lane + 64 * kk;
      inv[kk] = k < (ri.lenf & 0xFFFF) ? (int)pat_inv[(size_t)ri.pid * pattern_stride(T::S) + k] : 0;
    }
    write_row<K, 2, T::S>(slab + rr * T::S, ri.rowstart, ri.bidx, ri.pid, ri.lenf, inv, t.org[0] + u1,
                          t.org[1] + u2, 0, pat_fwd, values, b, kp);
  }
}

// ---------------------------------------------------------------------------
// launchers.  Tile shapes here must match tile_shape() in gdx_symbolic.cpp.
// ---------------------------------------------------------------------------
static unsigned grid_for(int ntiles) { return 8u * (unsigned)((ntiles + 7) / 8); }

int launch_coefficients(gdx_problem *p, hipStream_t st) {
  if (p->coef_cell) return launch_cell_coefficients(p, st);
  const KParams &kp = p->kp;
  int qnmax = kp.qn[0];
  for (int d = 1; d < 3; ++d) qnmax = qnmax > kp.qn[d] ? qnmax : kp.qn[d];
  if (kp.nmodes > 0) {
    dim3 g1((qnmax + 255) / 256, kp.nmodes, 3);
    hipLaunchKernelGGL(k_fourier_tables, g1, dim3(256), 0, st, p->dev.modes, kp, qnmax, p->dev.E);
  }
  const int bx = kp.qn[0] < 256 ? 64 * ((kp.qn[0] + 63) / 64) : 256;
  dim3 g2((kp.qn[0] + bx - 1) / bx, (kp.qn[1] + kCoefLines - 1) / kCoefLines, kp.qn[2]);
  hipLaunchKernelGGL(k_coefficients, g2, dim3(bx), 0, st, p->dev.modes, p->dev.E, kp, qnmax, p->dev.W, p->dev.F);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

template <int DIM, int K, int NQ, int R1, int R2, int R3>
static void rhs_run(gdx_problem *p, hipStream_t st) {
  const int nt = (int)p->tiles.size();
  hipLaunchKernelGGL((k_rhs<DIM, K, NQ, R1, R2, R3>), dim3(grid_for(nt)), dim3(256), 0, st, p->dev.tiles, nt,
                     p->dev.rowinfo, p->dev.F, p->kp, p->dev.b);
}

int launch_matrix3d(gdx_problem *p, hipStream_t st, int precision);

template <int K, int NQ, int R1, int R2>
static void run2(gdx_problem *p, hipStream_t st) {
  const int nt = (int)p->tiles.size();
  hipLaunchKernelGGL((k_assemble2d<K, NQ, R1, R2>), dim3(grid_for(nt)), dim3(256), 0, st, p->dev.tiles, nt,
                     p->dev.rowinfo, p->dev.W, p->dev.pat_fwd, p->dev.pat_inv, p->dev.values, p->dev.b, p->kp);
}

bool matrix_kernel_supported(int dim, int K, int NQ) {
  if (dim == 3) return (K == 2 && (NQ == 2 || NQ == 3)) || (K == 1 && NQ == 2);
  return dim == 2 && (K == 1 || K == 2) && (NQ == 2 || NQ == 3);
}

int launch_rhs(gdx_problem *p, hipStream_t st) {
  if (p->tiles.empty()) return GDX_OK;
  const int K = p->K, NQ = p->NQ;
  if (p->D == 3) {
    if (p->kvariant3d != 1) return GDX_OK;  // the z-marching and direct-store kernels write b themselves
    if (K == 2 && NQ == 3) rhs_run<3, 2, 3, 4, 8, 8>(p, st);
    else if (K == 2 && NQ == 2) rhs_run<3, 2, 2, 4, 8, 8>(p, st);
    else if (K == 1 && NQ == 2) rhs_run<3, 1, 2, 8, 8, 2>(p, st);
    else return GDX_ERR_UNSUPPORTED;
  } else {
    if (K == 2 && NQ == 3) rhs_run<2, 2, 3, 16, 16, 1>(p, st);
    else if (K == 2 && NQ == 2) rhs_run<2, 2, 2, 16, 16, 1>(p, st);
    else if (K == 1 && NQ == 2) rhs_run<2, 1, 2, 16, 16, 1>(p, st);
    else if (K == 1 && NQ == 3) rhs_run<2, 1, 3, 16, 16, 1>(p, st);
    else return GDX_ERR_UNSUPPORTED;
  }
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int launch_matrix(gdx_problem *p, hipStream_t st) {
  if (p->tiles.empty()) return GDX_OK;
  const int K = p->K, NQ = p->NQ;
  if (p->D == 3) {
    // the FP32 form exists for the z-marching and the direct-store kernels
    if (p->precision == GDX_PREC_FP32 && p->kvariant3d == 0) return launch_matrix3d(p, st, GDX_PREC_FP32);
    if (p->kvariant3d == 2) return launch_matrix_tile3d_direct(p, st);
    if (p->kvariant3d == 3) return launch_matrix_tile3d_wslab(p, st);
    return p->kvariant3d == 1 ? launch_matrix_tile3d(p, st) : launch_matrix3d(p, st, GDX_PREC_FP64);
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

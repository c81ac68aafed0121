// Learned two-grid preconditioner of libgdx for 3D Q1 problems on a cube of
// 2n x 2n x 2n cells with Dirichlet data on the whole boundary:
//
//   z  = w D^-1 r                          damped Jacobi pre-smoothing
//   rc = R (r - A z)                       full weighting (R = P^T, trilinear P)
//   ec = s N( D^-1 rc / s, log-diagonal )  network coarse solve, s = rms(D^-1 rc)
//   ec += w Dc^-1 (rc - Ac ec), k times     damped-Jacobi sweeps on the Galerkin
//                                          coarse operator Ac = P^T A P (27-point
//                                          stencil computed from the fine matrix)
//   z += a P ec, a = (P ec).s / (P ec).A(P ec) with s = r - A z: the
//                                          energy-optimal step along the correction
//   z += w D^-1 (r - A z)                  damped Jacobi post-smoothing (A z from
//                                          the two products above: no third SpMV)
//
// The coarse unknowns are the fine nodes with even indices; the (n-1)^3
// interior ones sit on the network's n^3 grid (the last plane is zero
// padding).  The network is trained on device from the problem's own matrix:
// a smooth random coarse error ec (white noise smoothed by repeated 3x3x3
// means), r = A P ec, target ec for the input built from r exactly as above
// (loss 1 - cos^2(N, ec): the preconditioner rescales the network's output to
// the coarse energy optimum, so only its direction matters).  Every reduction
// has a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "gdx_nn.h"

// Staging of r and z for gdx_nn_apply_noalloc: one file-scope array per GPU
// (each device that loads the library holds its own copy until the process
// ends).  Host code never passes the array's name to a kernel or a copy: it
// uses the per-device address that noalloc_setup obtains.
static constexpr size_t kNoAllocDoubles = (size_t)4 << 20;  // 32 MiB: r and z of up to 2,097,152 dofs
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace gdx {

namespace {

constexpr int kB = 256;
constexpr int kRed = 256;  // partial sums per reduction
constexpr int kS = 3 * kRed;  // scalar slots of 'red' after three partial-sum areas
constexpr double kOmega = 2.0 / 3.0;
constexpr double kOmegaC = 2.0 / 3.0;  // coarse sweeps

// A structured coarse level: m cells per direction, unknowns at the interior
// nodes 1..m-1 stored at [0, m-1) of an m^3 array (the last plane is padding).
struct Level {
  int m = 0;
  double *cst = nullptr;  // Galerkin operator as a 27-point stencil [27][m^3]
  bool own_cst = true;    // false for the network grid (its stencil is TwoGrid::cst)
  double *rhs = nullptr, *e = nullptr, *e2 = nullptr, *res = nullptr;  // [m^3]
};

struct TwoGrid {
  gdx_problem *prob = nullptr;
  int n = 0;       // network grid = coarse cells per direction
  int nf = 0;      // fine lattice extent per direction (2n + 1)
  int ng = 0;      // free dofs
  long long *gid_node = nullptr;
  double *flat = nullptr, *dlat = nullptr;   // fine lattice vectors
  double *t1 = nullptr, *t2 = nullptr, *t3 = nullptr;  // global-id temporaries
  double *rho = nullptr;                     // coarse scaled residual [n^3]
  double *red = nullptr;                     // reduction partials + scalars
  double *ec = nullptr, *ec2 = nullptr;      // training: coarse error [n^3]
  float *tgt = nullptr;
  const double *dinv = nullptr;
  double *fst = nullptr;   // fine matrix as a 27-point stencil on the lattice [27][nf^3]
  double *cst = nullptr;   // Galerkin coarse operator P^T A P as a stencil [27][n^3]
  double *rc = nullptr;    // coarse residual R s [n^3]
  double *ca = nullptr, *cb = nullptr;  // coarse correction (ping-pong) [n^3]
  int sweeps = 8;          // coarse damped-Jacobi sweeps after the network (two-grid form)
  bool use_net = true;     // false: the coarse correction starts from zero
  // multigrid form: a V-cycle from the network grid down to a grid of <= 8^3
  // cells replaces the coarse sweeps; lv[0] is the network grid
  bool mg = true;
  int nu = 1;              // damped-Jacobi sweeps before and after each coarse-grid correction
  int coarsest_sweeps = 24;
  std::vector<Level> lv;
};

TwoGrid *tg_of(gdx_nn *nn) { return reinterpret_cast<TwoGrid *>(nn->tg); }
inline unsigned nb(int64_t n) { return (unsigned)((n + kB - 1) / kB); }

__global__ void k_scatter(int ng, const long long *__restrict__ gn, const double *__restrict__ x, double *__restrict__ lat) {
  const int g = blockIdx.x * kB + threadIdx.x;
  if (g < ng) lat[gn[g]] = x[g];
}
__global__ void k_gather(int ng, const long long *__restrict__ gn, const double *__restrict__ lat, double *__restrict__ x) {
  const int g = blockIdx.x * kB + threadIdx.x;
  if (g < ng) x[g] = lat[gn[g]];
}

// rho[j] = dlat(2I) * sum_{|o|<=1} w(o) r(2I + o), I = j + 1, for interior coarse nodes; 0 on the padding plane
__global__ void k_restrict(int n, int nf, const double *__restrict__ flat, const double *__restrict__ dlat,
                           double *__restrict__ rho, double *__restrict__ rcoarse) {
  const int64_t tot = (int64_t)n * n * n;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  const int jx = (int)(j % n), jy = (int)((j / n) % n), jz = (int)(j / ((int64_t)n * n));
  if (jx == n - 1 || jy == n - 1 || jz == n - 1) {
    rho[j] = 0.0;
    rcoarse[j] = 0.0;
    return;
  }
  const int fx = 2 * (jx + 1), fy = 2 * (jy + 1), fz = 2 * (jz + 1);
  double acc = 0.0;
  for (int oz = -1; oz <= 1; ++oz)
    for (int oy = -1; oy <= 1; ++oy)
      for (int ox = -1; ox <= 1; ++ox) {
        const double w = (ox ? 0.5 : 1.0) * (oy ? 0.5 : 1.0) * (oz ? 0.5 : 1.0);
        acc = fma(w, flat[(fx + ox) + (int64_t)nf * ((fy + oy) + (int64_t)nf * (fz + oz))], acc);
      }
  rcoarse[j] = acc;
  rho[j] = acc * dlat[fx + (int64_t)nf * (fy + (int64_t)nf * fz)];
}

// fine matrix rows -> 27-point stencil on the lattice: fst[slot][node], slot of
// the column's offset (dx+1) + 3(dy+1) + 9(dz+1); Dirichlet columns are absent
// from the free x free matrix and stay zero (fst is cleared before).
//
// kFsRows consecutive rows per block.  When their entries are stored back to
// back, the block reads them coalesced (one entry per thread, its row found by
// bisection over the block's row starts), places each value in an LDS tile
// [slot][row] and then writes the tile plane by plane, so consecutive lanes
// store consecutive lattice nodes.  Other blocks walk their rows one thread
// each.  The slot comes from the node difference: b - a + (1 + nf + nf^2) has
// base-nf digits (dx+1, dy+1, dz+1) in 0..2, found by comparisons.
constexpr int kFsRows = 64;
__device__ __forceinline__ int stencil_slot(long long d, int nf) {
  const long long nf2 = (long long)nf * nf;
  long long t = d + 1 + nf + nf2;
  if (t < 0 || t >= 3 * nf2) return -1;
  const int dz1 = t >= 2 * nf2 ? 2 : (t >= nf2 ? 1 : 0);
  t -= dz1 * nf2;
  const int dy1 = t >= 2 * (long long)nf ? 2 : (t >= nf ? 1 : 0);
  t -= (long long)dy1 * nf;
  return t > 2 ? -1 : (int)t + 3 * dy1 + 9 * dz1;
}
__global__ __launch_bounds__(kB) void k_fine_stencil(int ng, int nf, const long long *__restrict__ rstart,
                                                     const int *__restrict__ rlen, const int *__restrict__ cgid,
                                                     const double *__restrict__ val,
                                                     const long long *__restrict__ gn, double *__restrict__ fst) {
  constexpr int CAP = 27 * kFsRows;
  __shared__ double tile[27][kFsRows];
  __shared__ long long rs[kFsRows], an[kFsRows];
  __shared__ int rl[kFsRows];
  __shared__ unsigned int present[kFsRows];
  const int t = threadIdx.x, g0 = blockIdx.x * kFsRows;
  const int nr = min(kFsRows, ng - g0);
  if (nr <= 0) return;  // uniform over the block
  const int64_t nfl = (int64_t)nf * nf * nf;
  if (t < nr) {
    rs[t] = rstart[g0 + t];
    rl[t] = rlen[g0 + t];
    an[t] = gn[g0 + t];
    present[t] = 0u;
  }
  __syncthreads();
  int ok = 1;
  if (t < nr - 1) ok = rs[t] + rl[t] == rstart[g0 + t + 1];
  const long long A = rs[0], B = rs[nr - 1] + rl[nr - 1];
  ok = __syncthreads_and(ok && (B - A) <= CAP);
  if (!ok) {
    if (t < nr) {
      const long long a = an[t], s0 = rs[t];
      for (int k = 0; k < rl[t]; ++k) {
        const int sl = stencil_slot(gn[cgid[s0 + k]] - a, nf);
        if (sl >= 0) fst[(int64_t)sl * nfl + a] = val[s0 + k];
      }
    }
    return;
  }
  const int m = (int)(B - A);
  // the thread's entries k = t + kB i: column ids and values, then their
  // lattice nodes, each group of loads issued together
  constexpr int IT = (CAP + kB - 1) / kB;
  int cg[IT];
  double vv[IT];
  long long gg[IT];
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int k = t + kB * i;
    cg[i] = k < m ? cgid[A + k] : 0;
    vv[i] = k < m ? val[A + k] : 0.0;
  }
#pragma unroll
  for (int i = 0; i < IT; ++i) gg[i] = t + kB * i < m ? gn[cg[i]] : 0;
#pragma unroll
  for (int i = 0; i < IT; ++i) {
    const int k = t + kB * i;
    if (k >= m) break;
    int lo = 0, hi = nr - 1;  // the row holding entry k: the last r with rs[r] - A <= k
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (rs[mid] - A <= k) lo = mid;
      else hi = mid - 1;
    }
    const int sl = stencil_slot(gg[i] - an[lo], nf);
    if (sl >= 0) {
      tile[sl][lo] = vv[i];
      atomicOr(&present[lo], 1u << sl);
    }
  }
  __syncthreads();
  for (int i = t; i < 27 * nr; i += kB) {
    const int sl = i / nr, r = i - sl * nr;
    if (present[r] >> sl & 1u) fst[(int64_t)sl * nfl + an[r]] = tile[sl][r];
  }
}

// Galerkin product, one thread per output entry (coarse node J, offset d):
//   Ac(J, J+d) = sum_{a = 2J+oa, b = 2(J+d)+ob} w(oa) A(a, b) w(ob)
// Per direction only the (oa, ob) with |2d + ob - oa| <= 1 meet the 27-point
// stencil of A: 7 pairs for d = 0, 3 for d = +-1, summed in a fixed order.
// The fine operator is a stencil [27][.] on a grid of stride S whose node I
// sits at index I - off (off = 0: the fine lattice, 1: a padded coarse level).
__constant__ signed char kGalPairs[3][7][2] = {
    {{-1, 0}, {-1, 1}, {0, 1}, {0, 0}, {0, 0}, {0, 0}, {0, 0}},                  // d = -1
    {{-1, -1}, {-1, 0}, {0, -1}, {0, 0}, {0, 1}, {1, 0}, {1, 1}},                // d = 0
    {{0, -1}, {1, -1}, {1, 0}, {0, 0}, {0, 0}, {0, 0}, {0, 0}}};                 // d = +1
// L lanes per output entry (a power of two <= 64, so a group sits in one wave)
// split its (oa, ob) pairs, flattened z-major, and sum them with a fixed
// shuffle tree; with L = 1 the order is the nested loops' order.  Small levels
// use many lanes: one lane's chain of up to 7^3 dependent loads would bound them.
template <int L>
__global__ __launch_bounds__(kB) void k_galerkin_e(int mc, int S, int off, int64_t fine_tot,
                                                   const double *__restrict__ af, double *__restrict__ cst) {
  const int64_t tot = (int64_t)mc * mc * mc;
  const int64_t tq = blockIdx.x * (int64_t)kB + threadIdx.x;
  const int64_t t = tq / L;
  const int sub = (int)(tq % L);
  if (t >= 27 * tot) return;  // whole lane groups: kB is a multiple of L
  const int k = (int)(t / tot);  // output slot: consecutive outputs share it (coalesced stores)
  const int64_t j = t - k * tot;
  const int jx = (int)(j % mc), jy = (int)((j / mc) % mc), jz = (int)(j / ((int64_t)mc * mc));
  const int dx = k % 3 - 1, dy = (k / 3) % 3 - 1, dz = k / 9 - 1;
  const int Jx = jx + 1, Jy = jy + 1, Jz = jz + 1;
  double acc = 0.0;
  const bool live = jx < mc - 1 && jy < mc - 1 && jz < mc - 1 && Jx + dx >= 1 && Jy + dy >= 1 && Jz + dz >= 1 &&
                    Jx + dx <= mc - 1 && Jy + dy <= mc - 1 && Jz + dz <= mc - 1;
  if (live && L == 1) {
    const int nx = dx ? 3 : 7, ny = dy ? 3 : 7, nz = dz ? 3 : 7;
    // y and x unrolled to their largest count (7) with the counts as
    // predicates, so that a z plane's loads are issued together; the sums
    // keep the nested loops' order
    for (int pz = 0; pz < nz; ++pz) {
      const int oaz = kGalPairs[dz + 1][pz][0], obz = kGalPairs[dz + 1][pz][1];
      const double wz = (oaz ? 0.5 : 1.0) * (obz ? 0.5 : 1.0);
      const int ez = 2 * dz + obz - oaz + 1;
      const int64_t az = 2 * Jz + oaz - off;
#pragma unroll
      for (int py = 0; py < 7; ++py) {
        if (py >= ny) break;
        const int oay = kGalPairs[dy + 1][py][0], oby = kGalPairs[dy + 1][py][1];
        const double wyz = wz * (oay ? 0.5 : 1.0) * (oby ? 0.5 : 1.0);
        const int ey = 2 * dy + oby - oay + 1;
        const int64_t ayz = (2 * Jy + oay - off) + (int64_t)S * az;
#pragma unroll
        for (int px = 0; px < 7; ++px) {
          if (px >= nx) break;
          const int oax = kGalPairs[dx + 1][px][0], obx = kGalPairs[dx + 1][px][1];
          const double w = wyz * (oax ? 0.5 : 1.0) * (obx ? 0.5 : 1.0);
          const int ex = 2 * dx + obx - oax + 1;
          const int64_t a = (2 * Jx + oax - off) + (int64_t)S * ayz;
          acc = fma(w, af[(int64_t)(ex + 3 * ey + 9 * ez) * fine_tot + a], acc);
        }
      }
    }
  } else if (live) {
    const int nx = dx ? 3 : 7, ny = dy ? 3 : 7, nz = dz ? 3 : 7;
    const int n = nx * ny * nz;
    for (int q = sub; q < n; q += L) {
      const int px = q % nx, py = (q / nx) % ny, pz = q / (nx * ny);
      const int oaz = kGalPairs[dz + 1][pz][0], obz = kGalPairs[dz + 1][pz][1];
      const int oay = kGalPairs[dy + 1][py][0], oby = kGalPairs[dy + 1][py][1];
      const int oax = kGalPairs[dx + 1][px][0], obx = kGalPairs[dx + 1][px][1];
      const double wz = (oaz ? 0.5 : 1.0) * (obz ? 0.5 : 1.0);
      const double wyz = wz * (oay ? 0.5 : 1.0) * (oby ? 0.5 : 1.0);
      const double w = wyz * (oax ? 0.5 : 1.0) * (obx ? 0.5 : 1.0);
      const int ex = 2 * dx + obx - oax + 1, ey = 2 * dy + oby - oay + 1, ez = 2 * dz + obz - oaz + 1;
      const int64_t az = 2 * Jz + oaz - off;
      const int64_t ayz = (2 * Jy + oay - off) + (int64_t)S * az;
      const int64_t a = (2 * Jx + oax - off) + (int64_t)S * ayz;
      acc = fma(w, af[(int64_t)(ex + 3 * ey + 9 * ez) * fine_tot + a], acc);
    }
  }
#pragma unroll
  for (int o = L / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, L);
  if (sub == 0) cst[k * tot + j] = acc;
}

// Galerkin product launch: lanes per output so that a level has about 2^20 threads
void galerkin(int mc, int S, int off, int64_t fine_tot, const double *af, double *cst, hipStream_t st) {
  const int64_t outs = 27 * (int64_t)mc * mc * mc;
  int L = 1;
  while (L < 64 && outs * L < (1 << 20)) L *= 2;
  const dim3 g((unsigned)((outs * L + kB - 1) / kB)), b(kB);
  switch (L) {
    case 1: hipLaunchKernelGGL(k_galerkin_e<1>, g, b, 0, st, mc, S, off, fine_tot, af, cst); break;
    case 2: hipLaunchKernelGGL(k_galerkin_e<2>, g, b, 0, st, mc, S, off, fine_tot, af, cst); break;
    case 4: hipLaunchKernelGGL(k_galerkin_e<4>, g, b, 0, st, mc, S, off, fine_tot, af, cst); break;
    case 8: hipLaunchKernelGGL(k_galerkin_e<8>, g, b, 0, st, mc, S, off, fine_tot, af, cst); break;
    case 16: hipLaunchKernelGGL(k_galerkin_e<16>, g, b, 0, st, mc, S, off, fine_tot, af, cst); break;
    case 32: hipLaunchKernelGGL(k_galerkin_e<32>, g, b, 0, st, mc, S, off, fine_tot, af, cst); break;
    default: hipLaunchKernelGGL(k_galerkin_e<64>, g, b, 0, st, mc, S, off, fine_tot, af, cst); break;
  }
}

// ec = s * out (network units -> coarse error), 0 on the padding plane; or 0 everywhere
__global__ void k_out_to_coarse(int n, const float *__restrict__ out, const double *__restrict__ s, int use,
                                double *__restrict__ ec) {
  const int64_t tot = (int64_t)n * n * n;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  const int jx = (int)(j % n), jy = (int)((j / n) % n), jz = (int)(j / ((int64_t)n * n));
  ec[j] = (use && jx < n - 1 && jy < n - 1 && jz < n - 1) ? (double)out[j] * *s : 0.0;
}

// b = Ac a on the network grid (padding plane 0)
__global__ void k_coarse_apply(int n, const double *__restrict__ cst, const double *__restrict__ a,
                               double *__restrict__ b) {
  const int64_t tot = (int64_t)n * n * n;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  const int jx = (int)(j % n), jy = (int)((j / n) % n), jz = (int)(j / ((int64_t)n * n));
  double r = 0.0;
  if (jx < n - 1 && jy < n - 1 && jz < n - 1)
    for (int k = 0; k < 27; ++k) {
      const int x = jx + k % 3 - 1, y = jy + (k / 3) % 3 - 1, z = jz + k / 9 - 1;
      if (x < 0 || y < 0 || z < 0 || x > n - 2 || y > n - 2 || z > n - 2) continue;
      r = fma(cst[k * tot + j], a[x + (int64_t)n * (y + (int64_t)n * z)], r);
    }
  b[j] = r;
}

__global__ void k_scale_by(int64_t tot, const double *__restrict__ alpha, double *__restrict__ a) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < tot) a[i] *= *alpha;
}

// partial sums of out.tgt, out.out, tgt.tgt (three areas of kRed)
__global__ __launch_bounds__(kB) void k_dot3f(int64_t tot, const float *__restrict__ a, const float *__restrict__ b,
                                             double *__restrict__ part) {
  __shared__ double s1[kB], s2[kB], s3[kB];
  double p1 = 0.0, p2 = 0.0, p3 = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < tot; i += (int64_t)gridDim.x * kB) {
    const double x = a[i], y = b[i];
    p1 = fma(x, y, p1);
    p2 = fma(x, x, p2);
    p3 = fma(y, y, p3);
  }
  s1[threadIdx.x] = p1;
  s2[threadIdx.x] = p2;
  s3[threadIdx.x] = p3;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
      s3[threadIdx.x] += s3[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1[0];
    part[kRed + blockIdx.x] = s2[0];
    part[2 * kRed + blockIdx.x] = s3[0];
  }
}

// sums of the three partial areas into out[0..3), and the loss
// 1 - (N.t)^2 / (|N|^2 |t|^2) into out[3]
__global__ __launch_bounds__(kB) void k_finish3(const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double s1[kB], s2[kB], s3[kB];
  double p1 = 0.0, p2 = 0.0, p3 = 0.0;
  for (int i = threadIdx.x; i < kRed; i += kB) {
    p1 += part[i];
    p2 += part[kRed + i];
    p3 += part[2 * kRed + i];
  }
  s1[threadIdx.x] = p1;
  s2[threadIdx.x] = p2;
  s3[threadIdx.x] = p3;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
      s3[threadIdx.x] += s3[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s1[0];
    out[1] = s2[0] > 0.0 ? s2[0] : 1e-300;
    out[2] = s3[0] > 0.0 ? s3[0] : 1e-300;
    out[3] = 1.0 - s1[0] * s1[0] / (out[1] * out[2]);
  }
}

// gradient of 1 - (N.t)^2 / (|N|^2 |t|^2) with respect to N:
// -2 (N.t) / (|N|^2 |t|^2) (t - (N.t) / |N|^2 N)
__global__ void k_cos_grad(int64_t tot, const float *__restrict__ out, const float *__restrict__ tgt,
                           const double *__restrict__ sums, float *__restrict__ dout) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i >= tot) return;
  const double nt = sums[0], nn = sums[1], tt = sums[2];
  dout[i] = (float)(-2.0 * nt / (nn * tt) * ((double)tgt[i] - nt / nn * (double)out[i]));
}

// one damped-Jacobi sweep on the coarse system: b = a + w Dc^-1 (rc - Ac a)
// r0 - sum over the interior neighbours of cst[k][j] x[neighbour k], in stencil
// order.  The 54 loads are predicated (zero outside the interior) and issued
// before the subtractions, so a small grid is not one chain of dependent
// loads; a skipped term subtracts 0 * 0, which leaves r unchanged exactly.
__device__ __forceinline__ double stencil_residual(double r0, int n, int64_t tot, int64_t j, int jx, int jy, int jz,
                                                   const double *__restrict__ cst,
                                                   const double *__restrict__ x) {
  double cv[27], xv[27];
#pragma unroll
  for (int k = 0; k < 27; ++k) {
    const int ox = k % 3 - 1, oy = (k / 3) % 3 - 1, oz = k / 9 - 1;
    const int X = jx + ox, Y = jy + oy, Z = jz + oz;
    const bool in = X >= 0 && Y >= 0 && Z >= 0 && X <= n - 2 && Y <= n - 2 && Z <= n - 2;
    cv[k] = in ? cst[k * tot + j] : 0.0;
    xv[k] = in ? x[j + ox + (int64_t)n * (oy + (int64_t)n * oz)] : 0.0;
  }
  double r = r0;
#pragma unroll
  for (int k = 0; k < 27; ++k) r -= cv[k] * xv[k];
  return r;
}

__global__ void k_coarse_sweep(int n, double w, const double *__restrict__ cst, const double *__restrict__ rc,
                               const double *__restrict__ a, double *__restrict__ b) {
  const int64_t tot = (int64_t)n * n * n;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  const int jx = (int)(j % n), jy = (int)((j / n) % n), jz = (int)(j / ((int64_t)n * n));
  if (jx == n - 1 || jy == n - 1 || jz == n - 1) {
    b[j] = 0.0;
    return;
  }
  const double r = stencil_residual(rc[j], n, tot, j, jx, jy, jz, cst, a);
  b[j] = a[j] + w * r / cst[13 * tot + j];
}

// fine lattice value of P ec (double coarse field on the network grid)
__global__ void k_prolong_d(int n, int nf, const double *__restrict__ ec, double *__restrict__ flat) {
  const int64_t tot = (int64_t)nf * nf * nf;
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i >= tot) return;
  const int x = (int)(i % nf), y = (int)((i / nf) % nf), z = (int)(i / ((int64_t)nf * nf));
  if (x == 0 || y == 0 || z == 0 || x == nf - 1 || y == nf - 1 || z == nf - 1) {
    flat[i] = 0.0;
    return;
  }
  double acc = 0.0;
  for (int kz = 0; kz < 1 + (z & 1); ++kz)
    for (int ky = 0; ky < 1 + (y & 1); ++ky)
      for (int kx = 0; kx < 1 + (x & 1); ++kx) {
        const int Ix = (x >> 1) + kx, Iy = (y >> 1) + ky, Iz = (z >> 1) + kz;
        if (Ix < 1 || Iy < 1 || Iz < 1 || Ix > n - 1 || Iy > n - 1 || Iz > n - 1) continue;
        const double w = ((x & 1) ? 0.5 : 1.0) * ((y & 1) ? 0.5 : 1.0) * ((z & 1) ? 0.5 : 1.0);
        acc = fma(w, ec[(Ix - 1) + (int64_t)n * ((Iy - 1) + (int64_t)n * (Iz - 1))], acc);
      }
  flat[i] = acc;
}

// ---------------------------------------------------------------------------
// multigrid on the structured coarse levels (padded interior storage)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool padding(int64_t j, int m, int &x, int &y, int &z) {
  x = (int)(j % m);
  y = (int)((j / m) % m);
  z = (int)(j / ((int64_t)m * m));
  return x == m - 1 || y == m - 1 || z == m - 1;
}

// res = rhs - A e
__global__ void k_resid(int m, const double *__restrict__ cst, const double *__restrict__ rhs,
                        const double *__restrict__ e, double *__restrict__ res) {
  const int64_t tot = (int64_t)m * m * m;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  int jx, jy, jz;
  if (padding(j, m, jx, jy, jz)) {
    res[j] = 0.0;
    return;
  }
  res[j] = stencil_residual(rhs[j], m, tot, j, jx, jy, jz, cst, e);
}

// first damped-Jacobi sweep from zero: e = w Dc^-1 rhs
__global__ void k_sweep0(int m, double w, const double *__restrict__ cst, const double *__restrict__ rhs,
                         double *__restrict__ e) {
  const int64_t tot = (int64_t)m * m * m;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  int jx, jy, jz;
  e[j] = padding(j, m, jx, jy, jz) ? 0.0 : w * rhs[j] / cst[13 * tot + j];
}

// full weighting from level mf to mf/2: rc[J] = sum_o w(o) r[2J + o]
__global__ void k_restrict_c(int mf, const double *__restrict__ rf, double *__restrict__ rc) {
  const int mc = mf / 2;
  const int64_t tot = (int64_t)mc * mc * mc;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  int jx, jy, jz;
  if (padding(j, mc, jx, jy, jz)) {
    rc[j] = 0.0;
    return;
  }
  const int fx = 2 * (jx + 1) - 1, fy = 2 * (jy + 1) - 1, fz = 2 * (jz + 1) - 1;  // storage of fine node 2J
  double acc = 0.0;
  for (int oz = -1; oz <= 1; ++oz)
    for (int oy = -1; oy <= 1; ++oy)
      for (int ox = -1; ox <= 1; ++ox) {
        const double w = (ox ? 0.5 : 1.0) * (oy ? 0.5 : 1.0) * (oz ? 0.5 : 1.0);
        acc = fma(w, rf[(fx + ox) + (int64_t)mf * ((fy + oy) + (int64_t)mf * (fz + oz))], acc);
      }
  rc[j] = acc;
}

// ef += P ec (trilinear, from level mf/2 to mf)
__global__ void k_prolong_add_c(int mf, const double *__restrict__ ec, double *__restrict__ ef) {
  const int mc = mf / 2;
  const int64_t tot = (int64_t)mf * mf * mf;
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i >= tot) return;
  int ix, iy, iz;
  if (padding(i, mf, ix, iy, iz)) return;
  const int x = ix + 1, y = iy + 1, z = iz + 1;  // fine node
  double acc = 0.0;
  for (int kz = 0; kz < 1 + (z & 1); ++kz)
    for (int ky = 0; ky < 1 + (y & 1); ++ky)
      for (int kx = 0; kx < 1 + (x & 1); ++kx) {
        const int Jx = (x >> 1) + kx, Jy = (y >> 1) + ky, Jz = (z >> 1) + kz;
        if (Jx < 1 || Jy < 1 || Jz < 1 || Jx > mc - 1 || Jy > mc - 1 || Jz > mc - 1) continue;
        const double w = ((x & 1) ? 0.5 : 1.0) * ((y & 1) ? 0.5 : 1.0) * ((z & 1) ? 0.5 : 1.0);
        acc = fma(w, ec[(Jx - 1) + (int64_t)mc * ((Jy - 1) + (int64_t)mc * (Jz - 1))], acc);
      }
  ef[i] += acc;
}

// coarsest level (m <= 8): damped-Jacobi sweeps from zero in one workgroup;
// the stencil is staged in LDS when it fits (m <= 6).  L lanes per grid point
// (the largest power of two with L * m^3 <= 1024) split the point's 27
// stencil terms and sum them with a fixed shuffle tree, so a sweep is not one
// lane's chain of 27 dependent LDS reads.
constexpr int kCoarsestMax = 8;
constexpr int kCoarsestLds = 6 * 6 * 6;
constexpr int kCoarsestThreads = 1024;
__host__ __device__ constexpr int coarsest_lanes(int tot) {
  int L = 32;
  while (L > 1 && L * tot > kCoarsestThreads) L >>= 1;
  return L;
}

template <int L>
__global__ __launch_bounds__(kCoarsestThreads) void k_coarsest(int m, double w, int sweeps,
                                                               const double *__restrict__ cstg,
                                                               const double *__restrict__ rhs,
                                                               double *__restrict__ e) {
  __shared__ double ea[512], eb[512], cl[27 * kCoarsestLds];
  const int tot = m * m * m;
  const int t = threadIdx.x, j = t / L, sub = t % L;  // the L lanes of a point share one wave
  int jx = 0, jy = 0, jz = 0;
  const bool live = j < tot && !padding(j, m, jx, jy, jz);
  const bool staged = tot <= kCoarsestLds;
  if (staged)
    for (int i = t; i < 27 * tot; i += blockDim.x) cl[i] = cstg[i];
  const double *cst = staged ? cl : cstg;
  if (sub == 0 && j < tot) ea[j] = 0.0;
  // the right-hand side and the diagonal are read once, not per sweep
  const double rj = live ? rhs[j] : 0.0;
  __syncthreads();
  const double dj = live ? cst[13 * tot + j] : 1.0;
  // this lane's stencil slots inside the interior, as a bit mask
  uint32_t nbr = 0;
  for (int k = sub; k < 27; k += L) {
    const int x = jx + k % 3 - 1, y = jy + (k / 3) % 3 - 1, z = jz + k / 9 - 1;
    if (live && x >= 0 && y >= 0 && z >= 0 && x <= m - 2 && y <= m - 2 && z <= m - 2) nbr |= 1u << k;
  }
  // this lane's stencil coefficients and neighbour offsets (slots k = sub + L i),
  // the same every sweep: read once into registers
  constexpr int KL = (27 + L - 1) / L;
  double cf[KL];
  int of[KL];
#pragma unroll
  for (int i = 0; i < KL; ++i) {
    const int k = sub + L * i;
    const bool use = k < 27 && (nbr >> k & 1u);
    cf[i] = use ? cst[k * tot + j] : 0.0;
    of[i] = use ? (k % 3 - 1) + m * ((k / 3) % 3 - 1) + m * m * (k / 9 - 1) : 0;
  }
  double *a = ea, *b = eb;
  for (int s = 0; s < sweeps; ++s) {
    double part = 0.0;
#pragma unroll
    for (int i = 0; i < KL; ++i) {
      const int k = sub + L * i;
      if (k >= 27 || !(nbr >> k & 1u)) continue;
      part = fma(cf[i], a[j + of[i]], part);
    }
    for (int o = L / 2; o > 0; o >>= 1) part += __shfl_xor(part, o, L);
    if (sub == 0 && j < tot) b[j] = live ? a[j] + w * (rj - part) / dj : 0.0;
    __syncthreads();
    double *tmp = a;
    a = b;
    b = tmp;
  }
  if (sub == 0 && j < tot) e[j] = a[j];
}

// partial sums of squares (and of a*b) over a grid-stride slice
__global__ __launch_bounds__(kB) void k_sumsq(int64_t n, const double *__restrict__ a, const float *__restrict__ fa,
                                             double *__restrict__ part) {
  __shared__ double sh[kB];
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < n; i += (int64_t)gridDim.x * kB) {
    const double v = a ? a[i] : (double)fa[i];
    s = fma(v, v, s);
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// out = sum of the partials (one block, fixed order); mode 0: out = rms over
// 'count' values, mode 1: out = plain sum
__global__ __launch_bounds__(kB) void k_finish(const double *__restrict__ part, int nparts, double count, int mode,
                                              double *__restrict__ out) {
  __shared__ double sh[kB];
  double s = 0.0;
  for (int i = threadIdx.x; i < nparts; i += kB) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double v = mode == 0 ? sqrt(sh[0] / count) : sh[0];
    if (mode == 0 && !(v > 0.0)) v = 1.0;  // zero residual: any scale
    *out = v;
  }
}

// network input channel 0 = rho / s
__global__ void k_input0(int64_t tot, const double *__restrict__ rho, const double *__restrict__ s, float *__restrict__ inp) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < tot) inp[i] = (float)(rho[i] / *s);
}

// channel 1 = log of the fine diagonal at the coarse node (scaled by h), centred
__global__ void k_input1(int n, int nf, double h, const double *__restrict__ dlat, float *__restrict__ inp1) {
  const int64_t tot = (int64_t)n * n * n;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  const int jx = (int)(j % n), jy = (int)((j / n) % n), jz = (int)(j / ((int64_t)n * n));
  float v = 0.0f;
  if (jx < n - 1 && jy < n - 1 && jz < n - 1) {
    const double d = dlat[2 * (jx + 1) + (int64_t)nf * (2 * (jy + 1) + (int64_t)nf * 2 * (jz + 1))];
    v = (float)(0.5 * (-log(d * h) - log(8.0 / 3.0)));
  }
  inp1[j] = v;
}

// fine lattice value of P ec at interior fine nodes: trilinear from the coarse
// nodes (ec = s * out on interior coarse nodes, 0 on the boundary)
__global__ void k_prolong(int n, int nf, const float *__restrict__ out, const double *__restrict__ s,
                          double *__restrict__ flat) {
  const int64_t tot = (int64_t)nf * nf * nf;
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i >= tot) return;
  const int x = (int)(i % nf), y = (int)((i / nf) % nf), z = (int)(i / ((int64_t)nf * nf));
  if (x == 0 || y == 0 || z == 0 || x == nf - 1 || y == nf - 1 || z == nf - 1) {
    flat[i] = 0.0;
    return;
  }
  const int nc = n;  // coarse cells: coarse nodes 0..nc, interior 1..nc-1
  double acc = 0.0;
  for (int kz = 0; kz < 2; ++kz) {
    const int Iz = (z >> 1) + kz * (z & 1);
    if (kz && !(z & 1)) break;
    for (int ky = 0; ky < 2; ++ky) {
      const int Iy = (y >> 1) + ky * (y & 1);
      if (ky && !(y & 1)) break;
      for (int kx = 0; kx < 2; ++kx) {
        const int Ix = (x >> 1) + kx * (x & 1);
        if (kx && !(x & 1)) break;
        if (Ix < 1 || Iy < 1 || Iz < 1 || Ix > nc - 1 || Iy > nc - 1 || Iz > nc - 1) continue;
        const double w = ((x & 1) ? 0.5 : 1.0) * ((y & 1) ? 0.5 : 1.0) * ((z & 1) ? 0.5 : 1.0);
        acc = fma(w, (double)out[(Ix - 1) + (int64_t)n * ((Iy - 1) + (int64_t)n * (Iz - 1))], acc);
      }
    }
  }
  flat[i] = acc * *s;
}

__global__ void k_jacobi0(int n, double w, const double *__restrict__ dinv, const double *__restrict__ r,
                          double *__restrict__ z) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) z[i] = w * dinv[i] * r[i];
}
// t = r - t (t held A z)
__global__ void k_rminus(int n, const double *__restrict__ r, double *__restrict__ t) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) t[i] = r[i] - t[i];
}
// z += c
__global__ void k_add(int n, double *__restrict__ z, const double *__restrict__ c) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) z[i] += c[i];
}
// z += w D^-1 (r - t), t = A z
__global__ void k_jacobi_post(int n, double w, const double *__restrict__ dinv, const double *__restrict__ r,
                              const double *__restrict__ t, double *__restrict__ z) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) z[i] = fma(w * dinv[i], r[i] - t[i], z[i]);
}

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// white noise on the interior coarse nodes of the network grid (0 on the padding)
__global__ void k_coarse_noise(int n, uint64_t seed, double *__restrict__ e) {
  const int64_t tot = (int64_t)n * n * n;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  const int jx = (int)(j % n), jy = (int)((j / n) % n), jz = (int)(j / ((int64_t)n * n));
  const uint64_t z = mix64(seed * 0x9E3779B97F4A7C15ull + (uint64_t)(j + 1) * 0xD1B54A32D192ED03ull);
  const double u = ((double)(z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  e[j] = (jx < n - 1 && jy < n - 1 && jz < n - 1) ? 2.0 * u - 1.0 : 0.0;
}

// one 3x3x3 mean over the interior coarse nodes (zero Dirichlet outside)
__global__ void k_box(int n, const double *__restrict__ a, double *__restrict__ b) {
  const int64_t tot = (int64_t)n * n * n;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  const int jx = (int)(j % n), jy = (int)((j / n) % n), jz = (int)(j / ((int64_t)n * n));
  if (jx == n - 1 || jy == n - 1 || jz == n - 1) {
    b[j] = 0.0;
    return;
  }
  double s = 0.0;
  for (int oz = -1; oz <= 1; ++oz)
    for (int oy = -1; oy <= 1; ++oy)
      for (int ox = -1; ox <= 1; ++ox) {
        const int x = jx + ox, y = jy + oy, z = jz + oz;
        if (x < 0 || y < 0 || z < 0 || x > n - 2 || y > n - 2 || z > n - 2) continue;
        s += a[x + (int64_t)n * (y + (int64_t)n * z)];
      }
  b[j] = s * (1.0 / 27.0);
}

__global__ void k_set(double *__restrict__ p, double v) { *p = v; }

// partial sums of a.b and c.d (grid-stride slices), part[0..kRed) and part[kRed..2kRed)
__global__ __launch_bounds__(kB) void k_dot2p(int n, const double *__restrict__ a, const double *__restrict__ b,
                                             const double *__restrict__ c, const double *__restrict__ d,
                                             double *__restrict__ part) {
  __shared__ double s1[kB], s2[kB];
  double p1 = 0.0, p2 = 0.0;
  for (int i = blockIdx.x * kB + threadIdx.x; i < n; i += gridDim.x * kB) {
    p1 = fma(a[i], b[i], p1);
    p2 = fma(c[i], d[i], p2);
  }
  s1[threadIdx.x] = p1;
  s2[threadIdx.x] = p2;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1[0];
    part[kRed + blockIdx.x] = s2[0];
  }
}

// alpha = (c.s) / (c.Ac): the energy-optimal step along the coarse correction
// (0 when the correction vanishes or is not finite)
__global__ __launch_bounds__(kB) void k_alpha(const double *__restrict__ part, double *__restrict__ alpha) {
  __shared__ double s1[kB], s2[kB];
  double p1 = 0.0, p2 = 0.0;
  for (int i = threadIdx.x; i < kRed; i += kB) {
    p1 += part[i];
    p2 += part[kRed + i];
  }
  s1[threadIdx.x] = p1;
  s2[threadIdx.x] = p2;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double a = s1[0] / s2[0];
    *alpha = (s2[0] > 0.0 && isfinite(a)) ? a : 0.0;
  }
}

// z += alpha c ; z += w D^-1 (s - alpha Ac)   (r - A z = s - alpha Ac after the coarse step)
__global__ void k_coarse_post(int n, double w, const double *__restrict__ alpha, const double *__restrict__ dinv,
                              const double *__restrict__ s, const double *__restrict__ c,
                              const double *__restrict__ ac, double *__restrict__ z) {
  const int i = blockIdx.x * kB + threadIdx.x;
  if (i < n) {
    const double a = *alpha;
    z[i] = fma(a, c[i], z[i]) + w * dinv[i] * fma(-a, ac[i], s[i]);
  }
}

// ec (double) -> network-output units for the prolongation (float) and the target ec / s
__global__ void k_ec_to_float(int64_t tot, const double *__restrict__ ec, float *__restrict__ outf) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < tot) outf[i] = (float)ec[i];
}
__global__ void k_target(int64_t tot, const double *__restrict__ ec, const double *__restrict__ s, float *__restrict__ tgt) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < tot) tgt[i] = (float)(ec[i] / *s);
}
// sum of squares / rms into a scalar slot of red
void rms(TwoGrid *T, const double *a, const float *fa, int64_t count, double *dst, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_sumsq, dim3(kRed), dim3(kB), 0, st, count, a, fa, T->red);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kB), 0, st, T->red, kRed, (double)count, mode, dst);
}

// network input from a fine residual in global-id order (s into red[kS]);
// with the network off only the coarse residual rc is formed
void make_input(gdx_nn *nn, TwoGrid *T, const double *r, hipStream_t st, bool for_net = true) {
  nn::Net &N = nn->net;
  const int64_t v = (int64_t)T->n * T->n * T->n;
  // restriction reads only interior lattice nodes, all of which the scatter
  // writes: no clearing (and no memset, so this can be captured in a graph)
  hipLaunchKernelGGL(k_scatter, dim3(nb(T->ng)), dim3(kB), 0, st, T->ng, T->gid_node, r, T->flat);
  hipLaunchKernelGGL(k_restrict, dim3(nb(v)), dim3(kB), 0, st, T->n, T->nf, T->flat, T->dlat, T->rho, T->rc);
  if (!for_net) return;
  rms(T, T->rho, nullptr, v, T->red + kS, 0, st);
  hipLaunchKernelGGL(k_input0, dim3(nb(v)), dim3(kB), 0, st, v, T->rho, T->red + kS, N.inp_f);
  N.input_from_f32(N.inp_f);
}

void mg_free(TwoGrid *T) {
  for (Level &L : T->lv) {
    void *ptrs[] = {L.own_cst ? L.cst : nullptr, L.rhs, L.e, L.e2, L.res};
    for (void *q : ptrs)
      if (q) (void)hipFree(q);
  }
  T->lv.clear();
}

// levels below the network grid: halve while the cell count is even and > 4
// (the coarsest grid has at most 8 cells per direction); buffers are kept
// across attachments to the same problem, the operators are recomputed
int mg_setup(TwoGrid *T, hipStream_t st) {
  if (T->lv.empty()) {
    std::vector<int> ms{T->n};
    while (ms.back() % 2 == 0 && ms.back() > 4) ms.push_back(ms.back() / 2);
    if (ms.back() > kCoarsestMax) ms.clear();  // no coarsest grid small enough: two-grid form only
    for (size_t l = 0; l < ms.size(); ++l) {
      Level L;
      L.m = ms[l];
      const size_t v = (size_t)L.m * L.m * L.m;
      T->lv.push_back(L);
      Level &M = T->lv.back();
      bool ok = hipMalloc(reinterpret_cast<void **>(&M.rhs), v * sizeof(double)) == hipSuccess &&
                hipMalloc(reinterpret_cast<void **>(&M.e), v * sizeof(double)) == hipSuccess &&
                hipMalloc(reinterpret_cast<void **>(&M.e2), v * sizeof(double)) == hipSuccess &&
                hipMalloc(reinterpret_cast<void **>(&M.res), v * sizeof(double)) == hipSuccess;
      if (l == 0) {
        M.cst = T->cst;
        M.own_cst = false;
      } else {
        ok = ok && hipMalloc(reinterpret_cast<void **>(&M.cst), 27 * v * sizeof(double)) == hipSuccess;
      }
      if (!ok) {
        mg_free(T);
        return GDX_ERR_ALLOC;
      }
    }
  }
  for (size_t l = 1; l < T->lv.size(); ++l) {
    const Level &F = T->lv[l - 1];
    galerkin(T->lv[l].m, F.m, 1, (int64_t)F.m * F.m * F.m, F.cst, T->lv[l].cst, st);
  }
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// V-cycle on level l for the right-hand side lv[l].rhs, from a zero guess:
// nu damped-Jacobi sweeps, coarse-grid correction, nu sweeps (the coarsest
// level: sweeps in one workgroup).  Returns the buffer holding the correction.
double *vcycle(TwoGrid *T, size_t l, hipStream_t st) {
  Level &L = T->lv[l];
  const int64_t v = (int64_t)L.m * L.m * L.m;
  if (l + 1 == T->lv.size()) {
    switch (coarsest_lanes(L.m * L.m * L.m)) {  // m <= 8: 2 to 32 lanes per point
#define GDX_COARSEST(NL)                                                                                           \
  case NL:                                                                                                         \
    hipLaunchKernelGGL(k_coarsest<NL>, dim3(1), dim3(kCoarsestThreads), 0, st, L.m, kOmegaC, T->coarsest_sweeps, \
                       L.cst, L.rhs, L.e);                                                                         \
    break;
      GDX_COARSEST(2) GDX_COARSEST(4) GDX_COARSEST(8) GDX_COARSEST(16) GDX_COARSEST(32)
#undef GDX_COARSEST
      default: break;
    }
    return L.e;
  }
  double *a = L.e, *b = L.e2;
  hipLaunchKernelGGL(k_sweep0, dim3(nb(v)), dim3(kB), 0, st, L.m, kOmegaC, L.cst, L.rhs, a);
  for (int k = 1; k < T->nu; ++k) {
    hipLaunchKernelGGL(k_coarse_sweep, dim3(nb(v)), dim3(kB), 0, st, L.m, kOmegaC, L.cst, L.rhs, a, b);
    std::swap(a, b);
  }
  hipLaunchKernelGGL(k_resid, dim3(nb(v)), dim3(kB), 0, st, L.m, L.cst, L.rhs, a, L.res);
  Level &C = T->lv[l + 1];
  const int64_t vc = (int64_t)C.m * C.m * C.m;
  hipLaunchKernelGGL(k_restrict_c, dim3(nb(vc)), dim3(kB), 0, st, L.m, L.res, C.rhs);
  const double *c = vcycle(T, l + 1, st);
  hipLaunchKernelGGL(k_prolong_add_c, dim3(nb(v)), dim3(kB), 0, st, L.m, c, a);
  for (int k = 0; k < T->nu; ++k) {
    hipLaunchKernelGGL(k_coarse_sweep, dim3(nb(v)), dim3(kB), 0, st, L.m, kOmegaC, L.cst, L.rhs, a, b);
    std::swap(a, b);
  }
  return a;
}

}  // namespace

gdx_problem *twogrid_problem(gdx_nn *nn) { return tg_of(nn) ? tg_of(nn)->prob : nullptr; }

int twogrid_export_coarse(gdx_nn *nn, double *h) {
  TwoGrid *T = tg_of(nn);
  const int64_t v = (int64_t)T->n * T->n * T->n;
  hipStream_t st = T->prob->ctx->stream;
  if (hipMemcpyAsync(h, T->cst, 27 * v * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GDX_ERR_HIP;
  return GDX_OK;
}

// level l of the V-cycle: its cell count m, and (h != NULL) its stencil [27][m^3]
int twogrid_export_level(gdx_nn *nn, int l, int *m, double *h) {
  TwoGrid *T = tg_of(nn);
  if (l < 0 || l >= (int)T->lv.size()) return GDX_ERR_ARG;
  const Level &L = T->lv[l];
  *m = L.m;
  if (!h) return GDX_OK;
  const int64_t v = (int64_t)L.m * L.m * L.m;
  hipStream_t st = T->prob->ctx->stream;
  if (hipMemcpyAsync(h, L.cst, 27 * v * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GDX_ERR_HIP;
  return GDX_OK;
}

void twogrid_options(gdx_nn *nn, int sweeps, bool use_net) {
  if (TwoGrid *T = tg_of(nn)) {
    T->sweeps = sweeps;
    T->use_net = use_net;
  }
}

void twogrid_multigrid(gdx_nn *nn, bool on, int nu) {
  if (TwoGrid *T = tg_of(nn)) {
    T->mg = on;
    T->nu = nu;
  }
}

int twogrid_levels(gdx_nn *nn) {
  TwoGrid *T = tg_of(nn);
  return T ? (int)T->lv.size() : 0;
}

void twogrid_free(gdx_nn *nn) {
  TwoGrid *T = tg_of(nn);
  if (!T) return;
  mg_free(T);
  void *ptrs[] = {T->gid_node, T->flat, T->dlat, T->t1, T->t2, T->t3, T->rho, T->red, T->ec, T->ec2, T->tgt,
                  T->fst, T->cst, T->rc, T->ca, T->cb};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete T;
  nn->tg = nullptr;
}

// bind the network to a problem's current matrix: transfer maps, the fine
// inverse diagonal on the lattice and the network's diagonal feature
int twogrid_attach(gdx_nn *nn, gdx_problem *p) {
  nn::Net &N = nn->net;
  const int n = N.n;
  if (p->D != 3 || p->K != 1) return GDX_ERR_UNSUPPORTED;
  if (p->desc.dir_kind != 1) return GDX_ERR_UNSUPPORTED;
  for (int d = 0; d < 3; ++d)
    if (p->kp.ncells[d] != 2 * n) return GDX_ERR_ARG;
  hipStream_t st = p->ctx->stream;
  TwoGrid *T = tg_of(nn);
  if (T && T->prob != p) {
    twogrid_free(nn);
    T = nullptr;
  }
  const int nf = 2 * n + 1;
  const int64_t v = (int64_t)n * n * n, nfl = (int64_t)nf * nf * nf;
  if (!T) {
    T = new TwoGrid;
    nn->tg = T;
    T->prob = p;
    T->n = n;
    T->nf = nf;
    T->ng = (int)p->ngdofs;
    // free global id -> fine lattice node (the problem's lattice: nf = 2n + 1
    // nodes per direction), built on the device when the numbering is there
    std::vector<long long> gid_node(p->dev_numbering ? 0 : std::max<int64_t>(p->ngdofs, 1), 0);
    for (const Part &P : p->P)
      for (int64_t i = 0; i < P.n_ldofs && !p->dev_numbering; ++i) {
        const int64_t l = P.ldof_lnode[i];
        const int64_t g0 = P.loc_lo[0] + l % P.lnn[0];
        const int64_t g1 = P.loc_lo[1] + (l / P.lnn[0]) % P.lnn[1];
        const int64_t g2 = P.loc_lo[2] + l / (P.lnn[0] * P.lnn[1]);
        gid_node[P.ldof_gid[i]] = g0 + (int64_t)nf * (g1 + (int64_t)nf * g2);
      }
    const size_t gsz = std::max<int64_t>(p->ngdofs, 1) * sizeof(double);
    bool ok = hipMalloc(reinterpret_cast<void **>(&T->gid_node), std::max<int64_t>(p->ngdofs, 1) * sizeof(long long)) ==
                  hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->flat), nfl * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->dlat), nfl * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->t1), gsz) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->t2), gsz) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->t3), gsz) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->rho), v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->red), (kS + 16) * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->ec), v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->ec2), v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->tgt), v * sizeof(float)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->fst), 27 * nfl * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->cst), 27 * v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->rc), v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->ca), v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->cb), v * sizeof(double)) == hipSuccess;
    if (!ok) {
      twogrid_free(nn);
      return GDX_ERR_ALLOC;
    }
    if (p->dev_numbering ? gid_node_device(p, T->gid_node) != GDX_OK
                         : hipMemcpy(T->gid_node, gid_node.data(), gid_node.size() * sizeof(long long),
                                     hipMemcpyHostToDevice) != hipSuccess)
      return GDX_ERR_HIP;
  }
  int ng = 0;
  int rc = solve_dinv(p, &T->dinv, &ng);
  if (rc) return rc;
  (void)hipMemsetAsync(T->dlat, 0, nfl * sizeof(double), st);
  hipLaunchKernelGGL(k_scatter, dim3(nb(T->ng)), dim3(kB), 0, st, T->ng, T->gid_node, T->dinv, T->dlat);
  hipLaunchKernelGGL(k_input1, dim3(nb(v)), dim3(kB), 0, st, n, nf, p->kp.h[0], T->dlat, N.inp_f + v);
  // fine stencil and the Galerkin coarse operator of the current matrix
  const long long *rs = nullptr;
  const int *rl = nullptr, *cg = nullptr;
  if ((rc = solve_csr_view(p, &rs, &rl, &cg))) return rc;
  if (hipMemsetAsync(T->fst, 0, 27 * nfl * sizeof(double), st) != hipSuccess) return GDX_ERR_HIP;
  hipLaunchKernelGGL(k_fine_stencil, dim3((unsigned)std::max(1, (T->ng + kFsRows - 1) / kFsRows)), dim3(kB), 0, st,
                     T->ng, nf, rs, rl, cg, p->dev.values, T->gid_node, T->fst);
  galerkin(n, nf, 0, nfl, T->fst, T->cst, st);
  if ((rc = mg_setup(T, st))) return rc;
  return hipStreamSynchronize(st) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// z = M r (device vectors in global-id order)
int twogrid_apply(void *ctx, const double *r, double *z, hipStream_t st) {
  gdx_nn *nn = reinterpret_cast<gdx_nn *>(ctx);
  TwoGrid *T = tg_of(nn);
  if (!T) return GDX_ERR_STATE;
  nn::Net &N = nn->net;
  gdx_problem *p = T->prob;
  const int ng = T->ng;
  const int64_t nfl = (int64_t)T->nf * T->nf * T->nf;
  hipLaunchKernelGGL(k_jacobi0, dim3(nb(ng)), dim3(kB), 0, st, ng, kOmega, T->dinv, r, z);
  solve_spmv_dev(p, z, T->t1, st);
  hipLaunchKernelGGL(k_rminus, dim3(nb(ng)), dim3(kB), 0, st, ng, r, T->t1);
  make_input(nn, T, T->t1, st, T->use_net);
  if (T->use_net) {
    int rc = N.forward_dev();
    if (rc) return rc;
  }
  const int64_t v = (int64_t)T->n * T->n * T->n;
  hipLaunchKernelGGL(k_out_to_coarse, dim3(nb(v)), dim3(kB), 0, st, T->n, N.out, T->red + kS, T->use_net ? 1 : 0,
                     T->ca);
  if (T->use_net) {
    // the network's direction, scaled to minimise the coarse energy error:
    // a1 = (ec.rc) / (ec.Ac ec)
    hipLaunchKernelGGL(k_coarse_apply, dim3(nb(v)), dim3(kB), 0, st, T->n, T->cst, T->ca, T->cb);
    hipLaunchKernelGGL(k_dot2p, dim3(kRed), dim3(kB), 0, st, (int)v, T->ca, T->rc, T->ca, T->cb, T->red);
    hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kB), 0, st, T->red, T->red + kS + 6);
    hipLaunchKernelGGL(k_scale_by, dim3(nb(v)), dim3(kB), 0, st, v, T->red + kS + 6, T->ca);
  }
  double *ea = T->ca, *eb = T->cb;
  if (T->mg && T->lv.size() > 1) {
    // V-cycle on the coarse residual of the (network's or zero) guess
    hipLaunchKernelGGL(k_resid, dim3(nb(v)), dim3(kB), 0, st, T->n, T->cst, T->rc, ea, T->lv[0].rhs);
    const double *d = vcycle(T, 0, st);
    hipLaunchKernelGGL(k_add, dim3(nb(v)), dim3(kB), 0, st, (int)v, ea, d);
  } else {
    for (int k = 0; k < T->sweeps; ++k) {
      hipLaunchKernelGGL(k_coarse_sweep, dim3(nb(v)), dim3(kB), 0, st, T->n, kOmegaC, T->cst, T->rc, ea, eb);
      std::swap(ea, eb);
    }
  }
  hipLaunchKernelGGL(k_prolong_d, dim3(nb(nfl)), dim3(kB), 0, st, T->n, T->nf, ea, T->flat);
  hipLaunchKernelGGL(k_gather, dim3(nb(ng)), dim3(kB), 0, st, ng, T->gid_node, T->flat, T->t2);  // c
  solve_spmv_dev(p, T->t2, T->t3, st);                                                             // A c
  hipLaunchKernelGGL(k_dot2p, dim3(kRed), dim3(kB), 0, st, ng, T->t2, T->t1, T->t2, T->t3, T->red);
  hipLaunchKernelGGL(k_alpha, dim3(1), dim3(kB), 0, st, T->red, T->red + kS + 5);
  hipLaunchKernelGGL(k_coarse_post, dim3(nb(ng)), dim3(kB), 0, st, ng, kOmega, T->red + kS + 5, T->dinv, T->t1,
                     T->t2, T->t3, z);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// one training step on a fresh sample; loss (relative squared error) to *loss
int twogrid_train_step(gdx_nn *nn, uint64_t seed, float lr, int smooth_passes, float *loss) {
  TwoGrid *T = tg_of(nn);
  if (!T) return GDX_ERR_STATE;
  nn::Net &N = nn->net;
  gdx_problem *p = T->prob;
  hipStream_t st = p->ctx->stream;
  const int n = T->n, ng = T->ng;
  const int64_t v = (int64_t)n * n * n, nfl = (int64_t)T->nf * T->nf * T->nf;
  // smooth random coarse error, unit rms
  hipLaunchKernelGGL(k_coarse_noise, dim3(nb(v)), dim3(kB), 0, st, n, seed, T->ec);
  for (int k = 0; k < smooth_passes; ++k) {
    hipLaunchKernelGGL(k_box, dim3(nb(v)), dim3(kB), 0, st, n, T->ec, T->ec2);
    std::swap(T->ec, T->ec2);
  }
  // r = A P ec
  hipLaunchKernelGGL(k_ec_to_float, dim3(nb(v)), dim3(kB), 0, st, v, T->ec, N.out);
  hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, st, T->red + kS + 2, 1.0);  // unit scale: P ec itself
  hipLaunchKernelGGL(k_prolong, dim3(nb(nfl)), dim3(kB), 0, st, n, T->nf, N.out, T->red + kS + 2, T->flat);
  hipLaunchKernelGGL(k_gather, dim3(nb(ng)), dim3(kB), 0, st, ng, T->gid_node, T->flat, T->t2);
  solve_spmv_dev(p, T->t2, T->t1, st);
  // input from r, target ec / s
  make_input(nn, T, T->t1, st);
  hipLaunchKernelGGL(k_target, dim3(nb(v)), dim3(kB), 0, st, v, T->ec, T->red + kS, T->tgt);
  int rc = N.forward_dev();
  if (rc) return rc;
  // scale-free loss 1 - cos^2(N, ec): the preconditioner rescales the
  // network's output optimally, so only its direction matters
  hipLaunchKernelGGL(k_dot3f, dim3(kRed), dim3(kB), 0, st, v, N.out, T->tgt, T->red);
  hipLaunchKernelGGL(k_finish3, dim3(1), dim3(kB), 0, st, T->red, T->red + kS + 8);
  hipLaunchKernelGGL(k_cos_grad, dim3(nb(v)), dim3(kB), 0, st, v, N.out, T->tgt, T->red + kS + 8, N.dout);
  if ((rc = N.backward_dev(false))) return rc;
  if ((rc = N.adam_step(lr, 1.0f))) return rc;
  double h = 0.0;
  if (hipMemcpyAsync(&h, T->red + kS + 11, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GDX_ERR_HIP;
  if (loss) *loss = (float)h;
  return GDX_OK;
}

// per-device address of the no-alloc staging array, set once with that device current
namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double *g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

int noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return GDX_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  if (g_noalloc_ptr[device]) return GDX_OK;
  int cur = 0;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return GDX_ERR_STATE;
  void *ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return GDX_ERR_HIP;
  g_noalloc_ptr[device] = static_cast<double *>(ptr);
  return GDX_OK;
}

double *noalloc_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_noalloc_mu);
  return device >= 0 && device < kMaxDevices ? g_noalloc_ptr[device] : nullptr;
}

}  // namespace gdx

// ---------------------------------------------------------------------------
// C ABI (include/gdx.h)
// ---------------------------------------------------------------------------
extern "C" {

int gdx_nn_attach(gdx_nn *nn, gdx_problem *prob) {
  if (!nn || !prob) {
    gdx::set_error("NULL argument");
    return GDX_ERR_ARG;
  }
  if (!nn->ctx) {
    gdx::set_error("the network's context was finalized");
    return GDX_ERR_STATE;
  }
  if (!prob->symbolic_done || !prob->ctx) {
    gdx::set_error("the problem needs gdx_symbolic and a device context");
    return GDX_ERR_STATE;
  }
  if (prob->ctx != nn->ctx) {  // the network's kernels run on its context's stream
    gdx::set_error("network and problem belong to different contexts");
    return GDX_ERR_ARG;
  }
  if (hipSetDevice(prob->ctx->device) != hipSuccess) return GDX_ERR_HIP;
  int rc = gdx::twogrid_attach(nn, prob);
  if (rc == GDX_ERR_UNSUPPORTED) gdx::set_error("the two-grid preconditioner is built for 3D Q1 with Dirichlet data on the whole boundary");
  else if (rc == GDX_ERR_ARG) gdx::set_error("the problem must have 2n x 2n x 2n cells for a network grid of n^3");
  else if (rc) gdx::set_error("attaching the network failed");
  return rc;
}

int gdx_nn_train_step(gdx_nn *nn, uint64_t seed, float lr, int32_t smooth_passes, float *loss) {
  if (!nn || !(lr > 0.0f) || smooth_passes < 0 || smooth_passes > 1000) {
    gdx::set_error("invalid argument");
    return GDX_ERR_ARG;
  }
  if (!nn->tg || !nn->ctx) {
    gdx::set_error("gdx_nn_attach has not run");
    return GDX_ERR_STATE;
  }
  if (hipSetDevice(nn->ctx->device) != hipSuccess) return GDX_ERR_HIP;
  return gdx::twogrid_train_step(nn, seed, lr, smooth_passes, loss);
}

int gdx_nn_set_options(gdx_nn *nn, int32_t coarse_sweeps, int32_t use_network) {
  if (!nn || coarse_sweeps < 0 || coarse_sweeps > 64) {
    gdx::set_error("invalid argument");
    return GDX_ERR_ARG;
  }
  if (!nn->tg) {
    gdx::set_error("gdx_nn_attach has not run");
    return GDX_ERR_STATE;
  }
  gdx::twogrid_options(nn, coarse_sweeps, use_network != 0);
  return GDX_OK;
}

int gdx_nn_coarse_operator(gdx_nn *nn, double *stencil) {
  if (!nn || !stencil) {
    gdx::set_error("NULL argument");
    return GDX_ERR_ARG;
  }
  if (!nn->tg || !nn->ctx) {
    gdx::set_error("gdx_nn_attach has not run");
    return GDX_ERR_STATE;
  }
  if (hipSetDevice(nn->ctx->device) != hipSuccess) return GDX_ERR_HIP;
  return gdx::twogrid_export_coarse(nn, stencil);
}

int gdx_nn_set_multigrid(gdx_nn *nn, int32_t enable, int32_t nu, int32_t *levels) {
  if (!nn || nu < 1 || nu > 8) {
    gdx::set_error("invalid argument");
    return GDX_ERR_ARG;
  }
  if (!nn->tg) {
    gdx::set_error("gdx_nn_attach has not run");
    return GDX_ERR_STATE;
  }
  gdx::twogrid_multigrid(nn, enable != 0, nu);
  if (levels) *levels = gdx::twogrid_levels(nn);
  return GDX_OK;
}

int gdx_nn_level_operator(gdx_nn *nn, int32_t level, int32_t *m, double *stencil) {
  if (!nn || !m) {
    gdx::set_error("NULL argument");
    return GDX_ERR_ARG;
  }
  if (!nn->tg || !nn->ctx) {
    gdx::set_error("gdx_nn_attach has not run");
    return GDX_ERR_STATE;
  }
  if (hipSetDevice(nn->ctx->device) != hipSuccess) return GDX_ERR_HIP;
  int mm = 0;
  const int rc = gdx::twogrid_export_level(nn, level, &mm, stencil);
  if (rc == GDX_ERR_ARG) gdx::set_error("no such level");
  *m = mm;
  return rc;
}

int gdx_nn_apply(gdx_nn *nn, const double *r, double *z) {
  if (!nn || !r || !z) {
    gdx::set_error("NULL argument");
    return GDX_ERR_ARG;
  }
  if (!nn->tg || !nn->ctx) {
    gdx::set_error("gdx_nn_attach has not run");
    return GDX_ERR_STATE;
  }
  if (hipSetDevice(nn->ctx->device) != hipSuccess) return GDX_ERR_HIP;
  gdx_problem *p = gdx::twogrid_problem(nn);
  const size_t bytes = std::max<int64_t>(p->ngdofs, 1) * sizeof(double);
  double *rd = nullptr, *zd = nullptr;
  hipStream_t st = p->ctx->stream;
  if (hipMalloc(reinterpret_cast<void **>(&rd), bytes) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&zd), bytes) != hipSuccess) {
    (void)hipFree(rd);
    return GDX_ERR_ALLOC;
  }
  int rc = hipMemcpyAsync(rd, r, p->ngdofs * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess ? GDX_OK
                                                                                                          : GDX_ERR_HIP;
  if (!rc) rc = gdx::twogrid_apply(nn, rd, zd, st);
  if (!rc && (hipMemcpyAsync(z, zd, p->ngdofs * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
              hipStreamSynchronize(st) != hipSuccess))
    rc = GDX_ERR_HIP;
  (void)hipFree(rd);
  (void)hipFree(zd);
  return rc;
}

int gdx_nn_apply_noalloc(gdx_nn *nn, const double *r, double *z) {
  if (!nn || !r || !z) {
    gdx::set_error("NULL argument");
    return GDX_ERR_ARG;
  }
  if (!nn->tg || !nn->ctx) {
    gdx::set_error("gdx_nn_attach has not run");
    return GDX_ERR_STATE;
  }
  const int dev = nn->ctx->device;
  if (hipSetDevice(dev) != hipSuccess) return GDX_ERR_HIP;
  gdx_problem *p = gdx::twogrid_problem(nn);
  const int64_t n = p->ngdofs;
  if (2 * n > (int64_t)kNoAllocDoubles) {
    gdx::set_error("gdx_nn_apply_noalloc: the problem has more than 2,097,152 free dofs (the staging limit)");
    return GDX_ERR_ARG;
  }
  int rc = gdx::noalloc_setup(dev);  // normally done by gdx_init on this device
  double *base = gdx::noalloc_scratch(dev);
  if (rc || !base) {
    gdx::set_error("gdx_nn_apply_noalloc: no device address for the staging array");
    return rc ? rc : GDX_ERR_HIP;
  }
  // r is copied in whole and z written in whole: nothing has to start at zero
  double *rd = base, *zd = base + n;
  hipStream_t st = p->ctx->stream;
  rc = n == 0 || hipMemcpyAsync(rd, r, n * sizeof(double), hipMemcpyHostToDevice, st) == hipSuccess ? GDX_OK
                                                                                                   : GDX_ERR_HIP;
  if (!rc) rc = gdx::twogrid_apply(nn, rd, zd, st);
  if (!rc && ((n > 0 && hipMemcpyAsync(z, zd, n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) ||
              hipStreamSynchronize(st) != hipSuccess))
    rc = GDX_ERR_HIP;
  if (rc) gdx::set_error("gdx_nn_apply_noalloc failed");
  return rc;
}

int gdx_fpcg(gdx_problem *prob, gdx_nn *nn, const double *rhs, double rtol, int32_t maxiter, double *xsol,
             int32_t *iters, double *relres, double *ms, int32_t *fellback) {
  if (!prob || !(rtol > 0.0) || maxiter < 1) {
    gdx::set_error("invalid argument");
    return GDX_ERR_ARG;
  }
  if (!prob->symbolic_done || !prob->ctx) {
    gdx::set_error("the problem needs gdx_symbolic and a device context");
    return GDX_ERR_STATE;
  }
  if (nn && (!nn->tg || gdx::twogrid_problem(nn) != prob)) {
    gdx::set_error("the network is not attached to this problem (gdx_nn_attach)");
    return GDX_ERR_STATE;
  }
  if (hipSetDevice(prob->ctx->device) != hipSuccess) return GDX_ERR_HIP;
  gdx::FpcgPrec pr{gdx::twogrid_apply, nn};
  int it = 0, fb = 0;
  int rc = gdx::solve_fpcg(prob, nn ? &pr : nullptr, rhs, rtol, maxiter, xsol, &it, relres, ms, &fb);
  if (iters) *iters = it;
  if (fellback) *fellback = fb;
  if (rc) gdx::set_error("flexible PCG failed");
  return rc;
}

}  // extern "C"


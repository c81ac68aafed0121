// Learned two-grid preconditioner of libgdx for 3D Q1 problems on a cube of
// 2n x 2n x 2n cells with Dirichlet data on the whole boundary:
//
//   z  = w D^-1 r                          damped Jacobi pre-smoothing
//   rc = R (r - A z)                       full weighting (R = P^T, trilinear P)
//   ec = s N( D^-1 rc / s, log-diagonal )  network coarse solve, s = rms(D^-1 rc)
//   z += P ec
//   z += w D^-1 (r - A z)                  damped Jacobi post-smoothing
//
// The coarse unknowns are the fine nodes with even indices; the (n-1)^3
// interior ones sit on the network's n^3 grid (the last plane is zero
// padding).  The network is trained on device from the problem's own matrix:
// a smooth random coarse error ec (white noise smoothed by repeated 3x3x3
// means), r = A P ec, target ec for the input built from r exactly as above
// (loss |N - ec/s|^2 / |ec/s|^2).  Every reduction has a fixed order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "gdx_nn.h"

namespace gdx {

namespace {

constexpr int kB = 256;
constexpr int kRed = 256;  // partial sums per reduction
constexpr double kOmega = 2.0 / 3.0;

struct TwoGrid {
  gdx_problem *prob = nullptr;
  int n = 0;       // network grid = coarse cells per direction
  int nf = 0;      // fine lattice extent per direction (2n + 1)
  int ng = 0;      // free dofs
  long long *gid_node = nullptr;
  double *flat = nullptr, *dlat = nullptr;   // fine lattice vectors
  double *t1 = nullptr, *t2 = nullptr;       // global-id temporaries
  double *rho = nullptr;                     // coarse scaled residual [n^3]
  double *red = nullptr;                     // reduction partials + scalars
  double *ec = nullptr, *ec2 = nullptr;      // training: coarse error [n^3]
  float *tgt = nullptr;
  const double *dinv = nullptr;
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
                           double *__restrict__ rho) {
  const int64_t tot = (int64_t)n * n * n;
  const int64_t j = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (j >= tot) return;
  const int jx = (int)(j % n), jy = (int)((j / n) % n), jz = (int)(j / ((int64_t)n * n));
  if (jx == n - 1 || jy == n - 1 || jz == n - 1) {
    rho[j] = 0.0;
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
  rho[j] = acc * dlat[fx + (int64_t)nf * (fy + (int64_t)nf * fz)];
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

// ec (double) -> network-output units for the prolongation (float) and the target ec / s
__global__ void k_ec_to_float(int64_t tot, const double *__restrict__ ec, float *__restrict__ outf) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < tot) outf[i] = (float)ec[i];
}
__global__ void k_target(int64_t tot, const double *__restrict__ ec, const double *__restrict__ s, float *__restrict__ tgt) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < tot) tgt[i] = (float)(ec[i] / *s);
}
// relative squared error: dout = 2 (out - tgt) / |tgt|^2, part = (out - tgt)^2 partial sums
__global__ __launch_bounds__(kB) void k_loss(int64_t tot, const float *__restrict__ out, const float *__restrict__ tgt,
                                            const double *__restrict__ tt, float *__restrict__ dout,
                                            double *__restrict__ part) {
  __shared__ double sh[kB];
  const double inv = 1.0 / *tt;
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x; i < tot; i += (int64_t)gridDim.x * kB) {
    const double d = (double)out[i] - (double)tgt[i];
    dout[i] = (float)(2.0 * d * inv);
    s = fma(d, d, s);
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] * inv;
}

// rms of rho into red[kRed] (scalar slot)
void rms(TwoGrid *T, const double *a, const float *fa, int64_t count, double *dst, int mode, hipStream_t st) {
  hipLaunchKernelGGL(k_sumsq, dim3(kRed), dim3(kB), 0, st, count, a, fa, T->red);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kB), 0, st, T->red, kRed, (double)count, mode, dst);
}

// network input from a fine residual in global-id order (s into red[kRed])
void make_input(gdx_nn *nn, TwoGrid *T, const double *r, hipStream_t st) {
  nn::Net &N = nn->net;
  const int64_t v = (int64_t)T->n * T->n * T->n, nfl = (int64_t)T->nf * T->nf * T->nf;
  (void)hipMemsetAsync(T->flat, 0, nfl * sizeof(double), st);
  hipLaunchKernelGGL(k_scatter, dim3(nb(T->ng)), dim3(kB), 0, st, T->ng, T->gid_node, r, T->flat);
  hipLaunchKernelGGL(k_restrict, dim3(nb(v)), dim3(kB), 0, st, T->n, T->nf, T->flat, T->dlat, T->rho);
  rms(T, T->rho, nullptr, v, T->red + kRed, 0, st);
  hipLaunchKernelGGL(k_input0, dim3(nb(v)), dim3(kB), 0, st, v, T->rho, T->red + kRed, N.inp_f);
  N.input_from_f32(N.inp_f);
}

}  // namespace

gdx_problem *twogrid_problem(gdx_nn *nn) { return tg_of(nn) ? tg_of(nn)->prob : nullptr; }

void twogrid_free(gdx_nn *nn) {
  TwoGrid *T = tg_of(nn);
  if (!T) return;
  void *ptrs[] = {T->gid_node, T->flat, T->dlat, T->t1, T->t2, T->rho, T->red, T->ec, T->ec2, T->tgt};
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
    std::vector<long long> gid_node(std::max<int64_t>(p->ngdofs, 1), 0);
    for (const Part &P : p->P)
      for (int64_t i = 0; i < P.n_ldofs; ++i) {
        const int64_t l = P.ldof_lnode[i];
        const int64_t g0 = P.loc_lo[0] + l % P.lnn[0];
        const int64_t g1 = P.loc_lo[1] + (l / P.lnn[0]) % P.lnn[1];
        const int64_t g2 = P.loc_lo[2] + l / (P.lnn[0] * P.lnn[1]);
        gid_node[P.ldof_gid[i]] = g0 + (int64_t)nf * (g1 + (int64_t)nf * g2);
      }
    const size_t gsz = std::max<int64_t>(p->ngdofs, 1) * sizeof(double);
    bool ok = hipMalloc(reinterpret_cast<void **>(&T->gid_node), gid_node.size() * sizeof(long long)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->flat), nfl * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->dlat), nfl * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->t1), gsz) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->t2), gsz) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->rho), v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->red), (kRed + 8) * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->ec), v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->ec2), v * sizeof(double)) == hipSuccess &&
              hipMalloc(reinterpret_cast<void **>(&T->tgt), v * sizeof(float)) == hipSuccess;
    if (!ok) {
      twogrid_free(nn);
      return GDX_ERR_ALLOC;
    }
    if (hipMemcpy(T->gid_node, gid_node.data(), gid_node.size() * sizeof(long long), hipMemcpyHostToDevice) !=
        hipSuccess)
      return GDX_ERR_HIP;
  }
  int ng = 0;
  int rc = solve_dinv(p, &T->dinv, &ng);
  if (rc) return rc;
  (void)hipMemsetAsync(T->dlat, 0, nfl * sizeof(double), st);
  hipLaunchKernelGGL(k_scatter, dim3(nb(T->ng)), dim3(kB), 0, st, T->ng, T->gid_node, T->dinv, T->dlat);
  hipLaunchKernelGGL(k_input1, dim3(nb(v)), dim3(kB), 0, st, n, nf, p->kp.h[0], T->dlat, N.inp_f + v);
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
  make_input(nn, T, T->t1, st);
  int rc = N.forward_dev();
  if (rc) return rc;
  hipLaunchKernelGGL(k_prolong, dim3(nb(nfl)), dim3(kB), 0, st, T->n, T->nf, N.out, T->red + kRed, T->flat);
  hipLaunchKernelGGL(k_gather, dim3(nb(ng)), dim3(kB), 0, st, ng, T->gid_node, T->flat, T->t2);
  hipLaunchKernelGGL(k_add, dim3(nb(ng)), dim3(kB), 0, st, ng, z, T->t2);
  solve_spmv_dev(p, z, T->t1, st);
  hipLaunchKernelGGL(k_jacobi_post, dim3(nb(ng)), dim3(kB), 0, st, ng, kOmega, T->dinv, r, T->t1, z);
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
  hipLaunchKernelGGL(k_set, dim3(1), dim3(1), 0, st, T->red + kRed + 2, 1.0);  // unit scale: P ec itself
  hipLaunchKernelGGL(k_prolong, dim3(nb(nfl)), dim3(kB), 0, st, n, T->nf, N.out, T->red + kRed + 2, T->flat);
  hipLaunchKernelGGL(k_gather, dim3(nb(ng)), dim3(kB), 0, st, ng, T->gid_node, T->flat, T->t2);
  solve_spmv_dev(p, T->t2, T->t1, st);
  // input from r, target ec / s
  make_input(nn, T, T->t1, st);
  hipLaunchKernelGGL(k_target, dim3(nb(v)), dim3(kB), 0, st, v, T->ec, T->red + kRed, T->tgt);
  int rc = N.forward_dev();
  if (rc) return rc;
  rms(T, nullptr, T->tgt, v, T->red + kRed + 3, 1, st);  // |tgt|^2
  hipLaunchKernelGGL(k_loss, dim3(kRed), dim3(kB), 0, st, v, N.out, T->tgt, T->red + kRed + 3, N.dout, T->red);
  hipLaunchKernelGGL(k_finish, dim3(1), dim3(kB), 0, st, T->red, kRed, 1.0, 1, T->red + kRed + 4);
  if ((rc = N.backward_dev(false))) return rc;
  if ((rc = N.adam_step(lr, 1.0f))) return rc;
  double h = 0.0;
  if (hipMemcpyAsync(&h, T->red + kRed + 4, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GDX_ERR_HIP;
  if (loss) *loss = (float)h;
  return GDX_OK;
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
  if (prob->ctx->device != nn->ctx->device) {
    gdx::set_error("network and problem live on different devices");
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


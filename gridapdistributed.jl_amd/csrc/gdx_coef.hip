// Cell-wise coefficients of libgdx: kappa given per mesh cell (x fastest),
// either uploaded by the caller or drawn on the device as a lognormal field
// kappa = exp(s g), g a unit-variance Gaussian random field with
// squared-exponential correlation exp(-r^2 / (2 l^2)): white noise convolved
// with a Gaussian kernel of standard deviation l / sqrt(2), normalised by the
// kernel's exact variance.  The noise is a counter-based hash of (seed, index)
// turned into normals by Box-Muller, so every field is reproducible from its
// seed on any device (and by the CPU restatement in the tests).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "gdx_internal.h"

namespace gdx {

namespace {

__host__ __device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// uniform in (0, 1]: 53 random bits
__host__ __device__ inline double unif(uint64_t seed, uint64_t i) {
  const uint64_t z = mix64(seed * 0x9E3779B97F4A7C15ull + (i + 1) * 0xD1B54A32D192ED03ull);
  return ((double)(z >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}

__global__ void k_noise(uint64_t seed, int64_t n, double *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double u1 = unif(seed, 2 * (uint64_t)i), u2 = unif(seed, 2 * (uint64_t)i + 1);
  out[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// valid convolution along one axis: in extents e[3], out extents e with
// e[axis] reduced by 2R; weights w[2R+1]
__global__ void k_gauss_axis(const double *__restrict__ in, double *__restrict__ out, int e0, int e1, int e2,
                             int axis, int R, const double *__restrict__ w) {
  const int o0 = axis == 0 ? e0 - 2 * R : e0, o1 = axis == 1 ? e1 - 2 * R : e1, o2 = axis == 2 ? e2 - 2 * R : e2;
  const int64_t n = (int64_t)o0 * o1 * o2;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int x = (int)(i % o0), y = (int)((i / o0) % o1), z = (int)(i / ((int64_t)o0 * o1));
  const int64_t s = axis == 0 ? 1 : axis == 1 ? e0 : (int64_t)e0 * e1;
  const double *src = in + x + (int64_t)e0 * (y + (int64_t)e1 * z);
  double acc = 0.0;
  for (int k = 0; k <= 2 * R; ++k) acc = fma(w[k], src[k * s], acc);
  out[i] = acc;
}

__global__ void k_lognormal(int64_t n, double *__restrict__ g, double scale) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) g[i] = exp(scale * g[i]);
}

// kappa and f at every quadrature point from the cell values
__global__ void k_cell_coef(KParams kp, const double *__restrict__ cellk, double *__restrict__ W,
                            double *__restrict__ F) {
  const int64_t n = (int64_t)kp.qn[0] * kp.qn[1] * kp.qn[2];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int qa = (int)(i % kp.qn[0]), qb = (int)((i / kp.qn[0]) % kp.qn[1]), qc = (int)(i / ((int64_t)kp.qn[0] * kp.qn[1]));
  const int64_t c = qa / kp.NQ + (int64_t)kp.ncells[0] * (qb / kp.NQ + (int64_t)kp.ncells[1] * (qc / kp.NQ));
  W[i] = kp.coef_const * cellk[c];
  F[i] = kp.rhs_const;
}

inline unsigned blocks_for(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int64_t num_cells(const gdx_problem *p) { return (int64_t)p->kp.ncells[0] * p->kp.ncells[1] * p->kp.ncells[2]; }

int cell_coef_alloc(gdx_problem *p) {
  if (p->dev.cellk) return GDX_OK;
  return hipMalloc(reinterpret_cast<void **>(&p->dev.cellk), std::max<int64_t>(num_cells(p), 1) * sizeof(double)) ==
                 hipSuccess
             ? GDX_OK
             : GDX_ERR_ALLOC;
}

int launch_cell_coefficients(gdx_problem *p, hipStream_t st) {
  if (!p->dev.cellk) return GDX_ERR_STATE;
  const int64_t nq = (int64_t)p->kp.qn[0] * p->kp.qn[1] * p->kp.qn[2];
  hipLaunchKernelGGL(k_cell_coef, dim3(blocks_for(nq)), dim3(256), 0, st, p->kp, p->dev.cellk, p->dev.W, p->dev.F);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// the Gaussian filter of standard deviation l / sqrt(2) in cells of each
// direction (the mesh may be anisotropic), truncated at 4 standard deviations
static void gauss_weights(double sd_cells, std::vector<double> &w, int &R) {
  R = std::max(1, (int)std::ceil(4.0 * sd_cells));
  w.assign(2 * R + 1, 0.0);
  double s = 0.0;
  for (int k = -R; k <= R; ++k) s += w[k + R] = std::exp(-0.5 * (k / sd_cells) * (k / sd_cells));
  for (double &v : w) v /= s;
}

int grf_cell_coefficient(gdx_problem *p, uint64_t seed, double corr_len, double log_scale) {
  int rc = cell_coef_alloc(p);
  if (rc) return rc;
  hipStream_t st = p->ctx->stream;
  const int D = p->D;
  int n[3], R[3];
  std::vector<double> w[3];
  for (int d = 0; d < 3; ++d) {
    n[d] = p->kp.ncells[d];
    if (d < D) gauss_weights(corr_len / std::sqrt(2.0) / p->kp.h[d], w[d], R[d]);
    else R[d] = 0, w[d].assign(1, 1.0);
  }
  // extended noise grid: valid convolution leaves exactly the cell grid
  int e[3] = {n[0] + 2 * R[0], n[1] + 2 * R[1], n[2] + 2 * R[2]};
  const int64_t ne = (int64_t)e[0] * e[1] * e[2];
  double *a = nullptr, *b = nullptr, *wd = nullptr;
  const int wmax = 2 * std::max(R[0], std::max(R[1], R[2])) + 1;
  if (hipMalloc(reinterpret_cast<void **>(&a), ne * sizeof(double)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&b), ne * sizeof(double)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&wd), 3 * wmax * sizeof(double)) != hipSuccess) {
    (void)hipFree(a);
    (void)hipFree(b);
    return GDX_ERR_ALLOC;
  }
  double var = 1.0;
  for (int d = 0; d < 3; ++d) {
    double s2 = 0.0;
    for (double v : w[d]) s2 += v * v;
    var *= s2;
    if (hipMemcpy(wd + d * wmax, w[d].data(), w[d].size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = GDX_ERR_HIP;
  }
  hipLaunchKernelGGL(k_noise, dim3(blocks_for(ne)), dim3(256), 0, st, seed, ne, a);
  double *src = a, *dst = b;
  for (int d = 0; d < 3; ++d) {
    if (R[d] == 0) continue;
    const int64_t no = (int64_t)e[0] * e[1] * e[2] / e[d] * (e[d] - 2 * R[d]);
    hipLaunchKernelGGL(k_gauss_axis, dim3(blocks_for(no)), dim3(256), 0, st, src, dst, e[0], e[1], e[2], d, R[d],
                       wd + d * wmax);
    e[d] -= 2 * R[d];
    std::swap(src, dst);
  }
  const int64_t nc = num_cells(p);
  // src now holds the cell grid
  hipLaunchKernelGGL(k_lognormal, dim3(blocks_for(nc)), dim3(256), 0, st, nc, src, log_scale / std::sqrt(var));
  if (hipMemcpyAsync(p->dev.cellk, src, nc * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
    rc = GDX_ERR_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) rc = GDX_ERR_HIP;
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(wd);
  if (hipGetLastError() != hipSuccess) rc = GDX_ERR_HIP;
  return rc;
}

}  // namespace gdx

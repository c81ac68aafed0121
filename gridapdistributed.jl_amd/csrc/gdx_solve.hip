// Solve side of libgdx: FP64 CSR SpMV over the owned rows of all parts and a
// Jacobi-preconditioned conjugate gradient, operating in place on the values
// the numeric assembly left in HBM.  The reference solves with A\b (sparse LU
// of the gathered system, README.md:76); these kernels replace that step.
//
// Global view: owned rows in global-id order.  Because global ids are
// rank-blocked and each part lists its owned rows in id order, row g of the
// global system is owned row (g - first_p) of part p; its values sit at
// rstart[g] in the concatenated values buffer and its column global ids in
// cgid (int32), built once.  Reductions are two-pass (per-block partials then
// one block in fixed order), so every solve is bit-for-bit repeatable.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cmath>
#include <vector>

#include "gdx_internal.h"
#include "gdx_nn.h"

namespace gdx {

constexpr int kRedBlocks = 1024;  // partial sums per reduction
constexpr int kBlk = 256;
constexpr int kSpmvRows = 128;  // rows per block of the block SpMV

// y[g] = sum_k val[rstart[g]+k] * x[cgid[rstart[g]+k]] ; LPR lanes per row
template <int LPR>
__global__ __launch_bounds__(kBlk) void k_spmv(int n, const long long *__restrict__ rstart,
                                               const int *__restrict__ rlen, const int *__restrict__ cgid,
                                               const double *__restrict__ val, const double *__restrict__ x,
                                               double *__restrict__ y) {
  const int lane = threadIdx.x % LPR;
  const int row = (blockIdx.x * kBlk + threadIdx.x) / LPR;
  double acc = 0.0;
  if (row < n) {
    const long long s = rstart[row];
    const int len = rlen[row];
    for (int k = lane; k < len; k += LPR) acc = fma(val[s + k], x[cgid[s + k]], acc);
  }
#pragma unroll
  for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
  if (row < n && lane == 0) y[row] = acc;
}

// Block form for short rows (Q1: at most 32 entries): RB consecutive rows per
// NT-thread block.  When the rows' values are stored back to back (own rows of
// one part), the block streams their values and column ids coalesced, each
// thread staging its U entries in registers before the gathers of x, and puts
// the products in LDS; LPR = NT / RB lanes per row then sum their row in a
// fixed order.  Other blocks (a part boundary or a ghost-row gap inside) read
// their rows directly with LPR lanes each.  A streamed block whose column ids
// fall in at most four windows of 16384 ids (its own part's rows and the ghost
// columns' ranges in up to three neighbour parts) reads them as 16-bit codes
// c16 = window << 14 | offset against the four bases bbase[block] (first base
// -1: 32-bit ids), 10 instead of 12 bytes per entry.  Values and codes, read
// once per product, are loaded non-temporally so that L2 keeps the gathered x.
template <int RB, int NT>
__global__ __launch_bounds__(NT) void k_spmv_blk(int n, const long long *__restrict__ rstart,
                                                 const int *__restrict__ rlen, const int *__restrict__ cgid,
                                                 const uint16_t *__restrict__ c16, const int4 *__restrict__ bbase,
                                                 const double *__restrict__ val, const double *__restrict__ x,
                                                 double *__restrict__ y) {
  constexpr int U = 8, CAP = U * NT, LPR = NT / RB;
  static_assert(NT % RB == 0 && RB <= NT, "whole rows per block");
  __shared__ double buf[CAP];
  __shared__ long long rs[RB];
  __shared__ int rl[RB];
  const int t = threadIdx.x, r0 = blockIdx.x * RB;
  const int nr = min(RB, n - r0);
  if (t < nr) {
    rs[t] = rstart[r0 + t];
    rl[t] = rlen[r0 + t];
  }
  __syncthreads();
  int ok = 1;
  if (t < nr - 1) ok = rs[t] + rl[t] == rstart[r0 + t + 1];
  const long long A = rs[0], B = rs[nr - 1] + rl[nr - 1];
  ok = __syncthreads_and(ok && (B - A) <= CAP);
  const int row = t / LPR, q = t % LPR;
  double acc = 0.0;
  if (ok) {
    const int m = (int)(B - A);
    const int4 bb = bbase[blockIdx.x];
    double vv[U];
    int cc[U];
    if (bb.x >= 0) {
#pragma unroll
      for (int i = 0; i < U; ++i) {
        const int k = t + NT * i;
        vv[i] = k < m ? __builtin_nontemporal_load(val + A + k) : 0.0;
        const int c = k < m ? (int)__builtin_nontemporal_load(c16 + A + k) : 0, w = c >> 14;
        cc[i] = (c & 0x3FFF) + (w == 0 ? bb.x : w == 1 ? bb.y : w == 2 ? bb.z : bb.w);
      }
    } else {
#pragma unroll
      for (int i = 0; i < U; ++i) {
        const int k = t + NT * i;
        vv[i] = k < m ? val[A + k] : 0.0;
        cc[i] = k < m ? cgid[A + k] : 0;
      }
    }
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const int k = t + NT * i;
      if (k < m) buf[k] = vv[i] * x[cc[i]];
    }
    __syncthreads();
    if (row < nr) {
      const int s = (int)(rs[row] - A), len = rl[row];
      for (int k = q; k < len; k += LPR) acc += buf[s + k];
    }
  } else if (row < nr) {
    const long long s = rs[row];
    const int len = rl[row];
    for (int k = q; k < len; k += LPR) acc = fma(val[s + k], x[cgid[s + k]], acc);
  }
#pragma unroll
  for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
  if (row < nr && q == 0) y[r0 + row] = acc;
}

// partial sums of up to two dot products: block i sums a grid-stride slice
__global__ __launch_bounds__(kBlk) void k_dot2(int n, const double *__restrict__ a, const double *__restrict__ b,
                                               const double *__restrict__ c, const double *__restrict__ d,
                                               double *__restrict__ part) {
  __shared__ double s1[kBlk], s2[kBlk];
  double p1 = 0.0, p2 = 0.0;
  for (int i = blockIdx.x * kBlk + threadIdx.x; i < n; i += gridDim.x * kBlk) {
    p1 = fma(a[i], b[i], p1);
    if (c) p2 = fma(c[i], d[i], p2);
  }
  s1[threadIdx.x] = p1;
  s2[threadIdx.x] = p2;
  __syncthreads();
  for (int o = kBlk / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[blockIdx.x] = s1[0];
    part[kRedBlocks + blockIdx.x] = s2[0];
  }
}

// finish the two reductions into out[0], out[1] (one block, fixed order)
__global__ __launch_bounds__(kBlk) void k_finish2(const double *__restrict__ part, int nb, double *__restrict__ out) {
  __shared__ double s1[kBlk], s2[kBlk];
  double p1 = 0.0, p2 = 0.0;
  for (int i = threadIdx.x; i < nb; i += kBlk) {
    p1 += part[i];
    p2 += part[kRedBlocks + i];
  }
  s1[threadIdx.x] = p1;
  s2[threadIdx.x] = p2;
  __syncthreads();
  for (int o = kBlk / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      s1[threadIdx.x] += s1[threadIdx.x + o];
      s2[threadIdx.x] += s2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = s1[0];
    out[1] = s2[0];
  }
}

__global__ void k_gather_b(int n, const long long *__restrict__ bidx, const double *__restrict__ bsrc,
                           double *__restrict__ b) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) b[i] = bsrc[bidx[i]];
}

__global__ void k_inv_diag(int n, const long long *__restrict__ dpos, const double *__restrict__ val,
                           double *__restrict__ dinv) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) dinv[i] = 1.0 / val[dpos[i]];
}

// r = b - Ax (Ax in r on entry), z = dinv r, p = z
__global__ void k_cg_init(int n, const double *__restrict__ b, const double *__restrict__ dinv,
                          double *__restrict__ r, double *__restrict__ z, double *__restrict__ p) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) {
    const double ri = b[i] - r[i];
    r[i] = ri;
    z[i] = dinv[i] * ri;
    p[i] = z[i];
  }
}

// alpha = rz / pq ; x += alpha p ; r -= alpha q ; z = dinv r
__global__ void k_cg_update(int n, const double *__restrict__ sc, double *__restrict__ x, double *__restrict__ r,
                            double *__restrict__ z, const double *__restrict__ p, const double *__restrict__ q,
                            const double *__restrict__ dinv) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) {
    const double alpha = sc[0] / sc[2];
    x[i] = fma(alpha, p[i], x[i]);
    const double ri = fma(-alpha, q[i], r[i]);
    r[i] = ri;
    z[i] = dinv[i] * ri;
  }
}

// p = z + (rz_new / rz_old) p
__global__ void k_cg_dir(int n, const double *__restrict__ sc, const double *__restrict__ z, double *__restrict__ p) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) p[i] = fma(sc[4] / sc[0], p[i], z[i]);
}

// scalar bookkeeping on the device: sc[0]=rz, sc[1]=rr, sc[2]=pq, sc[4]=rz_new, sc[5]=rr_new
__global__ void k_cg_roll(double *sc) {
  sc[0] = sc[4];
  sc[1] = sc[5];
}

namespace {

struct SolveData {
  int n = 0, lpr = 8;
  long long *rstart = nullptr, *bidx = nullptr, *dpos = nullptr;
  int *rlen = nullptr, *cgid = nullptr;
  // block SpMV column ids as 16-bit (window, offset) codes against four bases
  // per block (first base -1: the block keeps its 32-bit ids), and the SpMV's
  // bytes per launch
  uint16_t *c16 = nullptr;
  int4 *bbase = nullptr;
  double spmv_bytes = 0.0;
  double *x = nullptr, *b = nullptr, *r = nullptr, *z = nullptr, *p = nullptr, *q = nullptr, *dinv = nullptr;
  double *part = nullptr, *sc = nullptr;
};

}  // namespace

static SolveData *solve_data(gdx_problem *p) { return reinterpret_cast<SolveData *>(p->solve); }

// One solver iteration captured as a hipGraph and replayed: the iteration's
// kernels (SpMV, reductions, updates and the preconditioner) keep their
// arguments from one iteration to the next.  GDX_NO_GRAPHS=1 launches them
// eagerly instead (same kernels, same order).
namespace {
struct IterGraph {
  hipGraph_t g = nullptr;
  hipGraphExec_t exec = nullptr;
  ~IterGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (g) (void)hipGraphDestroy(g);
  }
};

bool graphs_enabled() {
  const char *e = std::getenv("GDX_NO_GRAPHS");
  return !(e && e[0] == '1');
}

// run one iteration: capture it on first use, replay afterwards
template <class F>
int run_iteration(hipStream_t st, IterGraph &G, F &&body) {
  if (!graphs_enabled()) return body();
  if (!G.exec) {
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess) return GDX_ERR_HIP;
    int rc = body();
    const hipError_t e = hipStreamEndCapture(st, &G.g);
    if (rc) return rc;
    if (e != hipSuccess || !G.g) return GDX_ERR_HIP;
    if (hipGraphInstantiate(&G.exec, G.g, nullptr, nullptr, 0) != hipSuccess) return GDX_ERR_HIP;
  }
  return hipGraphLaunch(G.exec, st) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}
}  // namespace

void solve_free(gdx_problem *p) {
  SolveData *s = solve_data(p);
  if (!s) return;
  void *ptrs[] = {s->rstart, s->bidx, s->dpos, s->rlen, s->cgid, s->c16, s->bbase, s->x, s->b, s->r, s->z, s->p, s->q, s->dinv,
                  s->part, s->sc};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete s;
  p->solve = nullptr;
}

static int solve_setup(gdx_problem *p) {
  if (p->solve) return GDX_OK;
  const int64_t n = p->ngdofs;
  if (n >= (int64_t)1 << 31) return GDX_ERR_UNSUPPORTED;
  if (int rc = host_fe(p)) return rc;
  if (int rc = host_csr(p)) return rc;
  std::vector<long long> rstart(n), bidx(n), dpos(n);
  std::vector<int> rlen(n), cgid(std::max<int64_t>(p->nnz_total, 1), 0);
  int64_t maxlen = 0;
  for (const Part &P : p->P) {
    for (int64_t o = 0; o < P.n_own; ++o) {
      const int64_t g = P.row_gid[o];
      rstart[g] = P.val_base + P.rowptr[o];
      rlen[g] = (int)(P.rowptr[o + 1] - P.rowptr[o]);
      bidx[g] = P.row_base + o;
      maxlen = std::max<int64_t>(maxlen, rlen[g]);
      dpos[g] = -1;
      for (int64_t k = P.rowptr[o]; k < P.rowptr[o + 1]; ++k) {
        const int64_t cg = P.col_gid[P.colidx[k]];
        cgid[P.val_base + k] = (int)cg;
        if (cg == g) dpos[g] = P.val_base + k;
      }
      if (dpos[g] < 0) return GDX_ERR_STATE;
    }
  }
  // 16-bit column codes of the block SpMV's streamed blocks (the kernel's own
  // test: rows stored back to back, at most 8 * 512 entries).  Windows of
  // 16384 ids are grown greedily over the block's entries; a block needing
  // more than four keeps its 32-bit ids.
  const int lpr = maxlen > 64 ? 32 : (maxlen > 32 ? 16 : 8);
  const int64_t nblk = (n + kSpmvRows - 1) / kSpmvRows;
  std::vector<uint16_t> c16;
  std::vector<int4> bbase;
  double spmv_bytes = 12.0 * (double)p->owned_nnz + 28.0 * (double)n;
  if (lpr == 8) {
    c16.assign(std::max<int64_t>(p->nnz_total, 1), 0);
    bbase.assign(std::max<int64_t>(nblk, 1), int4{-1, 0, 0, 0});
    spmv_bytes += 16.0 * (double)nblk;
    for (int64_t b = 0; b < nblk; ++b) {
      const int64_t r0 = b * kSpmvRows, r1 = std::min<int64_t>(n, r0 + kSpmvRows);
      bool ok = true;
      for (int64_t r = r0; r + 1 < r1 && ok; ++r) ok = rstart[r] + rlen[r] == rstart[r + 1];
      const long long A = rstart[r0], B = rstart[r1 - 1] + rlen[r1 - 1];
      if (!ok || B - A > 8 * 512 || B <= A) continue;
      int wlo[4], whi[4], nw = 0;
      for (long long k = A; k < B && nw <= 4; ++k) {
        const int c = cgid[k];
        int w = 0;
        while (w < nw && std::max(whi[w], c) - std::min(wlo[w], c) > 16383) ++w;
        if (w == nw) {
          if (nw == 4) {
            nw = 5;
            break;
          }
          wlo[w] = whi[w] = c;
          ++nw;
        }
        wlo[w] = std::min(wlo[w], c);
        whi[w] = std::max(whi[w], c);
      }
      if (nw > 4) continue;
      for (long long k = A; k < B; ++k) {
        const int c = cgid[k];
        int w = 0;
        while (c < wlo[w] || c > whi[w]) ++w;  // every id lies in the window that took it
        c16[k] = (uint16_t)(w << 14 | (c - wlo[w]));
      }
      bbase[b] = int4{wlo[0], nw > 1 ? wlo[1] : 0, nw > 2 ? wlo[2] : 0, nw > 3 ? wlo[3] : 0};
      spmv_bytes -= 2.0 * (double)(B - A);
    }
  }
  SolveData *s = new SolveData;
  p->solve = s;
  s->n = (int)n;
  s->lpr = lpr;
  s->spmv_bytes = spmv_bytes;
  auto up = [](auto **dst, const auto &v) -> hipError_t {
    hipError_t e = hipMalloc(reinterpret_cast<void **>(dst), std::max<size_t>(v.size(), 1) * sizeof(v[0]));
    if (e != hipSuccess) return e;
    return hipMemcpy(*dst, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice);
  };
  hipError_t e = hipSuccess;
  if ((e = up(&s->rstart, rstart)) || (e = up(&s->bidx, bidx)) || (e = up(&s->dpos, dpos)) ||
      (e = up(&s->rlen, rlen)) || (e = up(&s->cgid, cgid)) || (lpr == 8 && ((e = up(&s->c16, c16)) ||
                                                                           (e = up(&s->bbase, bbase))))) {
    solve_free(p);
    return GDX_ERR_ALLOC;
  }
  double **vecs[] = {&s->x, &s->b, &s->r, &s->z, &s->p, &s->q, &s->dinv};
  for (double **v : vecs)
    if (hipMalloc(reinterpret_cast<void **>(v), std::max<int64_t>(n, 1) * sizeof(double)) != hipSuccess) {
      solve_free(p);
      return GDX_ERR_ALLOC;
    }
  if (hipMalloc(reinterpret_cast<void **>(&s->part), 2 * kRedBlocks * sizeof(double)) != hipSuccess ||
      hipMalloc(reinterpret_cast<void **>(&s->sc), 8 * sizeof(double)) != hipSuccess) {
    solve_free(p);
    return GDX_ERR_ALLOC;
  }
  return GDX_OK;
}

static void spmv(gdx_problem *p, SolveData *s, const double *x, double *y, hipStream_t st) {
  const int rows_per_block = kBlk / s->lpr;
  const dim3 g((s->n + rows_per_block - 1) / rows_per_block);
  if (s->lpr == 8)  // rows of at most 32 entries
    hipLaunchKernelGGL((k_spmv_blk<kSpmvRows, 512>), dim3((s->n + kSpmvRows - 1) / kSpmvRows), dim3(512), 0, st, s->n,
                       s->rstart, s->rlen, s->cgid, s->c16, s->bbase, p->dev.values, x, y);
  else if (s->lpr == 16)
    hipLaunchKernelGGL(k_spmv<16>, g, dim3(kBlk), 0, st, s->n, s->rstart, s->rlen, s->cgid, p->dev.values, x, y);
  else
    hipLaunchKernelGGL(k_spmv<32>, g, dim3(kBlk), 0, st, s->n, s->rstart, s->rlen, s->cgid, p->dev.values, x, y);
}

static void dot2(SolveData *s, const double *a, const double *b, const double *c, const double *d, double *out,
                 hipStream_t st) {
  const int nb = std::min(kRedBlocks, std::max(1, (s->n + kBlk - 1) / kBlk));
  hipLaunchKernelGGL(k_dot2, dim3(nb), dim3(kBlk), 0, st, s->n, a, b, c, d, s->part);
  hipLaunchKernelGGL(k_finish2, dim3(1), dim3(kBlk), 0, st, s->part, nb, out);
}

// device time of one SpMV of the assembled matrix (the solver's kernel on
// its own vectors), averaged over reps launches between two hipEvents
int solve_spmv_timed(gdx_problem *p, int reps, double *ms, double *bytes) {
  int rc = solve_setup(p);
  if (rc) return rc;
  SolveData *s = solve_data(p);
  hipStream_t st = p->ctx->stream;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess) return GDX_ERR_HIP;
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return GDX_ERR_HIP;
  }
  spmv(p, s, s->x, s->q, st);  // warm-up
  float t = 0.0f;
  rc = hipEventRecord(e0, st) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
  for (int k = 0; k < reps && !rc; ++k) spmv(p, s, s->x, s->q, st);
  if (!rc && (hipGetLastError() != hipSuccess || hipEventRecord(e1, st) != hipSuccess ||
              hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess))
    rc = GDX_ERR_HIP;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (!rc) *ms = (double)t / reps;
  if (!rc && bytes) *bytes = s->spmv_bytes;
  return rc;
}

int solve_spmv(gdx_problem *p, const double *xh, double *yh) {
  int rc = solve_setup(p);
  if (rc) return rc;
  SolveData *s = solve_data(p);
  hipStream_t st = p->ctx->stream;
  if (hipMemcpyAsync(s->x, xh, s->n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) return GDX_ERR_HIP;
  spmv(p, s, s->x, s->q, st);
  if (hipMemcpyAsync(yh, s->q, s->n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess) return GDX_ERR_HIP;
  return hipStreamSynchronize(st) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// Jacobi-PCG on the assembled system, the solution left in s->x; stops when
// the recursive residual meets rtol ||b||, then computes the true FP64
// residual ||b - A x|| / ||b|| explicitly.  rhs: device vector in global-id
// order (nullptr = the assembled RHS; may be s->b itself); x0_given: start from
// s->x instead of 0.
static int pcg_run(gdx_problem *p, SolveData *s, double rtol, int maxiter, int *iters, double *relres,
                   const double *rhs = nullptr, bool x0_given = false) {
  hipStream_t st = p->ctx->stream;
  const int n = s->n;
  const dim3 g((n + kBlk - 1) / kBlk), b(kBlk);
  if (rhs) {
    if (rhs != s->b && hipMemcpyAsync(s->b, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return GDX_ERR_HIP;
  } else {
    hipLaunchKernelGGL(k_gather_b, g, b, 0, st, n, s->bidx, p->dev.b, s->b);
  }
  hipLaunchKernelGGL(k_inv_diag, g, b, 0, st, n, s->dpos, p->dev.values, s->dinv);
  if (x0_given) {
    spmv(p, s, s->x, s->r, st);  // r = A x0, turned into b - A x0 below
  } else {
    (void)hipMemsetAsync(s->x, 0, n * sizeof(double), st);
    (void)hipMemsetAsync(s->r, 0, n * sizeof(double), st);  // A x0 = 0
  }
  hipLaunchKernelGGL(k_cg_init, g, b, 0, st, n, s->b, s->dinv, s->r, s->z, s->p);
  dot2(s, s->r, s->z, s->b, s->b, s->sc, st);  // sc[0] = rz, sc[1] = bb
  double h[2];
  if (hipMemcpyAsync(h, s->sc, 2 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GDX_ERR_HIP;
  const double bnorm = std::sqrt(h[1]);
  int it = 0;
  const int check_every = 5;
  if (bnorm > 0.0) {
    IterGraph G;
    for (it = 1; it <= maxiter; ++it) {
      int rci = run_iteration(st, G, [&]() {
        spmv(p, s, s->p, s->q, st);
        dot2(s, s->p, s->q, nullptr, nullptr, s->sc + 2, st);  // sc[2] = pq
        hipLaunchKernelGGL(k_cg_update, g, b, 0, st, n, s->sc, s->x, s->r, s->z, s->p, s->q, s->dinv);
        dot2(s, s->r, s->z, s->r, s->r, s->sc + 4, st);  // sc[4] = rz_new, sc[5] = rr_new
        hipLaunchKernelGGL(k_cg_dir, g, b, 0, st, n, s->sc, s->z, s->p);
        hipLaunchKernelGGL(k_cg_roll, dim3(1), dim3(1), 0, st, s->sc);
        return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
      });
      if (rci) return rci;
      if (it % check_every == 0 || it == maxiter) {
        if (hipMemcpyAsync(h, s->sc, 2 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
          return GDX_ERR_HIP;
        if (std::sqrt(h[1]) <= rtol * bnorm) break;
      }
    }
    if (it > maxiter) it = maxiter;
  }
  // true residual
  spmv(p, s, s->x, s->q, st);
  hipLaunchKernelGGL(k_cg_init, g, b, 0, st, n, s->b, s->dinv, s->q, s->z, s->r);  // q = b - Ax
  dot2(s, s->q, s->q, nullptr, nullptr, s->sc + 6, st);
  if (hipMemcpyAsync(h, s->sc + 6, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GDX_ERR_HIP;
  if (iters) *iters = it;
  if (relres) *relres = bnorm > 0.0 ? std::sqrt(h[0]) / bnorm : 0.0;
  return GDX_OK;
}

int solve_pcg(gdx_problem *p, double rtol, int maxiter, double *xh, int *iters, double *relres, double *ms) {
  int rc = solve_setup(p);
  if (rc) return rc;
  SolveData *s = solve_data(p);
  hipStream_t st = p->ctx->stream;
  // timed: the solve, from the assembled system on the device to the solution
  // there (the host copy of the solution is outside)
  if (hipStreamSynchronize(st) != hipSuccess) return GDX_ERR_HIP;
  auto t0 = std::chrono::steady_clock::now();
  if ((rc = pcg_run(p, s, rtol, maxiter, iters, relres))) return rc;
  if (hipStreamSynchronize(st) != hipSuccess) return GDX_ERR_HIP;
  auto t1 = std::chrono::steady_clock::now();
  if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  if (xh && (hipMemcpyAsync(xh, s->x, s->n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess))
    return GDX_ERR_HIP;
  return GDX_OK;
}

// ---------------------------------------------------------------------------
// Flexible CG (Notay's FCG with one stored direction) for a preconditioner that
// may vary between iterations:
//   alpha = p.r / p.q ; x += alpha p ; r -= alpha q ; z = M(r)
//   p = z - (z.q / p.q) p
// sc[0] = p.q, sc[1] = p.r, sc[2] = z.q, sc[3] = r.r
// ---------------------------------------------------------------------------
__global__ void k_fcg_update(int n, const double *__restrict__ sc, double *__restrict__ x, double *__restrict__ r,
                             const double *__restrict__ p, const double *__restrict__ q) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) {
    const double alpha = sc[1] / sc[0];
    x[i] = fma(alpha, p[i], x[i]);
    r[i] = fma(-alpha, q[i], r[i]);
  }
}

__global__ void k_fcg_dir(int n, const double *__restrict__ sc, const double *__restrict__ z, double *__restrict__ p) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) p[i] = fma(-sc[2] / sc[0], p[i], z[i]);
}

__global__ void k_residual(int n, const double *__restrict__ b, const double *__restrict__ ax, double *__restrict__ r) {
  const int i = blockIdx.x * kBlk + threadIdx.x;
  if (i < n) r[i] = b[i] - ax[i];
}

// inverse diagonal of the current values (global-id order) and the device
// SpMV, for preconditioners built outside this file
int solve_dinv(gdx_problem *p, const double **dinv, int *n) {
  int rc = solve_setup(p);
  if (rc) return rc;
  SolveData *s = solve_data(p);
  const dim3 g((s->n + kBlk - 1) / kBlk), b(kBlk);
  hipLaunchKernelGGL(k_inv_diag, g, b, 0, p->ctx->stream, s->n, s->dpos, p->dev.values, s->dinv);
  *dinv = s->dinv;
  *n = s->n;
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// rows of the assembled matrix in global-id order: row g holds rlen[g] values
// at values[rstart[g]..] with column global ids cgid[rstart[g]..]
int solve_csr_view(gdx_problem *p, const long long **rstart, const int **rlen, const int **cgid) {
  int rc = solve_setup(p);
  if (rc) return rc;
  SolveData *s = solve_data(p);
  *rstart = s->rstart;
  *rlen = s->rlen;
  *cgid = s->cgid;
  return GDX_OK;
}

void solve_spmv_dev(gdx_problem *p, const double *x, double *y, hipStream_t st) {
  spmv(p, solve_data(p), x, y, st);
}

// FCG with preconditioner prec on the device RHS rhs (nullptr = the assembled
// RHS), the solution left in s->x; Jacobi-PCG when prec is null.  Falls back to
// Jacobi-PCG from the current iterate when the flexible iteration stops
// reducing the residual, or from zero when it produces a non-finite value.
static int fpcg_run(gdx_problem *p, SolveData *s, const FpcgPrec *prec, const double *rhs, double rtol, int maxiter,
                    int *iters, double *relres, int *fellback) {
  hipStream_t st = p->ctx->stream;
  const int n = s->n;
  const dim3 g((n + kBlk - 1) / kBlk), b(kBlk);
  if (fellback) *fellback = 0;
  if (!prec) return pcg_run(p, s, rtol, maxiter, iters, relres, rhs);
  int rc, it = 0;
  if (rhs) {
    if (hipMemcpyAsync(s->b, rhs, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) return GDX_ERR_HIP;
  } else {
    hipLaunchKernelGGL(k_gather_b, g, b, 0, st, n, s->bidx, p->dev.b, s->b);
  }
  hipLaunchKernelGGL(k_inv_diag, g, b, 0, st, n, s->dpos, p->dev.values, s->dinv);
  (void)hipMemsetAsync(s->x, 0, n * sizeof(double), st);
  if (hipMemcpyAsync(s->r, s->b, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) return GDX_ERR_HIP;
  dot2(s, s->b, s->b, nullptr, nullptr, s->sc + 6, st);
  double bb;
  if (hipMemcpyAsync(&bb, s->sc + 6, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return GDX_ERR_HIP;
  const double bnorm = std::sqrt(bb);
  bool fall = false, nonfinite = false;
  double last_check = bnorm;
  if (bnorm > 0.0) {
    if ((rc = prec->apply(prec->ctx, s->r, s->z, st))) return rc;
    if (hipMemcpyAsync(s->p, s->z, n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess) return GDX_ERR_HIP;
    IterGraph G;
    for (it = 1; it <= maxiter; ++it) {
      rc = run_iteration(st, G, [&]() {
        spmv(p, s, s->p, s->q, st);
        dot2(s, s->p, s->q, s->p, s->r, s->sc, st);  // sc[0] = pq, sc[1] = pr
        hipLaunchKernelGGL(k_fcg_update, g, b, 0, st, n, s->sc, s->x, s->r, s->p, s->q);
        int r2 = prec->apply(prec->ctx, s->r, s->z, st);
        if (r2) return r2;
        dot2(s, s->z, s->q, s->r, s->r, s->sc + 2, st);  // sc[2] = zq, sc[3] = rr
        hipLaunchKernelGGL(k_fcg_dir, g, b, 0, st, n, s->sc, s->z, s->p);
        return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
      });
      if (rc) return rc;
      // the residual is read back every iteration: a preconditioned iteration
      // costs far more than the read, and solves take few iterations
      {
        double h[2];
        if (hipMemcpyAsync(h, s->sc + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
          return GDX_ERR_HIP;
        const double rn = std::sqrt(h[1]);
        if (!std::isfinite(rn) || !std::isfinite(h[0])) {
          fall = true;
          nonfinite = true;
          break;
        }
        if (rn <= rtol * bnorm) break;
        if (it % 50 == 0) {  // stagnation guard
          if (rn > 0.999 * last_check) {
            fall = true;
            break;
          }
          last_check = rn;
        }
      }
    }
    if (it > maxiter) it = maxiter;
  }
  if (fall) {
    // Jacobi-PCG from the current iterate (or from zero if it is not finite);
    // s->b still holds the right-hand side
    if (fellback) *fellback = 1;
    int it2 = 0;
    if ((rc = pcg_run(p, s, rtol, std::max(1, maxiter - it), &it2, relres, s->b, !nonfinite))) return rc;
    it += it2;
  } else {
    // true residual
    spmv(p, s, s->x, s->q, st);
    hipLaunchKernelGGL(k_residual, g, b, 0, st, n, s->b, s->q, s->r);
    dot2(s, s->r, s->r, nullptr, nullptr, s->sc + 6, st);
    double h;
    if (hipMemcpyAsync(&h, s->sc + 6, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return GDX_ERR_HIP;
    if (relres) *relres = bnorm > 0.0 ? std::sqrt(h) / bnorm : 0.0;
  }
  if (iters) *iters = it;
  return GDX_OK;
}

// fpcg_run on a host RHS (nullptr = the assembled RHS), the solution copied to xh
int solve_fpcg(gdx_problem *p, const FpcgPrec *prec, const double *rhs_h, double rtol, int maxiter, double *xh,
               int *iters, double *relres, double *ms, int *fellback) {
  int rc = solve_setup(p);
  if (rc) return rc;
  SolveData *s = solve_data(p);
  hipStream_t st = p->ctx->stream;
  const int n = s->n;
  using clk = std::chrono::steady_clock;
  double *rhs_d = nullptr;
  if (rhs_h) {
    if (hipMalloc(reinterpret_cast<void **>(&rhs_d), std::max(n, 1) * sizeof(double)) != hipSuccess)
      return GDX_ERR_ALLOC;
    if (hipMemcpyAsync(rhs_d, rhs_h, n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) {
      (void)hipFree(rhs_d);
      return GDX_ERR_HIP;
    }
  }
  // timed: the solve, from the right-hand side resident on the device to the
  // solution there (the host copies of both are outside)
  rc = hipStreamSynchronize(st) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
  const auto t0 = clk::now();
  if (!rc) rc = fpcg_run(p, s, prec, rhs_d, rtol, maxiter, iters, relres, fellback);
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = GDX_ERR_HIP;
  if (ms) *ms = std::chrono::duration<double, std::milli>(clk::now() - t0).count();
  if (!rc && xh &&
      (hipMemcpyAsync(xh, s->x, n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
       hipStreamSynchronize(st) != hipSuccess))
    rc = GDX_ERR_HIP;
  if (rhs_d) {
    (void)hipStreamSynchronize(st);
    (void)hipFree(rhs_d);
  }
  return rc;
}

// x -= d
__global__ void k_sub(int n, double *__restrict__ x, const double *__restrict__ d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] -= d[i];
}

// Newton on the flux operator: r = residual at x (into the RHS), J = Jacobian
// (into the CSR values), J dx = r by Jacobi-PCG or two-grid FPCG, x -= dx.  All vectors stay on
// the device; the host reads one norm per iteration.
int solve_newton(gdx_problem *p, const gdx_newton_opts *o, double *xh, gdx_newton_report *rep) {
  int rc = solve_setup(p);
  if (rc) return rc;
  if ((rc = flux_setup(p))) return rc;
  SolveData *s = solve_data(p);
  hipStream_t st = p->ctx->stream;
  const int n = s->n;
  const dim3 g((n + kBlk - 1) / kBlk), b(kBlk);
  double *xs = flux_state(p);
  using clk = std::chrono::steady_clock;
  auto ms_since = [](clk::time_point t) { return std::chrono::duration<double, std::milli>(clk::now() - t).count(); };
  double ms_asm = 0.0, ms_sol = 0.0;
  if (xh) {
    if (hipMemcpyAsync(xs, xh, n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess) return GDX_ERR_HIP;
  } else if (n > 0 && hipMemsetAsync(xs, 0, n * sizeof(double), st) != hipSuccess) {
    return GDX_ERR_HIP;
  }
  // timed (ms_total): from the initial guess resident on the device to the
  // solution there (the host copies of x are outside)
  if (hipStreamSynchronize(st) != hipSuccess) return GDX_ERR_HIP;
  const auto t0 = clk::now();
  // assemble at xs and return ||r||
  auto assemble = [&](bool linear, double *rnorm) -> int {
    const auto ta = clk::now();
    int r = flux_assemble(p, xs, linear, o->jac_precision == GDX_PREC_FP32, st);
    if (r) return r;
    hipLaunchKernelGGL(k_gather_b, g, b, 0, st, n, s->bidx, p->dev.b, s->b);
    dot2(s, s->b, s->b, nullptr, nullptr, s->sc + 6, st);
    double h;
    if (hipMemcpyAsync(&h, s->sc + 6, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return GDX_ERR_HIP;
    *rnorm = std::sqrt(h);
    ms_asm += ms_since(ta);
    return GDX_OK;
  };
  // FPCG with the two-grid preconditioner, rebuilt from the current Jacobian
  // (fine diagonal and Galerkin coarse operator), or Jacobi-PCG
  const FpcgPrec prec{twogrid_apply, o->precond};
  int fallbacks = 0;
  auto correct = [&](int *li) -> int {
    const auto ts = clk::now();
    double rr;
    int r, fb = 0;
    if (o->precond) {
      if ((r = twogrid_attach(o->precond, p))) return r;
      r = fpcg_run(p, s, &prec, nullptr, o->lin_rtol, o->lin_maxiter, li, &rr, &fb);
      fallbacks += fb;
    } else {
      r = pcg_run(p, s, o->lin_rtol, o->lin_maxiter, li, &rr);
    }
    if (r) return r;
    hipLaunchKernelGGL(k_sub, g, b, 0, st, n, xs, s->x);
    if (hipStreamSynchronize(st) != hipSuccess) return GDX_ERR_HIP;
    ms_sol += ms_since(ts);
    return GDX_OK;
  };
  double res_ref = 0.0, res = 0.0;
  int lin_total = 0, li = 0;
  if ((rc = assemble(false, &res_ref))) return rc;
  if (o->init == GDX_NEWTON_INIT_POISSON) {
    double r0;
    if ((rc = assemble(true, &r0))) return rc;
    if (r0 > 0.0) {
      if ((rc = correct(&li))) return rc;
      lin_total += li;
    }
  }
  const double tol = std::max(o->atol, o->rtol * res_ref);
  int it = 0;
  bool conv = false;
  for (;; ++it) {
    if ((rc = assemble(false, &res))) return rc;
    if (!(res == res)) break;  // NaN: diverged
    if (res <= tol) {
      conv = true;
      break;
    }
    if (it == o->maxiter) break;
    if ((rc = correct(&li))) return rc;
    lin_total += li;
  }
  if (hipStreamSynchronize(st) != hipSuccess) return GDX_ERR_HIP;
  const double ms_total = ms_since(t0);
  if (xh && (hipMemcpyAsync(xh, xs, n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
             hipStreamSynchronize(st) != hipSuccess))
    return GDX_ERR_HIP;
  if (rep) {
    rep->iters = it;
    rep->lin_iters = lin_total;
    rep->converged = conv ? 1 : 0;
    rep->res_ref = res_ref;
    rep->res_final = res;
    rep->ms_total = ms_total;
    rep->ms_assembly = ms_asm;
    rep->ms_solve = ms_sol;
    rep->lin_fallbacks = fallbacks;
  }
  return GDX_OK;
}

}  // namespace gdx

// Device PVectors on a problem's row or column PRange and the PArrays
// operations the solve side of the reference uses on them:
//
//   consistent!(v)      ghost entries take their owner's value
//   assemble!(v)        ghost entries are added into their owner's entry, then zeroed
//   axpy!(a, x, y)      y = a x + y, then consistent!(y)   (src/Algebra.jl:42-49)
//   change_ghost(a, r)  own values onto another PRange, ghosts zero, optionally
//                       consistent!                        (src/Algebra.jl:270-290)
//   mul!(y, A, x)       consistent!(x), then every part's local CSR times its
//                       local x (own and ghost columns)
//
// A PVector holds every part's local vector (own entries, then ghost entries)
// concatenated in part order on the device: on the row PRange this is the
// layout of the assembled b.  Owned entries of part q hold consecutive global
// ids (rank-blocked numbering), so the owner slot of global id g is
// base[q] + g - first[q].  assemble! adds the ghost copies of an entry in
// ascending (part, local index) order, the order in which a sequential pass over
// the parts would receive them; the result does not depend on the launch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <new>
#include <vector>

#include "gdx_internal.h"

struct gdx_pvector {
  gdx_problem *p = nullptr;  // nullptr once the problem released its device memory
  int which = 0;             // GDX_PRANGE_ROWS or GDX_PRANGE_COLS
  int64_t n = 0;
  double *v = nullptr;
};

namespace gdx {

namespace {

constexpr int kB = 256;

struct RangeMaps {
  bool built = false;
  int64_t n = 0, nghost = 0, nasm = 0;
  long long *gdst = nullptr, *gsrc = nullptr;  // ghost slot <- owner slot
  long long *own = nullptr;                    // global id -> its owner slot
  long long *aown = nullptr, *aptr = nullptr, *asrc = nullptr;  // owner slot, ghost copies
};

struct PvData {
  RangeMaps r[2];
  bool mul_built = false;
  int lpr = 8;
  long long *rstart = nullptr, *rcb = nullptr;  // per row: first value, column base of its part
  int *rlen = nullptr, *colidx = nullptr;       // per row: length; per value: local column id
};

PvData *pvd(gdx_problem *p) { return reinterpret_cast<PvData *>(p->pv); }

__global__ void k_pv_gather(int64_t n, const long long *__restrict__ dst, const long long *__restrict__ src,
                            double *__restrict__ v) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < n) v[dst[i]] = v[src[i]];
}

__global__ void k_pv_zero(int64_t n, const long long *__restrict__ dst, double *__restrict__ v) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < n) v[dst[i]] = 0.0;
}

// owner entry += its ghost copies, in the order of the list
__global__ void k_pv_assemble(int64_t m, const long long *__restrict__ aown, const long long *__restrict__ aptr,
                              const long long *__restrict__ asrc, double *__restrict__ v) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i >= m) return;
  double acc = v[aown[i]];
  for (long long k = aptr[i]; k < aptr[i + 1]; ++k) acc += v[asrc[k]];
  v[aown[i]] = acc;
}

__global__ void k_pv_axpy(int64_t n, double a, const double *__restrict__ x, double *__restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < n) y[i] = fma(a, x[i], y[i]);
}

__global__ void k_pv_fill(int64_t n, double a, double *__restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (i < n) y[i] = a;
}

// own values of every global id from one layout to another
__global__ void k_pv_own_copy(int64_t ng, const long long *__restrict__ os, const long long *__restrict__ od,
                              const double *__restrict__ s, double *__restrict__ d) {
  const int64_t g = blockIdx.x * (int64_t)kB + threadIdx.x;
  if (g < ng) d[od[g]] = s[os[g]];
}

// local CSR of every part times the part's local x; LPR lanes per row
template <int LPR>
__global__ __launch_bounds__(kB) void k_pv_mul(int64_t nrows, const long long *__restrict__ rstart,
                                               const int *__restrict__ rlen, const long long *__restrict__ rcb,
                                               const int *__restrict__ colidx, const double *__restrict__ val,
                                               const double *__restrict__ x, double *__restrict__ y) {
  const int lane = threadIdx.x % LPR;
  const int64_t row = (blockIdx.x * (int64_t)kB + threadIdx.x) / LPR;
  double acc = 0.0;
  if (row < nrows) {
    const long long s = rstart[row], cb = rcb[row];
    const int len = rlen[row];
    for (int k = lane; k < len; k += LPR) acc = fma(val[s + k], x[cb + colidx[s + k]], acc);
  }
#pragma unroll
  for (int o = LPR / 2; o > 0; o >>= 1) acc += __shfl_xor(acc, o, LPR);
  if (row < nrows && lane == 0) y[row] = acc;
}

inline unsigned nblk(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + kB - 1) / kB); }

template <class T>
bool upload(T **dst, const std::vector<T> &src) {
  const size_t bytes = std::max<size_t>(src.size(), 1) * sizeof(T);
  if (hipMalloc(reinterpret_cast<void **>(dst), bytes) != hipSuccess) {
    *dst = nullptr;
    return false;
  }
  return src.empty() || hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice) == hipSuccess;
}

void free_range(RangeMaps &R) {
  void *ptrs[] = {R.gdst, R.gsrc, R.own, R.aown, R.aptr, R.asrc};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  R = RangeMaps{};
}

// local lengths, global ids and owners of one PRange of part r
void range_of(const Part &P, int which, int64_t &len, const std::vector<int64_t> *&gid,
              const std::vector<int32_t> *&owner) {
  len = which == GDX_PRANGE_ROWS ? P.nrows : P.ncols;
  gid = which == GDX_PRANGE_ROWS ? &P.row_gid : &P.col_gid;
  owner = which == GDX_PRANGE_ROWS ? &P.row_owner : &P.col_owner;
}

int build_range(gdx_problem *p, int which) {
  PvData *D = pvd(p);
  RangeMaps &R = D->r[which];
  if (R.built) return GDX_OK;
  if (int rc = host_fe(p)) return rc;
  const int np = p->nparts;
  std::vector<int64_t> base(np + 1, 0), first(np, 0);
  for (int r = 0; r < np; ++r) {
    int64_t len;
    const std::vector<int64_t> *gid;
    const std::vector<int32_t> *owner;
    range_of(p->P[r], which, len, gid, owner);
    base[r + 1] = base[r] + len;
    const Part &P = p->P[r];
    first[r] = P.n_own > 0 ? (*gid)[0] : 0;
    for (int64_t o = 0; o < P.n_own; ++o)
      if ((*gid)[o] != first[r] + o) {
        set_error("internal: owned global ids of a part are not consecutive");
        return GDX_ERR_STATE;
      }
  }
  std::vector<long long> own(std::max<int64_t>(p->ngdofs, 0), -1), gdst, gsrc;
  for (int r = 0; r < np; ++r)
    for (int64_t o = 0; o < p->P[r].n_own; ++o) {
      if (first[r] + o >= p->ngdofs) {
        set_error("internal: owned global id out of range");
        return GDX_ERR_STATE;
      }
      own[first[r] + o] = base[r] + o;
    }
  for (long long s : own)
    if (s < 0) {
      set_error("internal: a global id without an owner");
      return GDX_ERR_STATE;
    }
  // ghost slots and their owner slots, in ascending (part, local index) order
  std::vector<long long> cnt(std::max<int64_t>(p->ngdofs, 0) + 1, 0);
  for (int r = 0; r < np; ++r) {
    int64_t len;
    const std::vector<int64_t> *gid;
    const std::vector<int32_t> *owner;
    range_of(p->P[r], which, len, gid, owner);
    for (int64_t i = p->P[r].n_own; i < len; ++i) {
      const int64_t g = (*gid)[i];
      if (g < 0 || g >= p->ngdofs || own[g] < 0 || (*owner)[i] < 0 || (*owner)[i] >= np) {
        set_error("internal: ghost entry without an owner slot");
        return GDX_ERR_STATE;
      }
      gdst.push_back(base[r] + i);
      gsrc.push_back(own[g]);
      ++cnt[g + 1];
    }
  }
  // per global id with copies: its owner slot and the copies' slots, in the order above
  for (int64_t g = 0; g < p->ngdofs; ++g) cnt[g + 1] += cnt[g];
  std::vector<long long> fill(cnt.begin(), cnt.end() - 1), asrc(gdst.size());
  {
    size_t k = 0;
    for (int r = 0; r < np; ++r) {
      int64_t len;
      const std::vector<int64_t> *gid;
      const std::vector<int32_t> *owner;
      range_of(p->P[r], which, len, gid, owner);
      for (int64_t i = p->P[r].n_own; i < len; ++i, ++k) asrc[fill[(*gid)[i]]++] = gdst[k];
    }
  }
  std::vector<long long> aown, aptr{0};
  for (int64_t g = 0; g < p->ngdofs; ++g)
    if (cnt[g + 1] > cnt[g]) {
      aown.push_back(own[g]);
      aptr.push_back(cnt[g + 1]);
    }
  R.n = base[np];
  R.nghost = (int64_t)gdst.size();
  R.nasm = (int64_t)aown.size();
  if (!upload(&R.gdst, gdst) || !upload(&R.gsrc, gsrc) || !upload(&R.own, own) || !upload(&R.aown, aown) ||
      !upload(&R.aptr, aptr) || !upload(&R.asrc, asrc)) {
    free_range(R);
    set_error("device allocation failed for the PRange maps");
    return GDX_ERR_ALLOC;
  }
  R.built = true;
  return GDX_OK;
}

int build_mul(gdx_problem *p) {
  PvData *D = pvd(p);
  if (D->mul_built) return GDX_OK;
  int rc = build_range(p, GDX_PRANGE_COLS);
  if (rc) return rc;
  if ((rc = host_csr(p))) return rc;
  std::vector<long long> rstart, rcb;
  std::vector<int> rlen, colidx(std::max<int64_t>(p->nnz_total, 0));
  rstart.reserve(p->rows_total);
  int64_t cbase = 0, maxlen = 0;
  for (const Part &P : p->P) {
    for (int64_t r = 0; r < P.nrows; ++r) {
      rstart.push_back(P.val_base + P.rowptr[r]);
      rlen.push_back((int)(P.rowptr[r + 1] - P.rowptr[r]));
      rcb.push_back(cbase);
      maxlen = std::max<int64_t>(maxlen, P.rowptr[r + 1] - P.rowptr[r]);
    }
    std::copy(P.colidx.begin(), P.colidx.end(), colidx.begin() + P.val_base);
    cbase += P.ncols;
  }
  D->lpr = maxlen > 64 ? 32 : (maxlen > 32 ? 16 : 8);
  if (!upload(&D->rstart, rstart) || !upload(&D->rlen, rlen) || !upload(&D->rcb, rcb) ||
      !upload(&D->colidx, colidx)) {
    set_error("device allocation failed for the local CSR column ids");
    return GDX_ERR_ALLOC;
  }
  D->mul_built = true;
  return GDX_OK;
}

int live(const gdx_pvector *v) {
  if (!v) {
    set_error("PVector is NULL");
    return GDX_ERR_ARG;
  }
  if (!v->p || !v->p->ctx) {
    set_error("the PVector's problem released its device memory");
    return GDX_ERR_STATE;
  }
  return hipSetDevice(v->p->ctx->device) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int consistent(gdx_pvector *v, hipStream_t st) {
  const RangeMaps &R = pvd(v->p)->r[v->which];
  if (R.nghost > 0) hipLaunchKernelGGL(k_pv_gather, dim3(nblk(R.nghost)), dim3(kB), 0, st, R.nghost, R.gdst, R.gsrc, v->v);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int done(hipStream_t st) { return hipStreamSynchronize(st) == hipSuccess ? GDX_OK : GDX_ERR_HIP; }

}  // namespace

void pvector_release(gdx_problem *p) {
  for (gdx_pvector *v : p->pvecs) {
    if (v->v) (void)hipFree(v->v);
    v->v = nullptr;
    v->p = nullptr;
  }
  p->pvecs.clear();
  PvData *D = pvd(p);
  if (!D) return;
  free_range(D->r[0]);
  free_range(D->r[1]);
  void *ptrs[] = {D->rstart, D->rcb, D->rlen, D->colidx};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  delete D;
  p->pv = nullptr;
}

}  // namespace gdx

using gdx::set_error;

extern "C" {

int gdx_pvector_create(gdx_problem *p, int32_t which, gdx_pvector **out) {
  if (!p || !out) {
    set_error("NULL argument");
    return GDX_ERR_ARG;
  }
  *out = nullptr;
  if (which != GDX_PRANGE_ROWS && which != GDX_PRANGE_COLS) {
    set_error("which must be GDX_PRANGE_ROWS or GDX_PRANGE_COLS");
    return GDX_ERR_ARG;
  }
  if (!p->ctx || !p->symbolic_done) {
    set_error("the problem needs a device context and its symbolic phase");
    return GDX_ERR_STATE;
  }
  if (hipSetDevice(p->ctx->device) != hipSuccess) return GDX_ERR_HIP;
  if (!p->pv) p->pv = new (std::nothrow) gdx::PvData;
  if (!p->pv) return GDX_ERR_ALLOC;
  int rc = gdx::build_range(p, which);
  if (rc) return rc;
  gdx_pvector *v = new (std::nothrow) gdx_pvector;
  if (!v) return GDX_ERR_ALLOC;
  v->p = p;
  v->which = which;
  v->n = gdx::pvd(p)->r[which].n;
  if (hipMalloc(reinterpret_cast<void **>(&v->v), std::max<int64_t>(v->n, 1) * sizeof(double)) != hipSuccess) {
    delete v;
    set_error("device allocation failed for the PVector");
    return GDX_ERR_ALLOC;
  }
  hipStream_t st = p->ctx->stream;
  hipLaunchKernelGGL(gdx::k_pv_fill, dim3(gdx::nblk(v->n)), dim3(gdx::kB), 0, st, v->n, 0.0, v->v);
  p->pvecs.push_back(v);
  *out = v;
  return gdx::done(st);
}

int gdx_pvector_destroy(gdx_pvector *v) {
  if (!v) {
    set_error("PVector is NULL");
    return GDX_ERR_ARG;
  }
  if (v->p) {
    if (v->p->ctx) (void)hipStreamSynchronize(v->p->ctx->stream);
    gdx::unregister(v->p->pvecs, v);
    if (v->v) (void)hipFree(v->v);
  }
  delete v;
  return GDX_OK;
}

int gdx_pvector_length(gdx_pvector *v, int64_t *n) {
  if (!v || !n) {
    set_error("NULL argument");
    return GDX_ERR_ARG;
  }
  *n = v->n;
  return GDX_OK;
}

int gdx_pvector_device(gdx_pvector *v, double **data) {
  int rc = gdx::live(v);
  if (rc) return rc;
  if (!data) return GDX_ERR_ARG;
  *data = v->v;
  return GDX_OK;
}

int gdx_pvector_set(gdx_pvector *v, const double *h) {
  int rc = gdx::live(v);
  if (rc) return rc;
  if (!h && v->n > 0) return GDX_ERR_ARG;
  hipStream_t st = v->p->ctx->stream;
  if (v->n > 0 && hipMemcpyAsync(v->v, h, v->n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
    return GDX_ERR_HIP;
  return gdx::done(st);
}

int gdx_pvector_get(gdx_pvector *v, double *h) {
  int rc = gdx::live(v);
  if (rc) return rc;
  if (!h && v->n > 0) return GDX_ERR_ARG;
  hipStream_t st = v->p->ctx->stream;
  if (v->n > 0 && hipMemcpyAsync(h, v->v, v->n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    return GDX_ERR_HIP;
  return gdx::done(st);
}

int gdx_pvector_fill(gdx_pvector *v, double value) {
  int rc = gdx::live(v);
  if (rc) return rc;
  hipStream_t st = v->p->ctx->stream;
  hipLaunchKernelGGL(gdx::k_pv_fill, dim3(gdx::nblk(v->n)), dim3(gdx::kB), 0, st, v->n, value, v->v);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int gdx_pvector_copy_rhs(gdx_pvector *v) {
  int rc = gdx::live(v);
  if (rc) return rc;
  if (v->which != GDX_PRANGE_ROWS) {
    set_error("the assembled b lives on the row PRange");
    return GDX_ERR_ARG;
  }
  if (!v->p->dev.b || v->n != v->p->rows_total) {
    set_error("the problem has no assembled right-hand side");
    return GDX_ERR_STATE;
  }
  hipStream_t st = v->p->ctx->stream;
  if (v->n > 0 && hipMemcpyAsync(v->v, v->p->dev.b, v->n * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
    return GDX_ERR_HIP;
  return GDX_OK;
}

int gdx_pvector_consistent(gdx_pvector *v) {
  int rc = gdx::live(v);
  if (rc) return rc;
  return gdx::consistent(v, v->p->ctx->stream);
}

int gdx_pvector_assemble(gdx_pvector *v) {
  int rc = gdx::live(v);
  if (rc) return rc;
  const gdx::RangeMaps &R = gdx::pvd(v->p)->r[v->which];
  hipStream_t st = v->p->ctx->stream;
  if (R.nasm > 0)
    hipLaunchKernelGGL(gdx::k_pv_assemble, dim3(gdx::nblk(R.nasm)), dim3(gdx::kB), 0, st, R.nasm, R.aown, R.aptr,
                       R.asrc, v->v);
  if (R.nghost > 0)
    hipLaunchKernelGGL(gdx::k_pv_zero, dim3(gdx::nblk(R.nghost)), dim3(gdx::kB), 0, st, R.nghost, R.gdst, v->v);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int gdx_pvector_axpy(double alpha, gdx_pvector *x, gdx_pvector *y) {
  int rc = gdx::live(x);
  if (!rc) rc = gdx::live(y);
  if (rc) return rc;
  if (x->p != y->p || x->which != y->which) {
    set_error("axpy! needs two PVectors on the same PRange");
    return GDX_ERR_ARG;
  }
  hipStream_t st = y->p->ctx->stream;
  hipLaunchKernelGGL(gdx::k_pv_axpy, dim3(gdx::nblk(y->n)), dim3(gdx::kB), 0, st, y->n, alpha, x->v, y->v);
  return gdx::consistent(y, st);
}

int gdx_pvector_change_ghost(gdx_pvector *src, gdx_pvector *dst, int32_t make_consistent) {
  int rc = gdx::live(src);
  if (!rc) rc = gdx::live(dst);
  if (rc) return rc;
  if (src->p != dst->p || src == dst) {
    set_error("change_ghost needs two PVectors of the same problem");
    return GDX_ERR_ARG;
  }
  gdx_problem *p = dst->p;
  const gdx::RangeMaps &S = gdx::pvd(p)->r[src->which], &D = gdx::pvd(p)->r[dst->which];
  hipStream_t st = p->ctx->stream;
  if (D.nghost > 0) hipLaunchKernelGGL(gdx::k_pv_zero, dim3(gdx::nblk(D.nghost)), dim3(gdx::kB), 0, st, D.nghost, D.gdst, dst->v);
  if (p->ngdofs > 0)
    hipLaunchKernelGGL(gdx::k_pv_own_copy, dim3(gdx::nblk(p->ngdofs)), dim3(gdx::kB), 0, st, p->ngdofs, S.own, D.own,
                       src->v, dst->v);
  if (hipGetLastError() != hipSuccess) return GDX_ERR_HIP;
  return make_consistent ? gdx::consistent(dst, st) : GDX_OK;
}

int gdx_psparse_mul(gdx_problem *p, gdx_pvector *y, gdx_pvector *x) {
  if (!p) {
    set_error("problem is NULL");
    return GDX_ERR_ARG;
  }
  int rc = gdx::live(x);
  if (!rc) rc = gdx::live(y);
  if (rc) return rc;
  if (x->p != p || y->p != p || x->which != GDX_PRANGE_COLS || y->which != GDX_PRANGE_ROWS) {
    set_error("mul!(y, A, x) needs y on A's row PRange and x on its column PRange");
    return GDX_ERR_ARG;
  }
  if (!p->dev.values) {
    set_error("the problem has no device values");
    return GDX_ERR_STATE;
  }
  if ((rc = gdx::build_mul(p))) return rc;
  hipStream_t st = p->ctx->stream;
  if ((rc = gdx::consistent(x, st))) return rc;
  const gdx::PvData *D = gdx::pvd(p);
  const int64_t n = p->rows_total;
  if (n > 0) {
    const dim3 g((unsigned)((n * D->lpr + gdx::kB - 1) / gdx::kB)), b(gdx::kB);
    if (D->lpr == 8)
      hipLaunchKernelGGL(gdx::k_pv_mul<8>, g, b, 0, st, n, D->rstart, D->rlen, D->rcb, D->colidx, p->dev.values, x->v, y->v);
    else if (D->lpr == 16)
      hipLaunchKernelGGL(gdx::k_pv_mul<16>, g, b, 0, st, n, D->rstart, D->rlen, D->rcb, D->colidx, p->dev.values, x->v, y->v);
    else
      hipLaunchKernelGGL(gdx::k_pv_mul<32>, g, b, 0, st, n, D->rstart, D->rlen, D->rcb, D->colidx, p->dev.values, x->v, y->v);
  }
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

}  // extern "C"

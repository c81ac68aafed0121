// Batched numeric assembly of libgdx: many independent problems (uneven
// sizes and partitions) assembled by one graph launch.
//
//   gdx_batch_add       problems after their symbolic phase, on one context
//   gdx_batch_autotune  per 3D problem, times the z-marching matrix kernel
//                       with columns of 16, 32 and 64 z rows per workgroup
//                       (cells per block), the per-tile kernel and its
//                       direct-store and wave-slab forms, and keeps the fastest
//   gdx_batch_capture   records every problem's numeric launches (coefficient,
//                       RHS and matrix kernels; no memsets) into one hipGraph
//                       of independent branches: the problems are spread over
//                       kBranches streams (largest first, each to the branch
//                       with the least work), forked from and joined back into
//                       the context's stream, so small problems that fill few
//                       CUs run beside each other and beside the large ones
//   gdx_batch_launch    replays it on the context's stream
//
// The z-marching column length only splits the rows over workgroups: each
// row's values are computed from the same data in the same order, so those
// tilings give bit-identical matrices; the per-tile kernels contract in
// another order and agree to rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "gdx_internal.h"

struct gdx_batch {
  gdx_ctx *ctx = nullptr;
  std::vector<gdx_problem *> probs;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // capture-time fork/join streams and events (default priority)
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> ev;  // [0] fork, [1..] joins
};

namespace {

// branches of the batch graph: the context's stream plus kBranches - 1 side
// streams (the device runs at most 4 hardware queues per process by default)
constexpr int kBranches = 4;

int ensure_branches(gdx_batch *b) {
  while ((int)b->side.size() < kBranches - 1) {
    hipStream_t s;
    if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return GDX_ERR_HIP;
    b->side.push_back(s);
  }
  while ((int)b->ev.size() < kBranches) {
    hipEvent_t e;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return GDX_ERR_HIP;
    b->ev.push_back(e);
  }
  return GDX_OK;
}

void free_branches(gdx_batch *b) {
  for (hipStream_t s : b->side) (void)hipStreamDestroy(s);
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  b->side.clear();
  b->ev.clear();
}

int fail(int code, const char *msg) {
  gdx::set_error(msg);
  return code;
}

void drop_graph(gdx_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

}  // namespace

extern "C" {

int gdx_batch_create(gdx_ctx *ctx, gdx_batch **out) {
  if (!ctx || !out) return fail(GDX_ERR_ARG, "NULL argument");
  *out = new (std::nothrow) gdx_batch;
  if (!*out) return fail(GDX_ERR_ALLOC, "host allocation failed");
  (*out)->ctx = ctx;
  ctx->batches.push_back(*out);
  return GDX_OK;
}

int gdx_batch_add(gdx_batch *b, gdx_problem *p) {
  if (!b || !p) return fail(GDX_ERR_ARG, "NULL argument");
  if (!b->ctx) return fail(GDX_ERR_STATE, "the batch's context was finalized");
  if (p->ctx != b->ctx) return fail(GDX_ERR_ARG, "the problem belongs to another context");
  if (!p->symbolic_done) return fail(GDX_ERR_STATE, "gdx_symbolic has not run");
  if (std::find(b->probs.begin(), b->probs.end(), p) != b->probs.end())
    return fail(GDX_ERR_ARG, "the problem is already in the batch");
  drop_graph(b);
  b->probs.push_back(p);
  return GDX_OK;
}

int gdx_batch_autotune(gdx_batch *b, int32_t reps, int32_t *chosen) {
  if (!b || reps < 1 || reps > 100) return fail(GDX_ERR_ARG, "invalid argument");
  if (!b->ctx) return fail(GDX_ERR_STATE, "the batch's context was finalized");
  if (hipSetDevice(b->ctx->device) != hipSuccess) return fail(GDX_ERR_HIP, "hipSetDevice failed");
  hipStream_t st = b->ctx->stream;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return fail(GDX_ERR_HIP, "hipEventCreate failed");
  drop_graph(b);
  int rc = GDX_OK;
  for (size_t i = 0; i < b->probs.size() && !rc; ++i) {
    gdx_problem *p = b->probs[i];
    int best = p->tileR[2], best_var = p->kvariant3d;
    p->precision = GDX_PREC_FP64;  // the choice is made for (and applies to) the FP64 form
    p->hbm_patterns = false;       // and for the tiles' usual pattern placement
    if (p->D == 3 && p->desc.op_kind == GDX_OP_DIFFUSION) {
      float best_ms = 1e30f;
      // candidates: z-marching columns of 16, 32, 64 z rows; the per-tile kernel
      // (-1), its direct-store form (-2) and its wave-slab form (-3, Q2)
      for (int r3 : {16, 32, 64, -1, -2, -3}) {
        const int var = r3 < 0 ? -r3 : 0;
        if (var == 3 && !gdx::wslab_supported(p->K, p->NQ)) continue;
        if ((rc = gdx::set_kernel_variant(p, var, r3))) break;
        if ((rc = gdx::numeric_launches(p, st))) break;  // warm-up
        if (hipEventRecord(e0, st) != hipSuccess) rc = GDX_ERR_HIP;
        for (int k = 0; k < reps && !rc; ++k) rc = gdx::numeric_launches(p, st);
        float ms = 0.0f;
        if (rc || hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
          rc = rc ? rc : GDX_ERR_HIP;
          break;
        }
        if (ms < best_ms) best_ms = ms, best = p->tileR[2], best_var = var;
      }
      if (!rc) rc = gdx::set_kernel_variant(p, best_var, best);
    }
    if (chosen) chosen[i] = best_var == 0 ? best : -best_var;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) gdx::set_error("autotuning failed");
  return rc;
}

int gdx_batch_capture(gdx_batch *b) {
  if (!b) return fail(GDX_ERR_ARG, "NULL argument");
  if (!b->ctx) return fail(GDX_ERR_STATE, "the batch's context was finalized");
  if (b->probs.empty()) return fail(GDX_ERR_STATE, "the batch is empty");
  if (hipSetDevice(b->ctx->device) != hipSuccess) return fail(GDX_ERR_HIP, "hipSetDevice failed");
  hipStream_t st = b->ctx->stream;
  drop_graph(b);
  if (hipStreamSynchronize(st) != hipSuccess) return fail(GDX_ERR_HIP, "stream failed");
  if (ensure_branches(b)) return fail(GDX_ERR_HIP, "stream or event creation failed");
  // largest problems first, each to the least loaded branch (work = stored values);
  // within a branch the problems keep the batch order
  const int nb = std::min<int>(kBranches, (int)b->probs.size());
  std::vector<size_t> order(b->probs.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t x, size_t y) { return b->probs[x]->nnz_total > b->probs[y]->nnz_total; });
  std::vector<int> branch(b->probs.size(), 0);
  std::vector<double> load(nb, 0.0);
  for (size_t i : order) {
    const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    branch[i] = k;
    load[k] += (double)b->probs[i]->nnz_total + 1.0;
  }
  auto stream_of = [&](int k) { return k == 0 ? st : b->side[k - 1]; };
  if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) != hipSuccess)
    return fail(GDX_ERR_HIP, "hipStreamBeginCapture failed");
  int rc = GDX_OK;
  if (hipEventRecord(b->ev[0], st) != hipSuccess) rc = GDX_ERR_HIP;
  for (int k = 1; k < nb && !rc; ++k)
    if (hipStreamWaitEvent(stream_of(k), b->ev[0], 0) != hipSuccess) rc = GDX_ERR_HIP;
  for (size_t i = 0; i < b->probs.size() && !rc; ++i) rc = gdx::numeric_launches(b->probs[i], stream_of(branch[i]));
  // join every side branch that was forked, also after a failed launch, so the capture can end
  for (int k = 1; k < nb; ++k)
    if (hipEventRecord(b->ev[k], stream_of(k)) != hipSuccess || hipStreamWaitEvent(st, b->ev[k], 0) != hipSuccess)
      rc = rc ? rc : GDX_ERR_HIP;
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc || e != hipSuccess) {
    if (g) (void)hipGraphDestroy(g);
    return fail(rc ? rc : GDX_ERR_HIP, "graph capture failed");
  }
  b->graph = g;
  if (hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0) != hipSuccess) {
    drop_graph(b);
    return fail(GDX_ERR_HIP, "hipGraphInstantiate failed");
  }
  return GDX_OK;
}

int gdx_batch_launch(gdx_batch *b) {
  if (!b) return fail(GDX_ERR_ARG, "NULL argument");
  if (!b->ctx) return fail(GDX_ERR_STATE, "the batch's context was finalized");
  if (!b->exec) return fail(GDX_ERR_STATE, "gdx_batch_capture has not run since the batch changed");
  if (hipSetDevice(b->ctx->device) != hipSuccess) return fail(GDX_ERR_HIP, "hipSetDevice failed");
  return hipGraphLaunch(b->exec, b->ctx->stream) == hipSuccess ? GDX_OK : fail(GDX_ERR_HIP, "hipGraphLaunch failed");
}

int gdx_batch_destroy(gdx_batch *b) {
  if (!b) return fail(GDX_ERR_ARG, "NULL argument");
  if (gdx_ctx *ctx = b->ctx) {
    gdx::release_batch_device(b);
    gdx::unregister(ctx->batches, b);
  }
  delete b;
  return GDX_OK;
}

}  // extern "C"

namespace gdx {
void release_batch_device(gdx_batch *b) {
  if (!b->ctx) return;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  drop_graph(b);
  free_branches(b);
  b->probs.clear();
  b->ctx = nullptr;
}
void batch_invalidate_problem(gdx_ctx *ctx, gdx_problem *p) {
  for (gdx_batch *b : ctx->batches)
    if (std::find(b->probs.begin(), b->probs.end(), p) != b->probs.end()) drop_graph(b);
}
void batch_forget_problem(gdx_ctx *ctx, gdx_problem *p) {
  for (gdx_batch *b : ctx->batches)
    for (size_t i = 0; i < b->probs.size(); ++i)
      if (b->probs[i] == p) {
        drop_graph(b);
        b->probs.erase(b->probs.begin() + i);
        break;
      }
}
}  // namespace gdx

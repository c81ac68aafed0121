// Learned coarse solver of libgdx: network state (device buffers) and the
// host-side driver of its kernels (gdx_nn.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "gdx_internal.h"

struct gdx_nn;

namespace gdx {
namespace nn {

constexpr int kCin = 2, kC0 = 8, kC1 = 16, kC2 = 24, kLayers = 10;

void layer_shapes(int (*shape)[2]);  // [layer] = {outputs, inputs}
int64_t num_params();

struct Net {
  int n = 0;  // level-0 grid extent (multiple of 4)
  hipStream_t st = nullptr;
  int64_t np = 0;
  int64_t woff[kLayers], boff[kLayers];
  int adam_t = 0;
  float *theta = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *thb = nullptr;
  // the layers' weights again, as the forward convolutions read them: layer l
  // at pkoff[l], [output channel, padded to 16][k, padded to a multiple of
  // 32] with k = tap * inputs + input channel when the layer's inputs are a
  // multiple of 8, else input channel * 27 + tap; zeros in the padding;
  // re-made after every update
  uint16_t *wpk = nullptr;
  int64_t pkoff[kLayers + 1];
  // activations (BF16)
  uint16_t *a0 = nullptr, *x1a = nullptr, *x1 = nullptr, *p1 = nullptr, *x2a = nullptr, *x2 = nullptr, *p2 = nullptr,
           *x3a = nullptr, *x3 = nullptr, *u2 = nullptr, *y2a = nullptr, *y2 = nullptr, *u1 = nullptr, *y1a = nullptr;
  float *out = nullptr, *inp_f = nullptr, *dout = nullptr, *dinp = nullptr;
  float *gA0 = nullptr, *gB0 = nullptr, *gC0 = nullptr, *gA1 = nullptr, *gB1 = nullptr, *gC1 = nullptr, *gA2 = nullptr,
        *gB2 = nullptr;
  float *partial = nullptr, *lossbuf = nullptr;
  std::vector<void *> bufs;

  ~Net();
  void release();  // free every device buffer
  int alloc(void **p, size_t bytes);
  int init(int grid, uint64_t seed, hipStream_t s);
  int set_params(const float *host);
  int get_params(float *host);
  int input_from_f32(const float *dev_in);  // inp (FP32 [2][n^3], device) -> a0 (BF16)
  int forward_dev();                        // a0 -> out
  int backward_dev(bool want_dinp);         // dout -> grad (+ dinp)
  int adam_step(float lr, float clip);
  void pack_weights();  // thb -> wpk
};

}  // namespace nn
}  // namespace gdx

namespace gdx {
void twogrid_free(gdx_nn *nn);
int twogrid_attach(gdx_nn *nn, gdx_problem *p);
int twogrid_apply(void *nn, const double *r, double *z, hipStream_t st);
int twogrid_train_step(gdx_nn *nn, uint64_t seed, float lr, int smooth_passes, float *loss);
gdx_problem *twogrid_problem(gdx_nn *nn);
void twogrid_options(gdx_nn *nn, int sweeps, bool use_net);
void twogrid_multigrid(gdx_nn *nn, bool on, int nu);
int twogrid_levels(gdx_nn *nn);
int twogrid_export_level(gdx_nn *nn, int l, int *m, double *h);
int twogrid_export_coarse(gdx_nn *nn, double *host);
}  // namespace gdx

struct gdx_nn {
  gdx_ctx *ctx = nullptr;
  gdx::nn::Net net;
  void *tg = nullptr;  // two-grid preconditioner state (gdx_twogrid.hip)
};

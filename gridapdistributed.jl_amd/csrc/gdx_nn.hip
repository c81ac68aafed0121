// Learned coarse solver of libgdx: a small convolutional multilevel network
// (3 levels, channels 8/16/24, 3x3x3 convolutions, about 60k parameters)
// trained and applied on the device with these kernels only.
//
//   x1 = relu(conv(in)) ; x1 = relu(conv(x1))                       level 0 (n^3)
//   x2 = relu(conv(pool(x1))) ; x2 = relu(conv(x2))                  level 1
//   x3 = relu(conv(pool(x2))) ; x3 = relu(conv(x3))                  level 2
//   y2 = relu(conv(up(x3)) + x2) ; y2 = relu(conv(y2))               level 1
//   y1 = relu(conv(up(y2)) + x1) ; out = conv(y1)                    level 0
//
// Storage: weights and activations in BF16 (round to nearest even), every
// accumulation in FP32, gradients and the Adam state in FP32 with FP32 master
// weights.  Tensors are [channel][z][y][x].  Convolutions are tiled 8x8x4
// outputs per 256-thread workgroup with the input halo and the weights staged
// in LDS; the weight gradient is reduced per workgroup and then across
// workgroups in a fixed order, so training steps are deterministic.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "gdx_nn.h"

namespace gdx {
namespace nn {

namespace {

constexpr int kT = 256;  // threads per workgroup
constexpr int kTX = 8, kTY = 8, kTZ = 4, kHV = (kTX + 2) * (kTY + 2) * (kTZ + 2);
constexpr int kPartialBlocks = 256;

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }
__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u && (u & 0x7fffffu)) return (uint16_t)((u >> 16) | 0x40u);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__host__ __device__ inline int tiles_of(int n) { return ((n + kTX - 1) / kTX) * ((n + kTY - 1) / kTY) * ((n + kTZ - 1) / kTZ); }

__device__ __forceinline__ void tile_origin(int tile, int n, int &x0, int &y0, int &z0) {
  const int ntx = (n + kTX - 1) / kTX, nty = (n + kTY - 1) / kTY;
  x0 = (tile % ntx) * kTX;
  y0 = ((tile / ntx) % nty) * kTY;
  z0 = (tile / (ntx * nty)) * kTZ;
}

// MODE 0: forward, in = BF16 activations with KCI channels, weights W[KCO][KCI][27].
// MODE 1: backward data, in = FP32 gradient with KCI (= the layer's outputs)
//         channels, out = FP32 gradient of the layer's KCO inputs, weights
//         transposed with the taps mirrored.
// G workgroups share a tile, each computing KCO / G of its output channels
// (blockIdx.y picks the group): the small grids of the coarse levels have few
// tiles, and one workgroup per tile would leave most CUs idle.  Every output
// channel accumulates in the same order whatever G is.
template <int KCI, int KCO, int MODE, bool RELU, bool SKIP, bool OUTF32, bool ACCUM, int G = 1>
__global__ __launch_bounds__(kT) void k_conv3(int n, const void *__restrict__ in_, const uint16_t *__restrict__ w,
                                              const uint16_t *__restrict__ bias, const uint16_t *__restrict__ skip,
                                              void *__restrict__ out_) {
  static_assert(KCO % G == 0, "whole channel groups");
  constexpr int CH = KCI < 8 ? KCI : 8, KCG = KCO / G;
  const int cob = blockIdx.y * KCG;
  __shared__ float sx[CH * kHV];
  // weights [c][t][co]: the KCG weights of one input channel and tap are
  // contiguous, read as 128-bit LDS loads in the FMA loop
  __shared__ __align__(16) float sw[KCG * CH * 27];
  int x0, y0, z0;
  tile_origin(blockIdx.x, n, x0, y0, z0);
  const int tid = threadIdx.x, lx = tid & 7, ly = (tid >> 3) & 7, lz = tid >> 6;
  const int64_t n3 = (int64_t)n * n * n;
  float acc[KCG];
#pragma unroll
  for (int co = 0; co < KCG; ++co) acc[co] = MODE == 0 ? bf2f(bias[cob + co]) : 0.0f;
  for (int c0 = 0; c0 < KCI; c0 += CH) {
    __syncthreads();
    for (int i = tid; i < CH * kHV; i += kT) {
      const int c = i / kHV, h = i % kHV, hx = h % 10, hy = (h / 10) % 10, hz = h / 100;
      const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
      float v = 0.0f;
      if (c0 + c < KCI && gx >= 0 && gx < n && gy >= 0 && gy < n && gz >= 0 && gz < n) {
        const int64_t idx = (int64_t)(c0 + c) * n3 + ((int64_t)gz * n + gy) * n + gx;
        v = MODE == 0 ? bf2f(reinterpret_cast<const uint16_t *>(in_)[idx]) : reinterpret_cast<const float *>(in_)[idx];
      }
      sx[i] = v;
    }
    for (int i = tid; i < KCG * CH * 27; i += kT) {
      const int co = i / (CH * 27), r = i % (CH * 27), c = r / 27, t = r % 27;
      float v = 0.0f;
      if (c0 + c < KCI)
        v = MODE == 0 ? bf2f(w[((int64_t)(cob + co) * KCI + c0 + c) * 27 + t])
                      : bf2f(w[((int64_t)(c0 + c) * KCO + cob + co) * 27 + 26 - t]);
      sw[(c * 27 + t) * KCG + co] = v;
    }
    __syncthreads();
    for (int c = 0; c < CH; ++c) {
#pragma unroll
      for (int t = 0; t < 27; ++t) {
        const int dz = t / 9, dy = (t / 3) % 3, dx = t % 3;
        const float v = sx[c * kHV + (lz + dz) * 100 + (ly + dy) * 10 + lx + dx];
#pragma unroll
        for (int co = 0; co < KCG; ++co) acc[co] = fmaf(v, sw[(c * 27 + t) * KCG + co], acc[co]);
      }
    }
  }
  const int gx = x0 + lx, gy = y0 + ly, gz = z0 + lz;
  if (gx >= n || gy >= n || gz >= n) return;
  const int64_t vi = ((int64_t)gz * n + gy) * n + gx;
#pragma unroll
  for (int cg = 0; cg < KCG; ++cg) {
    const int co = cob + cg;
    float v = acc[cg];
    if (SKIP) v += bf2f(skip[co * n3 + vi]);
    if (RELU) v = v < 0.0f ? 0.0f : v;  // NaN passes, as max(v, 0) in the oracle
    if (OUTF32) {
      float *o = reinterpret_cast<float *>(out_) + co * n3 + vi;
      *o = ACCUM ? *o + v : v;
    } else {
      reinterpret_cast<uint16_t *>(out_)[co * n3 + vi] = f2bf(v);
    }
  }
}

// Forward convolution on the matrix cores: an implicit GEMM per tile of
// 8 x 8 x TZ outputs.  For a wave, M = its 16 TZ outputs (TZ blocks of 16: two
// x rows each), N = the output channels (blocks of 16, zero-padded), K =
// (input channel, tap) pairs (27 KCI, padded to a multiple of 32).  A
// fragments come from the BF16 input halo in LDS through a table of (channel,
// tap) offsets, B fragments from the weights staged in LDS as [output
// channel][k]; v_mfma_f32_16x16x32_bf16 accumulates in FP32.
// Lane layout of that instruction (checked on the device): lane l holds
// A[l % 16][8 (l / 16) + j] and B[8 (l / 16) + j][l % 16] for j < 8, and
// returns D[4 (l / 16) + i][l % 16] for i < 4.
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x4_t __attribute__((ext_vector_type(4)));

template <int KCI, int KCO, int TZ, int NBW>
struct ConvMfma {
  static constexpr int KT = KCI * 27, KP = (KT + 31) / 32 * 32, NB = (KCO + 15) / 16, NCO = NB * 16;
  static constexpr int HZ = TZ + 2, HV = 100 * HZ;
  // CM: k = tap KCI + channel and the halo stored [voxel][channel], so that
  // 8 consecutive k are one 16-byte LDS read; otherwise k = channel 27 + tap,
  // the halo [channel][voxel] followed by a zero block for the padded k
  static constexpr bool CM = KCI % 8 == 0;
  static constexpr int SX = CM ? KCI * HV : KCI * HV + HV;
  static constexpr int SW = NBW * 16 * KP;  // this workgroup's output channels of the packed weights
  static constexpr int ROWS = KCI * HZ * 10, RU = (ROWS + 255) / 256;  // halo rows of 10, per thread
  static constexpr int PO = 64 * TZ + 4;                                // output tile pitch per channel (floats)
  static constexpr int SXO = SX > 2 * NBW * 16 * PO ? SX : 2 * NBW * 16 * PO;  // halo, then output tile
};

__device__ __forceinline__ bool al16(const void *p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

// A workgroup of four waves computes the 8 x 8 x TZ outputs of a tile for NBW
// blocks of 16 output channels (blockIdx.y selects the first one); its 4 TZ
// voxel blocks of 16 (two x rows at one z) are split TZ per wave.  Weights
// come packed per layer ([output channel][k], both zero-padded) and are
// staged with 16-byte copies; on grids with n % 8 == 0 each halo row is one
// 16-byte load and two edge values, all of a thread's loads issued before its
// LDS stores, and the outputs go out through LDS as rows of 8 voxels.
template <int KCI, int KCO, bool RELU, bool SKIP, bool OUTF32, int TZ, int NBW>
__global__ __launch_bounds__(256) void k_conv3_mfma(int n, const uint16_t *__restrict__ in,
                                                    const uint16_t *__restrict__ wpk,
                                                    const uint16_t *__restrict__ bias,
                                                    const uint16_t *__restrict__ skip, void *__restrict__ out_) {
  using C = ConvMfma<KCI, KCO, TZ, NBW>;
  __shared__ __align__(16) uint16_t sx[C::SXO];
  __shared__ __align__(16) uint16_t sw[C::SW];
  __shared__ __align__(16) int koff[C::KP];
  const int ntx = (n + 7) / 8, nty = (n + 7) / 8;
  const int x0 = (blockIdx.x % ntx) * 8, y0 = ((blockIdx.x / ntx) % nty) * 8, z0 = (blockIdx.x / (ntx * nty)) * TZ;
  const int nb0 = blockIdx.y * NBW;
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int64_t n3 = (int64_t)n * n * n;
  const bool rows8 = (n & 7) == 0 && al16(in) && al16(out_) && (!SKIP || al16(skip));
  {
    const u16x8_t *src = reinterpret_cast<const u16x8_t *>(wpk + (int64_t)nb0 * 16 * C::KP);
    u16x8_t *dst = reinterpret_cast<u16x8_t *>(sw);
#pragma unroll 4
    for (int i = tid; i < C::SW / 8; i += 256) dst[i] = src[i];
  }
  if (rows8) {
    // halo rows of 10 (x0 + 8 <= n): row r is (c HZ + hz) 10 + hy, at sx[10 r]
    // (CM: (hz 10 + hy) KCI + c, at sx[10 KCI (r / KCI) + c] with stride KCI)
    u16x8_t mid[C::RU];
    uint16_t lf[C::RU], rt[C::RU];
#pragma unroll
    for (int u = 0; u < C::RU; ++u) {
      const int r = tid + 256 * u;
      mid[u] = u16x8_t{0, 0, 0, 0, 0, 0, 0, 0};
      lf[u] = rt[u] = 0;
      const int c = C::CM ? r % KCI : r / (C::HZ * 10);
      const int hy = C::CM ? (r / KCI) % 10 : r % 10, hz = C::CM ? r / (KCI * 10) : (r / 10) % C::HZ;
      const int gy = y0 - 1 + hy, gz = z0 - 1 + hz;
      if (r < C::ROWS && gy >= 0 && gy < n && gz >= 0 && gz < n) {
        const uint16_t *row = in + (int64_t)c * n3 + ((int64_t)gz * n + gy) * n + x0;
        mid[u] = *reinterpret_cast<const u16x8_t *>(row);
        if (x0 > 0) lf[u] = row[-1];
        if (x0 + 8 < n) rt[u] = row[8];
      }
    }
#pragma unroll
    for (int u = 0; u < C::RU; ++u) {
      const int r = tid + 256 * u;
      if (r < C::ROWS) {
        constexpr int st = C::CM ? KCI : 1;
        uint16_t *d = sx + (C::CM ? 10 * KCI * (r / KCI) + r % KCI : 10 * r);
        d[0] = lf[u];
#pragma unroll
        for (int j = 0; j < 8; ++j) d[st * (1 + j)] = mid[u][j];
        d[9 * st] = rt[u];
      }
    }
    if (!C::CM)
      for (int i = tid; i < C::HV; i += 256) sx[KCI * C::HV + i] = 0;
  } else {
    for (int i = tid; i < C::SX; i += 256) {
      const int c = C::CM ? i % KCI : i / C::HV, h = C::CM ? i / KCI : i % C::HV;
      const int hx = h % 10, hy = (h / 10) % 10, hz = h / 100;
      const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
      uint16_t v = 0;
      if (c < KCI && gx >= 0 && gx < n && gy >= 0 && gy < n && gz >= 0 && gz < n)
        v = in[(int64_t)c * n3 + ((int64_t)gz * n + gy) * n + gx];
      sx[i] = v;
    }
  }
  if (C::CM) {
    // per group of 8 k: the group's tap offset times KCI plus its first channel, -1 past KT
    for (int g = tid; g < C::KP / 8; g += 256) {
      const int t = 8 * g / KCI, c0 = 8 * g % KCI;
      koff[g] = t < 27 ? ((t / 9) * 100 + ((t / 3) % 3) * 10 + t % 3) * KCI + c0 : -1;
    }
  } else {
    for (int k = tid; k < C::KP; k += 256) {
      const int ci = k / 27, t = k % 27;
      koff[k] = k < C::KT ? ci * C::HV + (t / 9) * 100 + ((t / 3) % 3) * 10 + t % 3 : KCI * C::HV;
    }
  }
  __syncthreads();
  f32x4_t acc[TZ][NBW];
#pragma unroll
  for (int mb = 0; mb < TZ; ++mb)
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb) acc[mb][nb] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  const int m = lane & 15, kq = (lane >> 4) * 8;
  for (int kc = 0; kc < C::KP; kc += 32) {
    const int *ko = koff + kc + kq;
    bf16x8_t b[NBW];
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb)
      b[nb] = __builtin_bit_cast(bf16x8_t, *reinterpret_cast<const u16x8_t *>(sw + (nb * 16 + m) * C::KP + kc + kq));
    const int kg = C::CM ? koff[(kc + kq) >> 3] : 0, kgc = kg < 0 ? 0 : kg;
#pragma unroll
    for (int mb = 0; mb < TZ; ++mb) {
      const int q = wave * TZ + mb;  // voxel block: z = q / 4, rows 2 (q % 4) and 2 (q % 4) + 1
      const int voff = (q >> 2) * 100 + (2 * (q & 3) + (m >> 3)) * 10 + (m & 7);
      u16x8_t a;
      if (C::CM) {
        a = *reinterpret_cast<const u16x8_t *>(sx + voff * KCI + kgc);
        if (kg < 0) a = u16x8_t{0, 0, 0, 0, 0, 0, 0, 0};
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) a[j] = sx[ko[j] + voff];
      }
      const bf16x8_t av = __builtin_bit_cast(bf16x8_t, a);
#pragma unroll
      for (int nb = 0; nb < NBW; ++nb) acc[mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, b[nb], acc[mb][nb], 0, 0, 0);
    }
  }
  if (rows8) {
    // sums to LDS as [channel][z][y][x] (pitch PO), then rows of 8 voxels out
    __syncthreads();
    float *so = reinterpret_cast<float *>(sx);
#pragma unroll
    for (int nb = 0; nb < NBW; ++nb)
#pragma unroll
      for (int mb = 0; mb < TZ; ++mb) {
        const int q = wave * TZ + mb;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int mm = (lane >> 4) * 4 + i;
          so[(nb * 16 + m) * C::PO + (q >> 2) * 64 + (2 * (q & 3) + (mm >> 3)) * 8 + (mm & 7)] = acc[mb][nb][i];
        }
      }
    __syncthreads();
    for (int r = tid; r < NBW * 16 * TZ * 8; r += 256) {
      const int col = r / (TZ * 8), zl = (r / 8) % TZ, yl = r % 8;
      const int co = nb0 * 16 + col, gz = z0 + zl, gy = y0 + yl;
      if (co >= KCO || gz >= n || gy >= n) continue;
      const float *src = so + col * C::PO + zl * 64 + yl * 8;
      const f32x4_t lo = *reinterpret_cast<const f32x4_t *>(src), hi = *reinterpret_cast<const f32x4_t *>(src + 4);
      const int64_t o = (int64_t)co * n3 + ((int64_t)gz * n + gy) * n + x0;
      u16x8_t sk = u16x8_t{0, 0, 0, 0, 0, 0, 0, 0};
      if (SKIP) sk = *reinterpret_cast<const u16x8_t *>(skip + o);
      const float bco = bf2f(bias[co]);
      float v[8];
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        v[j] = (j < 4 ? lo[j] : hi[j - 4]) + bco;
        if (SKIP) v[j] += bf2f(sk[j]);
        if (RELU) v[j] = v[j] < 0.0f ? 0.0f : v[j];  // NaN passes, as max(v, 0) in the oracle
      }
      if (OUTF32) {
        f32x4_t *d = reinterpret_cast<f32x4_t *>(reinterpret_cast<float *>(out_) + o);
        d[0] = f32x4_t{v[0], v[1], v[2], v[3]};
        d[1] = f32x4_t{v[4], v[5], v[6], v[7]};
      } else {
        u16x8_t h;
#pragma unroll
        for (int j = 0; j < 8; ++j) h[j] = f2bf(v[j]);
        *reinterpret_cast<u16x8_t *>(reinterpret_cast<uint16_t *>(out_) + o) = h;
      }
    }
    return;
  }
#pragma unroll
  for (int nb = 0; nb < NBW; ++nb) {
    const int co = (nb0 + nb) * 16 + m;
    if (co >= KCO) continue;
    const float bco = bf2f(bias[co]);
#pragma unroll
    for (int mb = 0; mb < TZ; ++mb) {
      const int q = wave * TZ + mb, gz = z0 + (q >> 2);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int mm = (lane >> 4) * 4 + i, gx = x0 + (mm & 7), gy = y0 + 2 * (q & 3) + (mm >> 3);
        if (gx >= n || gy >= n || gz >= n) continue;
        const int64_t vi = ((int64_t)gz * n + gy) * n + gx;
        float v = acc[mb][nb][i] + bco;
        if (SKIP) v += bf2f(skip[co * n3 + vi]);
        if (RELU) v = v < 0.0f ? 0.0f : v;  // NaN passes, as max(v, 0) in the oracle
        if (OUTF32) reinterpret_cast<float *>(out_)[co * n3 + vi] = v;
        else reinterpret_cast<uint16_t *>(out_)[co * n3 + vi] = f2bf(v);
      }
    }
  }
}

// thb -> wpk for every layer (one thread per packed element)
struct PackTab {
  int64_t woff[kLayers], pkoff[kLayers + 1];
  int kt[kLayers], kp[kLayers], co[kLayers], ci[kLayers];  // ci: input channels, k tap-major if ci % 8 == 0
};

__global__ void k_pack_weights(PackTab T, const uint16_t *__restrict__ thb, uint16_t *__restrict__ wpk) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= T.pkoff[kLayers]) return;
  int l = 0;
  while (i >= T.pkoff[l + 1]) ++l;
  const int64_t j = i - T.pkoff[l];
  const int co = (int)(j / T.kp[l]), k = (int)(j % T.kp[l]);
  const int src = T.ci[l] % 8 == 0 ? (k % T.ci[l]) * 27 + k / T.ci[l] : k;  // thb holds [co][ci][tap]
  wpk[i] = (co < T.co[l] && k < T.kt[l]) ? thb[T.woff[l] + (int64_t)co * T.kt[l] + src] : (uint16_t)0;
}

// weight and bias gradient of a layer on the matrix cores: per workgroup the
// GEMM  dW[co][ci 27 + t] += g[co][v] x[ci][v + tap t]  over the voxels v of
// its tiles (8 x 8 x TZ), with a column of ones giving db[co].  M = output
// channels (blocks of 16), N = (input channel, tap) and the ones column (blocks
// of 16, every eighth block to each of the 8 waves), K = voxels, four per
// v_mfma_f32_16x16x4f32: g is FP32 and the BF16 activations are exact in FP32.
// Lane layout of that instruction (checked on the device): lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16], and returns D[4 (l / 16) + i][l % 16]
// for i < 4.  Partial sums per workgroup, partial[block][W (CO CI 27) | b
// (CO)], are reduced across workgroups in a fixed order (k_reduce_partials).
template <int CI, int CO, int TZ>
struct WgradMfma {
  static constexpr int NW = 8, NT = 64 * NW;  // waves and lanes per workgroup
  static constexpr int KT = CI * 27, NBT = (KT + 1 + 15) / 16, NBW = (NBT + NW - 1) / NW, MB = (CO + 15) / 16;
  static constexpr int HZ = TZ + 2, HV = 100 * HZ, SX = CI * HV, NV = 64 * TZ;
  static constexpr int XROWS = CI * HZ * 10, XRU = (XROWS + NT - 1) / NT;      // halo rows of 10, per thread
  static constexpr int GROWS = MB * 16 * TZ * 8, GRU = (GROWS + NT - 1) / NT;  // g rows of 8, per thread
};

template <int CI, int CO, int TZ>
__global__ __launch_bounds__(512) void k_wgrad_mfma(int n, const uint16_t *__restrict__ x, const float *__restrict__ g,
                                                    float *__restrict__ partial) {
  using C = WgradMfma<CI, CO, TZ>;
  __shared__ __align__(16) uint16_t sx[C::SX];
  __shared__ __align__(16) float sg[C::MB * 16 * C::NV];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int ntx = (n + 7) / 8, ntz = (n + TZ - 1) / TZ, ntiles = ntx * ntx * ntz;
  const int64_t n3 = (int64_t)n * n * n;
  const bool rows8 = (n & 7) == 0 && al16(x) && al16(g);
  // this lane's column in each of the wave's N blocks: halo offset of (ci, tap), -1 the ones column, -2 padding
  int kof[C::NBW];
#pragma unroll
  for (int j = 0; j < C::NBW; ++j) {
    const int nb = wave + C::NW * j, kk = nb * 16 + (lane & 15);
    const int ci = kk / 27, t = kk % 27;
    kof[j] = (nb >= C::NBT || kk > C::KT) ? -2
             : kk == C::KT                 ? -1
                                           : ci * C::HV + (t / 9) * 100 + ((t / 3) % 3) * 10 + t % 3;
  }
  f32x4_t acc[C::MB][C::NBW];
#pragma unroll
  for (int mb = 0; mb < C::MB; ++mb)
#pragma unroll
    for (int j = 0; j < C::NBW; ++j) acc[mb][j] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int x0 = (tile % ntx) * 8, y0 = ((tile / ntx) % ntx) * 8, z0 = (tile / (ntx * ntx)) * TZ;
    __syncthreads();  // the previous tile's reads are done
    if (rows8) {
      // rows of 8 voxels (x0 + 8 <= n): every load of the thread is issued before its LDS stores
      f32x4_t glo[C::GRU], ghi[C::GRU];
      u16x8_t mid[C::XRU];
      uint16_t lf[C::XRU], rt[C::XRU];
#pragma unroll
      for (int u = 0; u < C::GRU; ++u) {
        const int r = tid + C::NT * u, co = r / (TZ * 8), gz = z0 + (r / 8) % TZ, gy = y0 + r % 8;
        glo[u] = ghi[u] = f32x4_t{0.0f, 0.0f, 0.0f, 0.0f};
        if (r < C::GROWS && co < CO && gz < n && gy < n) {
          const f32x4_t *src = reinterpret_cast<const f32x4_t *>(g + co * n3 + ((int64_t)gz * n + gy) * n + x0);
          glo[u] = src[0];
          ghi[u] = src[1];
        }
      }
#pragma unroll
      for (int u = 0; u < C::XRU; ++u) {
        const int r = tid + C::NT * u, c = r / (C::HZ * 10), gz = z0 - 1 + (r / 10) % C::HZ, gy = y0 - 1 + r % 10;
        mid[u] = u16x8_t{0, 0, 0, 0, 0, 0, 0, 0};
        lf[u] = rt[u] = 0;
        if (r < C::XROWS && gy >= 0 && gy < n && gz >= 0 && gz < n) {
          const uint16_t *row = x + (int64_t)c * n3 + ((int64_t)gz * n + gy) * n + x0;
          mid[u] = *reinterpret_cast<const u16x8_t *>(row);
          if (x0 > 0) lf[u] = row[-1];
          if (x0 + 8 < n) rt[u] = row[8];
        }
      }
#pragma unroll
      for (int u = 0; u < C::GRU; ++u) {
        const int r = tid + C::NT * u;  // g row r: channel r / (8 TZ), voxels (r % (8 TZ)) 8 .. + 7 of the tile
        if (r < C::GROWS) {
          f32x4_t *d = reinterpret_cast<f32x4_t *>(sg + (r / (TZ * 8)) * C::NV + (r % (TZ * 8)) * 8);
          d[0] = glo[u];
          d[1] = ghi[u];
        }
      }
#pragma unroll
      for (int u = 0; u < C::XRU; ++u) {
        const int r = tid + C::NT * u;
        if (r < C::XROWS) {
          uint16_t *d = sx + 10 * r;
          d[0] = lf[u];
#pragma unroll
          for (int j = 0; j < 8; ++j) d[1 + j] = mid[u][j];
          d[9] = rt[u];
        }
      }
    } else {
      for (int i = tid; i < C::MB * 16 * C::NV; i += C::NT) {
        const int co = i / C::NV, v = i % C::NV;
        const int gx = x0 + (v & 7), gy = y0 + ((v >> 3) & 7), gz = z0 + (v >> 6);
        sg[i] = (co < CO && gx < n && gy < n && gz < n) ? g[co * n3 + ((int64_t)gz * n + gy) * n + gx] : 0.0f;
      }
      for (int i = tid; i < C::SX; i += C::NT) {
        const int c = i / C::HV, h = i % C::HV, hx = h % 10, hy = (h / 10) % 10, hz = h / 100;
        const int gx = x0 - 1 + hx, gy = y0 - 1 + hy, gz = z0 - 1 + hz;
        uint16_t v = 0;
        if (gx >= 0 && gx < n && gy >= 0 && gy < n && gz >= 0 && gz < n)
          v = x[(int64_t)c * n3 + ((int64_t)gz * n + gy) * n + gx];
        sx[i] = v;
      }
    }
    __syncthreads();
    for (int k0 = 0; k0 < C::NV; k0 += 4) {
      const int v = k0 + (lane >> 4);
      const int voff = (v >> 6) * 100 + ((v >> 3) & 7) * 10 + (v & 7);
      float a[C::MB];
#pragma unroll
      for (int mb = 0; mb < C::MB; ++mb) a[mb] = sg[(mb * 16 + (lane & 15)) * C::NV + v];
#pragma unroll
      for (int j = 0; j < C::NBW; ++j) {
        if (wave + C::NW * j >= C::NBT) break;  // padding blocks (wave-uniform)
        const int ko = kof[j];
        const float bv = ko >= 0 ? bf2f(sx[ko + voff]) : (ko == -1 ? 1.0f : 0.0f);
#pragma unroll
        for (int mb = 0; mb < C::MB; ++mb) acc[mb][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[mb], bv, acc[mb][j], 0, 0, 0);
      }
    }
  }
  float *dst = partial + (int64_t)blockIdx.x * (C::KT * CO + CO);
#pragma unroll
  for (int mb = 0; mb < C::MB; ++mb)
#pragma unroll
    for (int j = 0; j < C::NBW; ++j) {
      const int nb = wave + C::NW * j, kk = nb * 16 + (lane & 15);
      if (nb >= C::NBT || kk > C::KT) continue;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int co = mb * 16 + 4 * (lane >> 4) + i;
        if (co >= CO) continue;
        if (kk < C::KT) dst[co * C::KT + kk] = acc[mb][j][i];
        else dst[C::KT * CO + co] = acc[mb][j][i];
      }
    }
}

// out[e] = sum over the nb blocks of partial[b][e], in a fixed order: eight
// running sums over b = 8 i + j (so eight loads are in flight), then added j = 0..7
__global__ void k_reduce_partials(int nb, int len, const float *__restrict__ partial, float *__restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= len) return;
  float s[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  int b = 0;
  for (; b + 8 <= nb; b += 8)
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] += partial[(int64_t)(b + j) * len + e];
  for (int j = 0; b + j < nb; ++j) s[j] += partial[(int64_t)(b + j) * len + e];
  float t = 0.0f;
#pragma unroll
  for (int j = 0; j < 8; ++j) t += s[j];
  out[e] = t;
}

// 2x2x2 mean (BF16 -> BF16); n = fine extent
__global__ void k_pool(int C, int n, const uint16_t *__restrict__ in, uint16_t *__restrict__ out) {
  const int m = n / 2;
  const int64_t tot = (int64_t)C * m * m * m;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= tot) return;
  const int x = (int)(i % m), y = (int)((i / m) % m), z = (int)((i / ((int64_t)m * m)) % m), c = (int)(i / ((int64_t)m * m * m));
  const uint16_t *src = in + (int64_t)c * n * n * n;
  float s = 0.0f;
  for (int k = 0; k < 8; ++k)
    s += bf2f(src[((int64_t)(2 * z + (k >> 2)) * n + 2 * y + ((k >> 1) & 1)) * n + 2 * x + (k & 1)]);
  out[i] = f2bf(s * 0.125f);
}

// gradient through the 2x2x2 mean, plus an optional skip gradient: g_in (fine) = g_out/8 (+ skip)
__global__ void k_pool_bw(int C, int n, const float *__restrict__ gout, const float *__restrict__ skip,
                          float *__restrict__ gin) {
  const int64_t tot = (int64_t)C * n * n * n;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= tot) return;
  const int m = n / 2;
  const int x = (int)(i % n), y = (int)((i / n) % n), z = (int)((i / ((int64_t)n * n)) % n), c = (int)(i / ((int64_t)n * n * n));
  float v = gout[(((int64_t)c * m + z / 2) * m + y / 2) * m + x / 2] * 0.125f;
  if (skip) v += skip[i];
  gin[i] = v;
}

// nearest-neighbour x2 upsampling (BF16); n = fine extent
__global__ void k_up(int C, int n, const uint16_t *__restrict__ in, uint16_t *__restrict__ out) {
  const int64_t tot = (int64_t)C * n * n * n;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= tot) return;
  const int m = n / 2;
  const int x = (int)(i % n), y = (int)((i / n) % n), z = (int)((i / ((int64_t)n * n)) % n), c = (int)(i / ((int64_t)n * n * n));
  out[i] = in[(((int64_t)c * m + z / 2) * m + y / 2) * m + x / 2];
}

// gradient through the upsampling: coarse = sum of its 8 fine children
__global__ void k_up_bw(int C, int n, const float *__restrict__ gfine, float *__restrict__ gcoarse) {
  const int m = n / 2;
  const int64_t tot = (int64_t)C * m * m * m;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= tot) return;
  const int x = (int)(i % m), y = (int)((i / m) % m), z = (int)((i / ((int64_t)m * m)) % m), c = (int)(i / ((int64_t)m * m * m));
  const float *src = gfine + (int64_t)c * n * n * n;
  float s = 0.0f;
  for (int k = 0; k < 8; ++k) s += src[((int64_t)(2 * z + (k >> 2)) * n + 2 * y + ((k >> 1) & 1)) * n + 2 * x + (k & 1)];
  gcoarse[i] = s;
}

// g *= (y > 0): gradient through a ReLU whose output y was stored
__global__ void k_relu_mask(int64_t tot, float *__restrict__ g, const uint16_t *__restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < tot && !(bf2f(y[i]) > 0.0f)) g[i] = 0.0f;
}

__global__ void k_to_bf16(int64_t tot, const float *__restrict__ in, uint16_t *__restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < tot) out[i] = f2bf(in[i]);
}

__global__ void k_adam(int64_t np, float *__restrict__ th, uint16_t *__restrict__ thb, const float *__restrict__ g,
                       float *__restrict__ m, float *__restrict__ v, float lr, float b1, float b2, float eps,
                       float c1, float c2, float clip) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= np) return;
  const float gi = g[i] * clip;
  const float mi = b1 * m[i] + (1.0f - b1) * gi, vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float t = th[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  th[i] = t;
  thb[i] = f2bf(t);
}

inline unsigned nblk(int64_t n) { return (unsigned)((n + kT - 1) / kT); }

}  // namespace

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
void layer_shapes(int (*shape)[2]) {
  const int s[kLayers][2] = {{kC0, kCin}, {kC0, kC0}, {kC1, kC0}, {kC1, kC1}, {kC2, kC1},
                             {kC2, kC2}, {kC1, kC2}, {kC1, kC1}, {kC0, kC1}, {1, kC0}};
  for (int l = 0; l < kLayers; ++l) shape[l][0] = s[l][0], shape[l][1] = s[l][1];
}

int64_t num_params() {
  int s[kLayers][2];
  layer_shapes(s);
  int64_t np = 0;
  for (auto &l : s) np += (int64_t)l[0] * l[1] * 27 + l[0];
  return np;
}

Net::~Net() { release(); }

void Net::release() {
  for (void *q : bufs)
    if (q) (void)hipFree(q);
  bufs.clear();
}

int Net::alloc(void **p, size_t bytes) {
  if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) return GDX_ERR_ALLOC;
  bufs.push_back(*p);
  return GDX_OK;
}

int Net::init(int grid, uint64_t seed, hipStream_t s) {
  n = grid;
  st = s;
  np = num_params();
  int sh[kLayers][2];
  layer_shapes(sh);
  int64_t off = 0;
  for (int l = 0; l < kLayers; ++l) {
    woff[l] = off;
    off += (int64_t)sh[l][0] * sh[l][1] * 27;
    boff[l] = off;
    off += sh[l][0];
  }
  const int64_t v0 = (int64_t)n * n * n, v1 = v0 / 8, v2 = v1 / 8;
  int rc = GDX_OK;
  auto F = [&](float **p, int64_t cnt) { if (!rc) rc = alloc(reinterpret_cast<void **>(p), cnt * sizeof(float)); };
  auto B = [&](uint16_t **p, int64_t cnt) { if (!rc) rc = alloc(reinterpret_cast<void **>(p), cnt * sizeof(uint16_t)); };
  F(&theta, np); F(&grad, np); F(&m, np); F(&v, np); B(&thb, np);
  pkoff[0] = 0;
  for (int l = 0; l < kLayers; ++l)
    pkoff[l + 1] = pkoff[l] + (int64_t)(sh[l][0] + 15) / 16 * 16 * ((sh[l][1] * 27 + 31) / 32 * 32);
  B(&wpk, pkoff[kLayers]);
  B(&a0, kCin * v0); B(&x1a, kC0 * v0); B(&x1, kC0 * v0); B(&p1, kC0 * v1); B(&x2a, kC1 * v1); B(&x2, kC1 * v1);
  B(&p2, kC1 * v2); B(&x3a, kC2 * v2); B(&x3, kC2 * v2); B(&u2, kC2 * v1); B(&y2a, kC1 * v1); B(&y2, kC1 * v1);
  B(&u1, kC1 * v0); B(&y1a, kC0 * v0);
  F(&out, v0); F(&inp_f, kCin * v0); F(&dout, v0); F(&dinp, kCin * v0);
  // gradient scratch: the largest tensor of each level, three per level
  F(&gA0, kC1 * v0); F(&gB0, kC1 * v0); F(&gC0, kC1 * v0);
  F(&gA1, kC2 * v1); F(&gB1, kC2 * v1); F(&gC1, kC2 * v1);
  F(&gA2, kC2 * v2); F(&gB2, kC2 * v2);
  int64_t pmax = 0;
  for (int l = 0; l < kLayers; ++l) pmax = std::max<int64_t>(pmax, (int64_t)sh[l][0] * sh[l][1] * 27 + sh[l][0]);
  F(&partial, pmax * kPartialBlocks);
  F(&lossbuf, 2 * kPartialBlocks + 2);
  if (rc) return rc;
  // He-normal initialisation (fan-in 27 cin), last layer scaled by 0.1, zero biases
  std::vector<float> h(np, 0.0f);
  std::mt19937_64 gen(seed);
  for (int l = 0; l < kLayers; ++l) {
    std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (27.0f * sh[l][1])) * (l == kLayers - 1 ? 0.1f : 1.0f));
    for (int64_t i = 0; i < (int64_t)sh[l][0] * sh[l][1] * 27; ++i) h[woff[l] + i] = nd(gen);
  }
  return set_params(h.data());
}

int Net::set_params(const float *h) {
  if (hipMemcpyAsync(theta, h, np * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess) return GDX_ERR_HIP;
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(np)), dim3(kT), 0, st, np, theta, thb);
  pack_weights();
  if (hipMemsetAsync(m, 0, np * sizeof(float), st) != hipSuccess ||
      hipMemsetAsync(v, 0, np * sizeof(float), st) != hipSuccess)
    return GDX_ERR_HIP;
  adam_t = 0;
  return hipStreamSynchronize(st) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int Net::get_params(float *h) {
  if (hipMemcpyAsync(h, theta, np * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess) return GDX_ERR_HIP;
  return hipStreamSynchronize(st) == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// channel groups per tile: double while the grid has fewer than 512 workgroups
// and the output channels split evenly (at most 8 groups)
template <int KCO>
static int channel_groups(int n) {
  int g = 1;
  while (g < 8 && KCO % (2 * g) == 0 && tiles_of(n) * g < 512) g *= 2;
  return g;
}

template <int KCI, int KCO, int MODE, bool RELU, bool SKIP, bool OUTF32, bool ACCUM, int G>
static void conv_launch(const Net &N, int n, const void *in, const uint16_t *w, const uint16_t *b,
                        const uint16_t *skip, void *out) {
  if constexpr (KCO % G == 0)
    hipLaunchKernelGGL((k_conv3<KCI, KCO, MODE, RELU, SKIP, OUTF32, ACCUM, G>), dim3(tiles_of(n), G), dim3(kT), 0,
                       N.st, n, in, w, b, skip, out);
}

template <int KCI, int KCO, int MODE, bool RELU, bool SKIP, bool OUTF32, bool ACCUM>
static void conv_any(const Net &N, int n, const void *in, const uint16_t *w, const uint16_t *b, const uint16_t *skip,
                     void *out) {
  switch (channel_groups<KCO>(n)) {
    case 8: conv_launch<KCI, KCO, MODE, RELU, SKIP, OUTF32, ACCUM, 8>(N, n, in, w, b, skip, out); break;
    case 4: conv_launch<KCI, KCO, MODE, RELU, SKIP, OUTF32, ACCUM, 4>(N, n, in, w, b, skip, out); break;
    case 2: conv_launch<KCI, KCO, MODE, RELU, SKIP, OUTF32, ACCUM, 2>(N, n, in, w, b, skip, out); break;
    default: conv_launch<KCI, KCO, MODE, RELU, SKIP, OUTF32, ACCUM, 1>(N, n, in, w, b, skip, out); break;
  }
}

// forward convolutions on the matrix cores
template <int CI, int CO, bool RELU, bool SKIP, bool OUTF32>
static void conv_fwd(const Net &N, int n, const uint16_t *in, int layer, const uint16_t *skip, void *out) {
  constexpr int NB = (CO + 15) / 16;
  const uint16_t *w = N.wpk + N.pkoff[layer], *b = N.thb + N.boff[layer];
  const unsigned txy = (unsigned)(((n + 7) / 8) * ((n + 7) / 8));
  // 64^3 and up: tiles of 8 x 8 x 4 with every output channel; below, thinner
  // tiles and one workgroup per block of 16 output channels, for more workgroups
  if (n >= 64)
    hipLaunchKernelGGL((k_conv3_mfma<CI, CO, RELU, SKIP, OUTF32, 4, NB>), dim3(txy * ((n + 3) / 4)), dim3(256), 0,
                       N.st, n, in, w, b, skip, out);
  else if (n >= 32)
    hipLaunchKernelGGL((k_conv3_mfma<CI, CO, RELU, SKIP, OUTF32, 2, 1>), dim3(txy * ((n + 1) / 2), NB), dim3(256), 0,
                       N.st, n, in, w, b, skip, out);
  else
    hipLaunchKernelGGL((k_conv3_mfma<CI, CO, RELU, SKIP, OUTF32, 1, 1>), dim3(txy * n, NB), dim3(256), 0, N.st, n, in,
                       w, b, skip, out);
}

// dX of a layer (KCI = layer outputs, KCO = layer inputs)
template <int LCO, int LCI, bool ACCUM>
static void conv_bwd_data(const Net &N, int n, const float *g, int layer, float *dx) {
  conv_any<LCO, LCI, 1, false, false, true, ACCUM>(N, n, g, N.thb + N.woff[layer], nullptr, nullptr, dx);
}

template <int CI, int CO>
static void conv_wgrad(const Net &N, int n, const uint16_t *x, const float *g, int layer) {
  // 8 x 8 x 4 voxel tiles from 32^3 up, 8 x 8 x 1 below (more workgroups on the coarse grids)
  const int txy = ((n + 7) / 8) * ((n + 7) / 8);
  int nb;
  if (n >= 32) {
    nb = std::min(txy * ((n + 3) / 4), kPartialBlocks);
    hipLaunchKernelGGL((k_wgrad_mfma<CI, CO, 4>), dim3(nb), dim3(WgradMfma<CI, CO, 4>::NT), 0, N.st, n, x, g, N.partial);
  } else {
    nb = std::min(txy * n, kPartialBlocks);
    hipLaunchKernelGGL((k_wgrad_mfma<CI, CO, 1>), dim3(nb), dim3(WgradMfma<CI, CO, 1>::NT), 0, N.st, n, x, g, N.partial);
  }
  const int len = CO * CI * 27 + CO;
  hipLaunchKernelGGL(k_reduce_partials, dim3(nblk(len)), dim3(kT), 0, N.st, nb, len, N.partial, N.grad + N.woff[layer]);
}

// forward from the BF16 input already in a0
int Net::forward_dev() {
  const int n0 = n, n1 = n / 2, n2 = n / 4;
  const int64_t v0 = (int64_t)n0 * n0 * n0, v1 = v0 / 8;
  conv_fwd<kCin, kC0, true, false, false>(*this, n0, a0, 0, nullptr, x1a);
  conv_fwd<kC0, kC0, true, false, false>(*this, n0, x1a, 1, nullptr, x1);
  hipLaunchKernelGGL(k_pool, dim3(nblk(kC0 * v1)), dim3(kT), 0, st, kC0, n0, x1, p1);
  conv_fwd<kC0, kC1, true, false, false>(*this, n1, p1, 2, nullptr, x2a);
  conv_fwd<kC1, kC1, true, false, false>(*this, n1, x2a, 3, nullptr, x2);
  hipLaunchKernelGGL(k_pool, dim3(nblk(kC1 * v1 / 8)), dim3(kT), 0, st, kC1, n1, x2, p2);
  conv_fwd<kC1, kC2, true, false, false>(*this, n2, p2, 4, nullptr, x3a);
  conv_fwd<kC2, kC2, true, false, false>(*this, n2, x3a, 5, nullptr, x3);
  hipLaunchKernelGGL(k_up, dim3(nblk(kC2 * v1)), dim3(kT), 0, st, kC2, n1, x3, u2);
  conv_fwd<kC2, kC1, true, true, false>(*this, n1, u2, 6, x2, y2a);
  conv_fwd<kC1, kC1, true, false, false>(*this, n1, y2a, 7, nullptr, y2);
  hipLaunchKernelGGL(k_up, dim3(nblk(kC1 * v0)), dim3(kT), 0, st, kC1, n0, y2, u1);
  conv_fwd<kC1, kC0, true, true, false>(*this, n0, u1, 8, x1, y1a);
  conv_fwd<kC0, 1, false, false, true>(*this, n0, y1a, 9, nullptr, out);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

// backward of sum(dout * out) into grad (and dinp when want_dinp)
int Net::backward_dev(bool want_dinp) {
  const int n0 = n, n1 = n / 2, n2 = n / 4;
  const int64_t v0 = (int64_t)n0 * n0 * n0, v1 = v0 / 8, v2 = v1 / 8;
  // level 0, output layer
  conv_wgrad<kC0, 1>(*this, n0, y1a, dout, 9);
  conv_bwd_data<1, kC0, false>(*this, n0, dout, 9, gA0);  // d y1a
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC0 * v0)), dim3(kT), 0, st, kC0 * v0, gA0, y1a);  // gA0 = skip grad of x1
  conv_wgrad<kC1, kC0>(*this, n0, u1, gA0, 8);
  conv_bwd_data<kC0, kC1, false>(*this, n0, gA0, 8, gB0);  // d u1
  hipLaunchKernelGGL(k_up_bw, dim3(nblk(kC1 * v1)), dim3(kT), 0, st, kC1, n0, gB0, gA1);  // d y2
  // level 1, decoder
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC1 * v1)), dim3(kT), 0, st, kC1 * v1, gA1, y2);
  conv_wgrad<kC1, kC1>(*this, n1, y2a, gA1, 7);
  conv_bwd_data<kC1, kC1, false>(*this, n1, gA1, 7, gB1);  // d y2a
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC1 * v1)), dim3(kT), 0, st, kC1 * v1, gB1, y2a);  // gB1 = skip grad of x2
  conv_wgrad<kC2, kC1>(*this, n1, u2, gB1, 6);
  conv_bwd_data<kC1, kC2, false>(*this, n1, gB1, 6, gC1);  // d u2
  hipLaunchKernelGGL(k_up_bw, dim3(nblk(kC2 * v2)), dim3(kT), 0, st, kC2, n1, gC1, gA2);  // d x3
  // level 2
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC2 * v2)), dim3(kT), 0, st, kC2 * v2, gA2, x3);
  conv_wgrad<kC2, kC2>(*this, n2, x3a, gA2, 5);
  conv_bwd_data<kC2, kC2, false>(*this, n2, gA2, 5, gB2);  // d x3a
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC2 * v2)), dim3(kT), 0, st, kC2 * v2, gB2, x3a);
  conv_wgrad<kC1, kC2>(*this, n2, p2, gB2, 4);
  conv_bwd_data<kC2, kC1, false>(*this, n2, gB2, 4, gA2);  // d p2 (gA2 reused)
  // level 1, encoder
  hipLaunchKernelGGL(k_pool_bw, dim3(nblk(kC1 * v1)), dim3(kT), 0, st, kC1, n1, gA2, gB1, gA1);  // d x2
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC1 * v1)), dim3(kT), 0, st, kC1 * v1, gA1, x2);
  conv_wgrad<kC1, kC1>(*this, n1, x2a, gA1, 3);
  conv_bwd_data<kC1, kC1, false>(*this, n1, gA1, 3, gC1);  // d x2a
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC1 * v1)), dim3(kT), 0, st, kC1 * v1, gC1, x2a);
  conv_wgrad<kC0, kC1>(*this, n1, p1, gC1, 2);
  conv_bwd_data<kC1, kC0, false>(*this, n1, gC1, 2, gB1);  // d p1 (gB1 reused)
  // level 0, encoder
  hipLaunchKernelGGL(k_pool_bw, dim3(nblk(kC0 * v0)), dim3(kT), 0, st, kC0, n0, gB1, gA0, gC0);  // d x1
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC0 * v0)), dim3(kT), 0, st, kC0 * v0, gC0, x1);
  conv_wgrad<kC0, kC0>(*this, n0, x1a, gC0, 1);
  conv_bwd_data<kC0, kC0, false>(*this, n0, gC0, 1, gB0);  // d x1a
  hipLaunchKernelGGL(k_relu_mask, dim3(nblk(kC0 * v0)), dim3(kT), 0, st, kC0 * v0, gB0, x1a);
  conv_wgrad<kCin, kC0>(*this, n0, a0, gB0, 0);
  if (want_dinp) conv_bwd_data<kC0, kCin, false>(*this, n0, gB0, 0, dinp);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int Net::input_from_f32(const float *dev_in) {
  const int64_t tot = (int64_t)kCin * n * n * n;
  hipLaunchKernelGGL(k_to_bf16, dim3(nblk(tot)), dim3(kT), 0, st, tot, dev_in, a0);
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

int Net::adam_step(float lr, float clip) {
  ++adam_t;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)adam_t), c2 = 1.0f - std::pow(b2, (float)adam_t);
  hipLaunchKernelGGL(k_adam, dim3(nblk(np)), dim3(kT), 0, st, np, theta, thb, grad, m, v, lr, b1, b2, 1e-8f, c1, c2,
                     clip);
  pack_weights();
  return hipGetLastError() == hipSuccess ? GDX_OK : GDX_ERR_HIP;
}

void Net::pack_weights() {
  PackTab T;
  int sh[kLayers][2];
  layer_shapes(sh);
  for (int l = 0; l < kLayers; ++l) {
    T.woff[l] = woff[l];
    T.kt[l] = sh[l][1] * 27;
    T.kp[l] = (T.kt[l] + 31) / 32 * 32;
    T.co[l] = sh[l][0];
    T.ci[l] = sh[l][1];
  }
  for (int l = 0; l <= kLayers; ++l) T.pkoff[l] = pkoff[l];
  hipLaunchKernelGGL(k_pack_weights, dim3(nblk(pkoff[kLayers])), dim3(kT), 0, st, T, thb, wpk);
}

}  // namespace nn
}  // namespace gdx

// ---------------------------------------------------------------------------
// C ABI (include/gdx.h)
// ---------------------------------------------------------------------------
namespace {
int nn_fail(int code, const char *msg) {
  gdx::set_error(msg);
  return code;
}
}  // namespace

extern "C" {

int gdx_nn_create(gdx_ctx *ctx, int32_t n, uint64_t seed, gdx_nn **out) {
  if (!ctx || !out) return nn_fail(GDX_ERR_ARG, "NULL argument");
  *out = nullptr;
  if (n < 4 || n % 4 != 0 || n > 512) return nn_fail(GDX_ERR_ARG, "grid extent must be a multiple of 4 in [4, 512]");
  if (hipSetDevice(ctx->device) != hipSuccess) return nn_fail(GDX_ERR_HIP, "hipSetDevice failed");
  gdx_nn *h = new (std::nothrow) gdx_nn;
  if (!h) return nn_fail(GDX_ERR_ALLOC, "host allocation failed");
  h->ctx = ctx;
  int rc = h->net.init(n, seed, ctx->stream);
  if (rc) {
    delete h;
    return nn_fail(rc, "network allocation failed");
  }
  ctx->nns.push_back(h);
  *out = h;
  return GDX_OK;
}

int gdx_nn_num_params(gdx_nn *nn, int64_t *np) {
  if (!nn || !np) return nn_fail(GDX_ERR_ARG, "NULL argument");
  *np = nn->net.np;
  return GDX_OK;
}

int gdx_nn_save(gdx_nn *nn, float *theta) {
  if (!nn || !theta) return nn_fail(GDX_ERR_ARG, "NULL argument");
  if (!nn->ctx) return nn_fail(GDX_ERR_STATE, "the network's context was finalized");
  if (hipSetDevice(nn->ctx->device) != hipSuccess) return nn_fail(GDX_ERR_HIP, "hipSetDevice failed");
  return nn->net.get_params(theta);
}

int gdx_nn_load(gdx_nn *nn, const float *theta) {
  if (!nn || !theta) return nn_fail(GDX_ERR_ARG, "NULL argument");
  if (!nn->ctx) return nn_fail(GDX_ERR_STATE, "the network's context was finalized");
  for (int64_t i = 0; i < nn->net.np; ++i)
    if (!(theta[i] == theta[i]) || theta[i] == INFINITY || theta[i] == -INFINITY)
      return nn_fail(GDX_ERR_ARG, "parameters must be finite");
  if (hipSetDevice(nn->ctx->device) != hipSuccess) return nn_fail(GDX_ERR_HIP, "hipSetDevice failed");
  return nn->net.set_params(theta);
}

int gdx_nn_forward(gdx_nn *nn, const float *inp, float *out) {
  if (!nn || !inp || !out) return nn_fail(GDX_ERR_ARG, "NULL argument");
  if (!nn->ctx) return nn_fail(GDX_ERR_STATE, "the network's context was finalized");
  gdx::nn::Net &N = nn->net;
  if (hipSetDevice(nn->ctx->device) != hipSuccess) return nn_fail(GDX_ERR_HIP, "hipSetDevice failed");
  const int64_t v0 = (int64_t)N.n * N.n * N.n;
  if (hipMemcpyAsync(N.inp_f, inp, gdx::nn::kCin * v0 * sizeof(float), hipMemcpyHostToDevice, N.st) != hipSuccess)
    return nn_fail(GDX_ERR_HIP, "copy failed");
  int rc = N.input_from_f32(N.inp_f);
  if (!rc) rc = N.forward_dev();
  if (rc) return nn_fail(rc, "forward launch failed");
  if (hipMemcpyAsync(out, N.out, v0 * sizeof(float), hipMemcpyDeviceToHost, N.st) != hipSuccess ||
      hipStreamSynchronize(N.st) != hipSuccess)
    return nn_fail(GDX_ERR_HIP, "forward failed");
  return GDX_OK;
}

int gdx_nn_backward(gdx_nn *nn, const float *dout, float *grad, float *dinp) {
  if (!nn || !dout || !grad) return nn_fail(GDX_ERR_ARG, "NULL argument");
  if (!nn->ctx) return nn_fail(GDX_ERR_STATE, "the network's context was finalized");
  gdx::nn::Net &N = nn->net;
  if (hipSetDevice(nn->ctx->device) != hipSuccess) return nn_fail(GDX_ERR_HIP, "hipSetDevice failed");
  const int64_t v0 = (int64_t)N.n * N.n * N.n;
  if (hipMemcpyAsync(N.dout, dout, v0 * sizeof(float), hipMemcpyHostToDevice, N.st) != hipSuccess)
    return nn_fail(GDX_ERR_HIP, "copy failed");
  int rc = N.backward_dev(dinp != nullptr);
  if (rc) return nn_fail(rc, "backward launch failed");
  if (hipMemcpyAsync(grad, N.grad, N.np * sizeof(float), hipMemcpyDeviceToHost, N.st) != hipSuccess)
    return nn_fail(GDX_ERR_HIP, "copy failed");
  if (dinp && hipMemcpyAsync(dinp, N.dinp, gdx::nn::kCin * v0 * sizeof(float), hipMemcpyDeviceToHost, N.st) !=
                  hipSuccess)
    return nn_fail(GDX_ERR_HIP, "copy failed");
  if (hipStreamSynchronize(N.st) != hipSuccess) return nn_fail(GDX_ERR_HIP, "backward failed");
  return GDX_OK;
}

int gdx_nn_destroy(gdx_nn *nn) {
  if (!nn) return nn_fail(GDX_ERR_ARG, "NULL argument");
  if (gdx_ctx *ctx = nn->ctx) {
    gdx::release_nn_device(nn);
    gdx::unregister(ctx->nns, nn);
  }
  delete nn;
  return GDX_OK;
}

}  // extern "C"

namespace gdx {
void release_nn_device(gdx_nn *nn) {
  if (!nn->ctx) return;
  (void)hipSetDevice(nn->ctx->device);
  (void)hipStreamSynchronize(nn->ctx->stream);
  twogrid_free(nn);
  nn->net.release();
  nn->ctx = nullptr;
}
}  // namespace gdx


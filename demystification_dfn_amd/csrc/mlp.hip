// Step-size proposal network: 8 -> 64 -> 64 -> 1, GELU hidden layers, BF16
// weights with FP32 accumulation, trained with Adam on (features, accepted step)
// pairs.  No counterpart in the reference (its step size comes from rkqs alone,
// src/ode_solver/ode_bs.f90:125-174); the Cash-Karp error test still accepts or
// rejects every step the network proposes.
//
// Every matrix product runs on the matrix cores (v_mfma_f32_16x16x4f32: the BF16
// weights are exact in FP32, activations and deltas stay FP32).  One 64-lane
// workgroup takes tiles of 16 samples: the forward pass, the deltas and the
// tile's weight gradients in one kernel, the gradients of each workgroup summed
// in registers over its tiles and written as one partial; a second kernel sums
// the partials of a batch in workgroup order (the same gradient on every run) and
// applies Adam to the FP32 master weights, re-rounding the BF16 copy.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <new>
#include <numeric>
#include <random>
#include <vector>

#include "../../include/dfn.h"
#include "engine_common.hpp"

namespace dfn {

namespace {

constexpr int kIn = DFN_MLP_INPUTS, kHid = 64;
constexpr int oW1 = 0, oB1 = oW1 + kHid * kIn, oW2 = oB1 + kHid, oB2 = oW2 + kHid * kHid, oW3 = oB2 + kHid,
              oB3 = oW3 + kHid, kNp = oB3 + 1;
static_assert(kNp == DFN_MLP_PARAMS, "parameter count");

struct Norm {
  float mx[kIn], sx[kIn], my, sy;
};

// FP32 -> BF16, round to nearest even (finite inputs)
__host__ __device__ inline uint16_t to_bf16(float f) {
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

__host__ __device__ inline float from_bf16(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }

__device__ __forceinline__ float gelu_d(float z) {  // Phi(z) + z phi(z)
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

// weights (BF16 values as FP32) into LDS
__device__ __forceinline__ void load_weights(const uint16_t *__restrict__ w, float *sw) {
  for (int i = threadIdx.x; i < kNp; i += blockDim.x) sw[i] = from_bf16(w[i]);
  __syncthreads();
}

typedef float f4v __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v mfma4(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Matrix-core layout (FP32 16x16x4): lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; accumulator element e of lane l is C[4 (l / 16) + e][l % 16].
// Layers are computed transposed (weights as A, activations as B, samples along
// the 16 columns), so lane l = 16 q + s works on sample s of the tile: it holds
// the hidden units 16 t + 4 q + e (tile t, element e) of its sample, which are
// exactly the units it feeds to the next product when that product's k-steps run
// over unit 16 t + 4 q + e at step 4 t + e.
constexpr int kTile = 16, kPartStride = kNp + 3;  // partial: the parameters, then the loss (double)

// Forward pass of the tiles of one workgroup; with train != 0 also the deltas of
// the loss mean((y - t)^2) over the n samples of the batch, and the workgroup's
// gradient summed over its tiles into part[blockIdx.x * kPartStride ..] (the loss
// sum in double at kNp + 1).  Y: normalised outputs.
__global__ void __launch_bounds__(64) mlp_tile_kernel(const uint16_t *__restrict__ w, Norm nm,
                                                      const float *__restrict__ X, const float *__restrict__ T,
                                                      const int *__restrict__ idx, int n, int tiles_per_block,
                                                      int train, float *__restrict__ Y, float *__restrict__ part) {
  __shared__ float sw[kNp];
  __shared__ float sD2[kTile][65], sH1[kTile][65], sD1[kTile][65], sH2g[kTile][65], sX[kTile][9], sG[kTile];
  load_weights(w, sw);
  const int l = threadIdx.x, s = l & 15, q = l >> 4;
  f4v gW2[4][4], gW1[4];
  float gb1 = 0.0f, gb2 = 0.0f, gw3 = 0.0f, gb3 = 0.0f;
  double loss = 0.0;
  if (train) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      gW1[a] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int b = 0; b < 4; b++) gW2[a][b] = f4v{0.0f, 0.0f, 0.0f, 0.0f};
    }
  }
  for (int tt = 0; tt < tiles_per_block; tt++) {
    const int sample = (blockIdx.x * tiles_per_block + tt) * kTile + s;
    const bool valid = sample < n;
    const int src = valid ? (idx ? idx[sample] : sample) : 0;
    // normalised inputs k = q and q + 4 of this lane's sample
    float xk[2];
#pragma unroll
    for (int kk = 0; kk < 2; kk++) {
      const int k = q + 4 * kk;
      xk[kk] = valid ? (X[(int64_t)src * kIn + k] - nm.mx[k]) / nm.sx[k] : 0.0f;
    }
    // layer 1: z1 = W1 x + b1
    float z1[4][4], h1[4][4];
#pragma unroll
    for (int tm = 0; tm < 4; tm++) {
      f4v c = f4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kk = 0; kk < 2; kk++) c = mfma4(sw[oW1 + (16 * tm + s) * kIn + q + 4 * kk], xk[kk], c);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        z1[tm][e] = c[e] + sw[oB1 + 16 * tm + 4 * q + e];
        h1[tm][e] = gelu(z1[tm][e]);
      }
    }
    // layer 2: z2 = W2 h1 + b2
    float z2[4][4], h2[4][4];
#pragma unroll
    for (int tm = 0; tm < 4; tm++) {
      f4v c = f4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kk = 0; kk < 16; kk++)
        c = mfma4(sw[oW2 + (16 * tm + s) * kHid + 16 * (kk >> 2) + 4 * q + (kk & 3)], h1[kk >> 2][kk & 3], c);
#pragma unroll
      for (int e = 0; e < 4; e++) {
        z2[tm][e] = c[e] + sw[oB2 + 16 * tm + 4 * q + e];
        h2[tm][e] = gelu(z2[tm][e]);
      }
    }
    // output: y = w3 . h2 + b3, the four lanes of a sample summed
    float yp = 0.0f;
#pragma unroll
    for (int tm = 0; tm < 4; tm++)
#pragma unroll
      for (int e = 0; e < 4; e++) yp += sw[oW3 + 16 * tm + 4 * q + e] * h2[tm][e];
    yp += __shfl_xor(yp, 16);
    yp += __shfl_xor(yp, 32);
    const float y = yp + sw[oB3];
    if (!train) {
      if (valid && q == 0) Y[sample] = y;
      continue;
    }
    const float t = valid ? (T[src] - nm.my) / nm.sy : 0.0f;
    const float g = valid ? 2.0f * (y - t) / (float)n : 0.0f;  // d loss / d y
    if (valid && q == 0) {
      const double e = (double)y - (double)t;
      loss += e * e;
    }
    // output-layer and second-layer deltas: d2 = g w3 gelu'(z2)
    float d2[4][4];
#pragma unroll
    for (int tm = 0; tm < 4; tm++)
#pragma unroll
      for (int e = 0; e < 4; e++) d2[tm][e] = g * sw[oW3 + 16 * tm + 4 * q + e] * gelu_d(z2[tm][e]);
    // first-layer deltas: d1_j = gelu'(z1_j) sum_i W2[i][j] d2_i
#pragma unroll
    for (int tj = 0; tj < 4; tj++) {
      f4v c = f4v{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int kk = 0; kk < 16; kk++)
        c = mfma4(sw[oW2 + (16 * (kk >> 2) + 4 * q + (kk & 3)) * kHid + 16 * tj + s], d2[kk >> 2][kk & 3], c);
#pragma unroll
      for (int e = 0; e < 4; e++) sD1[s][16 * tj + 4 * q + e] = c[e] * gelu_d(z1[tj][e]);
    }
    // the tile's activations and deltas, sample-major, for the weight gradients
#pragma unroll
    for (int tm = 0; tm < 4; tm++)
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int u = 16 * tm + 4 * q + e;
        sD2[s][u] = d2[tm][e];
        sH1[s][u] = valid ? h1[tm][e] : 0.0f;
        sH2g[s][u] = g * h2[tm][e];
      }
#pragma unroll
    for (int kk = 0; kk < 2; kk++) sX[s][q + 4 * kk] = xk[kk];
    if (q == 0) sG[s] = g;
    __syncthreads();
    // dW2[i][j] += sum_s d2[s][i] h1[s][j]; dW1[j][k] += sum_s d1[s][j] x[s][k]
#pragma unroll
    for (int kk = 0; kk < 4; kk++) {
      const int sr = 4 * kk + q;  // the sample this lane supplies at step kk
      float hb[4];
#pragma unroll
      for (int tj = 0; tj < 4; tj++) hb[tj] = sH1[sr][16 * tj + s];
#pragma unroll
      for (int ti = 0; ti < 4; ti++) {
        const float da = sD2[sr][16 * ti + s];
#pragma unroll
        for (int tj = 0; tj < 4; tj++) gW2[ti][tj] = mfma4(da, hb[tj], gW2[ti][tj]);
      }
      const float xb = s < kIn ? sX[sr][s] : 0.0f;
#pragma unroll
      for (int tj = 0; tj < 4; tj++) gW1[tj] = mfma4(sD1[sr][16 * tj + s], xb, gW1[tj]);
    }
    // bias and output-layer gradients: lane u sums unit u over the tile's samples
#pragma unroll
    for (int sr = 0; sr < kTile; sr++) {
      gb1 += sD1[sr][l];
      gb2 += sD2[sr][l];
      gw3 += sH2g[sr][l];
    }
    if (l == 0)
#pragma unroll
      for (int sr = 0; sr < kTile; sr++) gb3 += sG[sr];
    __syncthreads();  // the LDS tiles are rewritten by the next tile
  }
  if (!train) return;
  float *P = part + (int64_t)blockIdx.x * kPartStride;
#pragma unroll
  for (int ti = 0; ti < 4; ti++)
#pragma unroll
    for (int tj = 0; tj < 4; tj++)
#pragma unroll
      for (int e = 0; e < 4; e++) P[oW2 + (16 * ti + 4 * q + e) * kHid + 16 * tj + s] = gW2[ti][tj][e];
#pragma unroll
  for (int tj = 0; tj < 4; tj++)
#pragma unroll
    for (int e = 0; e < 4; e++)
      if (s < kIn) P[oW1 + (16 * tj + 4 * q + e) * kIn + s] = gW1[tj][e];
  P[oB1 + l] = gb1;
  P[oB2 + l] = gb2;
  P[oW3 + l] = gw3;
  if (l == 0) P[oB3] = gb3;
  // the loss sum of the workgroup: the 16 lanes with q == 0 hold one sample each
  loss += __shfl_xor(loss, 1);
  loss += __shfl_xor(loss, 2);
  loss += __shfl_xor(loss, 4);
  loss += __shfl_xor(loss, 8);
  if (l == 0) __builtin_memcpy(P + kNp + 1, &loss, sizeof(double));
}

// grad[q] = sum of the workgroups' partials in workgroup order (thread kNp: the
// loss, divided by n); with lr > 0 also Adam on the FP32 master weights, which
// refreshes the BF16 copy
__global__ void __launch_bounds__(256) mlp_reduce_kernel(const float *__restrict__ part, int nparts, int n,
                                                         float *__restrict__ grad, double *__restrict__ loss,
                                                         float *__restrict__ wm, float *__restrict__ m,
                                                         float *__restrict__ v, uint16_t *__restrict__ wb, float lr,
                                                         float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > kNp) return;
  if (q == kNp) {
    double l = 0.0;
    for (int b = 0; b < nparts; b++) {
      double x;
      __builtin_memcpy(&x, part + (int64_t)b * kPartStride + kNp + 1, sizeof(double));
      l += x;
    }
    *loss = l / n;
    return;
  }
  float g = 0.0f;
  for (int b = 0; b < nparts; b++) g += part[(int64_t)b * kPartStride + q];
  grad[q] = g;
  if (lr > 0.0f) {
    const float mq = 0.9f * m[q] + 0.1f * g, vq = 0.999f * v[q] + 0.001f * g * g;
    m[q] = mq;
    v[q] = vq;
    const float wq = wm[q] - lr * (mq / c1) / (sqrtf(vq / c2) + 1e-8f);
    wm[q] = wq;
    wb[q] = to_bf16(wq);
  }
}

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

struct Mlp {
  hipStream_t stream = nullptr;
  Norm nm{};
  DMem wb, wm, m, v, grad, loss;
  // per-sample work buffers (sized for the largest call seen) and the gradient
  // partials of the workgroups of one launch
  DMem X, T, idx, Y, part;
  int64_t cap = 0, xcap = 0, pcap = 0;
  int64_t adam_t = 0;

  ~Mlp() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
  }

  // launch shape over n samples: tiles of 16, at most kMaxBlocks workgroups, each
  // taking the same number of consecutive tiles (fixed for a given n)
  static constexpr int kMaxBlocks = 1024;
  static void shape(int n, int &blocks, int &tiles_per_block) {
    const int tiles = (n + kTile - 1) / kTile;
    blocks = std::min(tiles, kMaxBlocks);
    tiles_per_block = (tiles + blocks - 1) / blocks;
    blocks = (tiles + tiles_per_block - 1) / tiles_per_block;
  }

  int init() {
    DFN_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    DFN_TRY(wb.alloc(sizeof(uint16_t) * kNp));
    DFN_TRY(wm.alloc(sizeof(float) * kNp));
    DFN_TRY(m.alloc(sizeof(float) * kNp));
    DFN_TRY(v.alloc(sizeof(float) * kNp));
    DFN_TRY(grad.alloc(sizeof(float) * kNp));
    DFN_TRY(loss.alloc(sizeof(double)));
    for (int j = 0; j < kIn; j++) { nm.mx[j] = 0.0f; nm.sx[j] = 1.0f; }
    nm.my = 0.0f;
    nm.sy = 1.0f;
    // He-style initialisation, fixed seed
    std::mt19937_64 rng(12345);
    std::normal_distribution<float> nd(0.0f, 1.0f);
    std::vector<float> p(kNp, 0.0f);
    for (int i = 0; i < kHid * kIn; i++) p[oW1 + i] = nd(rng) * std::sqrt(2.0f / kIn);
    for (int i = 0; i < kHid * kHid; i++) p[oW2 + i] = nd(rng) * std::sqrt(2.0f / kHid);
    for (int i = 0; i < kHid; i++) p[oW3 + i] = nd(rng) * std::sqrt(1.0f / kHid);
    return set_params(p.data(), nullptr);
  }

  int set_params(const float *p, const float *norm) {
    std::vector<uint16_t> b(kNp);
    for (int i = 0; i < kNp; i++) b[i] = to_bf16(p[i]);
    DFN_HIP(hipMemcpy(wm.p, p, sizeof(float) * kNp, hipMemcpyHostToDevice));
    DFN_HIP(hipMemcpy(wb.p, b.data(), sizeof(uint16_t) * kNp, hipMemcpyHostToDevice));
    DFN_HIP(hipMemset(m.p, 0, sizeof(float) * kNp));
    DFN_HIP(hipMemset(v.p, 0, sizeof(float) * kNp));
    adam_t = 0;
    if (norm) {
      for (int j = 0; j < kIn; j++) { nm.mx[j] = norm[j]; nm.sx[j] = norm[kIn + j]; }
      nm.my = norm[2 * kIn];
      nm.sy = norm[2 * kIn + 1];
    }
    return DFN_OK;
  }

  int ensure(int64_t n, int64_t nx) {
    if (n > cap) {
      DFN_TRY(Y.alloc(sizeof(float) * n));
      DFN_TRY(idx.alloc(sizeof(int) * n));
      cap = n;
    }
    if (nx > xcap) {
      DFN_TRY(X.alloc(sizeof(float) * nx * kIn));
      DFN_TRY(T.alloc(sizeof(float) * nx));
      xcap = nx;
    }
    if (pcap == 0) {
      DFN_TRY(part.alloc(sizeof(float) * (size_t)kMaxBlocks * kPartStride));
      pcap = kMaxBlocks;
    }
    return DFN_OK;
  }

  void forward(const float *dX, int n, float *dY, hipStream_t s) const {
    int blocks, tpb;
    shape(n, blocks, tpb);
    hipLaunchKernelGGL(mlp_tile_kernel, dim3(blocks), dim3(64), 0, s, wb.as<uint16_t>(), nm, dX, nullptr, nullptr, n,
                       tpb, 0, dY, nullptr);
  }

  // loss and gradient of the batch idx[0..n) (or the first n samples) of X/T on the
  // device; with lr > 0 followed by one Adam step
  void loss_grad(const int *didx, int n, float lr = 0.0f) {
    int blocks, tpb;
    shape(n, blocks, tpb);
    hipLaunchKernelGGL(mlp_tile_kernel, dim3(blocks), dim3(64), 0, stream, wb.as<uint16_t>(), nm, X.as<float>(),
                       T.as<float>(), didx, n, tpb, 1, nullptr, part.as<float>());
    float c1 = 1.0f, c2 = 1.0f;
    if (lr > 0.0f) {
      adam_t++;
      c1 = 1.0f - std::pow(0.9f, (float)adam_t);
      c2 = 1.0f - std::pow(0.999f, (float)adam_t);
    }
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3(nblk(kNp + 1, 256)), dim3(256), 0, stream, part.as<float>(), blocks, n,
                       grad.as<float>(), loss.as<double>(), wm.as<float>(), m.as<float>(), v.as<float>(),
                       wb.as<uint16_t>(), lr, c1, c2);
  }
};

int mlp_forward_device(Mlp *m, const float *dX, int n, float *dY, hipStream_t s) {
  // the ensemble's stream waits for nothing of the network's: the call is ordered on s
  m->forward(dX, n, dY, s);
  return hipGetLastError() == hipSuccess ? DFN_OK : DFN_ERR_HIP;
}

float mlp_output_scale(const Mlp *m, float y) { return m->nm.my + m->nm.sy * y; }

}  // namespace dfn

using dfn::Mlp;

struct dfn_mlp_s {
  Mlp m;
};

dfn::Mlp *dfn::mlp_of(dfn_mlp h) { return h ? &h->m : nullptr; }

#define API extern "C" __attribute__((visibility("default")))
#include "api_guard.hpp"

API int dfn_mlp_create(dfn_mlp *out) {
  DFN_API_BEGIN
  if (!out) return DFN_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return DFN_ERR_HIP;
  dfn_mlp h = new (std::nothrow) dfn_mlp_s;
  if (!h) return DFN_ERR_NOMEM;
  const int st = h->m.init();
  if (st != DFN_OK) { delete h; return st; }
  *out = h;
  return DFN_OK;
  DFN_API_END
}

API int dfn_mlp_destroy(dfn_mlp h) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  delete h;
  return DFN_OK;
  DFN_API_END
}

API int dfn_mlp_set_params(dfn_mlp h, const float *params, const float *norm) {
  DFN_API_BEGIN
  if (!h || !params) return DFN_ERR_ARG;
  if (norm) {  // the feature scales and the output scale must be positive
    for (int j = 0; j < DFN_MLP_INPUTS; j++)
      if (!(norm[DFN_MLP_INPUTS + j] > 0.0f)) return DFN_ERR_ARG;
    if (!(norm[2 * DFN_MLP_INPUTS + 1] > 0.0f)) return DFN_ERR_ARG;
  }
  return h->m.set_params(params, norm);
  DFN_API_END
}

API int dfn_mlp_get_params(dfn_mlp h, float *params, float *norm) {
  DFN_API_BEGIN
  if (!h) return DFN_ERR_ARG;
  DFN_HIP(hipStreamSynchronize(h->m.stream));
  if (params) DFN_HIP(hipMemcpy(params, h->m.wm.p, sizeof(float) * DFN_MLP_PARAMS, hipMemcpyDeviceToHost));
  if (norm) {
    for (int j = 0; j < DFN_MLP_INPUTS; j++) { norm[j] = h->m.nm.mx[j]; norm[DFN_MLP_INPUTS + j] = h->m.nm.sx[j]; }
    norm[2 * DFN_MLP_INPUTS] = h->m.nm.my;
    norm[2 * DFN_MLP_INPUTS + 1] = h->m.nm.sy;
  }
  return DFN_OK;
  DFN_API_END
}

API int dfn_mlp_forward(dfn_mlp h, const float *X, int64_t n, float *Y) {
  DFN_API_BEGIN
  if (!h || !X || !Y || n < 1 || n > INT32_MAX) return DFN_ERR_ARG;
  Mlp &m = h->m;
  DFN_TRY(m.ensure(n, n));
  DFN_HIP(hipMemcpyAsync(m.X.p, X, sizeof(float) * n * DFN_MLP_INPUTS, hipMemcpyHostToDevice, m.stream));
  m.forward(m.X.as<float>(), (int)n, m.Y.as<float>(), m.stream);
  DFN_HIP(hipGetLastError());
  std::vector<float> y(n);
  DFN_HIP(hipMemcpyAsync(y.data(), m.Y.p, sizeof(float) * n, hipMemcpyDeviceToHost, m.stream));
  DFN_HIP(hipStreamSynchronize(m.stream));
  for (int64_t s = 0; s < n; s++) Y[s] = dfn::mlp_output_scale(&m, y[s]);
  return DFN_OK;
  DFN_API_END
}

API int dfn_mlp_loss_grad(dfn_mlp h, const float *X, const float *T, int64_t n, double *loss, float *grad) {
  DFN_API_BEGIN
  if (!h || !X || !T || n < 1 || n > INT32_MAX) return DFN_ERR_ARG;
  Mlp &m = h->m;
  DFN_TRY(m.ensure(n, n));
  DFN_HIP(hipMemcpyAsync(m.X.p, X, sizeof(float) * n * DFN_MLP_INPUTS, hipMemcpyHostToDevice, m.stream));
  DFN_HIP(hipMemcpyAsync(m.T.p, T, sizeof(float) * n, hipMemcpyHostToDevice, m.stream));
  m.loss_grad(nullptr, (int)n);
  DFN_HIP(hipGetLastError());
  if (loss) DFN_HIP(hipMemcpyAsync(loss, m.loss.p, sizeof(double), hipMemcpyDeviceToHost, m.stream));
  if (grad) DFN_HIP(hipMemcpyAsync(grad, m.grad.p, sizeof(float) * DFN_MLP_PARAMS, hipMemcpyDeviceToHost, m.stream));
  DFN_HIP(hipStreamSynchronize(m.stream));
  return DFN_OK;
  DFN_API_END
}

API int dfn_mlp_train(dfn_mlp h, const float *X, const float *T, int64_t n, int32_t epochs, int32_t batch, float lr,
                      uint64_t seed, double *final_loss) {
  DFN_API_BEGIN
  if (!h || !X || !T || n < 1 || n > INT32_MAX || epochs < 0 || batch < 1 || !(lr > 0.0f)) return DFN_ERR_ARG;
  Mlp &m = h->m;
  const int b = (int)std::min<int64_t>(batch, n);
  DFN_TRY(m.ensure(b, n));
  DFN_HIP(hipMemcpyAsync(m.X.p, X, sizeof(float) * n * DFN_MLP_INPUTS, hipMemcpyHostToDevice, m.stream));
  DFN_HIP(hipMemcpyAsync(m.T.p, T, sizeof(float) * n, hipMemcpyHostToDevice, m.stream));
  std::vector<int> order(n);
  std::iota(order.begin(), order.end(), 0);
  std::mt19937_64 rng(seed);
  dfn::DMem didx;
  DFN_TRY(didx.alloc(sizeof(int) * n));
  for (int e = 0; e < epochs; e++) {
    std::shuffle(order.begin(), order.end(), rng);
    DFN_HIP(hipMemcpyAsync(didx.p, order.data(), sizeof(int) * n, hipMemcpyHostToDevice, m.stream));
    for (int64_t s0 = 0; s0 + b <= n; s0 += b) {  // whole batches
      m.loss_grad(didx.as<int>() + s0, b, lr);
    }
    DFN_HIP(hipGetLastError());
    DFN_HIP(hipStreamSynchronize(m.stream));  // the order buffer is reused next epoch
  }
  if (final_loss) {  // loss over the whole set (in batches), after training
    double tot = 0.0;
    int64_t cnt = 0;
    std::iota(order.begin(), order.end(), 0);
    DFN_HIP(hipMemcpyAsync(didx.p, order.data(), sizeof(int) * n, hipMemcpyHostToDevice, m.stream));
    for (int64_t s0 = 0; s0 < n; s0 += b) {
      const int nb = (int)std::min<int64_t>(b, n - s0);
      m.loss_grad(didx.as<int>() + s0, nb);
      double l = 0.0;
      DFN_HIP(hipMemcpyAsync(&l, m.loss.p, sizeof(double), hipMemcpyDeviceToHost, m.stream));
      DFN_HIP(hipStreamSynchronize(m.stream));
      tot += l * nb;
      cnt += nb;
    }
    *final_loss = tot / (double)cnt;
  }
  return DFN_OK;
  DFN_API_END
}

// Step-size proposal network: 8 -> 64 -> 64 -> 1, GELU hidden layers, BF16
// weights with FP32 accumulation, trained with Adam on (features, accepted step)
// pairs.  No counterpart in the reference (its step size comes from rkqs alone,
// src/ode_solver/ode_bs.f90:125-174); the Cash-Karp error test still accepts or
// rejects every step the network proposes.
//
// Kernels: the forward pass (one thread per sample, weights in LDS as FP32
// copies of the BF16 values), the backward pass in two sample-parallel kernels
// (output-layer and second-layer deltas, then first-layer deltas), the weight
// gradients as one thread per parameter summing its samples in order (the same
// gradient on every run), and Adam on FP32 master weights that re-rounds the
// BF16 copy the forward pass reads.
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

// Forward pass.  Raw features in (normalised here); normalised output y.  With
// train != 0 it also stores z1, h1 and the output-layer and second-layer deltas
// of the loss mean((y - t)^2) over the batch (t normalised here too).
__global__ void __launch_bounds__(256) mlp_forward_kernel(const uint16_t *__restrict__ w, Norm nm,
                                                          const float *__restrict__ X, const float *__restrict__ T,
                                                          const int *__restrict__ idx, int n, int train,
                                                          float *__restrict__ Y, float *__restrict__ Z1,
                                                          float *__restrict__ H1, float *__restrict__ H2,
                                                          float *__restrict__ D2, float *__restrict__ G) {
  __shared__ float sw[kNp];
  load_weights(w, sw);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int src = idx ? idx[s] : s;
  float x[kIn];
#pragma unroll
  for (int j = 0; j < kIn; j++) x[j] = (X[(int64_t)src * kIn + j] - nm.mx[j]) / nm.sx[j];
  float h1[kHid];
#pragma unroll
  for (int i = 0; i < kHid; i++) {
    float z = sw[oB1 + i];
#pragma unroll
    for (int j = 0; j < kIn; j++) z += sw[oW1 + i * kIn + j] * x[j];
    h1[i] = gelu(z);
    if (train) { Z1[(int64_t)s * kHid + i] = z; H1[(int64_t)s * kHid + i] = h1[i]; }
  }
  float y = sw[oB3];
  float z2s[kHid];
#pragma unroll 4
  for (int i = 0; i < kHid; i++) {
    float z = sw[oB2 + i];
#pragma unroll
    for (int j = 0; j < kHid; j++) z += sw[oW2 + i * kHid + j] * h1[j];
    z2s[i] = z;
    const float h2 = gelu(z);
    if (train) H2[(int64_t)s * kHid + i] = h2;
    y += sw[oW3 + i] * h2;
  }
  Y[s] = y;
  if (!train) return;
  const float t = (T[src] - nm.my) / nm.sy;
  const float g = 2.0f * (y - t) / (float)n;  // d loss / d y
  G[s] = g;
#pragma unroll 4
  for (int i = 0; i < kHid; i++) D2[(int64_t)s * kHid + i] = g * sw[oW3 + i] * gelu_d(z2s[i]);
}

// first-layer deltas: d1_j = gelu'(z1_j) * sum_i W2[i][j] d2_i
__global__ void __launch_bounds__(256) mlp_backward_kernel(const uint16_t *__restrict__ w, int n,
                                                           const float *__restrict__ Z1, const float *__restrict__ D2,
                                                           float *__restrict__ D1) {
  __shared__ float sw[kNp];
  load_weights(w, sw);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float d2[kHid];
#pragma unroll
  for (int i = 0; i < kHid; i++) d2[i] = D2[(int64_t)s * kHid + i];
#pragma unroll 4
  for (int j = 0; j < kHid; j++) {
    float a = 0.0f;
#pragma unroll
    for (int i = 0; i < kHid; i++) a += sw[oW2 + i * kHid + j] * d2[i];
    D1[(int64_t)s * kHid + j] = a * gelu_d(Z1[(int64_t)s * kHid + j]);
  }
}

// one thread per parameter: its gradient summed over the batch in sample order;
// thread kNp also sums the loss
__global__ void __launch_bounds__(256) mlp_grad_kernel(Norm nm, const float *__restrict__ X,
                                                       const float *__restrict__ T, const int *__restrict__ idx, int n,
                                                       const float *__restrict__ Y, const float *__restrict__ H1,
                                                       const float *__restrict__ H2, const float *__restrict__ D1,
                                                       const float *__restrict__ D2, const float *__restrict__ G,
                                                       float *__restrict__ grad, double *__restrict__ loss) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > kNp) return;
  float acc = 0.0f;
  if (q < oB1) {  // W1[i][j]: d1_i * x_j
    const int i = q / kIn, j = q % kIn;
    for (int s = 0; s < n; s++) {
      const int src = idx ? idx[s] : s;
      acc += D1[(int64_t)s * kHid + i] * ((X[(int64_t)src * kIn + j] - nm.mx[j]) / nm.sx[j]);
    }
  } else if (q < oW2) {  // b1
    for (int s = 0; s < n; s++) acc += D1[(int64_t)s * kHid + (q - oB1)];
  } else if (q < oB2) {  // W2[i][j]: d2_i * h1_j
    const int i = (q - oW2) / kHid, j = (q - oW2) % kHid;
    for (int s = 0; s < n; s++) acc += D2[(int64_t)s * kHid + i] * H1[(int64_t)s * kHid + j];
  } else if (q < oW3) {  // b2
    for (int s = 0; s < n; s++) acc += D2[(int64_t)s * kHid + (q - oB2)];
  } else if (q < oB3) {  // w3: g * h2_j
    for (int s = 0; s < n; s++) acc += G[s] * H2[(int64_t)s * kHid + (q - oW3)];
  } else if (q == oB3) {
    for (int s = 0; s < n; s++) acc += G[s];
  } else {  // the loss, in double
    double l = 0.0;
    for (int s = 0; s < n; s++) {
      const int src = idx ? idx[s] : s;
      const double e = (double)Y[s] - (double)((T[src] - nm.my) / nm.sy);
      l += e * e;
    }
    *loss = l / n;
    return;
  }
  grad[q] = acc;
}

// Adam on FP32 master weights; refreshes the BF16 copy
__global__ void __launch_bounds__(256) mlp_adam_kernel(float *__restrict__ wm, float *__restrict__ m,
                                                       float *__restrict__ v, const float *__restrict__ grad,
                                                       uint16_t *__restrict__ wb, float lr, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= kNp) return;
  const float g = grad[q];
  const float mq = 0.9f * m[q] + 0.1f * g, vq = 0.999f * v[q] + 0.001f * g * g;
  m[q] = mq;
  v[q] = vq;
  const float w = wm[q] - lr * (mq / c1) / (sqrtf(vq / c2) + 1e-8f);
  wm[q] = w;
  wb[q] = to_bf16(w);
}

inline int nblk(int64_t n, int b) { return (int)((n + b - 1) / b); }

}  // namespace

struct Mlp {
  hipStream_t stream = nullptr;
  Norm nm{};
  DMem wb, wm, m, v, grad, loss;
  // per-sample work buffers, sized for the largest batch seen
  DMem X, T, idx, Y, Z1, H1, H2, D1, D2, G;
  int64_t cap = 0, xcap = 0;
  int64_t adam_t = 0;

  ~Mlp() {
    if (stream) {
      (void)hipStreamSynchronize(stream);
      (void)hipStreamDestroy(stream);
    }
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
      const size_t h = sizeof(float) * (size_t)n * kHid;
      DFN_TRY(Y.alloc(sizeof(float) * n)); DFN_TRY(G.alloc(sizeof(float) * n)); DFN_TRY(idx.alloc(sizeof(int) * n));
      DFN_TRY(Z1.alloc(h)); DFN_TRY(H1.alloc(h)); DFN_TRY(H2.alloc(h)); DFN_TRY(D1.alloc(h)); DFN_TRY(D2.alloc(h));
      cap = n;
    }
    if (nx > xcap) {
      DFN_TRY(X.alloc(sizeof(float) * nx * kIn));
      DFN_TRY(T.alloc(sizeof(float) * nx));
      xcap = nx;
    }
    return DFN_OK;
  }

  void forward(const float *dX, int n, float *dY) {
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(nblk(n, 256)), dim3(256), 0, stream, wb.as<uint16_t>(), nm, dX,
                       nullptr, nullptr, n, 0, dY, nullptr, nullptr, nullptr, nullptr, nullptr);
  }

  // loss and gradient of the batch idx[0..n) (or the first n samples) of X/T on the device
  void loss_grad(const int *didx, int n) {
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(nblk(n, 256)), dim3(256), 0, stream, wb.as<uint16_t>(), nm,
                       X.as<float>(), T.as<float>(), didx, n, 1, Y.as<float>(), Z1.as<float>(), H1.as<float>(),
                       H2.as<float>(), D2.as<float>(), G.as<float>());
    hipLaunchKernelGGL(mlp_backward_kernel, dim3(nblk(n, 256)), dim3(256), 0, stream, wb.as<uint16_t>(), n,
                       Z1.as<float>(), D2.as<float>(), D1.as<float>());
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(nblk(kNp + 1, 256)), dim3(256), 0, stream, nm, X.as<float>(),
                       T.as<float>(), didx, n, Y.as<float>(), H1.as<float>(), H2.as<float>(), D1.as<float>(),
                       D2.as<float>(), G.as<float>(), grad.as<float>(), loss.as<double>());
  }

  void adam(float lr) {
    adam_t++;
    const float c1 = 1.0f - std::pow(0.9f, (float)adam_t), c2 = 1.0f - std::pow(0.999f, (float)adam_t);
    hipLaunchKernelGGL(mlp_adam_kernel, dim3(nblk(kNp, 256)), dim3(256), 0, stream, wm.as<float>(), m.as<float>(),
                       v.as<float>(), grad.as<float>(), wb.as<uint16_t>(), lr, c1, c2);
  }
};

int mlp_forward_device(Mlp *m, const float *dX, int n, float *dY, hipStream_t s) {
  // the ensemble's stream waits for nothing of the network's: the call is ordered on s
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(nblk(n, 256)), dim3(256), 0, s, m->wb.as<uint16_t>(), m->nm, dX,
                     nullptr, nullptr, n, 0, dY, nullptr, nullptr, nullptr, nullptr, nullptr);
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
  m.forward(m.X.as<float>(), (int)n, m.Y.as<float>());
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
      m.loss_grad(didx.as<int>() + s0, b);
      m.adam(lr);
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

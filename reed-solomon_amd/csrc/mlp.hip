// FP32 MLP cost model (12 -> 64 -> 64 -> 1, ReLU) trained and run on the GPU.
//
// It predicts log kernel time from shape / route / strip features so the
// autotuner measures only the top-ranked candidates.  Hand-written kernels:
//   k_forward   one thread per sample; weights staged in LDS; keeps the
//               hidden activations for the backward pass.
//   k_delta     per-sample error terms of the three layers (MSE loss).
//   k_grad      one thread per parameter: dW = sum_n delta_n * act_n, summed
//               over samples in a fixed order, so gradients are deterministic.
//   k_adam      Adam update, one thread per parameter.
// Parameter vector layout: W1[64][12], b1[64], W2[64][64], b2[64], W3[64], b3.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/trs_gpu.h"

namespace trs {
namespace {

constexpr int kIn = 12, kH = 64;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH,
              kW3 = kB2 + kH, kB3 = kW3 + kH, kParams = kB3 + 1;

__global__ __launch_bounds__(256) void k_forward(const float *__restrict__ P,
                                                 const float *__restrict__ X, int n,
                                                 float *__restrict__ H1, float *__restrict__ H2,
                                                 float *__restrict__ Y) {
  __shared__ float w[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float x[kIn];
#pragma unroll
  for (int i = 0; i < kIn; ++i) x[i] = X[size_t(s) * kIn + i];
  float h1[kH];
  for (int j = 0; j < kH; ++j) {
    float a = w[kB1 + j];
#pragma unroll
    for (int i = 0; i < kIn; ++i) a = fmaf(w[kW1 + j * kIn + i], x[i], a);
    h1[j] = a > 0.f ? a : 0.f;
    if (H1) H1[size_t(s) * kH + j] = h1[j];
  }
  float y = w[kB3];
  for (int j = 0; j < kH; ++j) {
    float a = w[kB2 + j];
    for (int i = 0; i < kH; ++i) a = fmaf(w[kW2 + j * kH + i], h1[i], a);
    const float h2 = a > 0.f ? a : 0.f;
    if (H2) H2[size_t(s) * kH + j] = h2;
    y = fmaf(w[kW3 + j], h2, y);
  }
  Y[s] = y;
}

// Error terms for loss L = (1/n) sum (y - t)^2:
//   d3 = 2 (y - t) / n;  d2_j = d3 W3_j [h2_j > 0];  d1_i = sum_j d2_j W2_ji [h1_i > 0].
__global__ __launch_bounds__(256) void k_delta(const float *__restrict__ P,
                                               const float *__restrict__ Y,
                                               const float *__restrict__ T,
                                               const float *__restrict__ H1,
                                               const float *__restrict__ H2, int n,
                                               float *__restrict__ D3, float *__restrict__ D2,
                                               float *__restrict__ D1) {
  __shared__ float w[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) w[i] = P[i];
  __syncthreads();
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const float d3 = 2.f * (Y[s] - T[s]) / float(n);
  D3[s] = d3;
  float d2[kH];
  for (int j = 0; j < kH; ++j) {
    d2[j] = H2[size_t(s) * kH + j] > 0.f ? d3 * w[kW3 + j] : 0.f;
    D2[size_t(s) * kH + j] = d2[j];
  }
  for (int i = 0; i < kH; ++i) {
    float a = 0.f;
    for (int j = 0; j < kH; ++j) a = fmaf(d2[j], w[kW2 + j * kH + i], a);
    D1[size_t(s) * kH + i] = H1[size_t(s) * kH + i] > 0.f ? a : 0.f;
  }
}

// One thread per parameter; sums over samples in index order.
__global__ __launch_bounds__(256) void k_grad(const float *__restrict__ X,
                                              const float *__restrict__ H1,
                                              const float *__restrict__ H2,
                                              const float *__restrict__ D3,
                                              const float *__restrict__ D2,
                                              const float *__restrict__ D1, int n,
                                              float *__restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  float g = 0.f;
  if (p < kB1) {  // W1[j][i]
    const int j = (p - kW1) / kIn, i = (p - kW1) % kIn;
    for (int s = 0; s < n; ++s) g = fmaf(D1[size_t(s) * kH + j], X[size_t(s) * kIn + i], g);
  } else if (p < kW2) {  // b1[j]
    const int j = p - kB1;
    for (int s = 0; s < n; ++s) g += D1[size_t(s) * kH + j];
  } else if (p < kB2) {  // W2[j][i]
    const int j = (p - kW2) / kH, i = (p - kW2) % kH;
    for (int s = 0; s < n; ++s) g = fmaf(D2[size_t(s) * kH + j], H1[size_t(s) * kH + i], g);
  } else if (p < kW3) {  // b2[j]
    const int j = p - kB2;
    for (int s = 0; s < n; ++s) g += D2[size_t(s) * kH + j];
  } else if (p < kB3) {  // W3[j]
    const int j = p - kW3;
    for (int s = 0; s < n; ++s) g = fmaf(D3[s], H2[size_t(s) * kH + j], g);
  } else {  // b3
    for (int s = 0; s < n; ++s) g += D3[s];
  }
  G[p] = g;
}

__global__ __launch_bounds__(256) void k_adam(float *__restrict__ P, const float *__restrict__ G,
                                              float *__restrict__ M, float *__restrict__ V,
                                              float lr, float b1, float b2, float eps, float c1,
                                              float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.f - b1) * g;
  const float v = b2 * V[p] + (1.f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

__global__ __launch_bounds__(256) void k_sqerr(const float *__restrict__ Y,
                                               const float *__restrict__ T, int n,
                                               float *__restrict__ out) {
  // single block, fixed-order reduction
  __shared__ float part[256];
  float a = 0.f;
  for (int s = threadIdx.x; s < n; s += blockDim.x) {
    const float d = Y[s] - T[s];
    a = fmaf(d, d, a);
  }
  part[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = part[0] / float(n);
}

}  // namespace
}  // namespace trs

using namespace trs;

struct trs_mlp {
  float *P = nullptr, *G = nullptr, *M = nullptr, *V = nullptr;
  int step = 0;
};

namespace {

int st_of(hipError_t e) {
  return e == hipSuccess ? TRS_OK : (e == hipErrorOutOfMemory ? TRS_ERR_OUT_OF_MEMORY : TRS_ERR_DEVICE);
}
#define MLP_TRY(x)                     \
  do {                                 \
    const hipError_t e_ = (x);         \
    if (e_ != hipSuccess) return st_of(e_); \
  } while (0)

// Device copies of a training / evaluation set plus activations.
struct Work {
  float *X = nullptr, *T = nullptr, *Y = nullptr, *H1 = nullptr, *H2 = nullptr, *D1 = nullptr,
        *D2 = nullptr, *D3 = nullptr, *loss = nullptr;
  ~Work() {
    for (float *p : {X, T, Y, H1, H2, D1, D2, D3, loss})
      if (p) (void)hipFree(p);
  }
  int alloc(int n, bool train) {
    MLP_TRY(hipMalloc(&X, size_t(n) * kIn * sizeof(float)));
    MLP_TRY(hipMalloc(&Y, size_t(n) * sizeof(float)));
    if (train) {
      MLP_TRY(hipMalloc(&T, size_t(n) * sizeof(float)));
      MLP_TRY(hipMalloc(&H1, size_t(n) * kH * sizeof(float)));
      MLP_TRY(hipMalloc(&H2, size_t(n) * kH * sizeof(float)));
      MLP_TRY(hipMalloc(&D1, size_t(n) * kH * sizeof(float)));
      MLP_TRY(hipMalloc(&D2, size_t(n) * kH * sizeof(float)));
      MLP_TRY(hipMalloc(&D3, size_t(n) * sizeof(float)));
      MLP_TRY(hipMalloc(&loss, sizeof(float)));
    }
    return TRS_OK;
  }
};

int grads(trs_mlp *m, Work &w, int n) {
  const dim3 blk(256), gs((n + 255) / 256), gp((kParams + 255) / 256);
  hipLaunchKernelGGL(k_forward, gs, blk, 0, nullptr, m->P, w.X, n, w.H1, w.H2, w.Y);
  hipLaunchKernelGGL(k_delta, gs, blk, 0, nullptr, m->P, w.Y, w.T, w.H1, w.H2, n, w.D3, w.D2,
                     w.D1);
  hipLaunchKernelGGL(k_grad, gp, blk, 0, nullptr, w.X, w.H1, w.H2, w.D3, w.D2, w.D1, n, m->G);
  hipLaunchKernelGGL(k_sqerr, dim3(1), blk, 0, nullptr, w.Y, w.T, n, w.loss);
  return st_of(hipGetLastError());
}

}  // namespace

extern "C" {

int trs_mlp_param_count(void) { return kParams; }

int trs_mlp_new(uint64_t seed, trs_mlp **out) {
  if (!out) return TRS_ERR_INVALID_ARGUMENT;
  auto *m = new trs_mlp();
  int st = TRS_OK;
  for (float **p : {&m->P, &m->G, &m->M, &m->V}) {
    if (hipMalloc(p, kParams * sizeof(float)) != hipSuccess) st = TRS_ERR_OUT_OF_MEMORY;
  }
  if (st) {
    trs_mlp_free(m);
    return st;
  }
  // He-uniform initial weights from a seeded splitmix64 stream, zero biases.
  std::vector<float> h(kParams, 0.f);
  auto uni = [&]() {
    seed += 0x9E3779B97F4A7C15ull;
    uint64_t z = seed;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return float((z >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
  };
  for (int i = 0; i < kH * kIn; ++i) h[kW1 + i] = uni() * std::sqrt(6.f / kIn);
  for (int i = 0; i < kH * kH; ++i) h[kW2 + i] = uni() * std::sqrt(6.f / kH);
  for (int i = 0; i < kH; ++i) h[kW3 + i] = uni() * std::sqrt(6.f / kH);
  std::vector<float> z(kParams, 0.f);
  if (hipMemcpy(m->P, h.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->M, z.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(m->V, z.data(), kParams * 4, hipMemcpyHostToDevice) != hipSuccess) {
    trs_mlp_free(m);
    return TRS_ERR_DEVICE;
  }
  *out = m;
  return TRS_OK;
}

void trs_mlp_free(trs_mlp *m) {
  if (!m) return;
  for (float *p : {m->P, m->G, m->M, m->V})
    if (p) (void)hipFree(p);
  delete m;
}

int trs_mlp_get_params(const trs_mlp *m, float *out, size_t n) {
  if (!m || !out || n != size_t(kParams)) return TRS_ERR_INVALID_ARGUMENT;
  MLP_TRY(hipMemcpy(out, m->P, kParams * 4, hipMemcpyDeviceToHost));
  return TRS_OK;
}

int trs_mlp_set_params(trs_mlp *m, const float *in, size_t n) {
  if (!m || !in || n != size_t(kParams)) return TRS_ERR_INVALID_ARGUMENT;
  MLP_TRY(hipMemcpy(m->P, in, kParams * 4, hipMemcpyHostToDevice));
  return TRS_OK;
}

int trs_mlp_gradients(trs_mlp *m, const float *X, const float *T, size_t n, float *grad_out,
                      float *loss_out) {
  if (!m || !X || !T || n == 0 || n > (1u << 30)) return TRS_ERR_INVALID_ARGUMENT;
  Work w;
  int st = w.alloc(int(n), true);
  if (st) return st;
  MLP_TRY(hipMemcpy(w.X, X, n * kIn * 4, hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(w.T, T, n * 4, hipMemcpyHostToDevice));
  if ((st = grads(m, w, int(n)))) return st;
  MLP_TRY(hipDeviceSynchronize());
  if (grad_out) MLP_TRY(hipMemcpy(grad_out, m->G, kParams * 4, hipMemcpyDeviceToHost));
  if (loss_out) MLP_TRY(hipMemcpy(loss_out, w.loss, 4, hipMemcpyDeviceToHost));
  return TRS_OK;
}

int trs_mlp_train(trs_mlp *m, const float *X, const float *T, size_t n, int steps, float lr,
                  float *loss_out) {
  if (!m || !X || !T || n == 0 || steps < 0 || n > (1u << 30)) return TRS_ERR_INVALID_ARGUMENT;
  Work w;
  int st = w.alloc(int(n), true);
  if (st) return st;
  MLP_TRY(hipMemcpy(w.X, X, n * kIn * 4, hipMemcpyHostToDevice));
  MLP_TRY(hipMemcpy(w.T, T, n * 4, hipMemcpyHostToDevice));
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  for (int it = 0; it < steps; ++it) {
    if ((st = grads(m, w, int(n)))) return st;
    ++m->step;
    const float c1 = 1.f - std::pow(b1, float(m->step));
    const float c2 = 1.f - std::pow(b2, float(m->step));
    hipLaunchKernelGGL(k_adam, dim3((kParams + 255) / 256), dim3(256), 0, nullptr, m->P, m->G,
                       m->M, m->V, lr, b1, b2, eps, c1, c2);
    MLP_TRY(hipGetLastError());
  }
  // loss at the final parameters
  if ((st = grads(m, w, int(n)))) return st;
  MLP_TRY(hipDeviceSynchronize());
  if (loss_out) MLP_TRY(hipMemcpy(loss_out, w.loss, 4, hipMemcpyDeviceToHost));
  return TRS_OK;
}

int trs_mlp_predict(const trs_mlp *m, const float *X, size_t n, float *Y) {
  if (!m || !X || !Y || n == 0 || n > (1u << 30)) return TRS_ERR_INVALID_ARGUMENT;
  Work w;
  int st = w.alloc(int(n), false);
  if (st) return st;
  MLP_TRY(hipMemcpy(w.X, X, n * kIn * 4, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_forward, dim3((n + 255) / 256), dim3(256), 0, nullptr, m->P, w.X, int(n),
                     nullptr, nullptr, w.Y);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpy(Y, w.Y, n * 4, hipMemcpyDeviceToHost));
  return TRS_OK;
}

}  // extern "C"

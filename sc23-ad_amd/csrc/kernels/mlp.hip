// The learned-parameter network on the GPU: forward pass, mean-squared-error
// gradient and Adam update, one thread per sample, parameters staged in LDS.
// Shape and parameter layout: csrc/host/learn.hpp (8 -> 32 -> 32 -> 2).
#include <hip/hip_runtime.h>

#include "launch.hpp"

namespace sc23ad::dev {
namespace {

constexpr int kIn = 8, kH = 32, kOut = 2;
constexpr int kOffW1 = 0, kOffB1 = kOffW1 + kH * kIn, kOffW2 = kOffB1 + kH,
              kOffB2 = kOffW2 + kH * kH, kOffW3 = kOffB2 + kH, kOffB3 = kOffW3 + kOut * kH,
              kNP = kOffB3 + kOut;
constexpr int kT = 256;

__device__ __forceinline__ float sigm(float z) { return 1.0f / (1.0f + expf(-z)); }

struct Act {
  float h1[kH], h2[kH], y[kOut];
};

__device__ __forceinline__ void forward(const float* P, const float* x, Act& a) {
#pragma unroll 4
  for (int j = 0; j < kH; ++j) {
    float z = P[kOffB1 + j];
    for (int i = 0; i < kIn; ++i) z += P[kOffW1 + j * kIn + i] * x[i];
    a.h1[j] = tanhf(z);
  }
#pragma unroll 4
  for (int j = 0; j < kH; ++j) {
    float z = P[kOffB2 + j];
    for (int i = 0; i < kH; ++i) z += P[kOffW2 + j * kH + i] * a.h1[i];
    a.h2[j] = tanhf(z);
  }
  for (int k = 0; k < kOut; ++k) {
    float z = P[kOffB3 + k];
    for (int i = 0; i < kH; ++i) z += P[kOffW3 + k * kH + i] * a.h2[i];
    a.y[k] = sigm(z);
  }
}

__global__ __launch_bounds__(kT) void k_mlp_forward(int B, const float* __restrict__ X,
                                                    const float* __restrict__ Pg,
                                                    float* __restrict__ Y) {
  __shared__ float P[kNP];
  for (int k = threadIdx.x; k < kNP; k += kT) P[k] = Pg[k];
  __syncthreads();
  const int s = blockIdx.x * kT + threadIdx.x;
  if (s >= B) return;
  float x[kIn];
  for (int i = 0; i < kIn; ++i) x[i] = X[s * kIn + i];
  Act a;
  forward(P, x, a);
  for (int k = 0; k < kOut; ++k) Y[s * kOut + k] = a.y[k];
}

// Accumulates (1/B) d/dP sum_s sum_k (y_sk - t_sk)^2 into G and the loss
// into *loss (both zeroed by the caller).
__global__ __launch_bounds__(kT) void k_mlp_grad(int B, const float* __restrict__ X,
                                                 const float* __restrict__ T,
                                                 const float* __restrict__ Pg, float* G,
                                                 float* loss) {
  __shared__ float P[kNP];
  __shared__ float g[kNP];
  __shared__ float lsum;
  for (int k = threadIdx.x; k < kNP; k += kT) {
    P[k] = Pg[k];
    g[k] = 0.0f;
  }
  if (threadIdx.x == 0) lsum = 0.0f;
  __syncthreads();
  const int s = blockIdx.x * kT + threadIdx.x;
  if (s < B) {
    float x[kIn];
    for (int i = 0; i < kIn; ++i) x[i] = X[s * kIn + i];
    Act a;
    forward(P, x, a);
    const float inv_b = 1.0f / static_cast<float>(B);
    float d3[kOut], l = 0.0f;
    for (int k = 0; k < kOut; ++k) {
      const float e = a.y[k] - T[s * kOut + k];
      l += e * e;
      d3[k] = 2.0f * e * inv_b * a.y[k] * (1.0f - a.y[k]);
    }
    atomicAdd(&lsum, l * inv_b);
    float d2[kH];
    for (int i = 0; i < kH; ++i) {
      float acc = 0.0f;
      for (int k = 0; k < kOut; ++k) acc += P[kOffW3 + k * kH + i] * d3[k];
      d2[i] = acc * (1.0f - a.h2[i] * a.h2[i]);
    }
    for (int k = 0; k < kOut; ++k) {
      atomicAdd(&g[kOffB3 + k], d3[k]);
      for (int i = 0; i < kH; ++i) atomicAdd(&g[kOffW3 + k * kH + i], d3[k] * a.h2[i]);
    }
    float d1[kH];
    for (int i = 0; i < kH; ++i) {
      float acc = 0.0f;
      for (int j = 0; j < kH; ++j) acc += P[kOffW2 + j * kH + i] * d2[j];
      d1[i] = acc * (1.0f - a.h1[i] * a.h1[i]);
    }
    for (int j = 0; j < kH; ++j) {
      atomicAdd(&g[kOffB2 + j], d2[j]);
      for (int i = 0; i < kH; ++i) atomicAdd(&g[kOffW2 + j * kH + i], d2[j] * a.h1[i]);
    }
    for (int j = 0; j < kH; ++j) {
      atomicAdd(&g[kOffB1 + j], d1[j]);
      for (int i = 0; i < kIn; ++i) atomicAdd(&g[kOffW1 + j * kIn + i], d1[j] * x[i]);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < kNP; k += kT) atomicAdd(&G[k], g[k]);
  if (threadIdx.x == 0) atomicAdd(loss, lsum);
}

__global__ __launch_bounds__(kT) void k_adam(int n, float* P, const float* G, float* M,
                                             float* V, float lr, float b1, float b2,
                                             float eps, float c1, float c2) {
  const int k = blockIdx.x * kT + threadIdx.x;
  if (k >= n) return;
  const float g = G[k];
  const float m = b1 * M[k] + (1.0f - b1) * g;
  const float v = b2 * V[k] + (1.0f - b2) * g * g;
  M[k] = m;
  V[k] = v;
  P[k] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

}  // namespace

int mlp_param_count() { return kNP; }

hipError_t launch_mlp_forward(int B, const float* X, const float* P, float* Y, hipStream_t s) {
  if (B <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_mlp_forward, dim3((B + kT - 1) / kT), dim3(kT), 0, s, B, X, P, Y);
  return hipGetLastError();
}

hipError_t launch_mlp_grad(int B, const float* X, const float* T, const float* P, float* G,
                           float* loss, hipStream_t s) {
  hipError_t e = hipMemsetAsync(G, 0, sizeof(float) * kNP, s);
  if (e != hipSuccess) return e;
  e = hipMemsetAsync(loss, 0, sizeof(float), s);
  if (e != hipSuccess || B <= 0) return e;
  hipLaunchKernelGGL(k_mlp_grad, dim3((B + kT - 1) / kT), dim3(kT), 0, s, B, X, T, P, G, loss);
  return hipGetLastError();
}

hipError_t launch_adam(int n, float* P, const float* G, float* M, float* V, float lr, int step,
                       hipStream_t s) {
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - powf(b1, static_cast<float>(step));
  const float c2 = 1.0f - powf(b2, static_cast<float>(step));
  hipLaunchKernelGGL(k_adam, dim3((n + kT - 1) / kT), dim3(kT), 0, s, n, P, G, M, V, lr, b1, b2,
                     1e-8f, c1, c2);
  return hipGetLastError();
}

}  // namespace sc23ad::dev

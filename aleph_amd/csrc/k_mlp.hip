// The planner's cost model: a 4-layer perceptron (kMlpIn features -> 3 x 64
// ReLU -> 1) in FP32, with hand-written forward, gradient and Adam kernels.
//
// Parameters live in one flat FP32 array (mlp_layout()): W1[in][64], b1,
// W2[64][64], b2, W3[64][64], b3, w4[64], b4.  A workgroup stages the whole
// parameter array in LDS (37 KiB) and each thread evaluates one sample with
// its activations in registers.  Gradients are accumulated per workgroup in
// LDS (float atomics), then added once per workgroup to the global gradient.
#include "common.hpp"
#include "kernels.hpp"

namespace aleph {
namespace {

constexpr int H = kMlpHidden;
constexpr int IN = kMlpIn;
constexpr int OW1 = 0, OB1 = OW1 + IN * H, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + H * H,
              OW4 = OB3 + H, OB4 = OW4 + H, NP = OB4 + 1;
static_assert(NP == kMlpParams, "parameter layout");

constexpr int kMlpThreads = 64;

__device__ __forceinline__ void stage_params(const float* __restrict__ p, float* __restrict__ s) {
  for (int i = threadIdx.x; i < NP; i += blockDim.x) s[i] = p[i];
  __syncthreads();
}

// h_out[j] = relu(b[j] + sum_i h_in[i] W[i][j])
template <int NI>
__device__ __forceinline__ void dense_relu(const float* __restrict__ W, const float* __restrict__ b,
                                           const float (&hin)[NI], float (&hout)[H]) {
#pragma unroll
  for (int j = 0; j < H; j++) hout[j] = b[j];
#pragma unroll
  for (int i = 0; i < NI; i++) {
    const float x = hin[i];
#pragma unroll
    for (int j = 0; j < H; j++) hout[j] = fmaf(x, W[i * H + j], hout[j]);
  }
#pragma unroll
  for (int j = 0; j < H; j++) hout[j] = fmaxf(hout[j], 0.f);
}

__device__ __forceinline__ float out_layer(const float* __restrict__ s, const float (&h3)[H]) {
  float y = s[OB4];
#pragma unroll
  for (int j = 0; j < H; j++) y = fmaf(h3[j], s[OW4 + j], y);
  return y;
}

__global__ void __launch_bounds__(kMlpThreads) k_mlp_forward(const float* __restrict__ params,
                                                             const float* __restrict__ X, uint32_t B,
                                                             float* __restrict__ out) {
  __shared__ float s[NP];
  stage_params(params, s);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B) return;
  float x[IN], h1[H], h2[H], h3[H];
#pragma unroll
  for (int k = 0; k < IN; k++) x[k] = X[(size_t)i * IN + k];
  dense_relu<IN>(s + OW1, s + OB1, x, h1);
  dense_relu<H>(s + OW2, s + OB2, h1, h2);
  dense_relu<H>(s + OW3, s + OB3, h2, h3);
  out[i] = out_layer(s, h3);
}

// Mean squared error over the batch; grads (global, zeroed by the caller)
// receives d(loss)/d(params), loss[0] the loss (atomically accumulated).
__global__ void __launch_bounds__(kMlpThreads) k_mlp_grad(const float* __restrict__ params,
                                                          const float* __restrict__ X, const float* __restrict__ y,
                                                          uint32_t B, float* __restrict__ grads,
                                                          float* __restrict__ loss) {
  __shared__ float s[NP];
  __shared__ float g[NP];
  for (int k = threadIdx.x; k < NP; k += blockDim.x) g[k] = 0.f;
  stage_params(params, s);
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < B) {
    float x[IN], h1[H], h2[H], h3[H];
#pragma unroll
    for (int k = 0; k < IN; k++) x[k] = X[(size_t)i * IN + k];
    dense_relu<IN>(s + OW1, s + OB1, x, h1);
    dense_relu<H>(s + OW2, s + OB2, h1, h2);
    dense_relu<H>(s + OW3, s + OB3, h2, h3);
    const float r = out_layer(s, h3) - y[i];
    atomicAdd(loss, r * r / (float)B);
    const float d4 = 2.f * r / (float)B;
    atomicAdd(&g[OB4], d4);
    // layer 3: d3[j] = d4 w4[j] [h3 > 0]
    float d[H];
#pragma unroll
    for (int j = 0; j < H; j++) {
      atomicAdd(&g[OW4 + j], d4 * h3[j]);
      d[j] = h3[j] > 0.f ? d4 * s[OW4 + j] : 0.f;
    }
    // W3, b3 from d3; d2 = (W3 d3) [h2 > 0]
    float e[H];
#pragma unroll
    for (int k = 0; k < H; k++) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < H; j++) {
        atomicAdd(&g[OW3 + k * H + j], h2[k] * d[j]);
        acc = fmaf(s[OW3 + k * H + j], d[j], acc);
      }
      e[k] = h2[k] > 0.f ? acc : 0.f;
    }
#pragma unroll
    for (int j = 0; j < H; j++) atomicAdd(&g[OB3 + j], d[j]);
    // W2, b2 from d2; d1 = (W2 d2) [h1 > 0]
#pragma unroll
    for (int k = 0; k < H; k++) {
      float acc = 0.f;
#pragma unroll
      for (int j = 0; j < H; j++) {
        atomicAdd(&g[OW2 + k * H + j], h1[k] * e[j]);
        acc = fmaf(s[OW2 + k * H + j], e[j], acc);
      }
      d[k] = h1[k] > 0.f ? acc : 0.f;
    }
#pragma unroll
    for (int j = 0; j < H; j++) atomicAdd(&g[OB2 + j], e[j]);
    // W1, b1 from d1
#pragma unroll
    for (int k = 0; k < IN; k++)
#pragma unroll
      for (int j = 0; j < H; j++) atomicAdd(&g[OW1 + k * H + j], x[k] * d[j]);
#pragma unroll
    for (int j = 0; j < H; j++) atomicAdd(&g[OB1 + j], d[j]);
  }
  __syncthreads();
  for (int k = threadIdx.x; k < NP; k += blockDim.x) atomicAdd(&grads[k], g[k]);
}

__global__ void __launch_bounds__(256) k_mlp_adam(float* __restrict__ params, const float* __restrict__ grads,
                                                  float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                                                  float b2, float eps, float c1, float c2) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= NP) return;
  const float gk = grads[k];
  const float mk = b1 * m[k] + (1.f - b1) * gk;
  const float vk = b2 * v[k] + (1.f - b2) * gk * gk;
  m[k] = mk;
  v[k] = vk;
  params[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

__global__ void k_mlp_zero(float* __restrict__ p, uint32_t n) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = 0.f;
}

}  // namespace

void launch_mlp_forward(hipStream_t s, const float* params, const float* X, uint32_t B, float* out) {
  if (B == 0) return;
  hipLaunchKernelGGL(k_mlp_forward, dim3((B + kMlpThreads - 1) / kMlpThreads), dim3(kMlpThreads), 0, s, params, X,
                     B, out);
  ALEPH_HIP_CHECK(hipGetLastError());
}

void launch_mlp_grad(hipStream_t s, const float* params, const float* X, const float* y, uint32_t B, float* grads,
                     float* loss) {
  hipLaunchKernelGGL(k_mlp_zero, dim3((kMlpParams + 1 + 255) / 256), dim3(256), 0, s, grads, (uint32_t)kMlpParams);
  ALEPH_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_zero, dim3(1), dim3(64), 0, s, loss, 1u);
  ALEPH_HIP_CHECK(hipGetLastError());
  if (B == 0) return;
  hipLaunchKernelGGL(k_mlp_grad, dim3((B + kMlpThreads - 1) / kMlpThreads), dim3(kMlpThreads), 0, s, params, X, y,
                     B, grads, loss);
  ALEPH_HIP_CHECK(hipGetLastError());
}

void launch_mlp_adam(hipStream_t s, float* params, const float* grads, float* m, float* v, float lr, float b1,
                     float b2, float eps, uint32_t t) {
  const float c1 = 1.f - powf(b1, (float)t), c2 = 1.f - powf(b2, (float)t);
  hipLaunchKernelGGL(k_mlp_adam, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, params, grads, m, v, lr, b1, b2,
                     eps, c1, c2);
  ALEPH_HIP_CHECK(hipGetLastError());
}

}  // namespace aleph

// The device library, part three: the small multilayer perceptron of the learned tile schedule.
// x (F features) -> relu(W1^T x + b1) (64) -> relu(W2^T h1 + b2) (64) -> w3 . h2 + b3, in FP32.
// Inference, the gradient of the mean squared error over a batch (forward and backward in one
// kernel, with a fixed-order reduction over workgroups), and an Adam step.  gfx950 only.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "common.h"

using ttdev::as_stream;
using ttdev::fail;

namespace {

constexpr int kHidden = 64;        // one wave per sample: lane j is hidden unit j
constexpr int kMaxFeatures = 64;
constexpr int kSamplesPerBlock = 4;  // four waves, one sample each
constexpr int kMlpThreads = kHidden * kSamplesPerBlock;
constexpr int kGradBlocks = 120;

// Offsets of the parameter vector: W1 [F][H], b1 [H], W2 [H][H], b2 [H], w3 [H], b3.
struct Layout {
  int64_t w1, b1, w2, b2, w3, b3, total;
  __host__ __device__ explicit Layout(int f) {
    w1 = 0;
    b1 = w1 + int64_t(f) * kHidden;
    w2 = b1 + kHidden;
    b2 = w2 + int64_t(kHidden) * kHidden;
    w3 = b2 + kHidden;
    b3 = w3 + kHidden;
    total = b3 + 1;
  }
};

// The sum of `v` over the 64 lanes of a wave, in a fixed order (every lane gets it).
__device__ float wave_sum(float v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
  return v;
}

__global__ __launch_bounds__(kMlpThreads) void mlp_forward_kernel(int f, int64_t batch, const float *__restrict__ x,
                                                                  const float *__restrict__ params,
                                                                  float *__restrict__ y) {
  __shared__ float xs[kSamplesPerBlock][kMaxFeatures];
  __shared__ float h1s[kSamplesPerBlock][kHidden];
  const Layout L(f);
  const int lane = threadIdx.x % kHidden, slot = threadIdx.x / kHidden;
  for (int64_t base = int64_t(blockIdx.x) * kSamplesPerBlock; base < batch;
       base += int64_t(gridDim.x) * kSamplesPerBlock) {
    const int64_t s = base + slot;
    const bool live = s < batch;
    if (live && lane < f) xs[slot][lane] = x[s * f + lane];
    __syncthreads();
    float a = params[L.b1 + lane];
    if (live)
      for (int i = 0; i < f; ++i) a += xs[slot][i] * params[L.w1 + int64_t(i) * kHidden + lane];
    h1s[slot][lane] = a > 0.0f ? a : 0.0f;
    __syncthreads();
    float b = params[L.b2 + lane];
    if (live)
      for (int i = 0; i < kHidden; ++i) b += h1s[slot][i] * params[L.w2 + int64_t(i) * kHidden + lane];
    const float h2 = b > 0.0f ? b : 0.0f;
    const float out = wave_sum(h2 * params[L.w3 + lane]) + params[L.b3];
    if (live && lane == 0) y[s] = out;
    __syncthreads();
  }
}

// Forward and backward of loss = mean over the batch of (y - t)^2.  Each workgroup takes samples
// blockIdx.x * 4 + slot, stepping by the grid, and accumulates its gradient in LDS (one copy per
// wave, summed at the end in wave order); partial[block] = (gradient, sum of squared errors).
__global__ __launch_bounds__(kMlpThreads) void mlp_gradient_kernel(int f, int64_t batch, const float *__restrict__ x,
                                                                   const float *__restrict__ target,
                                                                   const float *__restrict__ params,
                                                                   float *__restrict__ partial) {
  __shared__ float xs[kSamplesPerBlock][kMaxFeatures];
  __shared__ float h1s[kSamplesPerBlock][kHidden];
  __shared__ float g2s[kSamplesPerBlock][kHidden];
  __shared__ float acc_w1[kSamplesPerBlock][kMaxFeatures][kHidden];
  __shared__ float acc_w2[kSamplesPerBlock][kHidden][kHidden];
  const Layout L(f);
  const int lane = threadIdx.x % kHidden, slot = threadIdx.x / kHidden;
  for (int i = 0; i < kMaxFeatures; ++i) acc_w1[slot][i][lane] = 0.0f;
  for (int i = 0; i < kHidden; ++i) acc_w2[slot][i][lane] = 0.0f;
  float acc_b1 = 0.0f, acc_b2 = 0.0f, acc_w3 = 0.0f, acc_b3 = 0.0f, sse = 0.0f;
  const float scale = 2.0f / static_cast<float>(batch);
  for (int64_t base = int64_t(blockIdx.x) * kSamplesPerBlock; base < batch;
       base += int64_t(gridDim.x) * kSamplesPerBlock) {
    const int64_t s = base + slot;
    const bool live = s < batch;
    if (live && lane < f) xs[slot][lane] = x[s * f + lane];
    __syncthreads();
    float a = params[L.b1 + lane];
    for (int i = 0; i < f; ++i) a += (live ? xs[slot][i] : 0.0f) * params[L.w1 + int64_t(i) * kHidden + lane];
    const float h1 = a > 0.0f ? a : 0.0f;
    h1s[slot][lane] = h1;
    __syncthreads();
    float b = params[L.b2 + lane];
    for (int i = 0; i < kHidden; ++i) b += h1s[slot][i] * params[L.w2 + int64_t(i) * kHidden + lane];
    const float h2 = b > 0.0f ? b : 0.0f;
    const float y = wave_sum(h2 * params[L.w3 + lane]) + params[L.b3];
    const float err = live ? y - target[s] : 0.0f;
    const float dy = scale * err;
    sse += lane == 0 ? err * err : 0.0f;
    acc_w3 += dy * h2;
    acc_b3 += lane == 0 ? dy : 0.0f;
    const float g2 = b > 0.0f ? dy * params[L.w3 + lane] : 0.0f;
    g2s[slot][lane] = g2;
    acc_b2 += g2;
    for (int i = 0; i < kHidden; ++i) acc_w2[slot][i][lane] += h1s[slot][i] * g2;
    __syncthreads();
    // lane i: dh1_i = sum_j W2[i][j] g2_j
    float dh1 = 0.0f;
    for (int j = 0; j < kHidden; ++j) dh1 += params[L.w2 + int64_t(lane) * kHidden + j] * g2s[slot][j];
    const float g1 = a > 0.0f ? dh1 : 0.0f;
    acc_b1 += g1;
    for (int i = 0; i < f; ++i) acc_w1[slot][i][lane] += (live ? xs[slot][i] : 0.0f) * g1;
    __syncthreads();
  }
  // Sum the four waves' accumulators in wave order and write this block's partial gradient.
  float *out = partial + int64_t(blockIdx.x) * (L.total + 1);
  __shared__ float totals[kSamplesPerBlock][2];  // per wave: d b3, sum of squared errors
  const float wave_b3 = wave_sum(acc_b3), wave_sse = wave_sum(sse);
  if (lane == 0) {
    totals[slot][0] = wave_b3;
    totals[slot][1] = wave_sse;
  }
  __syncthreads();
  for (int64_t idx = threadIdx.x; idx < int64_t(f) * kHidden; idx += kMlpThreads) {
    const int i = static_cast<int>(idx / kHidden), j = static_cast<int>(idx % kHidden);
    float v = 0.0f;
    for (int w = 0; w < kSamplesPerBlock; ++w) v += acc_w1[w][i][j];
    out[L.w1 + idx] = v;
  }
  for (int64_t idx = threadIdx.x; idx < int64_t(kHidden) * kHidden; idx += kMlpThreads) {
    const int i = static_cast<int>(idx / kHidden), j = static_cast<int>(idx % kHidden);
    float v = 0.0f;
    for (int w = 0; w < kSamplesPerBlock; ++w) v += acc_w2[w][i][j];
    out[L.w2 + idx] = v;
  }
  __syncthreads();
  // The per-lane vectors b1, b2, w3: staged in turn through row 0 of each wave's (already written)
  // W2 accumulator, then summed in wave order.
  const float lane_values[3] = {acc_b1, acc_b2, acc_w3};
  const int64_t lane_offsets[3] = {L.b1, L.b2, L.w3};
  for (int k = 0; k < 3; ++k) {
    acc_w2[slot][0][lane] = lane_values[k];
    __syncthreads();
    if (threadIdx.x < kHidden) {
      float v = 0.0f;
      for (int w = 0; w < kSamplesPerBlock; ++w) v += acc_w2[w][0][threadIdx.x];
      out[lane_offsets[k] + threadIdx.x] = v;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    float b3 = 0.0f, loss = 0.0f;
    for (int w = 0; w < kSamplesPerBlock; ++w) {
      b3 += totals[w][0];
      loss += totals[w][1];
    }
    out[L.b3] = b3;
    out[L.total] = loss;
  }
}

// gradient[p] = sum over blocks of partial[block][p], in block order; loss = sum of squared errors / batch.
__global__ void mlp_reduce_kernel(int64_t total, int blocks, int64_t batch, const float *__restrict__ partial,
                                  float *__restrict__ gradient, float *__restrict__ loss) {
  for (int64_t p = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; p <= total; p += int64_t(gridDim.x) * blockDim.x) {
    float v = 0.0f;
    for (int b = 0; b < blocks; ++b) v += partial[int64_t(b) * (total + 1) + p];
    if (p < total)
      gradient[p] = v;
    else if (loss)
      *loss = v / static_cast<float>(batch);
  }
}

__global__ void adam_kernel(int64_t n, float *__restrict__ params, const float *__restrict__ grad, float *__restrict__ m,
                            float *__restrict__ v, float lr, float beta1, float beta2, float eps, float correct1,
                            float correct2) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
    const float g = grad[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * g;
    const float vi = beta2 * v[i] + (1.0f - beta2) * g * g;
    m[i] = mi;
    v[i] = vi;
    params[i] -= lr * (mi / correct1) / (sqrtf(vi / correct2) + eps);
  }
}

std::mutex g_partial_mutex;
std::vector<std::pair<float *, int64_t>> g_partial;  // per device: buffer, capacity in floats

int partial_buffer(int64_t floats, float **out) {
  int device = 0;
  TT_HIP(hipGetDevice(&device));
  std::lock_guard<std::mutex> lock(g_partial_mutex);
  if (g_partial.size() <= static_cast<size_t>(device)) g_partial.resize(static_cast<size_t>(device) + 1, {nullptr, 0});
  auto &slot = g_partial[static_cast<size_t>(device)];
  if (slot.second < floats) {
    if (slot.first) TT_HIP(hipFree(slot.first));
    slot = {nullptr, 0};
    TT_HIP(hipMalloc(reinterpret_cast<void **>(&slot.first), sizeof(float) * static_cast<size_t>(floats)));
    slot.second = floats;
  }
  *out = slot.first;
  return 0;
}

int check_shape(int features, int hidden) {
  if (hidden != kHidden) return fail("the cost model's hidden width is 64");
  if (features < 1 || features > kMaxFeatures) return fail("the cost model takes 1 to 64 features");
  return 0;
}

}  // namespace

extern "C" int64_t tt_mlp_param_count(int features, int hidden) {
  if (check_shape(features, hidden)) return -1;
  return Layout(features).total;
}

extern "C" int tt_mlp_forward(int features, int hidden, int64_t batch, const float *x, const float *params, float *y,
                              void *stream) {
  if (int rc = check_shape(features, hidden)) return rc;
  if (batch < 0) return fail("tt_mlp_forward: negative batch");
  if (batch == 0) return 0;
  if (!x || !params || !y) return fail("tt_mlp_forward: null buffer");
  const int64_t blocks = std::min<int64_t>((batch + kSamplesPerBlock - 1) / kSamplesPerBlock, 1024);
  hipLaunchKernelGGL(mlp_forward_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(kMlpThreads), 0, as_stream(stream),
                     features, batch, x, params, y);
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_mlp_gradient(int features, int hidden, int64_t batch, const float *x, const float *target,
                               const float *params, float *gradient, float *loss, void *stream) {
  if (int rc = check_shape(features, hidden)) return rc;
  if (batch <= 0) return fail("tt_mlp_gradient: empty batch");
  if (!x || !target || !params || !gradient) return fail("tt_mlp_gradient: null buffer");
  const Layout L(features);
  const int blocks = static_cast<int>(std::min<int64_t>((batch + kSamplesPerBlock - 1) / kSamplesPerBlock, kGradBlocks));
  float *partial = nullptr;
  if (int rc = partial_buffer(int64_t(blocks) * (L.total + 1), &partial)) return rc;
  hipLaunchKernelGGL(mlp_gradient_kernel, dim3(blocks), dim3(kMlpThreads), 0, as_stream(stream), features, batch, x,
                     target, params, partial);
  TT_LAUNCHED();
  hipLaunchKernelGGL(mlp_reduce_kernel, dim3(static_cast<uint32_t>((L.total + 1 + 255) / 256)), dim3(256), 0,
                     as_stream(stream), L.total, blocks, batch, partial, gradient, loss);
  TT_LAUNCHED();
  return 0;
}

extern "C" int tt_mlp_adam(int64_t n, float *params, const float *gradient, float *m, float *v, float lr, float beta1,
                           float beta2, float eps, int step, void *stream) {
  if (n < 0 || step < 1) return fail("tt_mlp_adam: bad size or step");
  if (n == 0) return 0;
  if (!params || !gradient || !m || !v) return fail("tt_mlp_adam: null buffer");
  const float c1 = 1.0f - std::pow(beta1, static_cast<float>(step));
  const float c2 = 1.0f - std::pow(beta2, static_cast<float>(step));
  hipLaunchKernelGGL(adam_kernel, dim3(static_cast<uint32_t>(std::min<int64_t>((n + 255) / 256, 1024))), dim3(256), 0,
                     as_stream(stream), n, params, gradient, m, v, lr, beta1, beta2, eps, c1, c2);
  TT_LAUNCHED();
  return 0;
}

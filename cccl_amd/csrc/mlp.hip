// Config-selector network: a 3-layer MLP (16 -> 64 -> 64 -> C, ReLU) that reads cheap
// input statistics of a sort call and predicts, per launch configuration, the log of its
// run time relative to the call's best; the selector takes the argmin.
//
// Hand-written gfx950 kernels for inference, training (forward + backward with a
// deterministic two-stage gradient reduction) and Adam.  Sizes are tiny (a few thousand
// parameters, batches of a few thousand calls), so this is plain fp32 VALU work in LDS;
// there is nothing here for MFMA to win.
//
// Parameter layout (floats): W1[16][64], b1[64], W2[64][64], b2[64], W3[64][C], b3[C].
// Loss: masked mean squared error sum(M * (Y - T)^2) / sum(M).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/cccl_amd/config_selector.h"

namespace cccl_amd {
namespace mlp {

constexpr int kIn      = 16;
constexpr int kHidden  = 64;
constexpr int kMaxOut  = 128;
constexpr int kSamples = 32; // samples per workgroup (keeps the LDS tiles at 51 KiB)
constexpr int kThreads = 256;

__host__ __device__ constexpr int param_count(int out)
{
  return kIn * kHidden + kHidden + kHidden * kHidden + kHidden + kHidden * out + out;
}

struct Offsets
{
  int w1, b1, w2, b2, w3, b3;
  __host__ __device__ explicit Offsets(int out)
      : w1(0)
      , b1(kIn * kHidden)
      , w2(b1 + kHidden)
      , b2(w2 + kHidden * kHidden)
      , w3(b2 + kHidden)
      , b3(w3 + kHidden * out)
  {}
};

struct StepParams
{
  const float* params;
  const float* x; // [batch][16]
  const float* target; // [batch][out] (train only)
  const float* mask; // [batch][out] (train only)
  float* y; // [batch][out] predictions (may be null when training)
  float* partial; // [blocks][P + 2]: per-block gradient sums, loss sum, mask sum (train only)
  int batch;
  int out;
  int train;
};

// One workgroup per kSamples samples: forward pass in LDS, then (train) the backward pass and
// this block's sums of the parameter gradients.
__global__ void __launch_bounds__(kThreads) mlp_step_kernel(StepParams p)
{
  __shared__ float s_x[kSamples][kIn];
  __shared__ float s_a1[kSamples][kHidden];
  __shared__ float s_a2[kSamples][kHidden];
  __shared__ float s_dy[kSamples][kMaxOut];
  __shared__ float s_dz2[kSamples][kHidden];
  __shared__ float s_dz1[kSamples][kHidden];
  __shared__ float s_red[2][kThreads];

  const int tid  = threadIdx.x;
  const int s0   = blockIdx.x * kSamples;
  const int out  = p.out;
  const Offsets o(out);
  const float* W = p.params;

  for (int i = tid; i < kSamples * kIn; i += kThreads)
  {
    const int s   = i / kIn;
    const int f   = i % kIn;
    s_x[s][f]     = (s0 + s < p.batch) ? p.x[size_t(s0 + s) * kIn + f] : 0.f;
  }
  __syncthreads();
  for (int i = tid; i < kSamples * kHidden; i += kThreads)
  {
    const int s = i / kHidden;
    const int h = i % kHidden;
    float z     = W[o.b1 + h];
#pragma unroll
    for (int f = 0; f < kIn; ++f)
    {
      z += s_x[s][f] * W[o.w1 + f * kHidden + h];
    }
    s_a1[s][h] = z > 0.f ? z : 0.f;
  }
  __syncthreads();
  for (int i = tid; i < kSamples * kHidden; i += kThreads)
  {
    const int s = i / kHidden;
    const int h = i % kHidden;
    float z     = W[o.b2 + h];
    for (int k = 0; k < kHidden; ++k)
    {
      z += s_a1[s][k] * W[o.w2 + k * kHidden + h];
    }
    s_a2[s][h] = z > 0.f ? z : 0.f;
  }
  __syncthreads();
  float loss = 0.f, msum = 0.f;
  for (int i = tid; i < kSamples * out; i += kThreads)
  {
    const int s = i / out;
    const int c = i % out;
    float y     = W[o.b3 + c];
    for (int k = 0; k < kHidden; ++k)
    {
      y += s_a2[s][k] * W[o.w3 + k * out + c];
    }
    const bool live = s0 + s < p.batch;
    if (live && p.y != nullptr)
    {
      p.y[size_t(s0 + s) * out + c] = y;
    }
    if (p.train)
    {
      const float m = live ? p.mask[size_t(s0 + s) * out + c] : 0.f;
      const float d = live ? (y - p.target[size_t(s0 + s) * out + c]) : 0.f;
      s_dy[s][c]    = m * d; // scaled by 2 / sum(M) in the reduction kernel
      loss += m * d * d;
      msum += m;
    }
  }
  if (!p.train)
  {
    return; // block-uniform
  }
  s_red[0][tid] = loss;
  s_red[1][tid] = msum;
  __syncthreads();

  // dZ2 = (dY W3^T) * [A2 > 0]
  for (int i = tid; i < kSamples * kHidden; i += kThreads)
  {
    const int s = i / kHidden;
    const int h = i % kHidden;
    float g     = 0.f;
    for (int c = 0; c < out; ++c)
    {
      g += s_dy[s][c] * W[o.w3 + h * out + c];
    }
    s_dz2[s][h] = s_a2[s][h] > 0.f ? g : 0.f;
  }
  __syncthreads();
  // dZ1 = (dZ2 W2^T) * [A1 > 0]
  for (int i = tid; i < kSamples * kHidden; i += kThreads)
  {
    const int s = i / kHidden;
    const int h = i % kHidden;
    float g     = 0.f;
    for (int k = 0; k < kHidden; ++k)
    {
      g += s_dz2[s][k] * W[o.w2 + h * kHidden + k];
    }
    s_dz1[s][h] = s_a1[s][h] > 0.f ? g : 0.f;
  }
  __syncthreads();

  // Per-block gradient sums over the block's samples, in a fixed order.
  const int P    = param_count(out);
  float* part    = p.partial + size_t(blockIdx.x) * (P + 2);
  for (int i = tid; i < P; i += kThreads)
  {
    float g = 0.f;
    if (i < o.b1)
    {
      const int f = i / kHidden, h = i % kHidden;
      for (int s = 0; s < kSamples; ++s)
      {
        g += s_x[s][f] * s_dz1[s][h];
      }
    }
    else if (i < o.w2)
    {
      const int h = i - o.b1;
      for (int s = 0; s < kSamples; ++s)
      {
        g += s_dz1[s][h];
      }
    }
    else if (i < o.b2)
    {
      const int k = (i - o.w2) / kHidden, h = (i - o.w2) % kHidden;
      for (int s = 0; s < kSamples; ++s)
      {
        g += s_a1[s][k] * s_dz2[s][h];
      }
    }
    else if (i < o.w3)
    {
      const int h = i - o.b2;
      for (int s = 0; s < kSamples; ++s)
      {
        g += s_dz2[s][h];
      }
    }
    else if (i < o.b3)
    {
      const int k = (i - o.w3) / out, c = (i - o.w3) % out;
      for (int s = 0; s < kSamples; ++s)
      {
        g += s_a2[s][k] * s_dy[s][c];
      }
    }
    else
    {
      const int c = i - o.b3;
      for (int s = 0; s < kSamples; ++s)
      {
        g += s_dy[s][c];
      }
    }
    part[i] = g;
  }
  // block loss and mask sums: fixed-order tree
  for (int w = kThreads / 2; w > 0; w >>= 1)
  {
    if (tid < w)
    {
      s_red[0][tid] += s_red[0][tid + w];
      s_red[1][tid] += s_red[1][tid + w];
    }
    __syncthreads();
  }
  if (tid == 0)
  {
    part[P]     = s_red[0][0];
    part[P + 1] = s_red[1][0];
  }
}

struct UpdateParams
{
  float* params;
  float* adam_m;
  float* adam_v;
  float* grad; // reduced gradient of the mean loss (output)
  float* loss; // [1] mean loss (output, may be null)
  const float* partial;
  int blocks;
  int P;
  float lr;
  float beta1;
  float beta2;
  float eps;
  int step; // 1-based Adam step; 0 = compute the gradient only
};

// Sums the per-block partials in block order (deterministic), scales to the gradient of
// the mean loss, and applies one Adam step.
__global__ void __launch_bounds__(kThreads) mlp_update_kernel(UpdateParams u)
{
  // total mask weight, needed by every parameter: every thread sums it the same way
  float msum = 0.f;
  for (int b = 0; b < u.blocks; ++b)
  {
    msum += u.partial[size_t(b) * (u.P + 2) + u.P + 1];
  }
  const float scale = msum > 0.f ? 2.f / msum : 0.f;
  const int i       = blockIdx.x * kThreads + threadIdx.x;
  if (i == 0 && u.loss != nullptr)
  {
    float l = 0.f;
    for (int b = 0; b < u.blocks; ++b)
    {
      l += u.partial[size_t(b) * (u.P + 2) + u.P];
    }
    u.loss[0] = msum > 0.f ? l / msum : 0.f;
  }
  if (i >= u.P)
  {
    return;
  }
  float g = 0.f;
  for (int b = 0; b < u.blocks; ++b)
  {
    g += u.partial[size_t(b) * (u.P + 2) + i];
  }
  g *= scale;
  u.grad[i] = g;
  if (u.step > 0)
  {
    const float m  = u.beta1 * u.adam_m[i] + (1.f - u.beta1) * g;
    const float v  = u.beta2 * u.adam_v[i] + (1.f - u.beta2) * g * g;
    u.adam_m[i]    = m;
    u.adam_v[i]    = v;
    const float mh = m / (1.f - powf(u.beta1, float(u.step)));
    const float vh = v / (1.f - powf(u.beta2, float(u.step)));
    u.params[i] -= u.lr * mh / (sqrtf(vh) + u.eps);
  }
}

} // namespace mlp
} // namespace cccl_amd

using namespace cccl_amd::mlp;

extern "C" {

CCCL_C_API int cccl_amd_mlp_param_count(int out)
{
  return (out >= 1 && out <= kMaxOut) ? param_count(out) : -1;
}

CCCL_C_API size_t cccl_amd_mlp_workspace_bytes(int batch, int out)
{
  if (batch <= 0 || out < 1 || out > kMaxOut)
  {
    return 0;
  }
  const size_t blocks = size_t((batch + kSamples - 1) / kSamples);
  return blocks * size_t(param_count(out) + 2) * sizeof(float);
}

CCCL_C_API int cccl_amd_mlp_forward(const float* params, const float* x, int batch, int out, float* y, CUstream stream)
{
  if (params == nullptr || x == nullptr || y == nullptr || batch < 0 || out < 1 || out > kMaxOut)
  {
    return hipErrorInvalidValue;
  }
  if (batch == 0)
  {
    return hipSuccess;
  }
  StepParams p{params, x, nullptr, nullptr, y, nullptr, batch, out, 0};
  const int blocks = (batch + kSamples - 1) / kSamples;
  hipLaunchKernelGGL(mlp_step_kernel, dim3(blocks), dim3(kThreads), 0, reinterpret_cast<hipStream_t>(stream), p);
  return int(hipGetLastError());
}

CCCL_C_API int cccl_amd_mlp_train_step(
  float* params,
  float* adam_m,
  float* adam_v,
  float* grad,
  float* loss,
  float* workspace,
  size_t workspace_bytes,
  const float* x,
  const float* target,
  const float* mask,
  int batch,
  int out,
  float lr,
  int step,
  CUstream stream)
{
  if (params == nullptr || grad == nullptr || workspace == nullptr || x == nullptr || target == nullptr
      || mask == nullptr || batch <= 0 || out < 1 || out > kMaxOut || (step > 0 && (adam_m == nullptr || adam_v == nullptr)))
  {
    return hipErrorInvalidValue;
  }
  if (workspace_bytes < cccl_amd_mlp_workspace_bytes(batch, out))
  {
    return hipErrorInvalidValue;
  }
  hipStream_t s    = reinterpret_cast<hipStream_t>(stream);
  const int blocks = (batch + kSamples - 1) / kSamples;
  StepParams p{params, x, target, mask, nullptr, workspace, batch, out, 1};
  hipLaunchKernelGGL(mlp_step_kernel, dim3(blocks), dim3(kThreads), 0, s, p);
  if (hipError_t e = hipGetLastError(); e != hipSuccess)
  {
    return int(e);
  }
  const int P = param_count(out);
  UpdateParams u{params, adam_m, adam_v, grad, loss, workspace, blocks, P, lr, 0.9f, 0.999f, 1e-8f, step};
  hipLaunchKernelGGL(mlp_update_kernel, dim3((P + kThreads - 1) / kThreads), dim3(kThreads), 0, s, u);
  return int(hipGetLastError());
}

} // extern "C"

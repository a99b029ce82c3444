// Training and batch inference of the BiCG warm-start network on gfx950 (FP32 VALU kernels).
//
// The arithmetic of the forward pass is the network definition in bex_mlp.h (bfloat16 rounding
// of inputs, weights and hidden activations, FP32 accumulation); the backward pass treats each
// rounding as the identity (straight-through) and differentiates the mean squared error of the
// factor y against the target column. Gradients are reduced deterministically: each block
// writes its partial sums, one reduction kernel adds the blocks in a fixed order.
#include <hip/hip_runtime.h>
#include <math.h>

#include "bex_capi.h"
#include "bex_mlp.h"

namespace bexm {
namespace {  // internal linkage: only the extern "C" entry points are exported

constexpr int H = BEX_MLP_HIDDEN;
constexpr int TS = 64;     // samples per block (one wave)
constexpr int RS = H + 1;  // padded LDS row stride of the per-sample hidden vectors

struct Lds {
  float w1[6 * H];  // bf16-rounded
  float w2[H * H];  // bf16-rounded, [input][hidden]
  float b2[H];
  float w3[H];      // bf16-rounded
  float b3;
  float in[TS][6];
  float h1[TS][RS];
  float z1[TS][RS];  // z1, then dL/dz1
  float h2[TS][RS];
  float z2[TS][RS];  // z2, then dL/dz2
  float dy[TS];
  float e2[TS];
};

__device__ inline void load_weights(Lds& L, const float* p) {
  for (int k = threadIdx.x; k < 6 * H; k += blockDim.x) L.w1[k] = bex_bf16_round(p[BEX_MLP_OFF_W1 + k]);
  for (int k = threadIdx.x; k < H * H; k += blockDim.x) L.w2[k] = bex_bf16_round(p[BEX_MLP_OFF_W2 + k]);
  for (int k = threadIdx.x; k < H; k += blockDim.x) {
    L.b2[k] = p[BEX_MLP_OFF_B2 + k];
    L.w3[k] = bex_bf16_round(p[BEX_MLP_OFF_W3 + k]);
  }
  if (threadIdx.x == 0) L.b3 = p[BEX_MLP_OFF_B3];
}

// Forward pass of sample s (thread s of the block) into the block's LDS rows; returns y.
__device__ inline float forward(Lds& L, int s, const float* row) {
  float in[6];
#pragma unroll
  for (int k = 0; k < 5; k++) in[k] = bex_bf16_round(row[k]);
  in[5] = 1.0f;
#pragma unroll
  for (int k = 0; k < 6; k++) L.in[s][k] = in[k];
  for (int i = 0; i < H; i++) {
    float z = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; k++) z = fmaf(L.w1[k * H + i], in[k], z);
    L.z1[s][i] = z;
    L.h1[s][i] = bex_bf16_round(z > 0.0f ? z : 0.0f);
  }
  float y = L.b3;
  for (int i = 0; i < H; i++) {
    float z = L.b2[i];
    for (int k = 0; k < H; k++) z = fmaf(L.w2[k * H + i], L.h1[s][k], z);
    L.z2[s][i] = z;
    const float h = bex_bf16_round(z > 0.0f ? z : 0.0f);
    L.h2[s][i] = h;
    y = fmaf(L.w3[i], h, y);
  }
  return y;
}

__global__ __launch_bounds__(TS) void loss_grad_kernel(const float* __restrict__ rows, int stride,
                                                        long long n, const float* __restrict__ p,
                                                        float* __restrict__ partial) {
  extern __shared__ __align__(16) unsigned char smem[];
  Lds& L = *reinterpret_cast<Lds*>(smem);
  load_weights(L, p);
  __syncthreads();
  const int s = threadIdx.x;
  const long long g = (long long)blockIdx.x * TS + s;
  const float inv_n = 1.0f / (float)n;
  if (g < n) {
    const float* row = rows + g * stride;
    const float y = forward(L, s, row);
    const float e = y - row[5];
    L.e2[s] = e * e;
    const float dy = 2.0f * e * inv_n;
    L.dy[s] = dy;
    // dL/dz2 and dL/dz1 of this sample.
    for (int i = 0; i < H; i++) L.z2[s][i] = (L.z2[s][i] > 0.0f) ? dy * L.w3[i] : 0.0f;
    for (int k = 0; k < H; k++) {
      float d = 0.0f;
      for (int i = 0; i < H; i++) d = fmaf(L.w2[k * H + i], L.z2[s][i], d);
      L.z1[s][k] = (L.z1[s][k] > 0.0f) ? d : 0.0f;
    }
  } else {
    // Padding samples contribute nothing.
#pragma unroll
    for (int k = 0; k < 6; k++) L.in[s][k] = 0.0f;
    for (int i = 0; i < H; i++) L.h1[s][i] = L.z1[s][i] = L.h2[s][i] = L.z2[s][i] = 0.0f;
    L.dy[s] = 0.0f;
    L.e2[s] = 0.0f;
  }
  __syncthreads();
  // Block partial sums over the TS samples, fixed order.
  float* out = partial + (long long)blockIdx.x * (BEX_MLP_NPARAM + 1);
  for (int q = threadIdx.x; q < BEX_MLP_NPARAM + 1; q += TS) {
    float acc = 0.0f;
    if (q < BEX_MLP_OFF_W2) {  // W1[k][i]
      const int k = q / H, i = q % H;
      for (int t = 0; t < TS; t++) acc = fmaf(L.in[t][k], L.z1[t][i], acc);
    } else if (q < BEX_MLP_OFF_B2) {  // W2[k][i]
      const int k = (q - BEX_MLP_OFF_W2) / H, i = (q - BEX_MLP_OFF_W2) % H;
      for (int t = 0; t < TS; t++) acc = fmaf(L.h1[t][k], L.z2[t][i], acc);
    } else if (q < BEX_MLP_OFF_W3) {  // b2[i]
      const int i = q - BEX_MLP_OFF_B2;
      for (int t = 0; t < TS; t++) acc += L.z2[t][i];
    } else if (q < BEX_MLP_OFF_B3) {  // w3[i]
      const int i = q - BEX_MLP_OFF_W3;
      for (int t = 0; t < TS; t++) acc = fmaf(L.h2[t][i], L.dy[t], acc);
    } else if (q == BEX_MLP_OFF_B3) {
      for (int t = 0; t < TS; t++) acc += L.dy[t];
    } else {  // sum of squared errors
      for (int t = 0; t < TS; t++) acc += L.e2[t];
    }
    out[q] = acc;
  }
}

__global__ void reduce_kernel(const float* __restrict__ partial, int nblocks,
                              float* __restrict__ grad, float* __restrict__ loss, float inv_n) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > BEX_MLP_NPARAM) return;
  float acc = 0.0f;
  for (int b = 0; b < nblocks; b++) acc += partial[(long long)b * (BEX_MLP_NPARAM + 1) + q];
  if (q < BEX_MLP_NPARAM)
    grad[q] = acc;
  else
    loss[0] = acc * inv_n;
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, float lr, float b1,
                            float b2, float eps, float c1, float c2) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= BEX_MLP_NPARAM) return;
  const float gq = g[q];
  const float mq = b1 * m[q] + (1.0f - b1) * gq;
  const float vq = b2 * v[q] + (1.0f - b2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / c1) / (sqrtf(vq / c2) + eps);
}

__global__ __launch_bounds__(TS) void predict_kernel(const float* __restrict__ rows, int stride,
                                                      long long n, const float* __restrict__ p,
                                                      float* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem[];
  Lds& L = *reinterpret_cast<Lds*>(smem);
  load_weights(L, p);
  __syncthreads();
  const long long g = (long long)blockIdx.x * TS + threadIdx.x;
  if (g < n) out[g] = forward(L, threadIdx.x, rows + g * stride);
}

int launch_status() {
  return (hipGetLastError() == hipSuccess) ? BEX_OK : BEX_E_HIP;
}

// The per-block LDS (~87 KB) is above the default dynamic limit; gfx950 provides 160 KB.
int allow_lds() {
  if (hipFuncSetAttribute(reinterpret_cast<const void*>(&loss_grad_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, sizeof(Lds)) != hipSuccess ||
      hipFuncSetAttribute(reinterpret_cast<const void*>(&predict_kernel),
                          hipFuncAttributeMaxDynamicSharedMemorySize, sizeof(Lds)) != hipSuccess)
    return BEX_E_HIP;
  return BEX_OK;
}

}  // namespace
}  // namespace bexm

extern "C" {

long long bex_mlp_workspace_floats(long long n) {
  if (n <= 0) return 0;
  return ((n + bexm::TS - 1) / bexm::TS) * (long long)(BEX_MLP_NPARAM + 1);
}

int bex_mlp_loss_grad(const float* rows, int stride, long long n, const float* params,
                      float* grad, float* loss, float* workspace, void* stream) {
  if (!rows || !params || !grad || !loss || !workspace || n <= 0 || stride < 6)
    return BEX_E_ARGUMENT;
  const long long nb = (n + bexm::TS - 1) / bexm::TS;
  if (nb > 0x7fffffff) return BEX_E_UNSUPPORTED;
  if (bexm::allow_lds() != BEX_OK) return BEX_E_HIP;
  hipStream_t st = reinterpret_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(bexm::loss_grad_kernel, dim3((unsigned)nb), dim3(bexm::TS),
                     sizeof(bexm::Lds), st, rows, stride, n, params, workspace);
  hipLaunchKernelGGL(bexm::reduce_kernel, dim3((BEX_MLP_NPARAM + 1 + 255) / 256), dim3(256), 0,
                     st, workspace, (int)nb, grad, loss, 1.0f / (float)n);
  return bexm::launch_status();
}

int bex_mlp_adam(float* params, const float* grad, float* m, float* v, int step, float lr,
                 float beta1, float beta2, float eps, void* stream) {
  if (!params || !grad || !m || !v || step < 1) return BEX_E_ARGUMENT;
  const float c1 = 1.0f - powf(beta1, (float)step);
  const float c2 = 1.0f - powf(beta2, (float)step);
  hipLaunchKernelGGL(bexm::adam_kernel, dim3((BEX_MLP_NPARAM + 255) / 256), dim3(256), 0,
                     reinterpret_cast<hipStream_t>(stream), params, grad, m, v, lr, beta1, beta2,
                     eps, c1, c2);
  return bexm::launch_status();
}

int bex_mlp_predict(const float* rows, int stride, long long n, const float* params, float* out,
                    void* stream) {
  if (!rows || !params || !out || n <= 0 || stride < 5) return BEX_E_ARGUMENT;
  const long long nb = (n + bexm::TS - 1) / bexm::TS;
  if (nb > 0x7fffffff) return BEX_E_UNSUPPORTED;
  if (bexm::allow_lds() != BEX_OK) return BEX_E_HIP;
  hipLaunchKernelGGL(bexm::predict_kernel, dim3((unsigned)nb), dim3(bexm::TS), sizeof(bexm::Lds),
                     reinterpret_cast<hipStream_t>(stream), rows, stride, n, params, out);
  return bexm::launch_status();
}

}  // extern "C"

// FP32 MLP of the learned direction / phase policy: 12 -> 32 -> 32 -> 3, ReLU, softmax
// cross-entropy.  Training (forward, backward, Adam) and batched inference kernels, and the
// C ABI entry points that drive them.  The same weights are evaluated inside the level
// closer (msc_kernels.hip: mlp_eval) with identical arithmetic order per output.
//
// Layout of a blob (kMlpBlob floats): W1[32][12] b1[32] W2[32][32] b2[32] W3[3][32] b3[3]
// mean[12] std[12].  Inputs are standardised as (x - mean) / std before layer 1.
//
// A training workgroup is one wave and 64 samples: each lane runs its sample's forward and
// backward pass in registers, writes activations and deltas to LDS, then the wave reduces
// the per-parameter sums over the 64 samples and adds them to the gradient with one atomic
// per parameter.  The matrices are far too small for MFMA tiles to pay.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <numeric>
#include <random>
#include <vector>

#include "msc_internal.h"

namespace msc {
namespace {

constexpr int kS = 64;  // samples per workgroup
constexpr int oW1 = 0, ob1 = oW1 + kMlpH * kMlpIn, oW2 = ob1 + kMlpH, ob2 = oW2 + kMlpH * kMlpH,
              oW3 = ob2 + kMlpH, ob3 = oW3 + kMlpOut * kMlpH, oMean = kMlpParams, oStd = oMean + kMlpIn;
static_assert(ob3 + kMlpOut == kMlpParams, "blob layout");

__device__ __forceinline__ void load_blob(float* sW, const float* blob) {
  for (int i = threadIdx.x; i < kMlpBlob; i += blockDim.x) sW[i] = blob[i];
  __syncthreads();
}

// Forward pass of one sample from LDS weights; returns logits, keeps activations.
__device__ __forceinline__ void forward(const float* sW, const float* xin, float* x, float* h1, float* h2, float* z) {
#pragma unroll
  for (int i = 0; i < kMlpIn; ++i) x[i] = (xin[i] - sW[oMean + i]) / sW[oStd + i];
#pragma unroll 4
  for (int o = 0; o < kMlpH; ++o) {
    float acc = sW[ob1 + o];
#pragma unroll
    for (int i = 0; i < kMlpIn; ++i) acc += sW[oW1 + o * kMlpIn + i] * x[i];
    h1[o] = fmaxf(acc, 0.0f);
  }
#pragma unroll 4
  for (int o = 0; o < kMlpH; ++o) {
    float acc = sW[ob2 + o];
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) acc += sW[oW2 + o * kMlpH + i] * h1[i];
    h2[o] = fmaxf(acc, 0.0f);
  }
#pragma unroll
  for (int o = 0; o < kMlpOut; ++o) {
    float acc = sW[ob3 + o];
#pragma unroll
    for (int i = 0; i < kMlpH; ++i) acc += sW[oW3 + o * kMlpH + i] * h2[i];
    z[o] = acc;
  }
}

__global__ void __launch_bounds__(kS) k_mlp_forward(const float* __restrict__ blob, const float* __restrict__ X,
                                                    int64_t count, float* __restrict__ logits) {
  __shared__ float sW[kMlpBlob];
  load_blob(sW, blob);
  int64_t s = (int64_t)blockIdx.x * kS + threadIdx.x;
  if (s >= count) return;
  float xin[kMlpIn], x[kMlpIn], h1[kMlpH], h2[kMlpH], z[kMlpOut];
  for (int i = 0; i < kMlpIn; ++i) xin[i] = X[s * kMlpIn + i];
  forward(sW, xin, x, h1, h2, z);
  for (int o = 0; o < kMlpOut; ++o) logits[s * kMlpOut + o] = z[o];
}

// Gradient of sum_b w_b * CE over samples idx[first .. first+count) (idx null: identity).
// acc[0] += loss sum, acc[1] += weight sum.
__global__ void __launch_bounds__(kS) k_mlp_grad(const float* __restrict__ blob, const float* __restrict__ X,
                                                 const int32_t* __restrict__ Y, const float* __restrict__ Wt,
                                                 const int32_t* __restrict__ idx, int64_t first, int64_t count,
                                                 float* __restrict__ grad, float* __restrict__ acc) {
  __shared__ float sW[kMlpBlob];
  __shared__ float sx[kS][kMlpIn + 1];
  __shared__ float sh1[kS][kMlpH + 1];
  __shared__ float sh2[kS][kMlpH + 1];
  __shared__ float sd1[kS][kMlpH + 1];
  __shared__ float sd2[kS][kMlpH + 1];
  __shared__ float sdz[kS][kMlpOut + 1];
  load_blob(sW, blob);
  const int l = threadIdx.x;
  const int64_t s = (int64_t)blockIdx.x * kS + l;
  const bool ok = s < count;
  float loss = 0.0f, wgt = 0.0f;
  float x[kMlpIn], h1[kMlpH], h2[kMlpH], z[kMlpOut], dz[kMlpOut], d2[kMlpH], d1[kMlpH];
  if (ok) {
    int64_t b = idx ? (int64_t)idx[first + s] : first + s;
    float xin[kMlpIn];
    for (int i = 0; i < kMlpIn; ++i) xin[i] = X[b * kMlpIn + i];
    forward(sW, xin, x, h1, h2, z);
    int y = Y[b];
    wgt = Wt ? Wt[b] : 1.0f;
    float zmax = fmaxf(z[0], fmaxf(z[1], z[2]));
    float e[kMlpOut], se = 0.0f;
    for (int o = 0; o < kMlpOut; ++o) {
      e[o] = expf(z[o] - zmax);
      se += e[o];
    }
    loss = wgt * (logf(se) - (z[y] - zmax));
    for (int o = 0; o < kMlpOut; ++o) dz[o] = wgt * (e[o] / se - (o == y ? 1.0f : 0.0f));
#pragma unroll 4
    for (int i = 0; i < kMlpH; ++i) {
      float a = 0.0f;
      for (int o = 0; o < kMlpOut; ++o) a += sW[oW3 + o * kMlpH + i] * dz[o];
      d2[i] = h2[i] > 0.0f ? a : 0.0f;
    }
#pragma unroll 4
    for (int i = 0; i < kMlpH; ++i) {
      float a = 0.0f;
      for (int o = 0; o < kMlpH; ++o) a += sW[oW2 + o * kMlpH + i] * d2[o];
      d1[i] = h1[i] > 0.0f ? a : 0.0f;
    }
  } else {
    for (int i = 0; i < kMlpIn; ++i) x[i] = 0.0f;
    for (int i = 0; i < kMlpH; ++i) h1[i] = h2[i] = d1[i] = d2[i] = 0.0f;
    for (int o = 0; o < kMlpOut; ++o) dz[o] = 0.0f;
  }
  for (int i = 0; i < kMlpIn; ++i) sx[l][i] = x[i];
  for (int i = 0; i < kMlpH; ++i) {
    sh1[l][i] = h1[i];
    sh2[l][i] = h2[i];
    sd1[l][i] = d1[i];
    sd2[l][i] = d2[i];
  }
  for (int o = 0; o < kMlpOut; ++o) sdz[l][o] = dz[o];
  // wave sums of loss and weight
  for (int o = 32; o > 0; o >>= 1) {
    loss += __shfl_xor(loss, o, 64);
    wgt += __shfl_xor(wgt, o, 64);
  }
  if (l == 0) {
    atomicAdd(&acc[0], loss);
    atomicAdd(&acc[1], wgt);
  }
  __syncthreads();
  for (int p = l; p < kMlpParams; p += kS) {
    float a = 0.0f;
    if (p < ob1) {
      int o = p / kMlpIn, i = p % kMlpIn;
      for (int t = 0; t < kS; ++t) a += sd1[t][o] * sx[t][i];
    } else if (p < oW2) {
      int o = p - ob1;
      for (int t = 0; t < kS; ++t) a += sd1[t][o];
    } else if (p < ob2) {
      int o = (p - oW2) / kMlpH, i = (p - oW2) % kMlpH;
      for (int t = 0; t < kS; ++t) a += sd2[t][o] * sh1[t][i];
    } else if (p < oW3) {
      int o = p - ob2;
      for (int t = 0; t < kS; ++t) a += sd2[t][o];
    } else if (p < ob3) {
      int o = (p - oW3) / kMlpH, i = (p - oW3) % kMlpH;
      for (int t = 0; t < kS; ++t) a += sdz[t][o] * sh2[t][i];
    } else {
      int o = p - ob3;
      for (int t = 0; t < kS; ++t) a += sdz[t][o];
    }
    atomicAdd(&grad[p], a);
  }
}

// Adam on the weights (first kMlpParams floats of the blob) with the gradient of the
// weighted-mean loss (grad / acc[1]).  Moves the batch loss into the epoch accumulators.
__global__ void k_adam(float* __restrict__ blob, float* __restrict__ grad, float* __restrict__ m,
                       float* __restrict__ v, float* __restrict__ acc, float lr, float b1, float b2, float eps,
                       float wd, float c1, float c2) {
  const float wsum = acc[1];
  const float inv = wsum > 0.0f ? 1.0f / wsum : 0.0f;
  for (int p = threadIdx.x; p < kMlpParams; p += blockDim.x) {
    float g = grad[p] * inv;
    float mm = b1 * m[p] + (1.0f - b1) * g;
    float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    float step = (mm / c1) / (sqrtf(vv / c2) + eps) + wd * blob[p];
    blob[p] -= lr * step;
    grad[p] = 0.0f;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    acc[2] += acc[0];
    acc[3] += acc[1];
    acc[0] = 0.0f;
    acc[1] = 0.0f;
  }
}

__global__ void k_zero_f32(float* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0.0f;
}

struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes ? bytes : 4); }
};

}  // namespace
}  // namespace msc

using namespace msc;

// ---- C ABI ----------------------------------------------------------------------------

extern "C" {

#define MLP_TRY(expr)                                            \
  do {                                                             \
    hipError_t _e = (expr);                                        \
    if (_e != hipSuccess) {                                        \
      (void)hipGetLastError();                                     \
      return _e == hipErrorOutOfMemory ? MSC_E_OOM : MSC_E_DEVICE; \
    }                                                              \
  } while (0)

static bool blob_ok(const float* blob) {
  for (int i = 0; i < kMlpIn; ++i)
    if (!(blob[kMlpParams + kMlpIn + i] > 0.0f)) return false;  // std > 0
  for (int i = 0; i < kMlpBlob; ++i)
    if (!std::isfinite(blob[i])) return false;
  return true;
}

int msc_policy_set_weights(msc_ctx* ctx, const float* blob, int32_t count) {
  if (!ctx || !blob || (count != kMlpBlob && count != 2 * kMlpBlob)) return MSC_E_INVALID;
  const int nets = count / kMlpBlob;
  for (int i = 0; i < nets; ++i)
    if (!blob_ok(blob + (size_t)i * kMlpBlob)) return MSC_E_INVALID;
  return msc::ctx_set_policy(ctx, blob, nets);
}

int msc_mlp_forward(msc_ctx* ctx, const float* blob, const float* x, int64_t count, float* logits_out) {
  if (!ctx || !blob || count < 0 || (count > 0 && (!x || !logits_out)) || !blob_ok(blob)) return MSC_E_INVALID;
  if (count == 0) return MSC_OK;
  hipStream_t s = nullptr;
  int device = 0;
  int rc = msc::ctx_stream(ctx, &s, &device);
  if (rc) return rc;
  MLP_TRY(hipSetDevice(device));
  DevBuf dB, dX, dZ;
  MLP_TRY(dB.alloc(sizeof(float) * kMlpBlob));
  MLP_TRY(dX.alloc(sizeof(float) * kMlpIn * count));
  MLP_TRY(dZ.alloc(sizeof(float) * kMlpOut * count));
  MLP_TRY(hipMemcpyAsync(dB.p, blob, sizeof(float) * kMlpBlob, hipMemcpyHostToDevice, s));
  MLP_TRY(hipMemcpyAsync(dX.p, x, sizeof(float) * kMlpIn * count, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_mlp_forward, dim3((unsigned)((count + kS - 1) / kS)), dim3(kS), 0, s, (const float*)dB.p,
                     (const float*)dX.p, count, (float*)dZ.p);
  MLP_TRY(hipGetLastError());
  MLP_TRY(hipMemcpyAsync(logits_out, dZ.p, sizeof(float) * kMlpOut * count, hipMemcpyDeviceToHost, s));
  MLP_TRY(hipStreamSynchronize(s));
  return MSC_OK;
}

static int check_labels(const int32_t* y, int64_t count) {
  for (int64_t i = 0; i < count; ++i)
    if (y[i] < 0 || y[i] >= kMlpOut) return MSC_E_INVALID;
  return MSC_OK;
}

int msc_mlp_grad(msc_ctx* ctx, const float* blob, const float* x, const int32_t* y, const float* w, int64_t count,
                 float* loss_out, float* grad_out) {
  if (!ctx || !blob || count < 0 || (count > 0 && (!x || !y)) || !grad_out || !blob_ok(blob)) return MSC_E_INVALID;
  if (check_labels(y, count)) return MSC_E_INVALID;
  hipStream_t s = nullptr;
  int device = 0;
  int rc = msc::ctx_stream(ctx, &s, &device);
  if (rc) return rc;
  MLP_TRY(hipSetDevice(device));
  DevBuf dB, dX, dY, dW, dG, dA;
  MLP_TRY(dB.alloc(sizeof(float) * kMlpBlob));
  MLP_TRY(dX.alloc(sizeof(float) * kMlpIn * count));
  MLP_TRY(dY.alloc(sizeof(int32_t) * count));
  MLP_TRY(dW.alloc(sizeof(float) * count));
  MLP_TRY(dG.alloc(sizeof(float) * kMlpParams));
  MLP_TRY(dA.alloc(sizeof(float) * 4));
  MLP_TRY(hipMemcpyAsync(dB.p, blob, sizeof(float) * kMlpBlob, hipMemcpyHostToDevice, s));
  if (count) {
    MLP_TRY(hipMemcpyAsync(dX.p, x, sizeof(float) * kMlpIn * count, hipMemcpyHostToDevice, s));
    MLP_TRY(hipMemcpyAsync(dY.p, y, sizeof(int32_t) * count, hipMemcpyHostToDevice, s));
    if (w) MLP_TRY(hipMemcpyAsync(dW.p, w, sizeof(float) * count, hipMemcpyHostToDevice, s));
  }
  hipLaunchKernelGGL(k_zero_f32, dim3(8), dim3(256), 0, s, (float*)dG.p, (int64_t)kMlpParams);
  hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, s, (float*)dA.p, (int64_t)4);
  if (count)
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)((count + kS - 1) / kS)), dim3(kS), 0, s, (const float*)dB.p,
                       (const float*)dX.p, (const int32_t*)dY.p, w ? (const float*)dW.p : nullptr,
                       (const int32_t*)nullptr, (int64_t)0, count, (float*)dG.p, (float*)dA.p);
  MLP_TRY(hipGetLastError());
  float acc[4];
  MLP_TRY(hipMemcpyAsync(grad_out, dG.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
  MLP_TRY(hipMemcpyAsync(acc, dA.p, sizeof(acc), hipMemcpyDeviceToHost, s));
  MLP_TRY(hipStreamSynchronize(s));
  if (loss_out) *loss_out = acc[0];
  return MSC_OK;
}

int msc_mlp_train(msc_ctx* ctx, float* blob, const float* x, const int32_t* y, const float* w, int64_t count,
                  const msc_train_opts* o, float* loss_curve_out) {
  if (!ctx || !blob || !o || count <= 0 || !x || !y || o->epochs < 1 || o->batch < 1 || !(o->lr > 0.0f) ||
      count > INT32_MAX || !blob_ok(blob))
    return MSC_E_INVALID;
  if (check_labels(y, count)) return MSC_E_INVALID;
  hipStream_t s = nullptr;
  int device = 0;
  int rc = msc::ctx_stream(ctx, &s, &device);
  if (rc) return rc;
  MLP_TRY(hipSetDevice(device));
  DevBuf dB, dX, dY, dW, dG, dM, dV, dA, dI;
  MLP_TRY(dB.alloc(sizeof(float) * kMlpBlob));
  MLP_TRY(dX.alloc(sizeof(float) * kMlpIn * count));
  MLP_TRY(dY.alloc(sizeof(int32_t) * count));
  MLP_TRY(dW.alloc(sizeof(float) * count));
  MLP_TRY(dG.alloc(sizeof(float) * kMlpParams));
  MLP_TRY(dM.alloc(sizeof(float) * kMlpParams));
  MLP_TRY(dV.alloc(sizeof(float) * kMlpParams));
  MLP_TRY(dA.alloc(sizeof(float) * 4));
  MLP_TRY(dI.alloc(sizeof(int32_t) * count));
  MLP_TRY(hipMemcpyAsync(dB.p, blob, sizeof(float) * kMlpBlob, hipMemcpyHostToDevice, s));
  MLP_TRY(hipMemcpyAsync(dX.p, x, sizeof(float) * kMlpIn * count, hipMemcpyHostToDevice, s));
  MLP_TRY(hipMemcpyAsync(dY.p, y, sizeof(int32_t) * count, hipMemcpyHostToDevice, s));
  if (w) MLP_TRY(hipMemcpyAsync(dW.p, w, sizeof(float) * count, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_zero_f32, dim3(8), dim3(256), 0, s, (float*)dG.p, (int64_t)kMlpParams);
  hipLaunchKernelGGL(k_zero_f32, dim3(8), dim3(256), 0, s, (float*)dM.p, (int64_t)kMlpParams);
  hipLaunchKernelGGL(k_zero_f32, dim3(8), dim3(256), 0, s, (float*)dV.p, (int64_t)kMlpParams);
  hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, s, (float*)dA.p, (int64_t)4);
  std::vector<int32_t> perm((size_t)count);
  std::iota(perm.begin(), perm.end(), 0);
  std::mt19937_64 gen(o->seed);
  int64_t step = 0;
  for (int ep = 0; ep < o->epochs; ++ep) {
    std::shuffle(perm.begin(), perm.end(), gen);
    MLP_TRY(hipMemcpyAsync(dI.p, perm.data(), sizeof(int32_t) * count, hipMemcpyHostToDevice, s));
    for (int64_t first = 0; first < count; first += o->batch) {
      int64_t nb = std::min<int64_t>(o->batch, count - first);
      ++step;
      hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)((nb + kS - 1) / kS)), dim3(kS), 0, s, (const float*)dB.p,
                         (const float*)dX.p, (const int32_t*)dY.p, w ? (const float*)dW.p : nullptr,
                         (const int32_t*)dI.p, first, nb, (float*)dG.p, (float*)dA.p);
      float c1 = 1.0f - std::pow(o->beta1, (float)step), c2 = 1.0f - std::pow(o->beta2, (float)step);
      hipLaunchKernelGGL(k_adam, dim3(1), dim3(256), 0, s, (float*)dB.p, (float*)dG.p, (float*)dM.p, (float*)dV.p,
                         (float*)dA.p, o->lr, o->beta1, o->beta2, o->eps, o->weight_decay, c1, c2);
    }
    MLP_TRY(hipGetLastError());
    float acc[4];
    MLP_TRY(hipMemcpyAsync(acc, dA.p, sizeof(acc), hipMemcpyDeviceToHost, s));
    MLP_TRY(hipStreamSynchronize(s));
    if (loss_curve_out) loss_curve_out[ep] = acc[3] > 0.0f ? acc[2] / acc[3] : 0.0f;
    hipLaunchKernelGGL(k_zero_f32, dim3(1), dim3(64), 0, s, (float*)dA.p, (int64_t)4);
  }
  MLP_TRY(hipMemcpyAsync(blob, dB.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, s));
  MLP_TRY(hipStreamSynchronize(s));
  return MSC_OK;
}

}  // extern "C"

// The learned warm start of stage two: a 12 -> 64 -> 64 -> 64 -> 2 GELU MLP
// mapping per-gene summary statistics to (log sigma^2, log phi).
//
// Parameters are one float blob (nh_mlp_param_count values):
//   feature mean[12], feature sd[12],
//   W1[64x12] b1[64]  W2[64x64] b2[64]  W3[64x64] b3[64]  W4[2x64] b4[2],
//   target mean[2], target sd[2]
// W is row-major (out x in).  The network predicts the correction to stage
// one's estimate: its target is (log sigma^2, log phi) minus features 10 and 11
// (stage one's log sigma^2 and log phi), standardised with the stored target
// mean and sd; features are standardised likewise.  Inference adds features
// 10 and 11 back.  The training loss is Huber (delta 1) on the standardised
// targets, so the genes whose stage-one estimate is far off do not dominate.
//
// Training runs on the device in FP32 with this file's kernels: dense forward
// (pre-activations kept), MSE loss and its gradient, backward through each
// layer (input gradients and weight gradients as separate kernels, each a
// fixed-order sum so a step is deterministic), Adam.  Inference converts the
// weights and activations to BF16 and runs every layer on the matrix cores
// (v_mfma_f32_16x16x16bf16_1k, FP32 accumulation), one wave per 16 genes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "../../include/nebula_hip.h"
#include "mlp.hpp"
#include "resident.hpp"

namespace nh {

namespace {

__device__ __forceinline__ float gelu(float x) {
  const float u = 0.7978845608f * (x + 0.044715f * x * x * x);
  return 0.5f * x * (1.0f + tanhf(u));
}

__device__ __forceinline__ float gelu_grad(float x) {
  const float u = 0.7978845608f * (x + 0.044715f * x * x * x);
  const float th = tanhf(u);
  return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * 0.7978845608f *
                                  (1.0f + 3.0f * 0.044715f * x * x);
}

// ---------------------------------------------------------------------------
// FP32 training kernels.  B rows per batch.

// in[b][i] = (feat[idx[b]][i] - mean[i]) / sd[i]; tgt likewise.
__global__ void mlp_gather_kernel(const float* __restrict__ feat, const float* __restrict__ tgt,
                                  const int* __restrict__ idx, int B,
                                  const float* __restrict__ norm, float* __restrict__ in,
                                  float* __restrict__ t) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B * (kMlpIn + kMlpOut)) return;
  const int b = g / (kMlpIn + kMlpOut), c = g % (kMlpIn + kMlpOut);
  const int r = idx[b];
  if (c < kMlpIn) {
    in[b * kMlpIn + c] = (feat[(size_t)r * kMlpIn + c] - norm[c]) / norm[kMlpIn + c];
  } else {
    const int j = c - kMlpIn;
    const float* tn = norm + kMlpNormFeat + kMlpNet;
    const float res = tgt[(size_t)r * kMlpOut + j] - feat[(size_t)r * kMlpIn + kMlpBase + j];
    t[b * kMlpOut + j] = (res - tn[j]) / tn[kMlpOut + j];
  }
}

// z = W in + b; a = gelu(z) (act) or z.
__global__ void mlp_dense_fwd_kernel(const float* __restrict__ in, int I,
                                     const float* __restrict__ W, const float* __restrict__ bias,
                                     int O, int B, int act, float* __restrict__ z,
                                     float* __restrict__ a) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B * O) return;
  const int b = g / O, o = g % O;
  float s = bias[o];
  for (int i = 0; i < I; ++i) s = fmaf(W[o * I + i], in[b * I + i], s);
  z[g] = s;
  a[g] = act ? gelu(s) : s;
}

// Huber loss (delta 1) per row over B rows and its gradient d = dL/dout
__global__ void mlp_loss_kernel(const float* __restrict__ out, const float* __restrict__ t, int B,
                                float* __restrict__ d, float* __restrict__ loss_row) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  float l = 0.0f;
  for (int j = 0; j < kMlpOut; ++j) {
    const float e = out[b * kMlpOut + j] - t[b * kMlpOut + j];
    const float ae = fabsf(e);
    d[b * kMlpOut + j] = fminf(fmaxf(e, -1.0f), 1.0f) / (float)B;
    l += ae <= 1.0f ? 0.5f * e * e : ae - 0.5f;
  }
  loss_row[b] = l / (float)B;
}

// one thread adds the rows in order
__global__ void mlp_sum_kernel(const float* __restrict__ v, int n, float* __restrict__ out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  float s = 0.0f;
  for (int i = 0; i < n; ++i) s += v[i];
  out[0] = s;
}

// dz_prev[b][i] = (sum_o dz[b][o] W[o][i]) * gelu'(z_prev[b][i])
__global__ void mlp_dense_bwd_in_kernel(const float* __restrict__ dz, int O,
                                        const float* __restrict__ W, int I, int B,
                                        const float* __restrict__ z_prev,
                                        float* __restrict__ dz_prev) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B * I) return;
  const int b = g / I, i = g % I;
  float s = 0.0f;
  for (int o = 0; o < O; ++o) s = fmaf(dz[b * O + o], W[o * I + i], s);
  dz_prev[g] = s * gelu_grad(z_prev[g]);
}

// gW[o][i] = sum_b dz[b][o] in[b][i]; gb[o] = sum_b dz[b][o]
__global__ void mlp_dense_bwd_w_kernel(const float* __restrict__ dz, int O,
                                       const float* __restrict__ in, int I, int B,
                                       float* __restrict__ gW, float* __restrict__ gb) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= O * I + O) return;
  float s = 0.0f;
  if (g < O * I) {
    const int o = g / I, i = g % I;
    for (int b = 0; b < B; ++b) s = fmaf(dz[b * O + o], in[b * I + i], s);
    gW[g] = s;
  } else {
    const int o = g - O * I;
    for (int b = 0; b < B; ++b) s += dz[b * O + o];
    gb[o] = s;
  }
}

__global__ void mlp_adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                                float* __restrict__ m, float* __restrict__ v, int n, float lr,
                                float b1, float b2, float eps, float bc1, float bc2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

inline unsigned blocks(size_t n, unsigned t) { return (unsigned)((n + t - 1) / t); }

// Device state of one forward/backward pass over up to `cap` rows.
struct MlpPass {
  DevBuf<float> in, t, z1, a1, z2, a2, z3, a3, out, d4, d3, d2, d1, loss_row, loss;
  size_t cap = 0;
  void reserve(size_t B) {
    if (B <= cap) return;
    in.alloc(B * kMlpIn);
    t.alloc(B * kMlpOut);
    z1.alloc(B * kMlpHidden);
    a1.alloc(B * kMlpHidden);
    z2.alloc(B * kMlpHidden);
    a2.alloc(B * kMlpHidden);
    z3.alloc(B * kMlpHidden);
    a3.alloc(B * kMlpHidden);
    out.alloc(B * kMlpOut);
    d4.alloc(B * kMlpOut);
    d3.alloc(B * kMlpHidden);
    d2.alloc(B * kMlpHidden);
    d1.alloc(B * kMlpHidden);
    loss_row.alloc(B);
    loss.alloc(1);
    cap = B;
  }
};

// Offsets of the trainable blocks inside the net part (after the feature norm).
struct NetLayout {
  static constexpr size_t W1 = 0, b1 = W1 + kMlpHidden * kMlpIn;
  static constexpr size_t W2 = b1 + kMlpHidden, b2 = W2 + kMlpHidden * kMlpHidden;
  static constexpr size_t W3 = b2 + kMlpHidden, b3 = W3 + kMlpHidden * kMlpHidden;
  static constexpr size_t W4 = b3 + kMlpHidden, b4 = W4 + kMlpOut * kMlpHidden;
  static constexpr size_t end = b4 + kMlpOut;
};
static_assert(NetLayout::end == kMlpNet, "layout");

// Forward and (optionally) backward on the rows already in P.in / P.t.
void pass(MlpPass& P, const float* net, float* grad, int B, hipStream_t st) {
  const unsigned T = 256;
  using L = NetLayout;
  hipLaunchKernelGGL(mlp_dense_fwd_kernel, dim3(blocks((size_t)B * kMlpHidden, T)), dim3(T), 0,
                     st, P.in.p, kMlpIn, net + L::W1, net + L::b1, kMlpHidden, B, 1, P.z1.p,
                     P.a1.p);
  hipLaunchKernelGGL(mlp_dense_fwd_kernel, dim3(blocks((size_t)B * kMlpHidden, T)), dim3(T), 0,
                     st, P.a1.p, kMlpHidden, net + L::W2, net + L::b2, kMlpHidden, B, 1, P.z2.p,
                     P.a2.p);
  hipLaunchKernelGGL(mlp_dense_fwd_kernel, dim3(blocks((size_t)B * kMlpHidden, T)), dim3(T), 0,
                     st, P.a2.p, kMlpHidden, net + L::W3, net + L::b3, kMlpHidden, B, 1, P.z3.p,
                     P.a3.p);
  hipLaunchKernelGGL(mlp_dense_fwd_kernel, dim3(blocks((size_t)B * kMlpOut, T)), dim3(T), 0, st,
                     P.a3.p, kMlpHidden, net + L::W4, net + L::b4, kMlpOut, B, 0, P.out.p,
                     P.out.p);
  hipLaunchKernelGGL(mlp_loss_kernel, dim3(blocks(B, T)), dim3(T), 0, st, P.out.p, P.t.p, B,
                     P.d4.p, P.loss_row.p);
  hipLaunchKernelGGL(mlp_sum_kernel, dim3(1), dim3(64), 0, st, P.loss_row.p, B, P.loss.p);
  if (!grad) return;
  hipLaunchKernelGGL(mlp_dense_bwd_w_kernel, dim3(blocks(kMlpOut * kMlpHidden + kMlpOut, T)),
                     dim3(T), 0, st, P.d4.p, kMlpOut, P.a3.p, kMlpHidden, B, grad + L::W4,
                     grad + L::b4);
  hipLaunchKernelGGL(mlp_dense_bwd_in_kernel, dim3(blocks((size_t)B * kMlpHidden, T)), dim3(T),
                     0, st, P.d4.p, kMlpOut, net + L::W4, kMlpHidden, B, P.z3.p, P.d3.p);
  hipLaunchKernelGGL(mlp_dense_bwd_w_kernel,
                     dim3(blocks(kMlpHidden * kMlpHidden + kMlpHidden, T)), dim3(T), 0, st,
                     P.d3.p, kMlpHidden, P.a2.p, kMlpHidden, B, grad + L::W3, grad + L::b3);
  hipLaunchKernelGGL(mlp_dense_bwd_in_kernel, dim3(blocks((size_t)B * kMlpHidden, T)), dim3(T),
                     0, st, P.d3.p, kMlpHidden, net + L::W3, kMlpHidden, B, P.z2.p, P.d2.p);
  hipLaunchKernelGGL(mlp_dense_bwd_w_kernel,
                     dim3(blocks(kMlpHidden * kMlpHidden + kMlpHidden, T)), dim3(T), 0, st,
                     P.d2.p, kMlpHidden, P.a1.p, kMlpHidden, B, grad + L::W2, grad + L::b2);
  hipLaunchKernelGGL(mlp_dense_bwd_in_kernel, dim3(blocks((size_t)B * kMlpHidden, T)), dim3(T),
                     0, st, P.d2.p, kMlpHidden, net + L::W2, kMlpHidden, B, P.z1.p, P.d1.p);
  hipLaunchKernelGGL(mlp_dense_bwd_w_kernel, dim3(blocks(kMlpHidden * kMlpIn + kMlpHidden, T)),
                     dim3(T), 0, st, P.d1.p, kMlpHidden, P.in.p, kMlpIn, B, grad + L::W1,
                     grad + L::b1);
}

// ---------------------------------------------------------------------------
// BF16 inference on the matrix cores.

typedef short short4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (short)(u >> 16);
}

constexpr int kPredWaves = 4;
constexpr int kK1 = 16;  // layer-1 inputs padded to one MFMA k-block

struct PredLds {
  short w1[kMlpHidden * kK1];             // [o][k], k >= 12 zero
  short w2[kMlpHidden * kMlpHidden];
  short w3[kMlpHidden * kMlpHidden];
  short w4[16 * kMlpHidden];              // rows >= 2 zero
  float b1[kMlpHidden], b2[kMlpHidden], b3[kMlpHidden], b4[16];
  float act[kPredWaves][2][16][kMlpHidden];
};

// One layer for the wave's 16 rows: out[16][16 t + j] = f(sum_k in[i][k] W[16t+j][k] + b).
// A fragment: lane l holds in[l % 16][kb + 4 (l / 16) + e]; B fragment:
// W[16 t + l % 16][kb + 4 (l / 16) + e]; D: lane l holds rows 4 (l / 16) + e of
// column l % 16.
template <int K, int TILES, bool ACT>
__device__ void mfma_layer(const float (*in)[kMlpHidden], const short* W, const float* bias,
                           float (*out)[kMlpHidden], int lane) {
  const int r = lane % 16, q = lane / 16;
  for (int t = 0; t < TILES; ++t) {
    float4_t d = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int kb = 0; kb < K; kb += 16) {
      short4_t a, b;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        a[e] = f2bf(in[r][kb + 4 * q + e]);
        b[e] = W[(16 * t + r) * K + kb + 4 * q + e];
      }
      d = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, d, 0, 0, 0);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const float z = d[e] + bias[16 * t + r];
      out[4 * q + e][16 * t + r] = ACT ? gelu(z) : z;
    }
  }
}

__global__ __launch_bounds__(kPredWaves * 64) void mlp_predict_kernel(
    const float* __restrict__ params, const float* __restrict__ feat, int n,
    float* __restrict__ out) {
  __shared__ PredLds L;
  const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
  const float* net = params + kMlpNormFeat;
  using NL = NetLayout;
  for (int i = tid; i < kMlpHidden * kK1; i += blockDim.x) {
    const int o = i / kK1, k = i % kK1;
    L.w1[i] = k < kMlpIn ? f2bf(net[NL::W1 + o * kMlpIn + k]) : (short)0;
  }
  for (int i = tid; i < kMlpHidden * kMlpHidden; i += blockDim.x) {
    L.w2[i] = f2bf(net[NL::W2 + i]);
    L.w3[i] = f2bf(net[NL::W3 + i]);
  }
  for (int i = tid; i < 16 * kMlpHidden; i += blockDim.x)
    L.w4[i] = i < kMlpOut * kMlpHidden ? f2bf(net[NL::W4 + i]) : (short)0;
  for (int i = tid; i < kMlpHidden; i += blockDim.x) {
    L.b1[i] = net[NL::b1 + i];
    L.b2[i] = net[NL::b2 + i];
    L.b3[i] = net[NL::b3 + i];
  }
  if (tid < 16) L.b4[tid] = tid < kMlpOut ? net[NL::b4 + tid] : 0.0f;
  // standardised features of the wave's 16 rows
  const int row0 = (blockIdx.x * kPredWaves + wave) * 16;
  for (int i = lane; i < 16 * kMlpHidden; i += 64) {
    const int r = i / kMlpHidden, c = i % kMlpHidden;
    float v = 0.0f;
    if (c < kMlpIn && row0 + r < n)
      v = (feat[(size_t)(row0 + r) * kMlpIn + c] - params[c]) / params[kMlpIn + c];
    L.act[wave][0][r][c] = v;
  }
  __syncthreads();
  mfma_layer<kK1, 4, true>(L.act[wave][0], L.w1, L.b1, L.act[wave][1], lane);
  __syncthreads();
  mfma_layer<kMlpHidden, 4, true>(L.act[wave][1], L.w2, L.b2, L.act[wave][0], lane);
  __syncthreads();
  mfma_layer<kMlpHidden, 4, true>(L.act[wave][0], L.w3, L.b3, L.act[wave][1], lane);
  __syncthreads();
  mfma_layer<kMlpHidden, 1, false>(L.act[wave][1], L.w4, L.b4, L.act[wave][0], lane);
  __syncthreads();
  const float* tn = params + kMlpNormFeat + kMlpNet;
  if (lane < 16 * kMlpOut) {
    const int r = lane / kMlpOut, j = lane % kMlpOut;
    if (row0 + r < n)
      out[(size_t)(row0 + r) * kMlpOut + j] = L.act[wave][0][r][j] * tn[kMlpOut + j] + tn[j] +
                                              feat[(size_t)(row0 + r) * kMlpIn + kMlpBase + j];
  }
}

}  // namespace

hipError_t launch_mlp_predict(const float* params, const float* feat, int n, float* out,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int rows = kPredWaves * 16;
  hipLaunchKernelGGL(mlp_predict_kernel, dim3((n + rows - 1) / rows), dim3(kPredWaves * 64), 0,
                     st, params, feat, n, out);
  return hipGetLastError();
}

void mlp_standardisation(const float* feat, const float* tgt, size_t n, float* params) {
  for (int c = 0; c < kMlpIn + kMlpOut; ++c) {
    const bool is_f = c < kMlpIn;
    const int j = is_f ? c : c - kMlpIn;
    double s = 0.0, ss = 0.0;
    for (size_t r = 0; r < n; ++r) {
      const double v = is_f ? feat[r * kMlpIn + j]
                            : (double)tgt[r * kMlpOut + j] - feat[r * kMlpIn + kMlpBase + j];
      s += v;
      ss += v * v;
    }
    const double mean = s / (double)std::max<size_t>(1, n);
    double var = ss / (double)std::max<size_t>(1, n) - mean * mean;
    const double sd = var > 1e-12 ? std::sqrt(var) : 1.0;
    float* dst = is_f ? params : params + kMlpNormFeat + kMlpNet;
    const int width = is_f ? kMlpIn : kMlpOut;
    dst[j] = (float)mean;
    dst[width + j] = (float)sd;
  }
}

}  // namespace nh

using namespace nh;

namespace {
void put_err(char* err, size_t len, const std::string& m) {
  if (err && len) {
    std::strncpy(err, m.c_str(), len - 1);
    err[len - 1] = 0;
  }
}
}  // namespace

extern "C" {

size_t nh_mlp_param_count(void) { return kMlpParams; }

int nh_mlp_loss_grad(const float* params, const float* features, const float* targets, size_t n,
                     int32_t device, float* loss, float* grads, char* err, size_t err_len) {
  if (!params || !features || !targets || !loss || !grads || n == 0 || n > (1u << 20)) {
    put_err(err, err_len, "invalid arguments");
    return NH_ERR_INVALID_ARGUMENT;
  }
  try {
    NH_HIP(hipSetDevice(device));
    hipStream_t st;
    NH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    DevBuf<float> dparams, dfeat, dtgt, dgrad;
    DevBuf<int> didx;
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    dparams.upload(params, kMlpParams, st);
    dfeat.upload(features, n * kMlpIn, st);
    dtgt.upload(targets, n * kMlpOut, st);
    didx.upload(idx.data(), n, st);
    dgrad.alloc(kMlpNet);
    MlpPass P;
    P.reserve(n);
    hipLaunchKernelGGL(mlp_gather_kernel, dim3(blocks(n * (kMlpIn + kMlpOut), 256)), dim3(256), 0,
                       st, dfeat.p, dtgt.p, didx.p, (int)n, dparams.p, P.in.p, P.t.p);
    pass(P, dparams.p + kMlpNormFeat, dgrad.p, (int)n, st);
    NH_HIP(hipGetLastError());
    NH_HIP(hipMemcpyAsync(loss, P.loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
    NH_HIP(hipMemcpyAsync(grads, dgrad.p, sizeof(float) * kMlpNet, hipMemcpyDeviceToHost, st));
    NH_HIP(hipStreamSynchronize(st));
    (void)hipStreamDestroy(st);
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  return NH_OK;
}

int nh_mlp_train(const float* features, const float* targets, size_t n, int32_t epochs,
                 int32_t batch, float lr, uint64_t seed, int32_t device, float* params,
                 float* final_loss, char* err, size_t err_len) {
  if (!features || !targets || !params || n < 2 || epochs < 1 || batch < 1 || !(lr > 0.0f)) {
    put_err(err, err_len, "invalid arguments");
    return NH_ERR_INVALID_ARGUMENT;
  }
  try {
    std::vector<float> h(kMlpParams, 0.0f);
    mlp_standardisation(features, targets, n, h.data());
    // Glorot-uniform weights, zero biases
    std::mt19937_64 rng(seed);
    auto fill = [&](size_t off, int out, int in) {
      const float a = std::sqrt(6.0f / (float)(in + out));
      std::uniform_real_distribution<float> U(-a, a);
      for (int i = 0; i < out * in; ++i) h[kMlpNormFeat + off + i] = U(rng);
    };
    fill(NetLayout::W1, kMlpHidden, kMlpIn);
    fill(NetLayout::W2, kMlpHidden, kMlpHidden);
    fill(NetLayout::W3, kMlpHidden, kMlpHidden);
    fill(NetLayout::W4, kMlpOut, kMlpHidden);
    NH_HIP(hipSetDevice(device));
    hipStream_t st;
    NH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    DevBuf<float> dparams, dfeat, dtgt, dgrad, dm, dv;
    DevBuf<int> didx;
    dparams.upload(h.data(), kMlpParams, st);
    dfeat.upload(features, n * kMlpIn, st);
    dtgt.upload(targets, n * kMlpOut, st);
    dgrad.alloc(kMlpNet);
    std::vector<float> zeros(kMlpNet, 0.0f);
    dm.upload(zeros.data(), kMlpNet, st);
    dv.upload(zeros.data(), kMlpNet, st);
    const size_t B = std::min<size_t>((size_t)batch, n);
    MlpPass P;
    P.reserve(std::max(B, n));
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    didx.alloc(n);
    float* net = dparams.p + kMlpNormFeat;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    long step = 0;
    for (int ep = 0; ep < epochs; ++ep) {
      std::shuffle(order.begin(), order.end(), rng);
      NH_HIP(hipMemcpyAsync(didx.p, order.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
      // cosine decay of the learning rate over the run
      for (size_t at = 0; at + B <= n; at += B) {
        ++step;
        const double frac = (double)step / ((double)epochs * (double)(n / B));
        const float lr_t = (float)(lr * 0.5 * (1.0 + std::cos(3.141592653589793 * frac)));
        hipLaunchKernelGGL(mlp_gather_kernel, dim3(blocks(B * (kMlpIn + kMlpOut), 256)),
                           dim3(256), 0, st, dfeat.p, dtgt.p, didx.p + at, (int)B, dparams.p,
                           P.in.p, P.t.p);
        pass(P, net, dgrad.p, (int)B, st);
        const float bc1 = 1.0f - std::pow(b1, (float)step), bc2 = 1.0f - std::pow(b2, (float)step);
        hipLaunchKernelGGL(mlp_adam_kernel, dim3(blocks(kMlpNet, 256)), dim3(256), 0, st, net,
                           dgrad.p, dm.p, dv.p, (int)kMlpNet, lr_t, b1, b2, eps, bc1, bc2);
      }
      NH_HIP(hipGetLastError());
    }
    // loss over every row
    std::iota(order.begin(), order.end(), 0);
    NH_HIP(hipMemcpyAsync(didx.p, order.data(), n * sizeof(int), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(mlp_gather_kernel, dim3(blocks(n * (kMlpIn + kMlpOut), 256)), dim3(256), 0,
                       st, dfeat.p, dtgt.p, didx.p, (int)n, dparams.p, P.in.p, P.t.p);
    pass(P, net, nullptr, (int)n, st);
    NH_HIP(hipGetLastError());
    float fl = 0.0f;
    NH_HIP(hipMemcpyAsync(&fl, P.loss.p, sizeof(float), hipMemcpyDeviceToHost, st));
    NH_HIP(hipMemcpyAsync(params, dparams.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost,
                          st));
    NH_HIP(hipStreamSynchronize(st));
    (void)hipStreamDestroy(st);
    if (final_loss) *final_loss = fl;
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  return NH_OK;
}

int nh_mlp_predict(const float* params, const float* features, size_t n, int32_t device,
                   float* out, char* err, size_t err_len) {
  if (!params || !features || !out || n > (1u << 30)) {
    put_err(err, err_len, "invalid arguments");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (n == 0) return NH_OK;
  try {
    NH_HIP(hipSetDevice(device));
    hipStream_t st;
    NH_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    DevBuf<float> dparams, dfeat, dout;
    dparams.upload(params, kMlpParams, st);
    dfeat.upload(features, n * kMlpIn, st);
    dout.alloc(n * kMlpOut);
    NH_HIP(launch_mlp_predict(dparams.p, dfeat.p, (int)n, dout.p, st));
    NH_HIP(hipMemcpyAsync(out, dout.p, sizeof(float) * n * kMlpOut, hipMemcpyDeviceToHost, st));
    NH_HIP(hipStreamSynchronize(st));
    (void)hipStreamDestroy(st);
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  return NH_OK;
}

}  // extern "C"

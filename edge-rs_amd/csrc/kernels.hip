// gfx950 kernels of the stage-two path and their host-side launchers.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>

#include "bobyqa.hpp"
#include "kernels.hpp"
#include "lgamma_pos.hpp"
#include "pml_fit.hpp"

namespace nh {

// ---------------------------------------------------------------------------
// One workgroup per request: a penalised fit at fixed (subject, cell)
// (edge-rs ResidentBatch::solve, src/gpu/nebula_gpu.rs:441, one fit per plane
// there; one per workgroup here).
template <int NB, int WAVES>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<NB>)))
void pml_batch_kernel(
    DeviceData dd, const DevRequest* req, const double* beta_init, const double* lw_init,
    DevReply* out, double* out_beta, double* out_lw, double* out_info, double* scratch,
    size_t stride, FitKnobs kn, int mode, FitCounters* ctr, int work_in_lds) {
  __shared__ FitLds<NB, WAVES> sh;
  __shared__ Fitter<NB, WAVES> f;
  extern __shared__ double nh_batch_dyn[];
  const int q = blockIdx.x;
  const int k = (int)dd.k;
  FitWork<NB> wk;
  wk.carve(work_in_lds ? nh_batch_dyn : scratch + (size_t)q * stride, k, WAVES);
  const DevRequest r = req[q];
  if (threadIdx.x < NB) sh.beta[threadIdx.x] = beta_init[(size_t)q * NB + threadIdx.x];
  for (int s = threadIdx.x; s < k; s += WAVES * 64)
    wk.lw[s] = lw_init ? lw_init[(size_t)q * k + s] : 0.0;
  if (threadIdx.x == 0) {
    const double exps = exp(r.subject);
    const double alpha = 1.0 / (exps - 1.0);
    const double lambda = 1.0 / (sqrt(exps) * (exps - 1.0));
    f.init(dd, &sh, wk, ctr, r.gene, r.cell, alpha, lambda);
  }
  __syncthreads();
  FitResult res = f.run(mode, r.ord & ~kOrdMeanStart, kn, false, (r.ord & kOrdMeanStart) != 0);
  if (threadIdx.x == 0) {
    DevReply o;
    o.ll = res.ll;
    o.ll_prev = res.ll_prev;
    o.log_det = res.log_det;
    o.second = res.second;
    o.iterations = res.iterations;
    o.backtracks = res.backtracks;
    o.finish_steps = res.finish_steps;
    o.exact = res.exact;
    out[q] = o;
  }
  if (threadIdx.x < NB) out_beta[(size_t)q * NB + threadIdx.x] = sh.beta[threadIdx.x];
  if (threadIdx.x < NB * NB) out_info[(size_t)q * NB * NB + threadIdx.x] = sh.vb2[threadIdx.x];
  for (int s = threadIdx.x; s < k; s += WAVES * 64) out_lw[(size_t)q * k + s] = wk.lw[s];
}

size_t fit_work_doubles(int nb, int k, int waves) {
  switch (nb) {
    case 1: return FitWork<1>::doubles(k, waves);
    case 2: return FitWork<2>::doubles(k, waves);
    case 3: return FitWork<3>::doubles(k, waves);
    case 4: return FitWork<4>::doubles(k, waves);
    case 5: return FitWork<5>::doubles(k, waves);
    case 6: return FitWork<6>::doubles(k, waves);
    case 7: return FitWork<7>::doubles(k, waves);
    default: return FitWork<8>::doubles(k, waves);
  }
}

template <int NB, int WAVES>
static hipError_t launch_batch_t(const BatchLaunch& L, hipStream_t st) {
  const size_t lds = L.work_in_lds ? L.stride * sizeof(double) : 0;
  hipLaunchKernelGGL((pml_batch_kernel<NB, WAVES>), dim3(L.n_req), dim3(WAVES * 64), lds, st,
                     L.dd, L.req, L.beta_init, L.lw_init, L.out, L.out_beta, L.out_lw,
                     L.out_info, L.scratch, L.stride, L.knobs, L.mode, L.counters,
                     L.work_in_lds);
  return hipGetLastError();
}

// 16 waves per gene only where a sweep fits 128 VGPRs (p <= 4).
template <int NB>
static hipError_t launch_batch_nb(const BatchLaunch& L, hipStream_t st) {
  switch (effective_waves(NB, L.waves)) {
    case 4: return launch_batch_t<NB, 4>(L, st);
    case 16:
      if constexpr (NB <= 4) return launch_batch_t<NB, 16>(L, st);
      [[fallthrough]];
    default: return launch_batch_t<NB, 8>(L, st);
  }
}

__global__ __launch_bounds__(256) void f64_math_kernel(const double* x, size_t n, int which,
                                                        double* out) {
  __shared__ double tab[kF64TabDoubles];
  fill_f64_tables(tab, (int)threadIdx.x);
  __syncthreads();
  const lds_double* t = (const lds_double*)tab;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
    const double v = x[i];
    out[i] = which == 0 ? fast_exp(v, t)
             : which == 1 ? fast_log(v, t)
             : which == 2 ? fast_rcp(v)
                          : lgamma_pos(v);
  }
}

hipError_t launch_f64_math(const double* x, size_t n, int which, double* out, hipStream_t st) {
  size_t blocks = (n + 255) / 256;
  blocks = blocks < 1 ? 1 : (blocks > 1024 ? 1024 : blocks);
  hipLaunchKernelGGL(f64_math_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x, n, which, out);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void widen_kernel(const float* x, size_t n, double* out) {
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
    out[i] = (double)x[i];
}

hipError_t launch_widen(const float* x, size_t n, double* out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  size_t blocks = (n + 255) / 256;
  blocks = blocks > 8192 ? 8192 : blocks;
  hipLaunchKernelGGL(widen_kernel, dim3((unsigned)blocks), dim3(256), 0, st, x, n, out);
  return hipGetLastError();
}

hipError_t launch_batch(const BatchLaunch& L, hipStream_t st) {
  if (L.n_req == 0) return hipSuccess;
  switch (L.dd.nb) {
    case 1: return launch_batch_nb<1>(L, st);
    case 2: return launch_batch_nb<2>(L, st);
    case 3: return launch_batch_nb<3>(L, st);
    case 4: return launch_batch_nb<4>(L, st);
    case 5: return launch_batch_nb<5>(L, st);
    case 6: return launch_batch_nb<6>(L, st);
    case 7: return launch_batch_nb<7>(L, st);
    case 8: return launch_batch_nb<8>(L, st);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// BOBYQA driven entirely on the device against built-in test objectives, so
// the device arithmetic can be checked against the nloptr traces.
__device__ static double test_objective(int problem, const double* x) {
  NH_STRICT_FP
  switch (problem) {
    case 0: {
      double a = x[1] - x[0] * x[0], b = 1.0 - x[0];
      return 100.0 * (a * a) + b * b;
    }
    case 1: {
      double a = x[0] - 3.0, b = x[1] + 1.0;
      return a * a + 10.0 * (b * b) + x[0] * x[1];
    }
    case 2: {
      double u = 10000.0 * x[0] + 30000.0 * x[1];
      double t = fabs(u - 2.0 * floor((u + 1.0) / 2.0));
      double a = x[0] - 0.25, b = x[1] + 0.125;
      return a * a + 2.0 * (b * b) + 0.5 * x[0] * x[1] + 0.5 * (t * t);
    }
    case 3: {
      double a = x[0] + x[1] - 1.0;
      return a * a;
    }
    default: {
      double a = x[0] - 1.0, b = x[1] - x[0] * x[0], c = x[2] - x[1], d = x[2] * x[2];
      return a * a + 100.0 * (b * b) + c * c + 0.5 * (d * d);
    }
  }
}

__global__ void bobyqa_trace_kernel(int problem, int n, const double* x0, const double* lo,
                                    const double* hi, double* trace, int max_trace,
                                    int* result) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Bobyqa<3> b;
  b.init(n, x0, lo, hi);
  int i = 0;
  while (b.want) {
    double v;
    {
      NH_STRICT_FP
      v = test_objective(problem, b.xask);
    }
    if (i < max_trace) {
      for (int j = 0; j < n; ++j) trace[(size_t)i * (n + 1) + j] = b.xask[j];
      trace[(size_t)i * (n + 1) + n] = v;
    }
    ++i;
    b.tell(v);
  }
  result[0] = b.status;
  result[1] = i;
}

// A plain device-to-device copy of n 16-byte words, each read and written
// once: the yardstick the HBM byte counters are scaled by.
__global__ void copy_words_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                  size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

hipError_t launch_copy_words(const void* src, void* dst, size_t n, hipStream_t st) {
  hipLaunchKernelGGL(copy_words_kernel, dim3(4096), dim3(256), 0, st, (const uint4*)src,
                     (uint4*)dst, n);
  return hipGetLastError();
}

hipError_t launch_bobyqa_trace(int problem, int n, const double* x0, const double* lo,
                               const double* hi, double* trace, int max_trace, int* result,
                               hipStream_t st) {
  hipLaunchKernelGGL(bobyqa_trace_kernel, dim3(1), dim3(64), 0, st, problem, n, x0, lo, hi,
                     trace, max_trace, result);
  return hipGetLastError();
}

}  // namespace nh

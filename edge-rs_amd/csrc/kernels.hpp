// Host-visible declarations of the device data layout and launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "pml_fit.hpp"

namespace nh {

// DevRequest::ord flag: a cold stage-two fit, whose log w start is set from
// the subjects' Poisson-limit optimum at the start's beta (Fitter::run).
constexpr uint32_t kOrdMeanStart = 1u << 16;

struct DevRequest {
  uint32_t gene;
  uint32_t ord;  // Laplace order, | kOrdMeanStart
  double subject;  // nebula's sigma[0]
  double cell;     // NB size gamma
};

struct DevReply {
  double ll, ll_prev, log_det, second;
  int iterations, backtracks, finish_steps, exact;
};

struct BatchLaunch {
  DeviceData dd;
  const DevRequest* req;
  const double* beta_init;  // n_req * nb
  const double* lw_init;    // n_req * k or null
  DevReply* out;
  double* out_beta;
  double* out_lw;
  double* out_info;
  double* scratch;     // global work area when it does not go in LDS
  size_t stride;       // work doubles per request
  int work_in_lds;     // 1: work arrays in dynamic LDS
  FitKnobs knobs;
  int mode;       // 0 f32 + f64 finish, 1 f64
  int waves;      // 4, 8 or 16
  int n_req;
  FitCounters* counters;
};

size_t fit_work_doubles(int nb, int k, int waves);
// Waves per gene a launch really uses: 16-wave workgroups need a sweep that
// fits 128 VGPRs, which holds up to 4 design columns.
inline int effective_waves(int nb, int waves) {
  if (waves >= 16) return nb <= 4 ? 16 : 8;
  return waves <= 4 ? 4 : 8;
}

// Largest per-workgroup work area placed in LDS (bytes); larger ones live in
// global memory.
#ifndef NH_WORK_LDS_BYTES
#define NH_WORK_LDS_BYTES (48 * 1024)
#endif
constexpr size_t kWorkLdsBytes = NH_WORK_LDS_BYTES;
hipError_t launch_batch(const BatchLaunch& L, hipStream_t st);
// out[i] = the FP64 sweep's exp / log / reciprocal (which 0 / 1 / 2) or the
// device objective's lgamma (3) of x[i].
hipError_t launch_f64_math(const double* x, size_t n, int which, double* out, hipStream_t st);
// out[i] = (double)x[i]
hipError_t launch_widen(const float* x, size_t n, double* out, hipStream_t st);
// dst[i] = src[i] for n 16-byte words
hipError_t launch_copy_words(const void* src, void* dst, size_t n, hipStream_t st);
hipError_t launch_bobyqa_trace(int problem, int n, const double* x0, const double* lo,
                               const double* hi, double* trace, int max_trace, int* result,
                               hipStream_t st);

}  // namespace nh

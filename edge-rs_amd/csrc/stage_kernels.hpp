#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pml_fit.hpp"

namespace nh {

struct PtmgRequest {
  uint32_t gene;
  uint32_t hess;  // 1: also the Hessian sums
};

// Waves per stage-one workgroup; a request's cells are cut into
// chunks * kPtmgWaves units, one wave each, and a workgroup runs one unit of
// kPtmgWaves requests (about `chunks` workgroups per request).
constexpr int kPtmgWaves = 4;

struct PtmgLaunch {
  DeviceData dd;
  // req, params and out may be mapped host memory: the kernels read each
  // request's entries once per workgroup and write out with whole-wave stores
  const PtmgRequest* req;
  const double* params;  // n_req * (nb + 2): [beta, sigma, phi]
  double* out;           // n_req * ptmg_out_doubles(nb, k)
  double* scratch;       // n_req * stride, device memory
  size_t stride;         // >= ptmg_scratch_doubles(nb, k, chunks * kPtmgWaves)
  int n_req;
  int chunks;            // units per request / kPtmgWaves
  bool any_hess;         // some request asks for the Hessian
  FitCounters* counters;
};

// Layout of one request's output: per subject s (PS = 4 + 4 nb doubles at
// s * PS): sum e, sum g e, sum (g e/w - e/w), sum g e^2/w, sum x e,
// sum g x e, sum x e (g/w - 1/w), sum x e g e/w; then per gene (at k * PS):
// sum y eta, sum (ym e - (gamma + y) ln w), sum ln w, sum g, sum 1/w,
// sum g/w, Gram[nb (nb + 1) / 2] (row-major upper triangle).
size_t ptmg_out_doubles(int nb, int k);
size_t ptmg_scratch_doubles(int nb, int k, int units);
hipError_t launch_ptmg(const PtmgLaunch& L, hipStream_t st);
// out[g * nb + j] = sum over gene g's positive counts of y * x64[cell * nb + j]
hipError_t launch_xty(const DeviceData& dd, size_t n_genes, double* out, hipStream_t st);

}  // namespace nh

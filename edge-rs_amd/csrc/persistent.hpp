#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.hpp"
#include "pml_fit.hpp"

namespace nh {

// One gene's stage-two search, as the persistent kernel reads it.
struct DevSearch {
  uint32_t gene;
  uint32_t ord;          // Laplace order of the first pass
  uint32_t kind;         // SearchKind: 2 both components, 1 the subject variance alone
  double start[2];       // stage one's (sigma, gamma); gamma is held for kind 1
  double beta_start[8];  // fixed effects every cold fit starts from (stage one's)
  uint64_t tail_off;     // count histogram slice (values != 1, 2)
  uint32_t tail_len;
  double n_pos, n_one, n_two;
};

struct DevSearchOut {
  double x[2];
  double f;
  int found, failed, evals, ord;
  int warm_rejected;  // learned start fell back to the cold search
};

struct PersistentArgs {
  DeviceData dd;
  const DevSearch* search;
  const uint32_t* order;  // work order, longest gene first
  int n_search;
  unsigned* queue;        // next slot
  const double* tail_val;
  const double* tail_mult;
  DevSearchOut* out;
  double* scratch;        // global work area when it does not go in LDS
  size_t stride;          // work doubles per workgroup
  int work_in_lds;        // 1: work arrays in dynamic LDS
  FitKnobs knobs;
  double lo[2], hi[2];
  uint32_t intercept;
  FitCounters* counters;
  const float* pred;      // learned starts: n_search x (log sigma^2, log phi), or null
  double warm_step;       // warm search's initial step relative to the predicted point
  double warm_xtol;       // warm search's rhoend / rhobeg
  int predict;            // warm inner fits start at the anchor's first-order
                          // prediction (Fitter::sensitivity)
};

size_t persistent_stride(int nb, int k, int waves);
hipError_t launch_search_trace(int problem, int kind, int ord, const double* start,
                               const double* lo, const double* hi, int inj_from, int inj_to,
                               int inj_order, int inj_kind, double* trace, int max_trace,
                               DevSearchOut* out, hipStream_t st);
hipError_t launch_persistent(const PersistentArgs& A, int waves, int grid, hipStream_t st);

}  // namespace nh

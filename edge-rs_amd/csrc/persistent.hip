// Stage two, batched engine: a persistent kernel whose workgroups pull genes
// off a device work queue (longest first) and run each gene's whole search on
// the device, BOBYQA included.
//
// The reference runs its stage-two searches in lockstep host rounds
// (src/gpu/stage_two.rs:362-544): every round waits for the slowest fit and
// for the host's f64 finish, so the slowest gene sets the round count and the
// host sets the pace.  Here lane 0 of a workgroup steps the gene's BOBYQA
// (the same stepper as the host's, edge-rs-amd/csrc/bobyqa.hpp, contraction
// off so the device arithmetic is NLopt's), the whole workgroup runs the
// penalised fit at each asked point (FP32 Newton + FP64 finish), assembles
// the profile objective in FP64 (VarianceObjective::assemble,
// src/sc/nebula.rs:1452-1477, lgamma tail over the gene's count histogram)
// and tells the stepper; a finished workgroup takes the next gene.  No host
// round trip until every gene is done.
//
// Search semantics are the host plain engine's: the order-one retry and the
// abandon-on-non-finite rule of StageTwoSearch (nebula.rs:1656-1741), inner
// fits warm-started from the gene's first finite fit (stage_two.rs:36-43).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <cmath>
#include <memory>
#include <numeric>
#include <vector>

#include "bobyqa.hpp"
#include "driver.hpp"
#include "pml_fit.hpp"
#include "persistent.hpp"
#include "mlp.hpp"
#include "lgamma_pos.hpp"
#include "resident.hpp"

namespace nh {

// Profile objective on the device, same arithmetic as host_common.hpp's.
__device__ static double device_objective(double subject, double cell, const FitResult& r,
                                          double n_cells, double k, double tail, double n_pos,
                                          double n_one, double n_two) {
  if (r.second < -1.0) return INFINITY;
  const double exps = exp(subject);
  const double alpha = 1.0 / (exps - 1.0);
  const double lambda = 1.0 / (sqrt(exps) * (exps - 1.0));
  const double base = isnan(r.ll) ? r.ll_prev : r.ll;
  double v = base + n_cells * cell * log(cell) + k * alpha * log(lambda) - k * lgamma_pos(alpha);
  v += tail - (n_pos - n_one - n_two) * lgamma_pos(cell) + (n_one + n_two) * log(cell) +
       n_two * log(cell + 1.0);
  v += -0.5 * r.log_det + log(1.0 + r.second);
  return isfinite(v) ? -v : INFINITY;
}

// The stepper runs on one lane between fits; kept out of line so the sweep
// loop's register allocation is not shaped by it.
__device__ __noinline__ static void bq_init(Bobyqa<2>* b, const double* x0, const double* lo,
                                            const double* hi) {
  b->init(2, x0, lo, hi);
}
__device__ __noinline__ static void bq_init_step(Bobyqa<2>* b, const double* x0, const double* lo,
                                                 const double* hi, const double* step,
                                                 double xtol) {
  b->init_with_step(2, x0, lo, hi, step, xtol, 1000);
}
__device__ __noinline__ static void bq_tell(Bobyqa<2>* b, double v) { b->tell(v); }

enum SearchPhase { PH_COLD = 0, PH_WARM = 1, PH_CHECK = 2 };

template <int NB, int WAVES>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<NB>)))
void stage_two_kernel(PersistentArgs A) {
  constexpr int NT = WAVES * 64;
  __shared__ FitLds<NB, WAVES> sh;
  __shared__ Bobyqa<2> bq;
  __shared__ double warm_beta[NB];
  __shared__ double red[NT / 64];
  __shared__ int s_slot, s_active, s_found, s_failed, s_have_warm, s_evals, s_order, s_ok;
  __shared__ int s_phase, s_wfailed, s_rejected, s_from_warm;
  __shared__ double s_subject, s_cell, s_value, s_wx[2], s_wf, s_cold[2], s_warm[2];
  __shared__ double s_anchor_f;  // objective at the inner fits' current start
  extern __shared__ double nh_s2_dyn[];
  const int tid = threadIdx.x;
  const DeviceData& dd = A.dd;
  const int k = (int)dd.k;
  FitWork<NB> wk;
  double* base = A.work_in_lds ? nh_s2_dyn : A.scratch + (size_t)blockIdx.x * A.stride;
  wk.carve(base, k, WAVES);
  double* warm_lw = base + FitWork<NB>::doubles(k, WAVES);
  const double lo[2] = {A.lo[0], A.lo[1]}, hi[2] = {A.hi[0], A.hi[1]};
  // (thread 0) starts a BOBYQA pass of the given phase at the current order
  auto begin_search = [&](int phase) {
    s_phase = phase;
    if (phase == PH_WARM) {
      // The initial step is a fraction of the predicted value, widened to
      // twice the distance to stage one's estimate where the two disagree
      // (a prediction far from stage one is less certain), at most to a
      // tenth of the predicted value: wider first steps cost more BOBYQA
      // stages than they save where the network is right and stage one is
      // not.  The final precision rhoend = warm_xtol * warm_step * x stays
      // the same whatever the step: xtol is scaled down when it is widened.
      double step[2];
      for (int i = 0; i < 2; ++i) {
        const double base = A.warm_step * s_warm[i];
        step[i] = fmax(base, fmin(2.0 * fabs(s_warm[i] - s_cold[i]), 0.1 * s_warm[i]));
        step[i] = fmin(fmax(step[i], 1e-3 * lo[i]), 0.45 * (hi[i] - lo[i]));
      }
      const double xtol = A.warm_xtol * fmin(1.0, A.warm_step * s_warm[0] / step[0]);
      bq_init_step(&bq, s_warm, lo, hi, step, xtol);
    } else {
      bq_init(&bq, s_cold, lo, hi);
    }
    s_active = bq.status != BOBYQA_BAD_BOX && bq.want;
  };
  for (;;) {
    if (tid == 0) s_slot = (int)atomicAdd(A.queue, 1u);
    __syncthreads();
    const int slot = s_slot;
    __syncthreads();
    if (slot >= A.n_search) break;
    const DevSearch S = A.search[A.order[slot]];
    if (tid == 0) {
      s_order = (int)S.ord;
      s_have_warm = 0;
      s_evals = 0;
      s_found = 0;
      s_failed = 0;
      s_rejected = 0;
      s_from_warm = 0;
      s_cold[0] = fmin(fmax(S.start[0], A.lo[0]), A.hi[0]);
      s_cold[1] = fmin(fmax(S.start[1], A.lo[1]), A.hi[1]);
      s_active = 0;
      if (A.pred) {
        const float* pr = A.pred + 2 * (size_t)A.order[slot];
        s_warm[0] = fmin(fmax(exp((double)pr[0]), A.lo[0]), A.hi[0]);
        s_warm[1] = fmin(fmax(exp(-(double)pr[1]), A.lo[1]), A.hi[1]);
        begin_search(PH_WARM);
      }
      if (!s_active) begin_search(PH_COLD);
    }
    __syncthreads();
    while (s_active) {
      // request for the asked point (the cold start when checking a warm result)
      if (tid == 0) {
        s_subject = s_phase == PH_CHECK ? s_cold[0] : bq.xask[0];
        s_cell = s_phase == PH_CHECK ? s_cold[1] : bq.xask[1];
      }
      __syncthreads();
      const double subject = s_subject, cell = s_cell;
      if (tid < NB) {
        double b = S.beta_start[tid];
        if (tid == (int)A.intercept) b -= subject / 2.0;
        sh.beta[tid] = s_have_warm ? warm_beta[tid] : b;
      }
      for (int s = tid; s < k; s += NT) wk.lw[s] = s_have_warm ? warm_lw[s] : 0.0;
      __syncthreads();
      const double exps = exp(subject);
      const double alpha = 1.0 / (exps - 1.0);
      const double lambda = 1.0 / (sqrt(exps) * (exps - 1.0));
      Fitter<NB, WAVES> f(dd, sh, wk, A.counters, S.gene, cell, alpha, lambda);
      const FitResult r = f.run(0, (unsigned)s_order, A.knobs, s_evals > 0);
      // finite point and the lgamma tail over the count histogram
      double part = 0.0;
      bool fin = true;
      for (int s = tid; s < k; s += NT) fin = fin && isfinite(wk.lw[s]);
      if (tid < NB) fin = fin && isfinite(sh.beta[tid]);
      for (uint32_t i = (uint32_t)tid; i < S.tail_len; i += NT)
        part += A.tail_mult[S.tail_off + i] * lgamma_pos(A.tail_val[S.tail_off + i] + cell);
      part = wave_sum(part);
      const bool wfin = __all(fin);
      if ((tid % 64) == 0) red[tid / 64] = wfin ? part : NAN;
      __syncthreads();
      if (tid == 0) {
        double tail = 0.0;
        bool allfin = true;
        for (int w = 0; w < WAVES; ++w) {
          if (isnan(red[w])) allfin = false;
          else tail += red[w];
        }
        double v = INFINITY;
        if (allfin)
          v = device_objective(subject, cell, r, (double)dd.n_cells, (double)k, tail, S.n_pos,
                               S.n_one, S.n_two);
        s_value = v;
        s_evals += 1;
        // Inner fits start from the gene's first finite fit (the cold search,
        // as edge-rs does); the warm search moves that start to its best
        // point so far, which its steps stay close to.
        s_ok = isfinite(v) && (!s_have_warm || (s_phase == PH_WARM && v < s_anchor_f));
        if (s_ok) s_anchor_f = v;
      }
      __syncthreads();
      if (s_ok) {
        if (tid < NB) warm_beta[tid] = sh.beta[tid];
        for (int s = tid; s < k; s += NT) warm_lw[s] = wk.lw[s];
      }
      __syncthreads();
      if (tid == 0) {
        if (s_ok) s_have_warm = 1;
        const double v = s_value;
        if (s_phase == PH_CHECK) {
          // keep the warm result unless the cold start point does better
          if (isfinite(v) && s_wf <= v) {
            s_active = 0;
            s_found = 1;
            s_failed = s_wfailed;
            s_from_warm = 1;
          } else {
            s_rejected = 1;
            s_order = (int)S.ord;
            s_have_warm = 0;
            begin_search(PH_COLD);
          }
        } else {
          bool pass_over = false, found = false;
          if (!isfinite(v)) {
            pass_over = true;
          } else {
            bq_tell(&bq, v);
            if (!bq.want) {
              pass_over = true;
              found = true;
            }
          }
          if (pass_over) {
            if (found || s_order == 1) {
              if (s_phase == PH_WARM && found) {
                s_wx[0] = bq.xbest_out[0];
                s_wx[1] = bq.xbest_out[1];
                s_wf = bq.fbest_out;
                s_wfailed = bq.status == BOBYQA_ROUNDOFF_LIMITED ? 1 : 0;
                s_phase = PH_CHECK;
              } else if (s_phase == PH_WARM) {
                s_rejected = 1;
                s_order = (int)S.ord;
                s_have_warm = 0;
                begin_search(PH_COLD);
              } else {
                s_active = 0;
                s_found = found ? 1 : 0;
                s_failed = bq.status == BOBYQA_ROUNDOFF_LIMITED ? 1 : 0;
              }
            } else {
              s_order = 1;
              begin_search(s_phase);
              if (!s_active && s_phase == PH_WARM) {
                s_rejected = 1;
                s_order = (int)S.ord;
                begin_search(PH_COLD);
              }
            }
          }
        }
      }
      __syncthreads();
    }
    if (tid == 0) {
      DevSearchOut o;
      o.x[0] = s_from_warm ? s_wx[0] : bq.xbest_out[0];
      o.x[1] = s_from_warm ? s_wx[1] : bq.xbest_out[1];
      o.f = s_from_warm ? s_wf : bq.fbest_out;
      o.found = s_found;
      o.failed = s_failed;
      o.evals = s_evals;
      o.ord = s_order;
      o.warm_rejected = s_rejected;
      A.out[A.order[slot]] = o;
    }
    __syncthreads();
  }
}

__global__ void reset_queue_kernel(unsigned* q) {
  if (threadIdx.x == 0) *q = 0u;
}

size_t persistent_stride(int nb, int k, int waves) { return fit_work_doubles(nb, k, waves) + (size_t)k; }

template <int NB, int WAVES>
static hipError_t launch_t(const PersistentArgs& A, int grid, hipStream_t st) {
  hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(64), 0, st, A.queue);
  const size_t lds = A.work_in_lds ? A.stride * sizeof(double) : 0;
  hipLaunchKernelGGL((stage_two_kernel<NB, WAVES>), dim3(grid), dim3(WAVES * 64), lds, st, A);
  return hipGetLastError();
}

template <int NB>
static hipError_t launch_nb(const PersistentArgs& A, int waves, int grid, hipStream_t st) {
  if (waves == 4) return launch_t<NB, 4>(A, grid, st);
  if constexpr (NB <= 4) {
    if (waves == 16) return launch_t<NB, 16>(A, grid, st);
  }
  return launch_t<NB, 8>(A, grid, st);
}

hipError_t launch_persistent(const PersistentArgs& A, int waves, int grid, hipStream_t st) {
  switch (A.dd.nb) {
    case 1: return launch_nb<1>(A, waves, grid, st);
    case 2: return launch_nb<2>(A, waves, grid, st);
    case 3: return launch_nb<3>(A, waves, grid, st);
    case 4: return launch_nb<4>(A, waves, grid, st);
    case 5: return launch_nb<5>(A, waves, grid, st);
    case 6: return launch_nb<6>(A, waves, grid, st);
    case 7: return launch_nb<7>(A, waves, grid, st);
    case 8: return launch_nb<8>(A, waves, grid, st);
    default: return hipErrorInvalidValue;
  }
}

// ---------------------------------------------------------------------------
// Host side.  Everything a batched stage-two run needs that depends only on
// the session's resident data, stage one's results and the launch shape --
// the search descriptors in work order (longest gene first), the count
// tails, the warm-start features, the device buffers and the captured
// queue-reset + persistent launch -- is built on the first run of a shape and
// kept in the session; a run then launches the network (learned engine) and
// replays the graph.
struct StageTwoPlan {
  int waves = 0;
  bool learned = false;
  double warm_step = 0.0, warm_xtol = 0.0;
  size_t n = 0;
  std::vector<uint32_t> idx_of;  // search -> kept index
  DevBuf<DevSearch> dsearch;
  DevBuf<uint32_t> dorder;
  DevBuf<unsigned> dqueue;
  DevBuf<double> dtv, dtm, dscr;
  DevBuf<float> dfeat, dpred;
  PinnedBuf<DevSearchOut> hout{false};  // the kernel writes each search's result here
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~StageTwoPlan() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

static std::shared_ptr<StageTwoPlan> build_stage_two_plan(const Prepared& P, Resident& R,
                                                          int waves, int grid,
                                                          const LearnedStart* learned) {
  const size_t G = P.kept.size(), nb = P.n_coef, k = P.k;
  const nebula_params_t& p = P.params;
  auto plan = std::make_shared<StageTwoPlan>();
  StageTwoPlan& S = *plan;
  S.waves = waves;
  S.learned = learned != nullptr;
  S.warm_step = learned ? learned->step : 0.0;
  S.warm_xtol = learned ? learned->xtol_rel : 0.0;
  std::vector<DevSearch> hs;
  std::vector<double> tv, tm;
  for (size_t i = 0; i < G; ++i) {
    if (P.plans[i].refit != 1) continue;
    const GenePlan& pl = P.plans[i];
    const GeneInfo& gi = P.genes[i];
    DevSearch d{};
    d.gene = (uint32_t)i;
    d.ord = pl.ord;
    d.start[0] = pl.sigma;
    d.start[1] = pl.gamma;
    for (size_t j = 0; j < nb; ++j) d.beta_start[j] = pl.beta_start[j];
    d.tail_off = tv.size();
    d.tail_len = (uint32_t)gi.tail.value.size();
    tv.insert(tv.end(), gi.tail.value.begin(), gi.tail.value.end());
    tm.insert(tm.end(), gi.tail.mult.begin(), gi.tail.mult.end());
    d.n_pos = gi.tail.n_pos;
    d.n_one = gi.tail.n_one;
    d.n_two = gi.tail.n_two;
    hs.push_back(d);
    S.idx_of.push_back((uint32_t)i);
  }
  const size_t n = hs.size();
  S.n = n;
  // longest first: work per evaluation ~ cells + positive counts
  std::vector<uint32_t> order(n);
  std::iota(order.begin(), order.end(), 0u);
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    const size_t na = P.gene_ptr[hs[a].gene + 1] - P.gene_ptr[hs[a].gene];
    const size_t nb2 = P.gene_ptr[hs[b].gene + 1] - P.gene_ptr[hs[b].gene];
    return na > nb2;
  });
  S.dsearch.upload(hs.data(), n, R.stream);
  S.dorder.upload(order.data(), n, R.stream);
  S.dtv.upload(tv.data(), tv.size(), R.stream);
  S.dtm.upload(tm.data(), tm.size(), R.stream);
  S.dqueue.alloc(1);
  S.hout.alloc(n);
  const size_t stride = persistent_stride((int)nb, (int)k, waves);
  const bool in_lds = stride * sizeof(double) <= kWorkLdsBytes;
  if (!in_lds) S.dscr.alloc(stride * (size_t)grid);
  PersistentArgs A;
  A.dd = R.view();
  A.search = S.dsearch.p;
  A.order = S.dorder.p;
  A.n_search = (int)n;
  A.queue = S.dqueue.p;
  A.tail_val = S.dtv.p;
  A.tail_mult = S.dtm.p;
  A.out = S.hout.dev;
  A.scratch = in_lds ? nullptr : S.dscr.p;
  A.stride = stride;
  A.work_in_lds = in_lds ? 1 : 0;
  A.knobs.eps = p.eps;
  A.knobs.noise_rel = 1e-8;
  A.knobs.max_iter = 50;
  A.knobs.max_backtrack = 10;
  A.knobs.decrement_stop = 1;
  A.lo[0] = p.min_subject;
  A.lo[1] = p.min_cell;
  A.hi[0] = p.max_subject;
  A.hi[1] = p.max_cell;
  A.intercept = (uint32_t)P.intercept;
  A.counters = R.counters.p;
  A.pred = nullptr;
  A.warm_step = 0.0;
  A.warm_xtol = 0.0;
  if (learned) {
    // features of the searched genes in search order; the network's
    // predictions land in dpred before each replay
    std::vector<float> hf(n * 12);
    for (size_t q = 0; q < n; ++q)
      std::memcpy(&hf[q * 12], &P.features[(size_t)S.idx_of[q] * 12], 12 * sizeof(float));
    S.dfeat.upload(hf.data(), hf.size(), R.stream);
    S.dpred.alloc(n * 2);
    A.pred = S.dpred.p;
    A.warm_step = learned->step;
    A.warm_xtol = learned->xtol_rel;
  }
  NH_HIP(hipStreamSynchronize(R.stream));
  NH_HIP(hipStreamBeginCapture(R.stream, hipStreamCaptureModeThreadLocal));
  hipError_t le = launch_persistent(A, waves, grid, R.stream);
  hipError_t ce = hipStreamEndCapture(R.stream, &S.graph);
  if (le != hipSuccess) throw HipError(std::string("launch_persistent: ") + hipGetErrorString(le));
  NH_HIP(ce);
  NH_HIP(hipGraphInstantiate(&S.exec, S.graph, nullptr, nullptr, 0));
  return plan;
}

void run_stage_two_batched(Prepared& P, Resident& R, const nebula_hip_options_t& o,
                           nebula_run_stats_t& st, const LearnedStart* learned) {
  const size_t G = P.kept.size(), nb = P.n_coef;
  size_t n = 0;
  for (size_t i = 0; i < G; ++i)
    if (P.plans[i].refit == 1) {
      P.plans[i].searched = true;
      ++n;
    }
  st.genes_searched = n;
  if (n == 0) return;
  int dev_cus = 0;
  NH_HIP(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, R.device));
  // tile shape: one gene per workgroup; wide workgroups when genes are few
  // waves per gene: the tuned or requested size; otherwise 16 (a whole CU per
  // gene) when there are no more genes than CUs, else 8 (two genes per CU
  // overlap one's serial steps with the other's sweeps)
  int waves = 8;
  if (o.wg_size >= 1024)
    waves = 16;
  else if (o.wg_size > 0 && o.wg_size <= 256)
    waves = 4;
  else if (o.wg_size == 0 && (int)n <= dev_cus)
    waves = 16;
  waves = effective_waves((int)nb, waves);
  const int per_cu = waves == 16 ? 1 : (waves == 8 ? 2 : 4);
  const int grid = (int)std::min<size_t>(n, (size_t)std::max(1, dev_cus * per_cu));
  std::shared_ptr<StageTwoPlan> plan;
  for (auto& c : R.stage_two_plans)
    if (c->waves == waves && c->learned == (learned != nullptr) &&
        (!learned || (c->warm_step == learned->step && c->warm_xtol == learned->xtol_rel)))
      plan = c;
  if (!plan) {
    const auto t0 = std::chrono::steady_clock::now();
    plan = build_stage_two_plan(P, R, waves, grid, learned);
    R.stage_two_plans.push_back(plan);
    st.t_plan = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  StageTwoPlan& S = *plan;
  if (S.n != n) throw HipError("stage-two plan out of date");
  if (learned) NH_HIP(launch_mlp_predict(learned->params, S.dfeat.p, (int)n, S.dpred.p, R.stream));
  NH_HIP(hipEventRecord(R.ev0, R.stream));
  NH_HIP(hipGraphLaunch(S.exec, R.stream));
  NH_HIP(hipEventRecord(R.ev1, R.stream));
  NH_HIP(hipStreamSynchronize(R.stream));
  float ms = 0.0f;
  NH_HIP(hipEventElapsedTime(&ms, R.ev0, R.ev1));
  R.kernel_ms += ms;
  R.launches += 1;
  st.rounds = 1;
  const DevSearchOut* ho = S.hout.p;
  for (size_t q = 0; q < n; ++q) {
    GenePlan& pl = P.plans[S.idx_of[q]];
    pl.found = ho[q].found != 0;
    pl.failed = ho[q].failed != 0;
    pl.v[0] = ho[q].x[0];
    pl.v[1] = ho[q].x[1];
    pl.objective = ho[q].f;
    pl.evaluations = ho[q].evals;
    pl.warm_rejected = ho[q].warm_rejected != 0;
    st.evaluations += (uint64_t)ho[q].evals;
  }
}

}  // namespace nh

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
// Search semantics are the host plain engine's (SearchControl, shared with
// it): BOBYQA over both components or NEBULA-LN's one-component simplex
// search, the order-one retry and the abandon-on-non-finite rules of
// StageTwoSearch (nebula.rs:1568-1951), inner fits warm-started from the
// gene's first finite fit (stage_two.rs:36-43).
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
#include "search_control.hpp"
#include "pml_fit.hpp"
#include "persistent.hpp"
#include "mlp.hpp"
#include "lgamma_pos.hpp"
#include "resident.hpp"

namespace nh {

// Largest first-order move of any coordinate (beta, log w) a warm start takes
// from the anchor fit; beyond it the start stays at the anchor's optimum.
// Starts moved by at most kSkipCap skip the FP32 loop.
#ifndef NH_PREDICT_CAP
#define NH_PREDICT_CAP 0.25
#endif
#ifndef NH_SKIP_CAP
#define NH_SKIP_CAP 0.003
#endif
constexpr double kPredictCap = NH_PREDICT_CAP;
constexpr double kSkipCap = NH_SKIP_CAP;

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

// The search control runs on one lane between fits; kept out of line so the
// sweep loop's register allocation is not shaped by it.  Its state is the
// kernels' __shared__ SearchControl (ds_ accesses, not flat).
__device__ __noinline__ static void ctl_begin(SearchControl* c, const DevSearch* S,
                                              const float* pred, const PersistentArgs* A) {
  NH_ASSUME_LDS(c);
  c->begin((int)S->kind, (int)S->ord, S->start, S->kind == kSearchBoth ? pred : nullptr, A->lo,
           A->hi, A->warm_step, A->warm_xtol);
}
__device__ __noinline__ static int ctl_tell(SearchControl* c, double v) {
  NH_ASSUME_LDS(c);
  return c->tell(v);
}
__device__ __noinline__ static void ctl_out(const SearchControl* c, DevSearchOut* o) {
  NH_ASSUME_LDS(c);
  double f;
  int ord;
  c->result(o->x, f, ord);
  o->f = f;
  o->found = c->found;
  o->failed = c->failed;
  o->evals = c->evals;
  o->ord = ord;
  o->warm_rejected = c->rejected;
}

template <int NB, int WAVES>
__global__ __launch_bounds__(WAVES * 64) __attribute__((amdgpu_waves_per_eu(kWavesPerEu<NB>)))
void stage_two_kernel(PersistentArgs A) {
  constexpr int NT = WAVES * 64;
  __shared__ FitLds<NB, WAVES> sh;
  __shared__ Fitter<NB, WAVES> f;  // the gene's fit object
  __shared__ SearchControl ctl;
  __shared__ double warm_beta[NB];
  __shared__ double warm_jb[2 * NB];  // d beta* / d (subject, gamma) at the anchor
  __shared__ double warm_vb[NB * NB];  // the anchor fit's fixed-effect curvature block
  __shared__ double red[NT / 64];
  __shared__ int s_slot, s_ok, s_pred;
  __shared__ double s_subject, s_cell, s_anchor[2];
  extern __shared__ double nh_s2_dyn[];
  const int tid = threadIdx.x;
  const DeviceData& dd = A.dd;
  const int k = (int)dd.k;
  FitWork<NB> wk;
  double* base = A.work_in_lds ? nh_s2_dyn : A.scratch + (size_t)blockIdx.x * A.stride;
  wk.carve(base, k, WAVES);
  double* warm_lw = base + FitWork<NB>::doubles(k, WAVES);
  double* warm_jl = warm_lw + k;  // 2 x k: d log w* / d (subject, gamma)
  double* warm_vwb = warm_jl + 2 * (size_t)k;  // k x NB: the anchor fit's vwb blocks
  for (;;) {
    if (tid == 0) s_slot = (int)atomicAdd(A.queue, 1u);
    __syncthreads();
    const int slot = s_slot;
    __syncthreads();
    if (slot >= A.n_search) break;
    const DevSearch S = A.search[A.order[slot]];
    if (tid == 0) s_pred = 0;
    if (tid == 0) {
      ctl_begin(&ctl, &A.search[A.order[slot]],
                A.pred ? A.pred + 2 * (size_t)A.order[slot] : nullptr, &A);
      // one fit object per gene: members set here, the variance components
      // once per evaluation
      f.init(dd, &sh, wk, A.counters, S.gene, 0.0, 0.0, 0.0);
    }
    __syncthreads();
    while (ctl.active) {
      if (tid == 0) ctl.point(s_subject, s_cell);
      __syncthreads();
      const double subject = s_subject, cell = s_cell;
      const int have_warm = ctl.have_warm;
      if (tid == 0) {
        // published by the barrier after the start values
        const double exps = exp(subject);
        f.gamma = cell;
        f.alpha = 1.0 / (exps - 1.0);
        f.lambda = 1.0 / (sqrt(exps) * (exps - 1.0));
      }
      // A warm fit starts from the anchor fit's optimum moved to first order
      // to this (subject, gamma), unless some coordinate would move by more
      // than kPredictCap (where the linear model is not trusted).
      const double dsub = subject - s_anchor[0], dcell = cell - s_anchor[1];
      double move = 0.0;
      if (have_warm && s_pred) {
        if (tid < NB) move = fabs(warm_jb[tid] * dsub + warm_jb[NB + tid] * dcell);
        for (int s = tid; s < k; s += NT)
          move = fmax(move, fabs(warm_jl[s] * dsub + warm_jl[k + s] * dcell));
      }
      const bool pred = have_warm && s_pred && !__syncthreads_or(move > kPredictCap);
      const bool near = pred && !__syncthreads_or(move > kSkipCap);
      if (tid < NB) {
        double b = S.beta_start[tid];
        if (tid == (int)A.intercept) b -= subject / 2.0;
        if (have_warm)
          b = warm_beta[tid] + (pred ? warm_jb[tid] * dsub + warm_jb[NB + tid] * dcell : 0.0);
        sh.beta[tid] = b;
      }
      for (int s = tid; s < k; s += NT)
        wk.lw[s] = have_warm ? warm_lw[s] + (pred ? warm_jl[s] * dsub + warm_jl[k + s] * dcell
                                                  : 0.0)
                             : 0.0;
      __syncthreads();
      // a predicted start goes straight to the FP64 finish, with the anchor
      // fit's curvature blocks
      const FitResult r = f.run(0, (unsigned)ctl.point_order(), A.knobs, ctl.evals > 0,
                                !have_warm, near ? warm_vwb : nullptr, near ? warm_vb : nullptr);
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
        s_ok = ctl_tell(&ctl, v);
      }
      __syncthreads();
      if (s_ok) {
        if (tid < NB) warm_beta[tid] = sh.beta[tid];
        for (int s = tid; s < k; s += NT) warm_lw[s] = wk.lw[s];
        const bool can = A.predict && r.predictable;
        if (can) {
          f.sensitivity(r.side, subject, warm_jb, warm_jl);
          const double* hs = f.side(r.side);
          for (int i = tid; i < k * NB; i += NT)
            warm_vwb[i] = hs[(i / NB) * Fitter<NB, WAVES>::SUB + 3 + i % NB];
          if (tid < NB * NB) warm_vb[tid] = sh.vb[r.side][tid];
        }
        if (tid == 0) {
          s_pred = can ? 1 : 0;
          s_anchor[0] = subject;
          s_anchor[1] = cell;
        }
      }
      __syncthreads();
    }
    if (tid == 0) ctl_out(&ctl, &A.out[A.order[slot]]);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// The persistent kernel's search control (the same ctl_begin / ctl_tell /
// ctl_out) driven on the device by a built-in objective instead of a fit, so
// the state machine's branches can be checked against the reference's cases.
// Objectives of (subject s, cell c): 0 (s-0.3)^2 + 0.01 (c-4)^2; 1 the valley
// (s+c-1)^2; 2 s; 3 (s-0.3)^2 + 2; 4 a shallow well at the lower bound and a
// deeper one at 0.05.  Injection: calls inj_from..inj_to (one-based) at an
// order >= inj_order return +inf (inj_kind 0), NaN (1) or the profile
// objective of a fit whose Laplace correction fell below -1 (2).
__device__ static double search_test_objective(int problem, double s, double c) {
  NH_STRICT_FP
  switch (problem) {
    case 0: {
      const double a = s - 0.3, b = c - 4.0;
      return a * a + 0.01 * (b * b);
    }
    case 1: {
      const double a = s + c - 1.0;
      return a * a;
    }
    case 2:
      return s;
    case 3: {
      const double a = s - 0.3;
      return a * a + 2.0;
    }
    default: {
      const double a = (s - 1e-4) / 2e-3, b = (s - 0.05) / 0.02;
      return -0.5 * exp(-(a * a)) - 1.0 * exp(-(b * b));
    }
  }
}

__global__ void search_trace_kernel(int problem, int kind, int ord, const double* start,
                                    const double* lo, const double* hi, int inj_from, int inj_to,
                                    int inj_order, int inj_kind, double* trace, int max_trace,
                                    DevSearchOut* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  __shared__ SearchControl ctl;
  DevSearch S{};
  S.kind = (uint32_t)kind;
  S.ord = (uint32_t)ord;
  S.start[0] = start[0];
  S.start[1] = start[1];
  PersistentArgs A{};
  A.lo[0] = lo[0];
  A.lo[1] = lo[1];
  A.hi[0] = hi[0];
  A.hi[1] = hi[1];
  ctl_begin(&ctl, &S, nullptr, &A);
  int call = 0;
  while (ctl.active) {
    double sub, cell;
    ctl.point(sub, cell);
    const int order = ctl.point_order();
    ++call;
    double v = search_test_objective(problem, sub, cell);
    if (call >= inj_from && call <= inj_to && order >= inj_order) {
      if (inj_kind == 0) {
        v = INFINITY;
      } else if (inj_kind == 1) {
        v = NAN;
      } else {
        FitResult r{};
        r.ll = -v;
        r.ll_prev = -v;
        r.second = -1.5;
        v = device_objective(sub, cell, r, 100.0, 4.0, 0.0, 10.0, 3.0, 2.0);
      }
    }
    if (call <= max_trace) {
      double* row = trace + (size_t)(call - 1) * 4;
      row[0] = sub;
      row[1] = cell;
      row[2] = (double)order;
      row[3] = v;
    }
    ctl_tell(&ctl, v);
  }
  ctl_out(&ctl, out);
}

hipError_t launch_search_trace(int problem, int kind, int ord, const double* start,
                               const double* lo, const double* hi, int inj_from, int inj_to,
                               int inj_order, int inj_kind, double* trace, int max_trace,
                               DevSearchOut* out, hipStream_t st) {
  hipLaunchKernelGGL(search_trace_kernel, dim3(1), dim3(64), 0, st, problem, kind, ord, start, lo,
                     hi, inj_from, inj_to, inj_order, inj_kind, trace, max_trace, out);
  return hipGetLastError();
}

__global__ void reset_queue_kernel(unsigned* q) {
  if (threadIdx.x == 0) *q = 0u;
}

// per workgroup: the fit's work arrays, the anchor's log w, its two
// sensitivity columns and its vwb curvature blocks
size_t persistent_stride(int nb, int k, int waves) {
  return fit_work_doubles(nb, k, waves) + (3 + (size_t)nb) * (size_t)k;
}

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
    if (P.plans[i].refit == 0) continue;
    const GenePlan& pl = P.plans[i];
    const GeneInfo& gi = P.genes[i];
    DevSearch d{};
    d.gene = (uint32_t)i;
    d.ord = pl.ord;
    d.kind = pl.refit == 1 ? kSearchBoth : kSearchSubject;
    d.start[0] = pl.sigma;
    d.start[1] = pl.gamma;
    for (size_t j = 0; j < nb; ++j) d.beta_start[j] = pl.beta_one[j];
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
  A.predict = 1;
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
    if (P.plans[i].refit != 0) {
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

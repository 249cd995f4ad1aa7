// Stage one on the device: the sums behind the marginal Poisson-gamma
// likelihood, its gradient and its Hessian (edge-rs src/sc/ptmg.rs:661-1252:
// linear_predictor, evaluate, evaluate_hessian / hessian_subject), one
// workgroup per (gene, point) request, FP64 throughout.
//
// Three passes over the gene's cells, each the coalesced 64-cell row walk of
// pml_fit.hpp with the positive counts merged through LDS:
//   1. e = exp(eta) per cell: per-subject sum of e, per-gene sum of y * eta
//      (exp, log and 1/w by the FP64 sweep's table-driven versions);
//   2. with ym = (Y_s + alpha) / (sum e + lambda) per subject: the gradient
//      sums (and the Hessian's per-subject vectors when asked);
//   3. Hessian only: the Gram  sum_i W_i x_i x_i'  with
//      W_i = e_i (ymm_s d42c_s - ym_s g_i + ym_s^2 g_i e_i / w_i), the single
//      subject-weighted combination the three curvature blocks of
//      hessian_subject enter in (ptmg.rs:1159-1171, 1186-1223).
// The host turns the sums into value, gradient and Hessian
// (gradient_parts, finish_gradient and the Hessian assembly).
#include <hip/hip_runtime.h>

#include "device_common.hpp"
#include "pml_fit.hpp"
#include "stage_kernels.hpp"

namespace nh {

template <int NB>
struct PtmgDims {
  static constexpr int NTRI = NB * (NB + 1) / 2;
  // per subject: csx, acc, hb2, b2, xf[NB], d41[NB], t1[NB], t2[NB]
  static constexpr int PS = 4 + 4 * NB;
  // per gene: linear, total, slpey, gstar, sum_tempa, sum_gt, gram[NTRI]
  static constexpr int PG = 6 + NTRI;
};

template <int NB>
__host__ __device__ size_t ptmg_out_doubles_t(int k) {
  return (size_t)k * PtmgDims<NB>::PS + PtmgDims<NB>::PG;
}

size_t ptmg_out_doubles(int nb, int k) {
  switch (nb) {
    case 1: return ptmg_out_doubles_t<1>(k);
    case 2: return ptmg_out_doubles_t<2>(k);
    case 3: return ptmg_out_doubles_t<3>(k);
    case 4: return ptmg_out_doubles_t<4>(k);
    case 5: return ptmg_out_doubles_t<5>(k);
    case 6: return ptmg_out_doubles_t<6>(k);
    case 7: return ptmg_out_doubles_t<7>(k);
    default: return ptmg_out_doubles_t<8>(k);
  }
}

size_t ptmg_scratch_doubles(int nb, int k, int waves) {
  // wave partials (k + waves) * PS, per-wave gene partials waves * PG
  const int ps = 4 + 4 * nb, pg = 6 + nb * (nb + 1) / 2;
  return (size_t)(k + waves) * ps + (size_t)waves * pg;
}

template <int NB, int WAVES>
struct PtmgWalk {
  static constexpr int NT = WAVES * kWave;
  const DeviceData& dd;
  uint64_t g0, g1;
  uint32_t blk;
  int k, wave, lane, tid;
  double (*ysc)[kWave];
  uint64_t* ip0;
  uint32_t* s0;

  __device__ uint32_t bb(int w) const { return min(dd.n_cells, (uint32_t)w * blk); }
  __device__ uint32_t be(int w) const { return min(dd.n_cells, (uint32_t)(w + 1) * blk); }

  __device__ void setup() {
    const uint32_t cb = bb(wave);
    uint64_t lo = g0, hi = g1;
    while (hi - lo > (uint64_t)kWave) {
      uint64_t step = (hi - lo + kWave - 1) / kWave;
      uint64_t pos = lo + (uint64_t)lane * step;
      bool pred = pos < hi && dd.gcell[pos] < cb;
      int cnt = popc64(wave_ballot(pred));
      if (cnt == 0) {
        hi = lo;
      } else {
        uint64_t nlo = lo + (uint64_t)(cnt - 1) * step + 1;
        uint64_t nhi = lo + (uint64_t)cnt * step;
        hi = nhi < hi ? nhi : hi;
        lo = nlo;
      }
    }
    uint64_t pos = lo + (uint64_t)lane;
    bool pred = pos < hi && dd.gcell[pos] < cb;
    uint64_t ip = lo + (uint64_t)popc64(wave_ballot(pred));
    int a = 0, b = k;
    while (b - a > 1) {
      int m = (a + b) / 2;
      if (dd.fid[m] <= cb) a = m; else b = m;
    }
    if (lane == 0) {
      ip0[wave] = ip;
      s0[wave] = (uint32_t)a;
    }
    __syncthreads();
  }

  // Calls cell(s, cell, y, row) for every cell of the wave's block (active
  // lanes only; row = the cell's packed design row, R doubles) and flush(s)
  // after each subject segment, uniformly across the wave.  The next 64 CSR
  // entries are held in registers and reloaded as soon as a tile has taken
  // its own, so their loads overlap the tile's arithmetic; every load is
  // issued at a clamped in-bounds index through global-address-space pointers.
  template <class Cell, class Flush>
  __device__ void walk(Cell&& cell, Flush&& flush) {
    constexpr int R = Dims<NB>::R;
    const uint32_t cb = bb(wave), ce = be(wave);
    if (cb >= ce) return;
    const gconst_double* rows = (const gconst_double*)dd.row64;
    const gconst_u32* gcell = (const gconst_u32*)dd.gcell;
    const gconst_double* gy = (const gconst_double*)dd.gy64;
    const uint64_t e_last = g1 > 0 ? g1 - 1 : 0;
    uint64_t ip = ip0[wave];
    int s = (int)s0[wave];
    uint64_t e = ip + (uint64_t)lane;
    uint64_t ec = e < g1 ? e : e_last;
    uint32_t wc = gcell[ec];
    double wy = gy[ec];
    wc = e < g1 ? wc : 0xffffffffu;
    for (; s < k && dd.fid[s] < ce; ++s) {
      const uint32_t lo = max(dd.fid[s], cb), hi = min(dd.fid[s + 1], ce);
      for (uint32_t base = lo; base < hi; base += kWave) {
        const uint32_t lim = min(base + (uint32_t)kWave, hi);
        const uint32_t cl = base + (uint32_t)lane;
        double xrow[R];
        load_row<double, R>(rows, cl < lim ? cl : lim - 1, xrow);
        // the window's entries inside [base, lim) form a prefix of it
        const bool in = wc < lim;
        const int m = popc64(wave_ballot(in));
        ysc[wave][lane] = 0.0;
        wave_lds_sync();
        if (in) ysc[wave][wc - base] = wy;
        wave_lds_sync();
        const double y = ysc[wave][lane];
        ip += (uint64_t)m;
        e = ip + (uint64_t)lane;
        ec = e < g1 ? e : e_last;
        wc = gcell[ec];
        wy = gy[ec];
        wc = e < g1 ? wc : 0xffffffffu;
        if (cl < lim) cell(s, cl, y, xrow);
      }
      flush(s);
    }
  }

  // Sum over a subject's wave partials (slot s + w), waves in order.
  __device__ double subject_total(const double* part, int stride, int s, int v) const {
    const uint32_t f0 = dd.fid[s], f1 = dd.fid[s + 1];
    const int w0 = (int)(f0 / blk), w1 = (int)((f1 - 1) / blk);
    double t = 0.0;
    for (int w = w0; w <= w1; ++w) t += part[(size_t)(s + w) * stride + v];
    return t;
  }
};

template <int NB, int WAVES>
__global__ __launch_bounds__(WAVES * 64) void ptmg_kernel(DeviceData dd, const PtmgRequest* req,
                                                          const double* params, double* out,
                                                          double* scratch, size_t sstride,
                                                          FitCounters* ctr) {
  using D = PtmgDims<NB>;
  constexpr int PS = D::PS, NTRI = D::NTRI;
  __shared__ double ysc[WAVES][kWave];
  __shared__ uint64_t ip0[WAVES];
  __shared__ uint32_t s0[WAVES];
  __shared__ double gpart[WAVES][6 + NTRI];
  __shared__ double f64tab[82];  // the FP64 sweep's exp / log tables
  const int q = blockIdx.x;
  const PtmgRequest r = req[q];
  const int k = (int)dd.k;
  const double* x = params + (size_t)q * (NB + 2);
  double* o = out + (size_t)q * ptmg_out_doubles_t<NB>(k);
  double* part = scratch + (size_t)q * sstride;

  PtmgWalk<NB, WAVES> W{dd};
  W.g0 = dd.gptr[r.gene];
  W.g1 = dd.gptr[r.gene + 1];
  {
    uint32_t per = (dd.n_cells + WAVES - 1) / WAVES;
    W.blk = (per + kWave - 1) / kWave * kWave;
  }
  W.k = k;
  W.tid = threadIdx.x;
  W.wave = threadIdx.x / kWave;
  W.lane = threadIdx.x % kWave;
  W.ysc = ysc;
  W.ip0 = ip0;
  W.s0 = s0;
  fill_f64_tables(f64tab, (int)threadIdx.x);  // setup() ends with a barrier
  W.setup();
  const lds_double* tab = (const lds_double*)f64tab;

  double beta[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) beta[j] = x[j];
  const double sigma = x[NB], gamma = x[NB + 1];
  const double exps = exp(sigma);
  const double alpha = 1.0 / (exps - 1.0);
  const double lambda = alpha / sqrt(exps);
  const double* tot_y = dd.gtot + (size_t)r.gene * k;

  // eta = log offset + x'beta from a packed row (non-intercept columns, then
  // the log offset), columns added in design order as x64 holds them
  const int icpt = (int)dd.icpt;
  auto row_eta = [&](const double* prow, double xr[NB]) {
    double eta = prow[NB - 1];
    int q = 0;
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      if (j == icpt) {
        xr[j] = 1.0;
        eta += beta[j];
      } else {
        xr[j] = prow[q];
        eta += prow[q] * beta[j];
        ++q;
      }
    }
    return eta;
  };

  // ---- pass 1: subject sums of e, gene sum of y * eta ----
  {
    double csx = 0.0, lin = 0.0;
    W.walk(
        [&](int, uint32_t, double y, const double* prow) {
          double xr[NB];
          const double eta = row_eta(prow, xr);
          if (y != 0.0) lin += eta * y;
          csx += fast_exp(eta, tab);
        },
        [&](int s) {
          const double t = wave_sum(csx);
          if (W.lane == 0) part[(size_t)(s + W.wave) * PS] = t;
          csx = 0.0;
        });
    const double l = wave_sum(lin);
    if (W.lane == 0) gpart[W.wave][0] = l;
    __syncthreads();
    for (int s = threadIdx.x; s < k; s += WAVES * 64) o[(size_t)s * PS + 0] = W.subject_total(part, PS, s, 0);
    if (threadIdx.x == 0) {
      double l2 = 0.0;
      for (int w = 0; w < WAVES; ++w) l2 += gpart[w][0];
      o[(size_t)k * PS + 0] = l2;
    }
    __syncthreads();
  }

  // ---- pass 2: gradient (and Hessian vector) sums ----
  const bool hess = r.hess != 0;
  {
    double acc[PS - 1];
    double gs[5];
#pragma unroll
    for (int v = 0; v < PS - 1; ++v) acc[v] = 0.0;
#pragma unroll
    for (int v = 0; v < 5; ++v) gs[v] = 0.0;
    double ym = 0.0;
    int cur_s = -1;
    W.walk(
        [&](int s, uint32_t, double y, const double* prow) {
          if (s != cur_s) {
            cur_s = s;
            ym = (tot_y[s] + alpha) / (o[(size_t)s * PS + 0] + lambda);
          }
          double xr[NB];
          const double eta = row_eta(prow, xr);
          const double e = fast_exp(eta, tab);
          const double we = ym * e;
          const double w = we + gamma;
          const double lw = fast_log(w, tab);
          const double iw = fast_rcp(w);
          const double g = (gamma + y) * iw;
          gs[0] += we - (gamma + y) * lw;  // total
          gs[1] += lw;                     // slpey
          gs[2] += g;                      // gstar
          acc[0] += g * e;                 // acc
          if (hess) {
            const double gt = g * iw;
            const double gte = gt * e;
            gs[3] += iw;
            gs[4] += gt;
            acc[1] += gte - e * iw;  // hb2
            acc[2] += gte * e;       // b2
            const double f1 = gt - iw;
#pragma unroll
            for (int a = 0; a < NB; ++a) {
              const double xe = xr[a] * e;
              acc[3 + a] += xe;
              acc[3 + NB + a] += g * xe;
              acc[3 + 2 * NB + a] += xe * f1;
              acc[3 + 3 * NB + a] += xe * gte;
            }
          } else {
#pragma unroll
            for (int a = 0; a < NB; ++a) {
              const double xe = xr[a] * e;
              acc[3 + a] += xe;
              acc[3 + NB + a] += g * xe;
            }
          }
        },
        [&](int s) {
#pragma unroll
          for (int v = 0; v < PS - 1; ++v) {
            const double t = wave_sum(acc[v]);
            if (W.lane == 0) part[(size_t)(s + W.wave) * PS + 1 + v] = t;
            acc[v] = 0.0;
          }
        });
#pragma unroll
    for (int v = 0; v < 5; ++v) {
      const double t = wave_sum(gs[v]);
      if (W.lane == 0) gpart[W.wave][1 + v] = t;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < k * (PS - 1); idx += WAVES * 64) {
      const int s = idx / (PS - 1), v = 1 + idx % (PS - 1);
      o[(size_t)s * PS + v] = W.subject_total(part, PS, s, v);
    }
    if (threadIdx.x < 5) {
      double t = 0.0;
      for (int w = 0; w < WAVES; ++w) t += gpart[w][1 + threadIdx.x];
      o[(size_t)k * PS + 1 + threadIdx.x] = t;
    }
    __syncthreads();
  }

  // ---- pass 3: the weighted Gram of the Hessian's beta block ----
  if (hess) {
    double gr[NTRI];
#pragma unroll
    for (int v = 0; v < NTRI; ++v) gr[v] = 0.0;
    double ym = 0.0, ymmd = 0.0;
    int cur_s = -1;
    W.walk(
        [&](int s, uint32_t, double y, const double* prow) {
          if (s != cur_s) {
            cur_s = s;
            const double csx = o[(size_t)s * PS + 0];
            const double mu = csx + lambda;
            const double ystar = tot_y[s] + alpha;
            ym = ystar / mu;
            const double ymm = ystar / mu / mu;
            ymmd = ymm * (o[(size_t)s * PS + 1] - csx);
          }
          double xr[NB];
          const double eta = row_eta(prow, xr);
          const double e = fast_exp(eta, tab);
          const double w = ym * e + gamma;
          const double iw = fast_rcp(w);
          const double g = (gamma + y) * iw;
          const double gte = g * iw * e;
          const double wt = e * ymmd - ym * (g * e) + ym * ym * (gte * e);
          int idx = 0;
#pragma unroll
          for (int a = 0; a < NB; ++a) {
            const double xa = wt * xr[a];
#pragma unroll
            for (int b = a; b < NB; ++b) gr[idx++] += xa * xr[b];
          }
        },
        [&](int) {});
#pragma unroll
    for (int v = 0; v < NTRI; ++v) {
      const double t = wave_sum(gr[v]);
      if (W.lane == 0) gpart[W.wave][6 + v] = t;
    }
    __syncthreads();
    for (int v = threadIdx.x; v < NTRI; v += WAVES * 64) {
      double t = 0.0;
      for (int w = 0; w < WAVES; ++w) t += gpart[w][6 + v];
      o[(size_t)k * PS + 6 + v] = t;
    }
  }
  if (threadIdx.x == 0) {
    atomicAdd(&ctr->sweeps_f64, hess ? 3ull : 2ull);
    atomicAdd(&ctr->cells_f64, (unsigned long long)dd.n_cells * (hess ? 3ull : 2ull));
    atomicAdd(&ctr->nnz_f64, (unsigned long long)(W.g1 - W.g0) * (hess ? 3ull : 2ull));
  }
}

template <int NB>
static hipError_t launch_ptmg_nb(const PtmgLaunch& L, hipStream_t st) {
  hipLaunchKernelGGL((ptmg_kernel<NB, 8>), dim3(L.n_req), dim3(8 * 64), 0, st, L.dd, L.req,
                     L.params, L.out, L.scratch, L.stride, L.counters);
  return hipGetLastError();
}

hipError_t launch_ptmg(const PtmgLaunch& L, hipStream_t st) {
  if (L.n_req == 0) return hipSuccess;
  switch (L.dd.nb) {
    case 1: return launch_ptmg_nb<1>(L, st);
    case 2: return launch_ptmg_nb<2>(L, st);
    case 3: return launch_ptmg_nb<3>(L, st);
    case 4: return launch_ptmg_nb<4>(L, st);
    case 5: return launch_ptmg_nb<5>(L, st);
    case 6: return launch_ptmg_nb<6>(L, st);
    case 7: return launch_ptmg_nb<7>(L, st);
    case 8: return launch_ptmg_nb<8>(L, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace nh

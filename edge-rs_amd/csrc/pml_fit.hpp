// One penalised NB fit (edge-rs opt_pml, src/sc/pml.rs:1614-1800) per HIP
// workgroup, for gfx950.
//
// Work split.  The workgroup's W waves cut the gene's cells into W contiguous
// blocks (64-cell aligned).  A wave walks its block subject by subject in rows
// of 64 consecutive cells, so design rows and log offsets arrive as coalesced
// loads.  The gene's positive counts are merged in on the fly: the wave keeps
// a cursor into the gene's sorted CSR row, ballots the entries that fall in the
// current row and scatters their counts into a 64-slot LDS row, so every cell
// is visited once with its own count.  That one pass is edge-rs's dense sweep
// at y = 0 plus its sparse correction (pml.rs:1360-1363, exact because every
// per-cell term is linear in y), without the second visit to positive cells.
//
// Precision.  Per cell the exp, log and reciprocal run in CT: float for the
// bulk Newton sweeps, double for the finishing sweeps.  Every accumulation is
// FP64 (on CDNA4 an FP64 FMA issues at the FP32 rate), so the long sums over
// 1e5..1e6 cells carry no summation error and the Schur complement
// vb - sum vwb vwb'/vw is formed from exact sums.  What FP32 leaves is the
// per-cell rounding of exp/log, which the FP64 finish removes: one f64 Newton
// step from the f32 optimum and the log-likelihood and log-determinant at the
// stepped point (edge-rs newton_finish, src/sc/pml.rs:807-1090).  If that step
// does not settle the fit, the whole Newton loop reruns in f64 from the f32
// point with opt_pml's exact control flow (opt_pml_from, pml.rs:637).
//
// Reductions.  Per subject, each wave reduces its lane partials with a fixed
// butterfly and lane 0 stores them; the workgroup then adds the (at most two
// or three) wave partials of a subject in wave order.  Results do not depend
// on timing.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "device_common.hpp"

namespace nh {

struct DeviceData {
  const float* x32;      // n_cells * p4, centred design, zero-padded columns
  const double* x64;     // n_cells * nb
  const float* off32;    // log offset
  const double* off64;
  const uint32_t* fid;   // k + 1 subject boundaries
  const uint64_t* gptr;  // n_genes + 1 CSR row starts
  const uint32_t* gcell; // positive cells, increasing within a gene
  const float* gy32;     // positive counts
  const double* gy64;
  const double* gtot;    // n_genes * k subject totals
  uint32_t n_cells, nb, p4, k;
};

struct FitKnobs {
  double eps;        // nebula's absolute stopping tolerance
  double noise_rel;  // f32 resolution floor relative to |loglik|
  int max_iter;      // Newton budget (50)
  int max_backtrack; // 10
};

// Counters every fit adds to (device-wide, relaxed atomics).
struct FitCounters {
  unsigned long long sweeps_f32, sweeps_f64;
  unsigned long long cells_f32, nnz_f32;  // summed over FP32 sweeps
  unsigned long long cells_f64, nnz_f64;  // summed over FP64 sweeps
};

enum SweepMode { SW_FULL = 0, SW_VALUE = 1, SW_LAPLACE = 2 };

template <int NB>
struct Dims {
  static constexpr int NTRI = NB * (NB + 1) / 2;
  static constexpr int V = 3 + 2 * NB + NTRI;  // resid, weight, llp, score, first, second
  static constexpr int SUB = 3 + NB;           // dw, vw, w, vwb[NB]
};

// Per-workgroup scratch in global memory.
template <int NB>
struct FitScratch {
  double* part;     // (k + W) * V   wave partials, slot (s + w)
  double* tot;      // k * V         per-subject totals
  double* sub[2];   // k * SUB       per side: dw, vw, w, vwb
  double* llsub;    // k
  double* lw;       // k  current log w
  double* new_lw;   // k  trial
  double* step_lw;  // k
  double* damp_lw;  // k
  double* vw_step;  // k  vw of the last Newton step (log-det, pml.rs:1767)
  __device__ void carve(double* base, int k, int waves) {
    constexpr int V = Dims<NB>::V;
    part = base;
    base += (size_t)(k + waves) * V;
    tot = base;
    base += (size_t)k * V;
    sub[0] = base;
    base += (size_t)k * Dims<NB>::SUB;
    sub[1] = base;
    base += (size_t)k * Dims<NB>::SUB;
    llsub = base;
    base += k;
    lw = base;
    base += k;
    new_lw = base;
    base += k;
    step_lw = base;
    base += k;
    damp_lw = base;
    base += k;
    vw_step = base;
  }
  __host__ __device__ static size_t doubles(int k, int waves) {
    return (size_t)(k + waves) * Dims<NB>::V + (size_t)k * Dims<NB>::V +
           2 * (size_t)k * Dims<NB>::SUB + 6 * (size_t)k;
  }
};

template <int NB, int WAVES>
struct FitLds {
  double beta[NB], new_beta[NB], step_beta[NB], damp_beta[NB];
  double db[2][NB];
  double vb[2][NB * NB];
  double vb2[NB * NB];
  double rhs[NB];
  double ll[2];
  double ysc[WAVES][kWave];
  uint64_t ip0[WAVES];
  uint32_t s0[WAVES];
  double scal[8];
  int flag[4];
};

// What one fit hands back.
struct FitResult {
  double ll, ll_prev, log_det, second;
  int iterations, backtracks, finish_steps, exact;  // exact: f64 Newton loop ran
};

// ---------------------------------------------------------------------------
// Eigen-style pivoted LDL' solve of a small symmetric system, in place
// (src/sc/pml.rs:1940-2021).  One thread.
template <int NB>
__device__ void ldlt_solve_dev(double* a, double* b) {
  const int n = NB;
  int perm[NB];
  double tmp[NB];
  for (int st = 0; st < n; ++st) {
    int piv = st;
    double best = fabs(a[st * n + st]);
    for (int i = st + 1; i < n; ++i) {
      double v = fabs(a[i * n + i]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
    perm[st] = piv;
    if (piv != st) {
      for (int j = 0; j < st; ++j) {
        double t = a[st * n + j];
        a[st * n + j] = a[piv * n + j];
        a[piv * n + j] = t;
      }
      for (int i = piv + 1; i < n; ++i) {
        double t = a[i * n + st];
        a[i * n + st] = a[i * n + piv];
        a[i * n + piv] = t;
      }
      double t = a[st * n + st];
      a[st * n + st] = a[piv * n + piv];
      a[piv * n + piv] = t;
      for (int i = st + 1; i < piv; ++i) {
        double u = a[i * n + st];
        a[i * n + st] = a[piv * n + i];
        a[piv * n + i] = u;
      }
    }
    if (st > 0) {
      for (int j = 0; j < st; ++j) tmp[j] = a[j * n + j] * a[st * n + j];
      double acc = 0.0;
      for (int j = 0; j < st; ++j) acc += a[st * n + j] * tmp[j];
      a[st * n + st] -= acc;
      for (int i = st + 1; i < n; ++i) {
        double ac = 0.0;
        for (int j = 0; j < st; ++j) ac += a[i * n + j] * tmp[j];
        a[i * n + st] -= ac;
      }
    }
    double d = a[st * n + st];
    if (d != 0.0)
      for (int i = st + 1; i < n; ++i) a[i * n + st] /= d;
  }
  for (int st = 0; st < n; ++st) {
    double t = b[st];
    b[st] = b[perm[st]];
    b[perm[st]] = t;
  }
  for (int i = 0; i < n; ++i) {
    double acc = b[i];
    for (int j = 0; j < i; ++j) acc -= a[i * n + j] * b[j];
    b[i] = acc;
  }
  for (int i = 0; i < n; ++i) {
    double d = a[i * n + i];
    b[i] = fabs(d) > 2.2250738585072014e-308 ? b[i] / d : 0.0;
  }
  for (int i = n - 1; i >= 0; --i) {
    double acc = b[i];
    for (int j = i + 1; j < n; ++j) acc -= a[j * n + i] * b[j];
    b[i] = acc;
  }
  for (int st = n - 1; st >= 0; --st) {
    double t = b[st];
    b[st] = b[perm[st]];
    b[perm[st]] = t;
  }
}

template <class CT>
__device__ __forceinline__ CT dexp(CT x);
template <>
__device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }
template <>
__device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <class CT>
__device__ __forceinline__ CT dlog(CT x);
template <>
__device__ __forceinline__ float dlog<float>(float x) { return __logf(x); }
template <>
__device__ __forceinline__ double dlog<double>(double x) { return log(x); }
template <class CT>
__device__ __forceinline__ CT drcp(CT x);
template <>
__device__ __forceinline__ float drcp<float>(float x) { return __frcp_rn(x); }
template <>
__device__ __forceinline__ double drcp<double>(double x) { return 1.0 / x; }

// ---------------------------------------------------------------------------
// The fit of one request, all threads of the workgroup cooperating.
template <int NB, int WAVES>
struct Fitter {
  static constexpr int NT = WAVES * kWave;
  static constexpr int V = Dims<NB>::V;
  static constexpr int SUB = Dims<NB>::SUB;
  static constexpr int NTRI = Dims<NB>::NTRI;

  const DeviceData& dd;
  FitLds<NB, WAVES>& sh;
  FitScratch<NB>& sc;
  FitCounters* counters;
  uint32_t gene;
  uint64_t g0, g1;   // CSR range
  double gamma, alpha, lambda;
  int k;
  uint32_t blk;      // cells per wave block
  int tid, wave, lane;

  __device__ Fitter(const DeviceData& d, FitLds<NB, WAVES>& s, FitScratch<NB>& c,
                    FitCounters* ctr, uint32_t g, double gam, double alp, double lam)
      : dd(d), sh(s), sc(c), counters(ctr), gene(g), gamma(gam), alpha(alp), lambda(lam) {
    k = (int)dd.k;
    g0 = dd.gptr[g];
    g1 = dd.gptr[g + 1];
    uint32_t per = (dd.n_cells + WAVES - 1) / WAVES;
    blk = (per + kWave - 1) / kWave * kWave;
    tid = threadIdx.x;
    wave = tid / kWave;
    lane = tid % kWave;
  }

  __device__ uint32_t block_begin(int w) const { return min(dd.n_cells, (uint32_t)w * blk); }
  __device__ uint32_t block_end(int w) const { return min(dd.n_cells, (uint32_t)(w + 1) * blk); }

  // Per wave: the first positive entry at or past the block start (64-ary
  // search over the gene's sorted cells) and the subject holding it.
  __device__ void setup() {
    uint32_t cb = block_begin(wave);
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
    // subject containing cb: last s with fid[s] <= cb
    int s = 0;
    {
      int a = 0, b = k;  // fid[a] <= cb < fid[b] (when cb < n_cells)
      while (b - a > 1) {
        int m = (a + b) / 2;
        if (dd.fid[m] <= cb) a = m; else b = m;
      }
      s = a;
    }
    if (lane == 0) {
      sh.ip0[wave] = ip;
      sh.s0[wave] = (uint32_t)s;
    }
    __syncthreads();
  }

  // One pass over the gene's cells at (beta in LDS `bsrc`, log w in global
  // `lwsrc`), leaving per-subject totals in sc.tot.
  template <class CT, int MODE>
  __device__ void sweep(const double* bsrc, const double* lwsrc) {
    constexpr int NV = MODE == SW_FULL ? V : (MODE == SW_VALUE ? 2 : 4);
    CT b[NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) b[j] = (CT)bsrc[j];
    const CT g = (CT)gamma;
    const uint32_t cb = block_begin(wave), ce = block_end(wave);
    uint64_t ip = sh.ip0[wave];
    int s = (int)sh.s0[wave];
    if (cb < ce) {
      for (; s < k && dd.fid[s] < ce; ++s) {
        const uint32_t lo = max(dd.fid[s], cb), hi = min(dd.fid[s + 1], ce);
        const CT lws = (CT)lwsrc[s];
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
        for (uint32_t base = lo; base < hi; base += kWave) {
          const uint32_t lim = min(base + (uint32_t)kWave, hi);
          // merge the row's positive counts through LDS
          uint64_t e = ip + (uint64_t)lane;
          uint32_t c = e < g1 ? dd.gcell[e] : 0xffffffffu;
          bool inrow = e < g1 && c < lim;
          int m = popc64(wave_ballot(inrow));
          sh.ysc[wave][lane] = 0.0;
          wave_lds_sync();
          if (inrow) sh.ysc[wave][c - base] = sizeof(CT) == 4 ? (double)dd.gy32[e] : dd.gy64[e];
          wave_lds_sync();
          const CT y = (CT)sh.ysc[wave][lane];
          ip += (uint64_t)m;
          const uint32_t cell = base + (uint32_t)lane;
          if (cell < lim) {
            CT x[NB];
            CT eta;
            if constexpr (sizeof(CT) == 4) {
              const float* row = dd.x32 + (size_t)cell * dd.p4;
#pragma unroll
              for (int j = 0; j < NB; ++j) x[j] = row[j];
              eta = dd.off32[cell];
            } else {
              const double* row = dd.x64 + (size_t)cell * NB;
#pragma unroll
              for (int j = 0; j < NB; ++j) x[j] = row[j];
              eta = dd.off64[cell];
            }
#pragma unroll
            for (int j = 0; j < NB; ++j) eta += x[j] * b[j];
            const CT ex = dexp<CT>(eta + lws);
            const CT t = ex + g;
            const CT inv = drcp<CT>(t);
            const CT u = ex * inv;
            const CT gs = g + y;
            if constexpr (MODE == SW_LAPLACE) {
              // (gamma + y) times the four derivative factors (pml.rs:1861-1874)
              double e64 = (double)ex, t64 = e64 + gamma, gs64 = gamma + (double)y;
              double p2 = e64 / (t64 * t64), p3 = p2 / t64, p4 = p3 / t64, p5 = p4 / t64;
              double e2 = e64 * e64;
              acc[0] += gs64 * p2;
              acc[1] += gs64 * (p3 * (gamma - e64));
              acc[2] += gs64 * (p4 * (gamma * gamma + e2 - (4.0 * gamma) * e64));
              acc[3] += gs64 * (p5 * (gamma * gamma * gamma - (11.0 * gamma * gamma) * e64 +
                                      (11.0 * gamma) * e2 - e2 * e64));
            } else {
              const CT lt = dlog<CT>(t);
              const double pw = (double)(gs * u * inv);
              const double term = (double)(y * eta) - (double)gs * (double)lt;
              if constexpr (MODE == SW_VALUE) {
                acc[0] += term;
                acc[1] += pw;
              } else {
                const double dres = (double)y - (double)gs * (double)u;
                acc[0] += dres;
                acc[1] += pw;
                acc[2] += term;
                double xd[NB];
#pragma unroll
                for (int j = 0; j < NB; ++j) xd[j] = (double)x[j];
#pragma unroll
                for (int a = 0; a < NB; ++a) {
                  acc[3 + a] += xd[a] * dres;
                  const double pa = pw * xd[a];
                  acc[3 + NB + a] += pa;
#pragma unroll
                  for (int bb = a; bb < NB; ++bb)
                    acc[3 + 2 * NB + a * NB - a * (a - 1) / 2 + (bb - a)] += pa * xd[bb];
                }
              }
            }
          }
        }
        // flush subject s: wave sum, slot s + wave
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          double r = wave_sum(acc[v]);
          if (lane == 0) sc.part[(size_t)(s + wave) * V + v] = r;
        }
      }
    }
    __syncthreads();
    // per-subject totals, waves added in order
    for (int idx = tid; idx < k * NV; idx += NT) {
      const int ss = idx / NV, v = idx % NV;
      const uint32_t f0 = dd.fid[ss], f1 = dd.fid[ss + 1];
      const int w0 = (int)(f0 / blk), w1 = (int)((f1 - 1) / blk);
      double t = 0.0;
      for (int w = w0; w <= w1; ++w) t += sc.part[(size_t)(ss + w) * V + v];
      sc.tot[(size_t)ss * V + v] = t;
    }
    if (tid == 0) {
      if (sizeof(CT) == 4) {
        atomicAdd(&counters->sweeps_f32, 1ull);
        atomicAdd(&counters->cells_f32, (unsigned long long)dd.n_cells);
        atomicAdd(&counters->nnz_f32, (unsigned long long)(g1 - g0));
      } else {
        atomicAdd(&counters->sweeps_f64, 1ull);
        atomicAdd(&counters->cells_f64, (unsigned long long)dd.n_cells);
        atomicAdd(&counters->nnz_f64, (unsigned long long)(g1 - g0));
      }
    }
    __syncthreads();
  }

  // Gradient, curvature blocks and penalised log-likelihood of side `sd` from
  // sc.tot (pml.rs:1564-1589), log w in `lwsrc`.
  __device__ void finalize(int sd, const double* lwsrc) {
    const double* tot_y = dd.gtot + (size_t)gene * k;
    double* sub = sc.sub[sd];
    for (int s = tid; s < k; s += NT) {
      const double* t = sc.tot + (size_t)s * V;
      const double lws = lwsrc[s];
      const double w = exp(lws);
      sub[s * SUB + 0] = t[0] + (alpha - lambda * w);
      sub[s * SUB + 1] = gamma * t[1] + lambda * w;
      sub[s * SUB + 2] = w;
#pragma unroll
      for (int a = 0; a < NB; ++a) sub[s * SUB + 3 + a] = gamma * t[3 + NB + a];
      sc.llsub[s] = t[2] + lws * tot_y[s] + (alpha * lws - lambda * w);
    }
    __syncthreads();
    // db (NB), vb (NTRI), ll: one thread each, subjects in order
    for (int idx = tid; idx < NB + NTRI + 1; idx += NT) {
      double acc = 0.0;
      if (idx < NB) {
        for (int s = 0; s < k; ++s) acc += sc.tot[(size_t)s * V + 3 + idx];
        sh.db[sd][idx] = acc;
      } else if (idx < NB + NTRI) {
        const int q = idx - NB;
        int a = 0, r = q;
        while (r >= NB - a) {
          r -= NB - a;
          ++a;
        }
        const int bb = a + r;
        for (int s = 0; s < k; ++s) acc += gamma * sc.tot[(size_t)s * V + 3 + 2 * NB + q];
        sh.vb[sd][a * NB + bb] = acc;
        sh.vb[sd][bb * NB + a] = acc;
      } else {
        for (int s = 0; s < k; ++s) acc += sc.llsub[s];
        sh.ll[sd] = acc;
      }
    }
    __syncthreads();
  }

  // Newton step from side `sd` (pml.rs:1664-1710): Schur complement, LDL'
  // solve, per-subject steps; writes sh.new_beta and sc.new_lw.
  __device__ void newton_step(int sd) {
    const double* sub = sc.sub[sd];
    for (int idx = tid; idx < NTRI + NB; idx += NT) {
      if (idx < NTRI) {
        int a = 0, r = idx;
        while (r >= NB - a) {
          r -= NB - a;
          ++a;
        }
        const int bb = a + r;
        double acc = 0.0;
        for (int s = 0; s < k; ++s)
          acc += sub[s * SUB + 3 + a] * sub[s * SUB + 3 + bb] / sub[s * SUB + 1];
        const double v = sh.vb[sd][a * NB + bb] - acc;
        sh.vb2[a * NB + bb] = v;
        sh.vb2[bb * NB + a] = v;
      } else {
        const int j = idx - NTRI;
        double acc = 0.0;
        for (int s = 0; s < k; ++s)
          acc += sub[s * SUB + 3 + j] * (sub[s * SUB + 0] / sub[s * SUB + 1]);
        sh.rhs[j] = sh.db[sd][j] - acc;
      }
    }
    __syncthreads();
    if (tid == 0) {
      double f[NB * NB];
#pragma unroll
      for (int i = 0; i < NB * NB; ++i) f[i] = sh.vb2[i];
      double st[NB];
#pragma unroll
      for (int i = 0; i < NB; ++i) st[i] = sh.rhs[i];
      ldlt_solve_dev<NB>(f, st);
#pragma unroll
      for (int i = 0; i < NB; ++i) {
        sh.step_beta[i] = st[i];
        sh.new_beta[i] = sh.beta[i] + st[i];
        sh.damp_beta[i] = 1.0;
      }
    }
    __syncthreads();
    for (int s = tid; s < k; s += NT) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) acc += sub[s * SUB + 3 + j] * sh.step_beta[j];
      const double dwvw = sub[s * SUB + 0] / sub[s * SUB + 1];
      const double stp = dwvw - acc / sub[s * SUB + 1];
      sc.step_lw[s] = stp;
      sc.new_lw[s] = sc.lw[s] + stp;
      sc.damp_lw[s] = 1.0;
      sc.vw_step[s] = sub[s * SUB + 1];
    }
    __syncthreads();
  }

  // Per-coordinate damping of the trial point (pml.rs:1729-1755).
  __device__ void damp(double min_step) {
    const double cut = 40.0;
    if (tid < NB) {
      const int j = tid;
      const double d = sh.step_beta[j];
      if (d < cut && d > -cut) {
        sh.damp_beta[j] /= 2.0;
        sh.new_beta[j] = sh.beta[j] + d * sh.damp_beta[j];
      } else if (d > 0.0) {
        sh.new_beta[j] = sh.beta[j] + min_step;
      } else {
        sh.new_beta[j] = sh.beta[j] - min_step;
      }
    }
    for (int s = tid; s < k; s += NT) {
      const double d = sc.step_lw[s];
      if (d < cut && d > -cut) {
        sc.damp_lw[s] /= 2.0;
        sc.new_lw[s] = sc.lw[s] + d * sc.damp_lw[s];
      } else if (d > 0.0) {
        sc.new_lw[s] = sc.lw[s] + min_step;
      } else {
        sc.new_lw[s] = sc.lw[s] - min_step;
      }
    }
    __syncthreads();
  }

  __device__ void accept() {
    if (tid < NB) sh.beta[tid] = sh.new_beta[tid];
    for (int s = tid; s < k; s += NT) sc.lw[s] = sc.new_lw[s];
    __syncthreads();
  }

  // max |gradient| over db and dw of side sd (pml.rs:1720-1724)
  __device__ double max_grad(int sd) {
    if (tid == 0) {
      double m = 0.0;
      for (int j = 0; j < NB; ++j) m = fmax(m, fabs(sh.db[sd][j]));
      double mw = 0.0;
      for (int s = 0; s < k; ++s) mw = fmax(mw, fabs(sc.sub[sd][s * SUB + 0]));
      sh.scal[0] = fmax(m, mw) > 0.01 ? 1.0 : 0.0;
    }
    __syncthreads();
    return sh.scal[0];
  }

  // The Newton loop of optimise (pml.rs:1650-1766) with sweeps in CT.  In
  // f64 (noise_rel == 0) this is opt_pml's control flow exactly.  In f32 the
  // comparisons allow the resolution floor noise_rel * |ll|.  Starts by
  // sweeping the current point into side 0.  Returns the side holding the
  // final point; out gets ll, ll_prev, iterations, backtracks.
  template <class CT>
  __device__ int newton_loop(const FitKnobs& kn, double noise_rel, FitResult& out) {
    int cur = 0;
    sweep<CT, SW_FULL>(sh.beta, sc.lw);
    finalize(cur, sc.lw);
    double ll = sh.ll[cur], ll_prev = 0.0, likdif = 0.0;
    int step = 0, backtracks = 0;
    while (step == 0 ||
           (likdif > (noise_rel > 0.0 ? fmax(kn.eps, noise_rel * fabs(ll)) : kn.eps) &&
            step < kn.max_iter)) {
      ++step;
      newton_step(cur);
      const int tr = 1 - cur;
      ll_prev = ll;
      sweep<CT, SW_FULL>(sh.new_beta, sc.new_lw);
      finalize(tr, sc.new_lw);
      ll = sh.ll[tr];
      likdif = ll - ll_prev;
      const double noise = noise_rel * fabs(ll_prev);
      backtracks = 0;
      double min_step = 40.0;
      while (likdif < -noise || isinf(ll) || (noise_rel > 0.0 && isnan(ll))) {
        ++backtracks;
        min_step /= 2.0;
        if (backtracks > kn.max_backtrack) {
          likdif = 0.0;
          ll = ll_prev;
          if (max_grad(cur) > 0.0) ++backtracks;
          break;
        }
        damp(min_step);
        sweep<CT, SW_FULL>(sh.new_beta, sc.new_lw);
        finalize(tr, sc.new_lw);
        ll = sh.ll[tr];
        likdif = ll - ll_prev;
      }
      accept();
      cur = tr;
    }
    out.ll = ll;
    out.ll_prev = ll_prev;
    out.iterations = step;
    out.backtracks = backtracks;
    return cur;
  }

  // FP64 finish from the current point (newton_finish, pml.rs:807-1090):
  // Newton steps without damping, log-likelihood and log-determinant at the
  // stepped point.  Returns false when a step worsens the objective beyond
  // rounding or leaves it not finite, or the budget runs out.
  __device__ bool finish(const FitKnobs& kn, FitResult& out) {
    int cur = 0;
    sweep<double, SW_FULL>(sh.beta, sc.lw);
    finalize(cur, sc.lw);
    for (int it = 1; it <= kn.max_iter; ++it) {
      const double ll_prev = sh.ll[cur];
      newton_step(cur);
      const int tr = 1 - cur;
      sweep<double, SW_FULL>(sh.new_beta, sc.new_lw);
      finalize(tr, sc.new_lw);
      const double ll = sh.ll[tr];
      const double likdif = ll - ll_prev;
      if (isnan(likdif) || likdif < -1e-12 * fabs(ll)) return false;
      accept();
      cur = tr;
      if (likdif <= kn.eps) {
        // log-det at the stepped point
        if (tid == 0) {
          double ld = 0.0;
          for (int s = 0; s < k; ++s) ld += log(fabs(sc.sub[cur][s * SUB + 1]));
          sh.scal[1] = ld;
        }
        __syncthreads();
        out.ll = ll;
        out.ll_prev = ll_prev;
        out.log_det = sh.scal[1];
        out.finish_steps = it;
        out.second = 0.0;
        // information at the final point for the reply
        newton_info(cur);
        return true;
      }
    }
    return false;
  }

  // Schur complement of side sd into sh.vb2 (no step), for the reply.
  __device__ void newton_info(int sd) {
    const double* sub = sc.sub[sd];
    for (int idx = tid; idx < NTRI; idx += NT) {
      int a = 0, r = idx;
      while (r >= NB - a) {
        r -= NB - a;
        ++a;
      }
      const int bb = a + r;
      double acc = 0.0;
      for (int s = 0; s < k; ++s)
        acc += sub[s * SUB + 3 + a] * sub[s * SUB + 3 + bb] / sub[s * SUB + 1];
      const double v = sh.vb[sd][a * NB + bb] - acc;
      sh.vb2[a * NB + bb] = v;
      sh.vb2[bb * NB + a] = v;
    }
    __syncthreads();
  }

  // log-det from the last Newton step's vw (pml.rs:1767-1771)
  __device__ double log_det_step() {
    if (tid == 0) {
      double ld = 0.0;
      for (int s = 0; s < k; ++s) ld += log(fabs(sc.vw_step[s]));
      sh.scal[2] = ld;
    }
    __syncthreads();
    return sh.scal[2];
  }

  // Laplace correction at the current point (pml.rs:1828-1919), f64.
  __device__ double laplace(unsigned ord) {
    sweep<double, SW_LAPLACE>(sh.beta, sc.lw);
    if (tid == 0) {
      const bool high = ord > 2;
      double third_acc = 0.0, fourth_acc = 0.0, fifth_acc = 0.0;
      for (int s = 0; s < k; ++s) {
        const double* t = sc.tot + (size_t)s * V;
        const double prior = lambda * exp(sc.lw[s]);
        const double vw = gamma * t[0] + prior;
        const double third = gamma * t[1] + prior;
        third_acc += third * third / (vw * vw * vw);
        if (high) {
          const double fourth = gamma * t[2] + prior;
          const double fifth = gamma * t[3] + prior;
          const double vs = vw * vw;
          fourth_acc += fourth / vs;
          fifth_acc += fifth * third / (vs * vs);
        }
      }
      double second = 5.0 * third_acc / 24.0;
      if (high) {
        second -= fourth_acc / 8.0;
        second += 7.0 * fifth_acc / 48.0;
      }
      sh.scal[3] = second;
    }
    __syncthreads();
    return sh.scal[3];
  }

  // The whole fit.  mode 0: f32 Newton then f64 finish (stage two); mode 1:
  // opt_pml in f64 from the given point (stage three).  beta in sh.beta and
  // log w in sc.lw on entry.
  __device__ FitResult run(int mode, unsigned ord, const FitKnobs& kn) {
    FitResult r{};
    r.finish_steps = 0;
    r.exact = 0;
    setup();
    bool done = false;
    if (mode == 0) {
      FitResult f32r{};
      newton_loop<float>(kn, kn.noise_rel, f32r);
      r.iterations = f32r.iterations;
      r.backtracks = f32r.backtracks;
      done = finish(kn, r);
    }
    if (!done) {
      FitResult e{};
      newton_loop<double>(kn, 0.0, e);
      r.ll = e.ll;
      r.ll_prev = e.ll_prev;
      r.iterations += e.iterations;
      r.backtracks = e.backtracks;
      r.log_det = log_det_step();
      r.exact = 1;
    }
    r.second = ord > 1 ? laplace(ord) : 0.0;
    return r;
  }
};

}  // namespace nh

// Stage one on the device: the sums behind the marginal Poisson-gamma
// likelihood, its gradient and its Hessian (edge-rs src/sc/ptmg.rs:661-1252:
// linear_predictor, evaluate, evaluate_hessian / hessian_subject) for a batch
// of (gene, point) requests, FP64 throughout.  Each request's cells are cut
// into units (one wave each, whole 64-cell rows) spread over `chunks`
// workgroups, so a round with few requests left still fills the device.
//
// Three passes over the gene's cells, one launch each, each the coalesced
// 64-cell row walk of pml_fit.hpp with the positive counts merged through LDS;
// every unit writes its own partial sums, and a pass reads the subject totals
// the previous launch left:
//   1. e = exp(eta) per cell: per-subject sum of e, per-gene sum of y * eta
//      (exp, log and 1/w by the FP64 sweep's table-driven versions);
//   2. with ym = (Y_s + alpha) / (sum e + lambda) per subject: the gradient
//      sums (sum g e, sum x e, sum g x e per subject; the gene's sums of
//      ym e - (gamma + y) ln w, ln w and g);
//   3. Hessian only: the Hessian's per-subject vectors (hb2, b2, t1, t2),
//      the gene's sums of 1/w and g/w, and the Gram  sum_i W_i x_i x_i'  with
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

// Cells per lane in one tile of a pass's walk (tile = 64 x kPtmgTile cells).
constexpr int kPtmgTile = 2;

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

size_t ptmg_scratch_doubles(int nb, int k, int units) {
  // per-(unit, subject) partials in slots s + unit: (k + units) * PS, then the
  // per-unit gene partials: units * PG, then the request's output block as
  // the reduction assembles it, before one coalesced copy to `out`
  const int ps = 4 + 4 * nb, pg = 6 + nb * (nb + 1) / 2;
  return (size_t)(k + units) * ps + (size_t)units * pg + ptmg_out_doubles(nb, k);
}

// Cells per unit: the cell range split into `units` blocks, each a whole
// number of 64-cell rows.
__host__ __device__ inline uint32_t ptmg_block(uint32_t n_cells, int units) {
  const uint32_t per = (n_cells + (uint32_t)units - 1) / (uint32_t)units;
  return (per + kWave - 1) / kWave * kWave;
}

// Sum of subject s's unit partials (slot s + u for every unit u that holds
// some of its cells) at value v, added in unit order with the loads issued
// four at a time; the pass kernels and the reduction call the same function,
// so every reader of a subject total sees the same bits.
__device__ inline double ptmg_subject_total(const DeviceData& dd, const double* part, int stride,
                                            uint32_t blk, int s, int v) {
  const uint32_t f0 = dd.fid[s], f1 = dd.fid[s + 1];
  if (f1 <= f0) return 0.0;
  const int u0 = (int)(f0 / blk), u1 = (int)((f1 - 1) / blk);
  const double* p = part + (size_t)s * stride + v;
  double t = 0.0;
  int u = u0;
  for (; u + 3 <= u1; u += 4) {
    const double a = p[(size_t)u * stride], b = p[(size_t)(u + 1) * stride];
    const double c = p[(size_t)(u + 2) * stride], d = p[(size_t)(u + 3) * stride];
    t += a;
    t += b;
    t += c;
    t += d;
  }
  for (; u <= u1; ++u) t += p[(size_t)u * stride];
  return t;
}

template <int NB, int WAVES>
struct PtmgWalk {
  const DeviceData& dd;
  uint64_t g0, g1;
  uint32_t blk;
  int k, unit, wave, lane;
  double (*ysc)[kWave * kPtmgTile];
  uint64_t* ip0;
  uint32_t* s0;

  __device__ uint32_t bb() const { return min(dd.n_cells, (uint32_t)unit * blk); }
  __device__ uint32_t be() const { return min(dd.n_cells, (uint32_t)(unit + 1) * blk); }

  // First CSR entry at or after the unit's first cell (64-way search over the
  // gene's positive cells) and the subject that holds that cell.
  __device__ void setup() {
    const uint32_t cb = bb();
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
#pragma unroll
    for (int u = 0; u < kPtmgTile; ++u) ysc[wave][lane + kWave * u] = 0.0;
    __syncthreads();
  }

  // Calls begin(s) before and flush(s) after each subject segment of the
  // unit's block, uniformly across the wave, and cell(s, cell, y, row) for
  // every cell of the segment (active lanes only; row = the cell's packed
  // design row, R doubles).  Tiles of 64 x kPtmgTile cells, kPtmgTile per
  // lane, as in the stage-two sweeps: the rows come from a scalar base with
  // 32-bit offsets, and the next 64 x kPtmgTile CSR entries (32-bit indices
  // relative to the wave's first entry) are held in registers and reloaded as
  // soon as a tile has scattered its own into the wave's LDS row, so their
  // loads overlap the tile's arithmetic; every load is issued at a clamped
  // in-bounds index.  The LDS row starts zeroed (setup) and each tile zeroes
  // what it read.
  template <class Begin, class Cell, class Flush>
  __device__ void walk(Begin&& begin, Cell&& cell, Flush&& flush) {
    constexpr int R = Dims<NB>::R64;
    constexpr int U = kPtmgTile;
    constexpr uint32_t TILE = kWave * U;
    const uint32_t cb = bb(), ce = be();
    if (cb >= ce) return;
    gconst_char* const rows = (gconst_char*)dd.row64;
    const uint64_t e_last = g1 > 0 ? g1 - 1 : 0;
    const uint64_t ip = uni(ip0[wave]);
    const uint64_t wbase = ip < e_last ? ip : e_last;
    gconst_char* const gc = (gconst_char*)((gconst_u32*)dd.gcell + wbase);
    gconst_char* const gyb = (gconst_char*)((gconst_double*)dd.gy64 + wbase);
    const uint32_t n_rel = (uint32_t)(g1 - wbase);
    const uint32_t c_last = n_rel > 0 ? n_rel - 1 : 0;
    uint32_t rel = (uint32_t)(ip - wbase);
    auto entry = [&](uint32_t e, uint32_t& c, double& y) {
      const uint32_t ec = e < n_rel ? e : c_last;
      c = *(gconst_u32*)(gc + (ec << 2));
      y = *(gconst_double*)(gyb + (ec << 3));
      c = e < n_rel ? c : 0xffffffffu;
    };
    lds_double* const ys = (lds_double*)&ysc[wave][0];
    uint32_t wc[U];
    double wy[U];
#pragma unroll
    for (int j = 0; j < U; ++j) entry(rel + (uint32_t)(lane + kWave * j), wc[j], wy[j]);
    for (int s = (int)uni(s0[wave]); s < k && dd.fid[s] < ce; ++s) {
      const uint32_t lo = max(dd.fid[s], cb), hi = min(dd.fid[s + 1], ce);
      begin(s);
      for (uint32_t base = lo; base < hi; base += TILE) {
        const uint32_t tend = min(base + TILE, hi);
        double xr[U][R];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t cl = base + (uint32_t)(lane + kWave * u);
          load_row_at<double, R>(rows, cl < tend ? cl : tend - 1, xr[u]);
        }
        // the window's entries inside [base, tend) form a prefix of it
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < U; ++j) {
          const bool in = wc[j] < tend;
          if (in) ys[wc[j] - base] = wy[j];
          m += (uint32_t)popc64(wave_ballot(in));
        }
        rel += m;
#pragma unroll
        for (int j = 0; j < U; ++j) entry(rel + (uint32_t)(lane + kWave * j), wc[j], wy[j]);
        __builtin_amdgcn_wave_barrier();
        double yv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) yv[u] = ys[lane + kWave * u];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < U; ++u) ys[lane + kWave * u] = 0.0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t cl = base + (uint32_t)(lane + kWave * u);
          if (cl < tend) cell(s, cl, yv[u], xr[u]);
        }
      }
      flush(s);
    }
  }
};

// One pass over one unit of the cells for WAVES requests: workgroup b takes
// unit b % units of requests (b / units) * WAVES + wave, one per wave, so the
// waves of a workgroup walk the same design rows (read once from L2 for all
// of them); with units a multiple of 8, every workgroup on a given unit runs
// on the same XCD (workgroups go to XCDs round robin), whose L2 then keeps
// that share of the rows for the whole round.
// Pass 1 writes the per-(unit, subject) sums of e and the unit's sum of
// y * eta; pass 2 reads the subject totals of e (every unit of the subject,
// written by the pass-1 launch) and writes the gradient / Hessian-vector
// sums; pass 3 (Hessian requests only) reads the totals of e and g e and
// writes the unit's weighted Gram.
template <int NB, int WAVES, int PASS>
__global__ __launch_bounds__(WAVES * 64) void ptmg_pass_kernel(DeviceData dd, const PtmgRequest* req,
                                                               const double* params,
                                                               double* scratch, size_t sstride,
                                                               int chunks, int n_req) {
  using D = PtmgDims<NB>;
  constexpr int PS = D::PS, PG = D::PG, NTRI = D::NTRI;
  __shared__ double ysc[WAVES][kWave * kPtmgTile];
  __shared__ uint64_t ip0[WAVES];
  __shared__ uint32_t s0[WAVES];
  __shared__ double f64tab[kF64TabDoubles];  // the FP64 sweep's exp / log tables
  const int units = chunks * WAVES;
  // wave-uniform: keeps the request's values in scalar registers
  const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kWave);
  const int q = ((int)blockIdx.x / units) * WAVES + wave;
  const bool have = q < n_req;
  const PtmgRequest r = have ? req[q] : PtmgRequest{0u, 0u};
  const bool hess = r.hess != 0;
  // a wave without work still takes part in the set-up barrier
  const bool active = have && (PASS != 3 || hess);
  const int k = (int)dd.k;
  const double* x = params + (size_t)(have ? q : 0) * (NB + 2);
  double* part = scratch + (size_t)(have ? q : 0) * sstride;
  double* gpart = part + (size_t)(k + units) * PS;

  PtmgWalk<NB, WAVES> W{dd};
  W.g0 = dd.gptr[r.gene];
  W.g1 = dd.gptr[r.gene + 1];
  W.blk = ptmg_block(dd.n_cells, units);
  W.k = k;
  W.wave = wave;
  W.lane = (int)threadIdx.x % kWave;
  W.unit = (int)blockIdx.x % units;
  W.ysc = ysc;
  W.ip0 = ip0;
  W.s0 = s0;
  fill_f64_tables(f64tab, (int)threadIdx.x);  // setup() ends with a barrier
  W.setup();
  if (!active) return;  // no barrier follows
  const lds_double* tab = (const lds_double*)f64tab;
  const int u = W.unit;
  const bool lead = W.lane == 0;

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
  // (stored column j before the intercept, j - 1 after it: constant indices
  // and a select, so the row stays in registers)
  auto row_eta = [&](const double* prow, double xr[NB]) {
    double eta = prow[NB - 1];
#pragma unroll
    for (int j = 0; j < NB; ++j) {
      const double before = j < NB - 1 ? prow[j] : 0.0;
      const double after = j > 0 ? prow[j - 1] : 0.0;
      xr[j] = j < icpt ? before : (j == icpt ? 1.0 : after);
      eta += xr[j] * beta[j];
    }
    return eta;
  };

  if (PASS == 1) {
    // ---- subject sums of e, sum of y * eta ----
    double csx = 0.0, lin = 0.0;
    W.walk(
        [&](int) {},
        [&](int, uint32_t, double y, const double* prow) {
          double xr[NB];
          const double eta = row_eta(prow, xr);
          if (y != 0.0) lin += eta * y;
          csx += fast_exp(eta, tab);
        },
        [&](int s) {
          const double t = wave_sum(csx);
          if (lead) part[(size_t)(s + u) * PS] = t;
          csx = 0.0;
        });
    const double l = wave_sum(lin);
    if (lead) gpart[(size_t)u * PG] = l;
  } else if (PASS == 2) {
    // ---- gradient sums: sum g e, sum x e, sum g x e per subject ----
    constexpr int NA = 1 + 2 * NB;
    double acc[NA];
    double gs[3];
#pragma unroll
    for (int v = 0; v < NA; ++v) acc[v] = 0.0;
#pragma unroll
    for (int v = 0; v < 3; ++v) gs[v] = 0.0;
    double ym = 0.0;
    W.walk(
        [&](int s) {
          ym = (tot_y[s] + alpha) / (ptmg_subject_total(dd, part, PS, W.blk, s, 0) + lambda);
        },
        [&](int, uint32_t, double y, const double* prow) {
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
#pragma unroll
          for (int a = 0; a < NB; ++a) {
            const double xe = xr[a] * e;
            acc[1 + a] += xe;
            acc[1 + NB + a] += g * xe;
          }
        },
        [&](int s) {
          // slots 1 (acc) and 4 .. 4 + 2 NB (xf, d41)
#pragma unroll
          for (int v = 0; v < NA; ++v) {
            const double t = wave_sum(acc[v]);
            if (lead) part[(size_t)(s + u) * PS + (v == 0 ? 1 : 3 + v)] = t;
            acc[v] = 0.0;
          }
        });
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double t = wave_sum(gs[v]);
      if (lead) gpart[(size_t)u * PG + 1 + v] = t;
    }
  } else {
    // ---- the Hessian sums: per subject hb2, b2, t1, t2; the weighted Gram
    // of the beta block; sum 1/w and sum g/w ----
    constexpr int NH = 2 + 2 * NB;
    double hacc[NH];
    double gr[NTRI];
    double hs[2] = {0.0, 0.0};
#pragma unroll
    for (int v = 0; v < NH; ++v) hacc[v] = 0.0;
#pragma unroll
    for (int v = 0; v < NTRI; ++v) gr[v] = 0.0;
    double ym = 0.0, ymmd = 0.0;
    W.walk(
        [&](int s) {
          const double csx = ptmg_subject_total(dd, part, PS, W.blk, s, 0);
          const double mu = csx + lambda;
          const double ystar = tot_y[s] + alpha;
          ym = ystar / mu;
          const double ymm = ystar / mu / mu;
          ymmd = ymm * (ptmg_subject_total(dd, part, PS, W.blk, s, 1) - csx);
        },
        [&](int, uint32_t, double y, const double* prow) {
          double xr[NB];
          const double eta = row_eta(prow, xr);
          const double e = fast_exp(eta, tab);
          const double we = ym * e;
          const double w = we + gamma;
          const double iw = fast_rcp(w);
          const double g = (gamma + y) * iw;
          const double gt = g * iw;
          const double gte = gt * e;
          hs[0] += iw;
          hs[1] += gt;
          hacc[0] += gte - e * iw;  // hb2
          hacc[1] += gte * e;       // b2
          const double f1 = gt - iw;
          const double wt = e * ymmd - ym * (g * e) + ym * ym * (gte * e);
          int idx = 0;
#pragma unroll
          for (int a = 0; a < NB; ++a) {
            const double xe = xr[a] * e;
            hacc[2 + a] += xe * f1;        // t1
            hacc[2 + NB + a] += xe * gte;  // t2
            const double xa = wt * xr[a];
#pragma unroll
            for (int b = a; b < NB; ++b) gr[idx++] += xa * xr[b];
          }
        },
        [&](int s) {
          // slots 2, 3 (hb2, b2) and 4 + 2 NB .. 4 + 4 NB (t1, t2)
#pragma unroll
          for (int v = 0; v < NH; ++v) {
            const double t = wave_sum(hacc[v]);
            if (lead) part[(size_t)(s + u) * PS + (v < 2 ? 2 + v : 2 + 2 * NB + v)] = t;
            hacc[v] = 0.0;
          }
        });
#pragma unroll
    for (int v = 0; v < 2; ++v) {
      const double t = wave_sum(hs[v]);
      if (lead) gpart[(size_t)u * PG + 4 + v] = t;
    }
#pragma unroll
    for (int v = 0; v < NTRI; ++v) {
      const double t = wave_sum(gr[v]);
      if (lead) gpart[(size_t)u * PG + 6 + v] = t;
    }
  }
}

// Per request: the subject totals and the gene sums over every unit, in unit
// order, assembled in device memory and then copied to the request's output
// block with whole-wave stores (the output may live in host memory).
template <int NB>
__global__ __launch_bounds__(256) void ptmg_reduce_kernel(DeviceData dd, const PtmgRequest* req,
                                                          double* out, double* scratch,
                                                          size_t sstride, int units,
                                                          FitCounters* ctr) {
  using D = PtmgDims<NB>;
  constexpr int PS = D::PS, PG = D::PG;
  const int q = (int)blockIdx.x, k = (int)dd.k;
  const PtmgRequest r = req[q];
  const bool hess = r.hess != 0;
  const uint32_t blk = ptmg_block(dd.n_cells, units);
  const size_t ostride = ptmg_out_doubles_t<NB>(k);
  double* part = scratch + (size_t)q * sstride;
  const double* gpart = part + (size_t)(k + units) * PS;
  double* o = part + (size_t)(k + units) * PS + (size_t)units * PG;
  const int wave = (int)threadIdx.x / kWave, lane = (int)threadIdx.x % kWave;
  const int nw = (int)blockDim.x / kWave;
  for (int idx = (int)threadIdx.x; idx < k * PS; idx += (int)blockDim.x) {
    const int v = idx % PS;
    // hb2, b2, t1, t2 come from pass 3, which gradient-only requests skip
    const bool hess_only = v == 2 || v == 3 || v >= 4 + 2 * NB;
    o[idx] = (hess || !hess_only) ? ptmg_subject_total(dd, part, PS, blk, idx / PS, v) : 0.0;
  }
  // gene sums: one wave per value, lanes over the units, then a wave sum
  const int ng = hess ? PG : 6;
  for (int v = wave; v < ng; v += nw) {
    double t = 0.0;
    if (hess || v < 4)  // sum 1/w and sum g/w come from pass 3
      for (int w = lane; w < units; w += kWave) t += gpart[(size_t)w * PG + v];
    t = wave_sum(t);
    if (lane == 0) o[(size_t)k * PS + v] = t;
  }
  __syncthreads();
  double* dst = out + (size_t)q * ostride;
  const size_t nout = (size_t)k * PS + (size_t)ng;
  for (size_t i = threadIdx.x; i < nout; i += blockDim.x) dst[i] = o[i];
  if (threadIdx.x == 0) {
    const unsigned long long passes = hess ? 3ull : 2ull;
    atomicAdd(&ctr->sweeps_f64, passes);
    atomicAdd(&ctr->cells_f64, (unsigned long long)dd.n_cells * passes);
    atomicAdd(&ctr->nnz_f64, (unsigned long long)(dd.gptr[r.gene + 1] - dd.gptr[r.gene]) * passes);
  }
}

// Per gene: the sum over its positive counts of y times the centred design
// row (the constant term of stage one's beta gradient, ptmg.rs's x'y), one
// workgroup per gene; threads stride over the gene's counts and the block
// adds the partials in a fixed order.
template <int NB>
__global__ __launch_bounds__(256) void xty_kernel(DeviceData dd, double* out) {
  const uint32_t g = blockIdx.x;
  const uint64_t a = dd.gptr[g], b = dd.gptr[g + 1];
  double acc[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) acc[j] = 0.0;
  for (uint64_t e = a + threadIdx.x; e < b; e += 256) {
    const double y = dd.gy64[e];
    const double* x = dd.x64 + (size_t)dd.gcell[e] * NB;
#pragma unroll
    for (int j = 0; j < NB; ++j) acc[j] += x[j] * y;
  }
  __shared__ double red[NB][4];
  const int wave = (int)threadIdx.x / kWave, lane = (int)threadIdx.x % kWave;
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    const double t = wave_sum(acc[j]);
    if (lane == 0) red[j][wave] = t;
  }
  __syncthreads();
  if (threadIdx.x < NB) {
    double t = 0.0;
    for (int w = 0; w < 4; ++w) t += red[threadIdx.x][w];
    out[(size_t)g * NB + threadIdx.x] = t;
  }
}

hipError_t launch_xty(const DeviceData& dd, size_t n_genes, double* out, hipStream_t st) {
  if (n_genes == 0) return hipSuccess;
  const dim3 grid((unsigned)n_genes), block(256);
  switch (dd.nb) {
    case 1: hipLaunchKernelGGL(xty_kernel<1>, grid, block, 0, st, dd, out); break;
    case 2: hipLaunchKernelGGL(xty_kernel<2>, grid, block, 0, st, dd, out); break;
    case 3: hipLaunchKernelGGL(xty_kernel<3>, grid, block, 0, st, dd, out); break;
    case 4: hipLaunchKernelGGL(xty_kernel<4>, grid, block, 0, st, dd, out); break;
    case 5: hipLaunchKernelGGL(xty_kernel<5>, grid, block, 0, st, dd, out); break;
    case 6: hipLaunchKernelGGL(xty_kernel<6>, grid, block, 0, st, dd, out); break;
    case 7: hipLaunchKernelGGL(xty_kernel<7>, grid, block, 0, st, dd, out); break;
    case 8: hipLaunchKernelGGL(xty_kernel<8>, grid, block, 0, st, dd, out); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <int NB>
static hipError_t launch_ptmg_nb(const PtmgLaunch& L, hipStream_t st) {
  constexpr int W = kPtmgWaves;
  // (requests / W, rounded up) x units workgroups, W requests per workgroup
  const int groups = (L.n_req + W - 1) / W;
  const dim3 grid((unsigned)(groups * L.chunks * W)), block(W * 64);
  hipLaunchKernelGGL((ptmg_pass_kernel<NB, W, 1>), grid, block, 0, st, L.dd, L.req, L.params,
                     L.scratch, L.stride, L.chunks, L.n_req);
  hipLaunchKernelGGL((ptmg_pass_kernel<NB, W, 2>), grid, block, 0, st, L.dd, L.req, L.params,
                     L.scratch, L.stride, L.chunks, L.n_req);
  if (L.any_hess)
    hipLaunchKernelGGL((ptmg_pass_kernel<NB, W, 3>), grid, block, 0, st, L.dd, L.req, L.params,
                       L.scratch, L.stride, L.chunks, L.n_req);
  hipLaunchKernelGGL((ptmg_reduce_kernel<NB>), dim3((unsigned)L.n_req), dim3(256), 0, st, L.dd,
                     L.req, L.out, L.scratch, L.stride, L.chunks * W, L.counters);
  return hipGetLastError();
}

hipError_t launch_ptmg(const PtmgLaunch& L, hipStream_t st) {
  if (L.n_req == 0) return hipSuccess;
  if (L.chunks < 1 || L.stride < ptmg_scratch_doubles((int)L.dd.nb, (int)L.dd.k, L.chunks * kPtmgWaves))
    return hipErrorInvalidValue;
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

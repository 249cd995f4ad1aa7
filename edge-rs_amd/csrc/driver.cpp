// nebula_sparse_hip: edge-rs's nebula_sparse_gpu call shape
// (src/gpu/stage_two.rs:233-265) with stages one, two and three on gfx950.
//
// Host work is validation, preparation and the small per-gene control logic
// (search steppers, Newton decisions, O(k nb^2) assembly of device sums);
// every pass over a gene's cells runs on the device.
//
// Set-up restates nebula_sparse_with (src/sc/nebula.rs:494-669): subject runs
// (:2138), offset summary (ptmg.rs:1827), centred design (ptmg.rs:1748),
// intercept detection, cell-level columns (ptmg.rs:1962), subject totals
// (ptmg.rs:1596), method override (:612-617) and the expression filter
// (:621-636).  Finishing restates finish_gene (:918-1000) and assemble
// (:2233-2277).
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "../../include/nebula_hip.h"
#include "bobyqa.hpp"
#include "driver.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "resident.hpp"
#include "search_control.hpp"
#include "stage_kernels.hpp"
#include "stage_one.hpp"

namespace nh {

namespace {

using Clock = std::chrono::steady_clock;
double secs(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double>(b - a).count();
}

void put_err(char* err, size_t len, const std::string& m) {
  if (err && len) {
    std::strncpy(err, m.c_str(), len - 1);
    err[len - 1] = 0;
  }
}

template <class T>
T* dup(const std::vector<T>& v) {
  T* p = (T*)std::malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
  if (!p) api_fail(NH_ERR_GPU, "out of host memory");
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}

// Smallest eigenvalue by cyclic Jacobi (src/sc/pml.rs:2084-2133).
double min_eigenvalue(const double* m, size_t n) {
  if (n == 1) return m[0];
  std::vector<double> a(m, m + n * n);
  double frob = 0.0;
  for (double v : a) frob += v * v;
  const double eps = 2.220446049250313e-16;
  const double tol = std::max(frob * eps * eps, 2.2250738585072014e-308);
  for (int sw = 0; sw < 30; ++sw) {
    double off = 0.0;
    for (size_t p = 0; p < n; ++p)
      for (size_t q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
    if (off <= tol) break;
    for (size_t p = 0; p < n; ++p)
      for (size_t q = p + 1; q < n; ++q) {
        double apq = a[p * n + q];
        if (apq == 0.0) continue;
        double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        double sg = std::signbit(theta) ? -1.0 : 1.0;
        double t = sg / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (size_t r = 0; r < n; ++r) {
          double arp = a[r * n + p], arq = a[r * n + q];
          a[r * n + p] = c * arp - s * arq;
          a[r * n + q] = s * arp + c * arq;
        }
        for (size_t r = 0; r < n; ++r) {
          double apr = a[p * n + r], aqr = a[q * n + r];
          a[p * n + r] = c * apr - s * aqr;
          a[q * n + r] = s * apr + c * aqr;
        }
      }
  }
  double mn = INFINITY;
  for (size_t i = 0; i < n; ++i) mn = std::min(mn, a[i * n + i]);
  return mn;
}

bool cholesky_inverse(const std::vector<double>& a, size_t n, std::vector<double>& inv) {
  inv.assign(n * n, 0.0);
  std::vector<double> col(n), sol;
  for (size_t j = 0; j < n; ++j) {
    std::fill(col.begin(), col.end(), 0.0);
    col[j] = 1.0;
    if (!chol_solve(a, n, col, sol)) return false;
    for (size_t i = 0; i < n; ++i) inv[i * n + j] = sol[i];
  }
  return true;
}

}  // namespace

int fit_convergence(const DevReply& r, const double* information, size_t nb, bool at_bound,
                    int initial) {
  int code = initial;
  if (code == NH_CONV_SUCCESS) {
    if (at_bound)
      code = NH_CONV_AT_BOUND;
    else if (std::isnan(r.ll))
      code = NH_CONV_NOT_FINITE;
    else if (r.iterations == kFitMaxIter)
      code = NH_CONV_MAX_ITER;
    else if (r.backtracks == kFitMaxBacktrack + 1)
      code = NH_CONV_CRITICAL_POINT;
    else if (r.backtracks == kFitMaxBacktrack + 2)
      code = NH_CONV_STALLED;
  }
  if (nb > 1 && min_eigenvalue(information, nb) < 1e-8) code = NH_CONV_SINGULAR;
  return code;
}

// ---------------------------------------------------------------------------
// ptmg value / gradient / Hessian from the device sums (ptmg.rs:288-478,
// 742-896, 1061-1252).
void assemble_ptmg(const Prepared& P, const GeneInfo& gi, const double* params, const double* o,
                   bool hess, double& value, std::vector<double>& grad, std::vector<double>& h) {
  const size_t nb = P.n_coef, k = P.k, np = nb + 2;
  const size_t PS = 4 + 4 * nb;
  const double sigma = params[nb], gamma = params[nb + 1];
  const double exps = std::exp(sigma), m1 = exps - 1.0, exps_s = std::sqrt(exps);
  const double alpha = 1.0 / m1, lambda = alpha / exps_s, sq = m1 * m1;
  const double log_lambda = std::log(lambda), log_gamma = std::log(gamma);
  const double alpha_pr = -exps / sq;
  const double lambda_pr = (1.0 - 3.0 * exps) / (2.0 * exps_s * sq);
  const double alpha_dpr = 2.0 * exps * exps / (sq * m1) - exps / sq;
  const double lambda_dpr = -3.0 * exps / (2.0 * exps_s * sq) +
                            (3.0 * exps - 1.0) / (4.0 * exps_s * sq) +
                            (3.0 * exps - 1.0) * exps_s / (sq * m1);
  const double* G = o + k * PS;
  double total = G[0];
  std::vector<double> mlog(k), ym(k), ymm(k), imu(k), d42c(k);
  for (size_t s = 0; s < k; ++s) {
    const double csx = o[s * PS + 0];
    const double ystar = gi.totals[s] + alpha;
    const double mu = csx + lambda;
    const double lm = std::log(mu);
    mlog[s] = lm;
    ym[s] = ystar / mu;
    ymm[s] = ystar / mu / mu;
    imu[s] = 1.0 / mu;
    total -= ystar * lm;
    d42c[s] = o[s * PS + 1] - csx;
  }
  total += (double)k * alpha * log_lambda;
  total += (double)P.n_cells * gamma * log_gamma;
  total += G[1];
  const double slpey = G[2], gstar_sum = G[3], sum_tempa = G[4], sum_gt = G[5];
  auto XF = [&](size_t s, size_t j) { return o[s * PS + 4 + j]; };
  auto D41 = [&](size_t s, size_t j) { return o[s * PS + 4 + nb + j]; };
  auto T1 = [&](size_t s, size_t j) { return o[s * PS + 4 + 2 * nb + j]; };
  auto T2 = [&](size_t s, size_t j) { return o[s * PS + 4 + 3 * nb + j]; };
  // gradient_parts
  std::vector<double> ymm_d(k), hbl0(k);
  for (size_t s = 0; s < k; ++s) ymm_d[s] = ymm[s] * d42c[s];
  std::vector<double> gr(np, 0.0);
  for (size_t j = 0; j < nb; ++j) {
    double acc = 0.0;
    for (size_t s = 0; s < k; ++s) acc += XF(s, j) * ymm_d[s] - D41(s, j) * ym[s];
    gr[j] = acc + gi.xty[j];
  }
  double smlog = 0, sym = 0, dbim = 0, sum_ymm_d = 0;
  for (size_t s = 0; s < k; ++s) {
    smlog += mlog[s];
    sym += ym[s];
    hbl0[s] = d42c[s] * imu[s];
  }
  for (size_t s = 0; s < k; ++s) dbim += hbl0[s];
  for (size_t s = 0; s < k; ++s) sum_ymm_d += ymm_d[s];
  const double ldm = log_lambda * (double)k - smlog;
  const double adlmy = exps_s * (double)k - sym;
  gr[nb] = -alpha_pr * dbim + lambda_pr * sum_ymm_d + alpha_pr * ldm + lambda_pr * adlmy;
  gr[nb + 1] = log_gamma * (double)P.n_cells + (double)P.n_cells - slpey - gstar_sum;
  // gamma-function terms (ptmg.rs:404-478)
  double npos = 0, lsum = 0, dsum = 0, tsum = 0;
  for (size_t s = 0; s < k; ++s)
    if (gi.totals[s] > 0.0) {
      npos += 1.0;
      lsum += std::lgamma(gi.totals[s] + alpha);
      dsum += digamma(gi.totals[s] + alpha);
      if (hess) tsum += trigamma(gi.totals[s] + alpha);
    }
  const double sv = lsum - npos * std::lgamma(alpha);
  const double sd = dsum - npos * digamma(alpha);
  const double st = hess ? tsum - npos * trigamma(alpha) : 0.0;
  double lb = 0, db = 0, tb = 0;
  for (size_t i = 0; i < gi.tail.value.size(); ++i) {
    const double yv = gi.tail.value[i] + gamma, mm = gi.tail.mult[i];
    lb += mm * std::lgamma(yv);
    db += mm * digamma(yv);
    if (hess) tb += mm * trigamma(yv);
  }
  const double nbig = gi.tail.n_pos - gi.tail.n_one - gi.tail.n_two;
  const double ns = gi.tail.n_one + gi.tail.n_two;
  const double cv = lb - nbig * std::lgamma(gamma) + ns * log_gamma + gi.tail.n_two * std::log(gamma + 1.0);
  const double cd = db - nbig * digamma(gamma) + ns / gamma + gi.tail.n_two / (gamma + 1.0);
  const double ct = hess ? tb - nbig * trigamma(gamma) - ns / (gamma * gamma) -
                               gi.tail.n_two / ((gamma + 1.0) * (gamma + 1.0))
                         : 0.0;
  value = -total - sv - cv;
  grad.assign(np, 0.0);
  for (size_t j = 0; j < np; ++j) grad[j] = -gr[j];
  grad[nb] -= alpha_pr * sd;
  grad[nb + 1] -= cd;
  if (!hess) {
    h.clear();
    return;
  }
  const double* gram = G + 6;
  h.assign(np * np, 0.0);
  const double lr = lambda_pr / lambda;
  double hs = (alpha_dpr * log_lambda + 2.0 * alpha_pr * lr - alpha * lr * lr +
               alpha / lambda * lambda_dpr) *
              (double)k;
  double s_hbl0_imu = 0, s_ymmd_imu = 0, s_imu = 0, s_ym = 0, s_ymm = 0;
  for (size_t s = 0; s < k; ++s) {
    s_hbl0_imu += hbl0[s] * imu[s];
    s_ymmd_imu += ymm_d[s] * imu[s];
  }
  hs -= dbim * alpha_dpr - 2.0 * alpha_pr * lambda_pr * s_hbl0_imu - lambda_dpr * sum_ymm_d +
        2.0 * lambda_pr * lambda_pr * s_ymmd_imu;
  for (size_t s = 0; s < k; ++s) {
    s_imu += imu[s];
    s_ym += ym[s];
    s_ymm += ymm[s];
  }
  hs -= 2.0 * alpha_pr * lambda_pr * s_imu + alpha_dpr * smlog + lambda_dpr * s_ym -
        lambda_pr * lambda_pr * s_ymm;
  std::vector<double> hsb(nb), hgb(nb), apm(k);
  for (size_t j = 0; j < nb; ++j) {
    double a1 = 0, a2 = 0, a3 = 0, a4 = 0;
    for (size_t s = 0; s < k; ++s) {
      a1 += D41(s, j) * imu[s];
      a2 += D41(s, j) * ymm[s];
      a3 += XF(s, j) * hbl0[s] * imu[s];
      a4 += XF(s, j) * hbl0[s] * ymm[s];
    }
    hsb[j] = -alpha_pr * a1 + lambda_pr * a2 - (-alpha_pr * a3 + 2.0 * lambda_pr * a4);
  }
  for (size_t s = 0; s < k; ++s) apm[s] = alpha_pr * imu[s] - lambda_pr * ymm[s];
  const double hg = (double)P.n_cells / gamma - 2.0 * sum_tempa + sum_gt;
  for (size_t j = 0; j < nb; ++j) {
    double a = 0, b = 0;
    for (size_t s = 0; s < k; ++s) {
      a += T1(s, j) * ym[s];
      b += XF(s, j) * o[s * PS + 2] * ymm[s];
    }
    hgb[j] = a - b;
  }
  double hsg = 0.0;
  for (size_t s = 0; s < k; ++s) hsg += apm[s] * o[s * PS + 2];
  for (size_t j = 0; j < nb; ++j) {
    double a = 0, b = 0;
    for (size_t s = 0; s < k; ++s) {
      a += T2(s, j) * apm[s] * ym[s];
      b += XF(s, j) * apm[s] * o[s * PS + 3] * ymm[s];
    }
    hsb[j] += a - b;
  }
  {
    double add = 0.0;
    for (size_t s = 0; s < k; ++s) add += o[s * PS + 3] * apm[s] * apm[s];
    hs += add;
  }
  size_t tri = 0;
  for (size_t p = 0; p < nb; ++p)
    for (size_t q = p; q < nb; ++q, ++tri) {
      double acc = gram[tri];
      for (size_t s = 0; s < k; ++s) {
        const double xp = XF(s, p), xq = XF(s, q);
        const double yyi = ym[s] * ym[s] * imu[s];
        acc += (D41(s, p) * xq + D41(s, q) * xp) * ymm[s];
        acc -= yyi * (T2(s, p) * xq + T2(s, q) * xp);
        const double cross = xp * xq;
        acc -= cross * ymm[s];
        acc += o[s * PS + 3] * yyi * imu[s] * cross;
        acc -= 2.0 * cross * ymm_d[s] * imu[s];
      }
      h[p * np + q] = acc;
      h[q * np + p] = acc;
    }
  h[nb * np + nb] = hs;
  h[(nb + 1) * np + nb + 1] = hg;
  h[nb * np + nb + 1] = hsg;
  h[(nb + 1) * np + nb] = hsg;
  for (size_t j = 0; j < nb; ++j) {
    h[nb * np + j] = hsb[j];
    h[j * np + nb] = hsb[j];
    h[(nb + 1) * np + j] = hgb[j];
    h[j * np + nb + 1] = hgb[j];
  }
  for (double& v : h) v = -v;
  h[nb * np + nb] -= alpha_dpr * sd + alpha_pr * alpha_pr * st;
  h[(nb + 1) * np + nb + 1] -= ct;
}

// digamma / trigamma: upward recurrence to 20 then the asymptotic series
// (Abramowitz & Stegun 6.3.18, 6.4.12), the method of src/numeric/gamma.rs.
double digamma(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return NAN;
  double acc = 0.0;
  while (x < 20.0) {
    acc += 1.0 / x;
    x += 1.0;
  }
  const double z2 = 1.0 / (x * x);
  const double c[8] = {1.0 / 12.0,  -1.0 / 120.0,     1.0 / 252.0, -1.0 / 240.0,
                       1.0 / 132.0, -691.0 / 32760.0, 1.0 / 12.0,  -3617.0 / 8160.0};
  double s = 0.0;
  for (int i = 7; i >= 0; --i) s = c[i] + z2 * s;
  return std::log(x) - 0.5 / x - s * z2 - acc;
}

double trigamma(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return NAN;
  double acc = 0.0;
  while (x < 21.0) {
    acc += 1.0 / (x * x);
    x += 1.0;
  }
  const double z2 = 1.0 / (x * x);
  const double b[8] = {1.0 / 6.0,  -1.0 / 30.0,     1.0 / 42.0, -1.0 / 30.0,
                       5.0 / 66.0, -691.0 / 2730.0, 7.0 / 6.0,  -3617.0 / 510.0};
  double s = 0.0;
  for (int i = 7; i >= 0; --i) s = b[i] + z2 * s;
  return 1.0 / x + 0.5 * z2 + s * z2 / x + acc;
}

// ---------------------------------------------------------------------------
// Whether the device sweeps read the run's zero-count tables: option 2
// always, 0 never, 1 when the groups average at least kTableCells cells --
// a lookup is a chain of dependent loads (group, panel index, coefficients)
// that costs more than sweeping a few hundred cells (200 genes x 5,000 cells
// in groups of 500: 10.3 ms tabled against 6.7 ms dense; groups of 5,000:
// 3.1 against 3.8 ms).
constexpr size_t kTableCells = 4096;
bool use_tables(const ZeroTables& z, int32_t option) {
  if (option == 0 || z.groups.empty()) return false;
  if (option == 2) return true;
  return z.tabled_cells >= kTableCells * z.groups.size();
}

Prepared prepare(size_t n_genes, size_t n_cells, const uint32_t* indptr, const uint32_t* indices,
                 const double* data, const uint64_t* subject_id, const double* design,
                 size_t n_coef, const double* offset, const nebula_params_t& p) {
  const double los[2] = {p.min_subject, p.min_cell}, his[2] = {p.max_subject, p.max_cell};
  for (int i = 0; i < 2; ++i) {
    if (!(std::isfinite(los[i]) && los[i] > 0.0)) api_fail(NH_ERR_MUST_BE_POSITIVE, "min." + std::to_string(i));
    if (!(std::isfinite(his[i]) && his[i] > los[i]))
      api_fail(NH_ERR_INVALID_BOUNDS, "invalid bounds for variance component " + std::to_string(i));
  }
  if (!(std::isfinite(p.eps) && p.eps > 0.0)) api_fail(NH_ERR_MUST_BE_POSITIVE, "eps");
  if (p.mincp == 0) api_fail(NH_ERR_MUST_BE_POSITIVE, "mincp");
  if (p.reml)
    api_fail(NH_ERR_INVALID_ARGUMENT,
             "The GPU NEBULA path does not implement `reml`; use the CPU path.");
  if (n_genes == 0 || n_cells == 0) api_fail(NH_ERR_EMPTY_COUNTS, "empty counts");
  if (n_cells < 2) api_fail(NH_ERR_INVALID_ARGUMENT, "NEBULA needs more than one cell in the count matrix.");
  if (n_coef == 0) api_fail(NH_ERR_MUST_BE_POSITIVE, "n_coef");
  if (n_coef > 8)
    api_fail(NH_ERR_INVALID_ARGUMENT, "the device fit takes at most 8 design columns; got " +
                                          std::to_string(n_coef));
  if (!indptr || !subject_id || !design) api_fail(NH_ERR_INVALID_ARGUMENT, "null input");
  if (indptr[0] != 0) api_fail(NH_ERR_MALFORMED_SPARSE, "indptr must start at zero");
  if (indptr[n_genes] > 0 && (!indices || !data))
    api_fail(NH_ERR_INVALID_ARGUMENT, "null indices or data with stored entries");
  // Cells' subject runs (checked for contiguity further down, after the
  // counts, as before), then one pass over the counts, genes in parallel:
  // structure and values checked, and for a valid gene its subject totals
  // and positive count.  The first offending gene in order decides, structure
  // before values, as a gene-by-gene scan followed by a pass over the values
  // would report.
  std::vector<uint32_t> run_of(n_cells, 0);
  size_t n_runs = 1;
  for (size_t i = 1; i < n_cells; ++i) {
    if (subject_id[i] != subject_id[i - 1]) ++n_runs;
    run_of[i] = (uint32_t)(n_runs - 1);
  }
  std::vector<double> totals(n_genes * n_runs, 0.0);
  std::vector<size_t> positives(n_genes, 0);
  {
    enum : uint8_t { kOk, kIndptr, kRange, kOrder };
    std::vector<uint8_t> shape(n_genes, kOk), value(n_genes, 0);
    parallel_for(n_genes, 0, [&](size_t g) {
      if (indptr[g + 1] < indptr[g]) {
        shape[g] = kIndptr;
        return;
      }
      // each subject's total in a register until the gene leaves that
      // subject (cells increase, so each is visited once, in order)
      size_t np_ = 0, run = 0;
      double t = 0.0;
      bool bad = false;
      for (uint32_t i = indptr[g]; i < indptr[g + 1]; ++i) {
        if (indices[i] >= n_cells) {
          shape[g] = kRange;
          return;
        }
        if (i > indptr[g] && indices[i] <= indices[i - 1]) {
          shape[g] = kOrder;
          return;
        }
        const double y = data[i];
        if (!(std::isfinite(y) && y >= 0.0)) bad = true;
        if (y > 0.0 && !bad) {
          const size_t r = run_of[indices[i]];
          if (r != run) {
            totals[g * n_runs + run] = t;
            t = 0.0;
            run = r;
          }
          t += y;
          ++np_;
        }
      }
      totals[g * n_runs + run] = t;
      value[g] = bad ? 1 : 0;
      positives[g] = np_;
    }, 16);
    for (size_t g = 0; g < n_genes; ++g) {
      if (shape[g] == kIndptr) api_fail(NH_ERR_MALFORMED_SPARSE, "indptr must be non-decreasing");
      if (shape[g] == kRange) api_fail(NH_ERR_MALFORMED_SPARSE, "cell index out of range");
      if (shape[g] == kOrder)
        api_fail(NH_ERR_MALFORMED_SPARSE, "cell indices must increase within a gene");
    }
    for (size_t g = 0; g < n_genes; ++g)
      if (value[g])
        api_fail(NH_ERR_INVALID_ARGUMENT, "The counts hold a value that is not a non-negative finite number.");
  }
  Prepared P;
  P.n_genes = n_genes;
  P.n_cells = n_cells;
  P.n_coef = n_coef;
  P.params = p;
  P.fid.push_back(0);
  {
    // a subject seen in two runs: sort the runs' labels and look for a
    // repeat (O(K log K)); the message names the first label, in cell
    // order, whose run is not its first
    std::vector<uint64_t> label{subject_id[0]};
    for (size_t i = 1; i < n_cells; ++i)
      if (subject_id[i] != subject_id[i - 1]) {
        label.push_back(subject_id[i]);
        P.fid.push_back(i);
      }
    P.fid.push_back(n_cells);
    std::vector<uint64_t> sorted = label;
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end()) {
      std::unordered_set<uint64_t> seen;
      for (uint64_t l : label)
        if (!seen.insert(l).second)
          api_fail(NH_ERR_SUBJECTS_NOT_CONTIGUOUS,
                   "subject " + std::to_string(l) + " is not one contiguous run");
    }
  }
  P.k = P.fid.size() - 1;
  if (P.k != n_runs) api_fail(NH_ERR_INVALID_ARGUMENT, "internal: subject runs disagree");
  if (P.k < 2) api_fail(NH_ERR_TOO_FEW_SUBJECTS, "need at least 2 subjects, got " + std::to_string(P.k));
  P.log_offset.assign(n_cells, 0.0);
  double mean_offset = 1.0, cv = 0.0;
  if (offset) {
    for (size_t i = 0; i < n_cells; ++i)
      if (!(std::isfinite(offset[i]) && offset[i] > 0.0)) api_fail(NH_ERR_MUST_BE_POSITIVE, "offset");
    double s = 0.0;
    for (size_t i = 0; i < n_cells; ++i) s += offset[i];
    mean_offset = s / (double)n_cells;
    double ss = 0.0;
    for (size_t i = 0; i < n_cells; ++i) ss += (offset[i] - mean_offset) * (offset[i] - mean_offset);
    cv = std::sqrt(ss / (double)n_cells) / mean_offset;
    for (size_t i = 0; i < n_cells; ++i) P.log_offset[i] = std::log(offset[i]);
  }
  P.log_mean_offset = std::log(mean_offset);
  P.cv2 = cv * cv;
  // centre_design, row-major passes over the design (each column's sums
  // still run over the cells in order)
  P.centred.assign(n_cells * n_coef, 0.0);
  P.sds.assign(n_coef, 0.0);
  {
    std::vector<double> mean(n_coef, 0.0), var(n_coef, 0.0), sd(n_coef, 0.0);
    for (size_t i = 0; i < n_cells; ++i)
      for (size_t j = 0; j < n_coef; ++j) mean[j] += design[i * n_coef + j];
    for (size_t j = 0; j < n_coef; ++j) mean[j] /= (double)n_cells;
    for (size_t i = 0; i < n_cells; ++i)
      for (size_t j = 0; j < n_coef; ++j) {
        const double d = design[i * n_coef + j] - mean[j];
        P.centred[i * n_coef + j] = d;
        var[j] += d * d;
      }
    // per column: 0 scale by the sd, 1 the intercept (all ones), 2 a zero column
    std::vector<int> mode(n_coef, 0);
    for (size_t j = 0; j < n_coef; ++j) {
      sd[j] = std::sqrt(var[j] / (double)n_cells);
      if (sd[j] > 0.0) {
        P.sds[j] = sd[j];
      } else if (design[j] != 0.0) {
        mode[j] = 1;
        P.sds[j] = sd[j];
      } else {
        mode[j] = 2;
        P.sds[j] = -1.0;
      }
    }
    for (size_t i = 0; i < n_cells; ++i)
      for (size_t j = 0; j < n_coef; ++j) {
        if (mode[j] == 0)
          P.centred[i * n_coef + j] /= sd[j];
        else if (mode[j] == 1)
          P.centred[i * n_coef + j] = 1.0;
      }
  }
  for (double s : P.sds)
    if (s < 0.0) api_fail(NH_ERR_INVALID_ARGUMENT, "Some predictors are a zero vector.");
  std::vector<size_t> constant;
  for (size_t j = 0; j < n_coef; ++j)
    if (P.sds[j] == 0.0) constant.push_back(j);
  if (constant.empty()) api_fail(NH_ERR_MISSING_INTERCEPT, "the design has no intercept column");
  if (constant.size() > 1) api_fail(NH_ERR_INVALID_ARGUMENT, "More than one predictor has zero variation.");
  P.intercept = constant[0];
  P.cells_per_subject = (double)n_cells / (double)P.k;
  P.method = P.cells_per_subject < 30.0 ? NH_METHOD_HL : p.method;
  // zero-count tables (src/sc/zeros.rs:219-278, built as nebula.rs:641 builds
  // them): where groups of >= 32 cells sharing a design row within a subject
  // hold half the cells, every device sweep reads them, and stage one runs
  // the quasi-Newton, as edge-rs does
  {
    const auto tz = std::chrono::steady_clock::now();
    P.tables = build_zero_tables(P.centred.data(), P.log_offset.data(), n_cells, n_coef,
                                 reinterpret_cast<const uint64_t*>(P.fid.data()), P.k, 0, P.zt);
    P.t_tables =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - tz).count();
  }
  // cell-level columns (ptmg.rs:1962-2011)
  for (size_t j = 0; j < n_coef; ++j) {
    bool varies = false;
    for (size_t s = 0; s < P.k && !varies; ++s) {
      double first = P.centred[P.fid[s] * n_coef + j];
      for (size_t i = P.fid[s]; i < P.fid[s + 1]; ++i)
        if (P.centred[i * n_coef + j] != first) {
          varies = true;
          break;
        }
    }
    if (varies) P.cell_columns.push_back(j);
  }
  // filter, on the subject totals of the first pass (the subjects are the
  // contiguous runs it used)
  for (size_t g = 0; g < n_genes; ++g) {
    double total = 0.0;
    for (size_t s = 0; s < P.k; ++s) total += totals[g * P.k + s];
    const size_t stored = indptr[g + 1] - indptr[g];
    bool pooled = total / (double)n_cells > p.cpc && stored >= p.mincp;
    if (pooled && p.min_subjects > 0) {
      size_t expressing = 0;
      for (size_t s = 0; s < P.k; ++s)
        if (totals[g * P.k + s] / (double)(P.fid[s + 1] - P.fid[s]) > p.cpc) ++expressing;
      pooled = expressing >= p.min_subjects;
    }
    if (pooled) P.kept.push_back(g);
  }
  if (P.kept.empty()) api_fail(NH_ERR_NO_GENES_AFTER_FILTERING, "no genes survived the expression filter");
  // per kept gene, one more pass over its counts (genes in parallel, each
  // gene's sums in its own CSR order): summaries, count histogram and the
  // kept genes' CSR of positive counts, first touched by the threads that
  // fill it
  const size_t KG = P.kept.size();
  P.genes.resize(KG);
  P.gene_ptr.assign(KG + 1, 0);
  for (size_t idx = 0; idx < KG; ++idx)
    P.gene_ptr[idx + 1] = P.gene_ptr[idx] + positives[P.kept[idx]];
  P.cells.resize(P.gene_ptr[KG]);
  P.counts.resize(P.gene_ptr[KG]);
  parallel_for(KG, 0, [&](size_t idx) {
    const size_t g = P.kept[idx];
    GeneInfo& gi = P.genes[idx];
    double sum = 0.0;
    // count histogram: whole counts below kDense in an array, any other value
    // in a map; read back in increasing value
    constexpr size_t kDense = 4096;
    std::vector<double> dense(kDense, 0.0);
    size_t top = 0;
    std::map<double, double> other;
    size_t o = P.gene_ptr[idx];
    // counters in registers, stored once: neighbouring genes' records share
    // cache lines and run on different threads
    double n_pos = 0.0, n_one = 0.0, n_two = 0.0;
    for (uint32_t i = indptr[g]; i < indptr[g + 1]; ++i)
      if (data[i] > 0.0) {
        const double y = data[i];
        P.cells[o] = indices[i];
        P.counts[o] = y;
        ++o;
        sum += y;
        n_pos += 1.0;
        if (y == 1.0) {
          n_one += 1.0;
        } else if (y == 2.0) {
          n_two += 1.0;
        } else if (y < (double)kDense && y == std::floor(y)) {
          dense[(size_t)y] += 1.0;
          top = std::max(top, (size_t)y);
        } else {
          other[y] += 1.0;
        }
      }
    auto it = other.begin();
    auto put = [&](double v, double m) {
      gi.tail.value.push_back(v);
      gi.tail.mult.push_back(m);
    };
    for (size_t v = 3; v <= top; ++v) {
      if (dense[v] == 0.0) continue;
      for (; it != other.end() && it->first < (double)v; ++it) put(it->first, it->second);
      put((double)v, dense[v]);
    }
    for (; it != other.end(); ++it) put(it->first, it->second);
    gi.tail.n_pos = n_pos;
    gi.tail.n_one = n_one;
    gi.tail.n_two = n_two;
    gi.mean_count = sum / (double)n_cells;
    gi.totals.assign(totals.begin() + g * P.k, totals.begin() + (g + 1) * P.k);
  });
  return P;
}

// ---------------------------------------------------------------------------
// Stage one for every kept gene: coroutine searches, one ptmg launch per round.
void run_stage_one(Prepared& P, Resident& R, int host_threads, uint64_t& rounds) {
  const size_t G = P.kept.size(), nb = P.n_coef, np = nb + 2, k = P.k;
  const nebula_params_t& p = P.params;
  std::vector<S1Result> res(G);
  std::vector<S1Task> tasks;
  tasks.reserve(G);
  P.plans.resize(G);
  for (size_t i = 0; i < G; ++i) {
    GenePlan& pl = P.plans[i];
    const double lmc = std::log(P.genes[i].mean_count);
    pl.ord = P.genes[i].mean_count * P.cells_per_subject < 3.0 ? 3 : 1;
    std::vector<double> start(np, 0.0), lo(np, -100.0), hi(np, 100.0);
    start[P.intercept] = lmc - P.log_mean_offset;
    start[nb] = 1.0;
    start[nb + 1] = 1.0;
    lo[nb] = p.min_subject;
    lo[nb + 1] = p.min_cell;
    hi[nb] = p.max_subject;
    hi[nb + 1] = p.max_cell;
    pl.beta_start.assign(nb, 0.0);
    pl.beta_start[P.intercept] = lmc - P.log_mean_offset;
    tasks.push_back(stage_one_search(start, lo, hi, P.tables, nb, &res[i]));
  }
  {
    // each gene's x'y over its positive counts, on the device
    DevBuf<double> dx;
    dx.alloc(G * nb);
    NH_HIP(launch_xty(R.view(), G, dx.p, R.stream));
    std::vector<double> hx(G * nb);
    if (G) NH_HIP(hipMemcpyAsync(hx.data(), dx.p, G * nb * sizeof(double), hipMemcpyDeviceToHost,
                                 R.stream));
    NH_HIP(hipStreamSynchronize(R.stream));
    for (size_t i = 0; i < G; ++i)
      P.genes[i].xty.assign(hx.begin() + i * nb, hx.begin() + (i + 1) * nb);
  }
  const size_t out_stride = ptmg_out_doubles((int)nb, (int)k);
  // Workgroups per request: four per CU over the round's requests, with at
  // least 8 rows of 64 cells per unit.
  int dev_cus = 0;
  NH_HIP(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, R.device));
  const size_t target_wgs = 4 * (size_t)std::max(1, dev_cus);
  const size_t max_chunks = std::max<size_t>(1, P.n_cells / (kPtmgWaves * 512));

  // Two groups of genes (alternate kept genes) take turns: while the device
  // evaluates one group's requests, the host assembles the other group's
  // results and resumes its searches up to their next requests.  Requests,
  // points and results stay in mapped host memory that the kernels read and
  // write in place, so a round puts only kernel launches on the stream.
  struct Group {
    std::vector<size_t> genes, live;
    DevBuf<double> dscratch;
    PinnedBuf<PtmgRequest> hreq;
    PinnedBuf<double> hparams, hout{false};
    hipEvent_t done = nullptr;
    bool inflight = false;
    ~Group() {
      if (done) (void)hipEventDestroy(done);
    }
  };
  Group grp[2];
  for (size_t i = 0; i < G; ++i) grp[i % 2].genes.push_back(i);
  for (Group& g : grp) NH_HIP(hipEventCreateWithFlags(&g.done, hipEventDisableTiming));
  const int threads =
      host_threads > 0 ? host_threads : (int)std::min(8u, std::max(1u, std::thread::hardware_concurrency()));
  WorkPool pool(threads - 1);

  auto submit = [&](Group& g) {
    g.live.clear();
    for (size_t i : g.genes)
      if (tasks[i].waiting()) g.live.push_back(i);
    const size_t n = g.live.size();
    if (n == 0) return;
    // an even count makes the units per request a multiple of 8, so every
    // workgroup on a given unit lands on the same XCD
    size_t chunks = std::min(max_chunks, std::max<size_t>(1, (target_wgs + n - 1) / n));
    if (chunks % 2 == 1 && chunks + 1 <= max_chunks) ++chunks;
    const size_t sc_stride = ptmg_scratch_doubles((int)nb, (int)k, (int)chunks * kPtmgWaves);
    g.hreq.alloc(n);
    g.hparams.alloc(n * np);
    g.hout.alloc(n * out_stride);
    bool any_hess = false;
    for (size_t q = 0; q < n; ++q) {
      auto& pr = tasks[g.live[q]].h.promise();
      g.hreq.p[q] = PtmgRequest{(uint32_t)g.live[q], pr.req.hess ? 1u : 0u};
      any_hess = any_hess || pr.req.hess;
      std::copy(pr.req.x.begin(), pr.req.x.end(), g.hparams.p + q * np);
    }
    g.dscratch.alloc(n * sc_stride);
    PtmgLaunch L;
    L.dd = R.view();
    L.req = g.hreq.dev;
    L.params = g.hparams.dev;
    L.out = g.hout.dev;
    L.scratch = g.dscratch.p;
    L.stride = sc_stride;
    L.n_req = (int)n;
    L.chunks = (int)chunks;
    L.any_hess = any_hess;
    L.counters = R.counters.p;
    NH_HIP(launch_ptmg(L, R.stream));
    NH_HIP(hipEventRecord(g.done, R.stream));
    g.inflight = true;
    ++rounds;
  };
  auto collect = [&](Group& g) {
    NH_HIP(hipEventSynchronize(g.done));
    g.inflight = false;
    // every search is its own coroutine: assemble and resume in parallel
    pool.run(g.live.size(), [&](size_t q) {
      auto& task = tasks[g.live[q]];
      auto& pr = task.h.promise();
      EvalRes er;
      assemble_ptmg(P, P.genes[g.live[q]], g.hparams.p + q * np, g.hout.p + q * out_stride,
                    pr.req.hess, er.f, er.g, er.h);
      pr.res = std::move(er);
      task.h.resume();
    });
  };
  submit(grp[0]);
  submit(grp[1]);
  while (grp[0].inflight || grp[1].inflight) {
    for (Group& g : grp)
      if (g.inflight) {
        collect(g);
        submit(g);
      }
  }
  for (size_t i = 0; i < G; ++i) {
    GenePlan& pl = P.plans[i];
    pl.convergence = res[i].failed ? 0 : 1;
    pl.sigma = res[i].x[nb];
    pl.gamma = res[i].x[nb + 1];
    pl.beta_one = pl.beta_start;
    bool beta_fin = true;
    for (size_t j = 0; j < nb; ++j) beta_fin = beta_fin && std::isfinite(res[i].x[j]);
    if (beta_fin) pl.beta_one.assign(res[i].x.begin(), res[i].x.begin() + nb);
    if (P.method == NH_METHOD_HL) {
      pl.refit = 1;
    } else {
      double cv2p;
      if (P.cell_columns.empty()) {
        cv2p = P.cv2;
      } else {
        std::vector<double> mu(P.n_cells);
        double mean = 0.0;
        for (size_t c = 0; c < P.n_cells; ++c) {
          double eta = P.log_offset[c];
          for (size_t j : P.cell_columns) eta += P.centred[c * nb + j] * res[i].x[j];
          mu[c] = std::exp(eta);
          mean += mu[c];
        }
        mean /= (double)P.n_cells;
        double ss = 0.0;
        for (double v : mu) ss += (v - mean) * (v - mean);
        cv2p = mean <= 0.0 ? 0.0 : ss / (double)P.n_cells / (mean * mean);
      }
      const double gni = P.cells_per_subject * pl.gamma;
      if (gni < p.cutoff_cell || pl.convergence == 0 || std::isnan(cv2p)) {
        pl.refit = 1;
      } else {
        const double kappa_obs = gni / (1.0 + cv2p);
        const bool weak = kappa_obs < 20.0 || (kappa_obs < p.kappa && pl.sigma < 8.0 / kappa_obs);
        pl.refit = weak ? 2 : 0;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Warm-start features of every kept gene, from its counts and stage one's
// estimate (no counterpart in edge-rs):
//   0 log(sum y / sum offset)           1 positive fraction nnz / N
//   2 log(variance of y over cells)     3 log method-of-moments NB dispersion
//   4-6 10/50/90 % quantiles of the per-subject log rate log((Y_s + 1/2) / O_s)
//       minus feature 0                 7 their standard deviation
//   8 log nnz                           9 log(N / K)
//   10 log sigma^2 of stage one         11 log phi = -log gamma of stage one
void compute_features(Prepared& P) {
  const size_t G = P.kept.size(), k = P.k;
  const nebula_params_t& p = P.params;
  std::vector<double> osub(k, 0.0);
  double otot = 0.0;
  for (size_t s = 0; s < k; ++s) {
    for (size_t c = P.fid[s]; c < P.fid[s + 1]; ++c) osub[s] += std::exp(P.log_offset[c]);
    otot += osub[s];
  }
  const double n = (double)P.n_cells;
  P.features.assign(G * 12, 0.0f);
  std::vector<double> lr(k);
  for (size_t i = 0; i < G; ++i) {
    double sum = 0.0, sumsq = 0.0;
    for (uint64_t e = P.gene_ptr[i]; e < P.gene_ptr[i + 1]; ++e) {
      sum += P.counts[e];
      sumsq += P.counts[e] * P.counts[e];
    }
    const double nnz = (double)(P.gene_ptr[i + 1] - P.gene_ptr[i]);
    const double mean = sum / n, var = std::max(sumsq / n - mean * mean, 0.0);
    const double rate = std::log(std::max(sum, 0.5) / otot);
    const double mom = std::min(std::max((var - mean) / (mean * mean), 1e-4), 1e3);
    for (size_t s = 0; s < k; ++s) lr[s] = std::log((P.genes[i].totals[s] + 0.5) / osub[s]) - rate;
    std::vector<double> srt(lr);
    std::sort(srt.begin(), srt.end());
    auto quant = [&](double q) {
      const double at = q * (double)(k - 1);
      const size_t lo = (size_t)std::floor(at), hi = std::min(lo + 1, k - 1);
      return srt[lo] + (at - (double)lo) * (srt[hi] - srt[lo]);
    };
    double m = 0.0, ss = 0.0;
    for (double v : lr) m += v;
    m /= (double)k;
    for (double v : lr) ss += (v - m) * (v - m);
    const GenePlan& pl = P.plans[i];
    const double sig = std::min(std::max(pl.sigma, p.min_subject), p.max_subject);
    const double gam = std::min(std::max(pl.gamma, p.min_cell), p.max_cell);
    const double f[12] = {rate,
                          nnz / n,
                          std::log(var + 1e-8),
                          std::log(mom),
                          quant(0.1),
                          quant(0.5),
                          quant(0.9),
                          std::sqrt(ss / (double)k),
                          std::log(std::max(nnz, 1.0)),
                          std::log(n / (double)k),
                          std::log(sig),
                          -std::log(gam)};
    for (int j = 0; j < 12; ++j) P.features[i * 12 + j] = (float)f[j];
  }
}

// ---------------------------------------------------------------------------
// Stage two, plain engine: one search per gene on the host (SearchControl:
// BOBYQA over both components, or the subject variance alone for NEBULA-LN's
// weak genes), all live searches' points in one fit launch per round (edge-rs
// search_all, stage_two.rs:362-544, without the cohorts); a one-component
// search's polish stencil goes out as its three points in one round.  Inner
// fits start from the gene's first finite fit (stage_two.rs:36-43).
void run_stage_two_plain(Prepared& P, Resident& R, int waves, nebula_run_stats_t& st) {
  const size_t G = P.kept.size(), nb = P.n_coef, k = P.k;
  const nebula_params_t& p = P.params;
  struct Search {
    SearchControl c;
    std::vector<double> warm_beta, warm_lw;
  };
  std::vector<Search> S(G);
  const double lo[2] = {p.min_subject, p.min_cell}, hi[2] = {p.max_subject, p.max_cell};
  auto finish = [&](size_t i) {
    GenePlan& pl = P.plans[i];
    const SearchControl& c = S[i].c;
    pl.found = c.found != 0;
    pl.failed = c.failed != 0;
    int ord;
    c.result(pl.v, pl.objective, ord);
    pl.evaluations = c.evals;
  };
  for (size_t i = 0; i < G; ++i) {
    GenePlan& pl = P.plans[i];
    if (pl.refit == 0) continue;
    const double start[2] = {pl.sigma, pl.gamma};
    S[i].c.begin(pl.refit == 1 ? kSearchBoth : kSearchSubject, (int)pl.ord, start, nullptr, lo,
                 hi, 0.0, 0.0);
    pl.searched = true;
    if (!S[i].c.active) finish(i);
    ++st.genes_searched;
  }
  std::vector<DevRequest> req;
  std::vector<double> bi, lwi;
  std::vector<size_t> who;
  std::vector<DevReply> out;
  std::vector<double> ob, olw;
  for (;;) {
    req.clear();
    bi.clear();
    lwi.clear();
    who.clear();
    for (size_t i = 0; i < G; ++i) {
      SearchControl& c = S[i].c;
      if (!c.active) continue;
      // every point the search waits for: one, or a polish stencil's three
      const int n_pts = c.kind == kSearchSubject ? c.ss.n_ask : 1;
      for (int q = 0; q < n_pts; ++q) {
        double subject, cell;
        if (c.kind == kSearchSubject) {
          subject = c.ss.ask[q];
          cell = c.cold[1];
        } else {
          c.point(subject, cell);
        }
        const bool cold = S[i].warm_beta.empty();
        req.push_back(DevRequest{(uint32_t)i,
                                 (uint32_t)c.point_order() | (cold ? kOrdMeanStart : 0u),
                                 subject, cell});
        if (!cold) {
          bi.insert(bi.end(), S[i].warm_beta.begin(), S[i].warm_beta.end());
          lwi.insert(lwi.end(), S[i].warm_lw.begin(), S[i].warm_lw.end());
        } else {
          std::vector<double> b0 = P.plans[i].beta_one;
          b0[P.intercept] -= subject / 2.0;
          bi.insert(bi.end(), b0.begin(), b0.end());
          lwi.insert(lwi.end(), k, 0.0);
        }
        who.push_back(i);
      }
    }
    if (who.empty()) break;
    R.solve(req, bi, &lwi, p.eps, 0, waves, out, ob, olw, nullptr, true);
    ++st.rounds;
    for (size_t q = 0; q < who.size(); ++q) {
      const size_t i = who[q];
      Search& s = S[i];
      const DevReply& r = out[q];
      bool fin = true;
      for (size_t j = 0; j < nb; ++j) fin = fin && std::isfinite(ob[q * nb + j]);
      for (size_t j = 0; j < k; ++j) fin = fin && std::isfinite(olw[q * k + j]);
      double v = INFINITY;
      if (fin)
        v = profile_objective(req[q].subject, req[q].cell, r.ll, r.ll_prev, r.log_det, r.second,
                              (double)P.n_cells, (double)k, P.genes[i].tail);
      ++st.evaluations;
      // a stencil's later points were asked with the start this round began
      // with; the first finite fit of the gene becomes the start after it
      const bool was_cold = s.warm_beta.empty();
      if (s.c.tell(v) && was_cold) {
        s.warm_beta.assign(ob.begin() + q * nb, ob.begin() + (q + 1) * nb);
        s.warm_lw.assign(olw.begin() + q * k, olw.begin() + (q + 1) * k);
      }
      if (!s.c.active) finish(i);
    }
  }
}

// ---------------------------------------------------------------------------
// Stage three (finish_gene, nebula.rs:918-1000): one f64 opt_pml per gene on
// the device at the chosen variance components, then the covariance.
void run_stage_three(Prepared& P, Resident& R, int waves, nebula_fit_t* out) {
  const size_t G = P.kept.size(), nb = P.n_coef;
  const nebula_params_t& p = P.params;
  std::vector<DevRequest> req(G);
  std::vector<double> bi(G * nb);
  std::vector<int> initial(G);
  std::vector<double> sig(G), gam(G);
  for (size_t i = 0; i < G; ++i) {
    GenePlan& pl = P.plans[i];
    double sigma = pl.sigma, gamma = pl.gamma;
    int conv = pl.convergence;
    if (pl.refit != 0 && pl.searched) {
      if (pl.found) {
        sigma = pl.v[0];
        if (pl.refit == 1) gamma = pl.v[1];
        conv = pl.failed ? NH_CONV_OUTER_FAILED : NH_CONV_SUCCESS;
      } else {
        conv = NH_CONV_OUTER_FAILED;
      }
    }
    sig[i] = sigma;
    gam[i] = gamma;
    initial[i] = conv;
    req[i] = DevRequest{(uint32_t)i, 1u, sigma, gamma};
    for (size_t j = 0; j < nb; ++j) bi[i * nb + j] = pl.beta_start[j];
    bi[i * nb + P.intercept] -= sigma / 2.0;
  }
  std::vector<DevReply> rep;
  std::vector<double> ob, olw, oinfo;
  R.solve(req, bi, nullptr, p.eps, 1, waves, rep, ob, olw, &oinfo);
  const size_t packed = nb * (nb + 1) / 2;
  std::vector<double> coef(G * nb), se(G * nb), cov(G * packed), subj(G), cell(G);
  std::vector<int32_t> convs(G);
  std::vector<uint8_t> at_floor(G);
  std::vector<uint64_t> gidx(G);
  std::vector<double> scale = P.sds;
  scale[P.intercept] = 1.0;
  for (size_t i = 0; i < G; ++i) {
    const DevReply& r = rep[i];
    const double* info = &oinfo[i * nb * nb];
    const bool at_bound = sig[i] == p.max_subject || gam[i] == p.min_cell;
    int c2 = fit_convergence(r, info, nb, at_bound, initial[i]);
    std::vector<double> infov(info, info + nb * nb), inv;
    std::vector<double> gcov(packed), gse(nb);
    if (c2 != NH_CONV_SINGULAR && cholesky_inverse(infov, nb, inv)) {
      for (size_t j = 0; j < nb; ++j)
        for (size_t a = 0; a <= j; ++a) gcov[j * (j + 1) / 2 + a] = inv[a * nb + j];
      for (size_t j = 0; j < nb; ++j) gse[j] = std::sqrt(inv[j * nb + j]);
    } else {
      c2 = NH_CONV_SINGULAR;
      std::fill(gcov.begin(), gcov.end(), NAN);
      std::fill(gse.begin(), gse.end(), NAN);
    }
    for (size_t j = 0; j < nb; ++j) {
      coef[i * nb + j] = ob[i * nb + j] / scale[j];
      se[i * nb + j] = gse[j] / scale[j];
    }
    size_t t = 0;
    for (size_t j = 0; j < nb; ++j)
      for (size_t a = 0; a <= j; ++a, ++t) cov[i * packed + t] = gcov[j * (j + 1) / 2 + a] / (scale[a] * scale[j]);
    subj[i] = sig[i];
    cell[i] = 1.0 / gam[i];
    convs[i] = c2;
    at_floor[i] = sig[i] <= p.min_subject * (1.0 + 1e-4) ? 1 : 0;
    gidx[i] = P.kept[i];
  }
  out->n_genes = G;
  out->n_coef = nb;
  out->coefficients = dup(coef);
  out->covariance = dup(cov);
  out->se = dup(se);
  out->subject_overdispersion = dup(subj);
  out->cell_overdispersion = dup(cell);
  out->convergence = dup(convs);
  out->sigma_at_bound = dup(at_floor);
  out->gene_index = dup(gidx);
}

}  // namespace nh

using namespace nh;

extern "C" {

// A prepared run: set-up, upload and stage one done; stage two can be run
// (and timed) any number of times, then finished into a NebulaFit.
struct nh_session {
  Prepared P;
  Resident R;
  nebula_hip_options_t o;
  DevBuf<float> mlp;  // warm-start MLP parameters on the device
  bool has_mlp = false;
  int32_t tuned_wg = 0;  // autotuned threads per gene of the batched engines (0: none)
  double t_setup = 0, t_upload = 0, t_stage_one = 0;
  uint64_t s1_rounds = 0;
};

static int waves_of(const nebula_hip_options_t& o) {
  return o.wg_size >= 1024 ? 16 : (o.wg_size > 0 && o.wg_size <= 256 ? 4 : 8);
}

int nh_session_create(size_t n_genes, size_t n_cells, const uint32_t* indptr,
                      const uint32_t* indices, const double* data, const uint64_t* subject_id,
                      const double* design, size_t n_coef, const double* offset,
                      const nebula_params_t* params, const nebula_hip_options_t* options,
                      nh_session_t** out, char* err, size_t err_len) {
  if (!out) {
    put_err(err, err_len, "null output");
    return NH_ERR_INVALID_ARGUMENT;
  }
  *out = nullptr;
  nebula_params_t p;
  nebula_params_default(&p);
  if (params) p = *params;
  std::unique_ptr<nh_session> S(new nh_session);
  nebula_hip_options_default(&S->o);
  if (options) S->o = *options;
  try {
    const auto t0 = Clock::now();
    S->P = prepare(n_genes, n_cells, indptr, indices, data, subject_id, design, n_coef, offset, p);
    const auto t1 = Clock::now();
    S->R.upload(n_cells, n_coef, S->P.k, S->P.centred.data(), S->P.log_offset.data(),
                reinterpret_cast<const uint64_t*>(S->P.fid.data()), S->P.kept.size(),
                S->P.gene_ptr.data(), S->P.cells.data(), S->P.counts.data(), S->o.device);
    if (S->P.tables && use_tables(S->P.zt, S->o.zero_tables)) S->R.set_tables(S->P.zt);
    const auto t2 = Clock::now();
    run_stage_one(S->P, S->R, S->o.host_threads, S->s1_rounds);
    compute_features(S->P);
    if (S->o.mlp_weights) {
      S->mlp.upload(S->o.mlp_weights, nh_mlp_param_count(), S->R.stream);
      NH_HIP(hipStreamSynchronize(S->R.stream));
      S->has_mlp = true;
    }
    const auto t3 = Clock::now();
    S->t_setup = secs(t0, t1);
    S->t_upload = secs(t1, t2);
    S->t_stage_one = secs(t2, t3);
  } catch (const ApiError& e) {
    put_err(err, err_len, e.message);
    return e.code;
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  *out = S.release();
  return NH_OK;
}

size_t nh_session_genes(const nh_session_t* s) { return s ? s->P.kept.size() : 0; }

int nh_session_features(const nh_session_t* s, float* out) {
  if (!s || !out) return NH_ERR_INVALID_ARGUMENT;
  std::memcpy(out, s->P.features.data(), s->P.features.size() * sizeof(float));
  return NH_OK;
}

int nh_session_warm_rejected(const nh_session_t* s, uint8_t* out) {
  if (!s || !out) return NH_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < s->P.plans.size(); ++i) out[i] = s->P.plans[i].warm_rejected ? 1 : 0;
  return NH_OK;
}

int nh_session_set_mlp(nh_session_t* s, const float* params, char* err, size_t err_len) {
  if (!s || !params) {
    put_err(err, err_len, "null argument");
    return NH_ERR_INVALID_ARGUMENT;
  }
  try {
    s->mlp.upload(params, nh_mlp_param_count(), s->R.stream);
    NH_HIP(hipStreamSynchronize(s->R.stream));
    s->has_mlp = true;
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  return NH_OK;
}

int nh_session_stage_one(const nh_session_t* s, double* sigma_gamma, int32_t* ord) {
  if (!s) return NH_ERR_INVALID_ARGUMENT;
  for (size_t i = 0; i < s->P.plans.size(); ++i) {
    if (sigma_gamma) {
      sigma_gamma[2 * i] = s->P.plans[i].sigma;
      sigma_gamma[2 * i + 1] = s->P.plans[i].gamma;
    }
    if (ord) ord[i] = (int32_t)s->P.plans[i].ord;
  }
  return NH_OK;
}

int nh_session_stage_two(nh_session_t* s, int32_t engine, double* minimisers,
                         double* objectives, int32_t* evaluations,
                         nebula_run_stats_t* stats, char* err, size_t err_len) {
  if (!s) {
    put_err(err, err_len, "null session");
    return NH_ERR_INVALID_ARGUMENT;
  }
  nebula_run_stats_t st;
  std::memset(&st, 0, sizeof(st));
  try {
    for (GenePlan& pl : s->P.plans) {
      pl.searched = pl.found = pl.failed = false;
      pl.v[0] = pl.v[1] = 0.0;
      pl.evaluations = 0;
      pl.objective = 0.0;
      pl.warm_rejected = false;
    }
    s->R.reset_counters();
    s->R.kernel_ms = 0.0;
    s->R.launches = 0;
    if (engine == NH_ENGINE_LEARNED && !s->has_mlp)
      api_fail(NH_ERR_INVALID_ARGUMENT, "the learned engine needs MLP weights");
    LearnedStart ls;
    ls.params = s->mlp.p;
    if (s->o.learned_step > 0.0) ls.step = s->o.learned_step;
    if (s->o.learned_xtol > 0.0) ls.xtol_rel = s->o.learned_xtol;
    const auto t3 = Clock::now();
    nebula_hip_options_t ob = s->o;
    if (ob.wg_size == 0 && s->tuned_wg > 0) ob.wg_size = s->tuned_wg;
    if (engine == NH_ENGINE_BATCHED || engine == NH_ENGINE_LEARNED)
      run_stage_two_batched(s->P, s->R, ob, st, engine == NH_ENGINE_LEARNED ? &ls : nullptr);
    else
      run_stage_two_plain(s->P, s->R, waves_of(s->o), st);
    const auto t4 = Clock::now();
    FitCounters c = s->R.read_counters();
    st.sweeps_f32 = c.sweeps_f32;
    st.sweeps_f64 = c.sweeps_f64;
    st.cell_sweeps = c.cells_f32;
    st.nnz_sweeps = c.nnz_f32;
    st.cell_sweeps_f64 = c.cells_f64;
    st.nnz_sweeps_f64 = c.nnz_f64;
    st.fit_kernel_ms = s->R.kernel_ms;
    st.fit_kernel_launches = s->R.launches;
    st.t_setup = s->t_setup;
    st.t_upload = s->t_upload;
    st.t_stage_one = s->t_stage_one;
    st.t_stage_two = secs(t3, t4);
    st.zero_groups = s->R.tabled ? s->R.zgroup.n : 0;
    st.t_tables = s->P.t_tables;
    for (size_t i = 0; i < s->P.plans.size(); ++i) {
      const GenePlan& pl = s->P.plans[i];
      const bool ok = pl.searched && pl.found;
      if (minimisers) {
        minimisers[2 * i] = ok ? pl.v[0] : pl.sigma;
        minimisers[2 * i + 1] = ok ? pl.v[1] : pl.gamma;
      }
      if (objectives) objectives[i] = ok ? pl.objective : NAN;
      if (evaluations) evaluations[i] = pl.evaluations;
      if (pl.warm_rejected) ++st.warm_fallbacks;
    }
  } catch (const ApiError& e) {
    put_err(err, err_len, e.message);
    return e.code;
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  if (stats) *stats = st;
  return NH_OK;
}

int nh_session_autotune(nh_session_t* s, int32_t engine, int32_t* wg_size, double* ms, char* err,
                        size_t err_len) {
  if (!s) {
    put_err(err, err_len, "null session");
    return NH_ERR_INVALID_ARGUMENT;
  }
  const int32_t sizes[3] = {256, 512, 1024};
  double best = INFINITY;
  int32_t pick = 0;
  int seen[3] = {0, 0, 0};
  for (int c = 0; c < 3; ++c) {
    // sizes that launch the same kernel as an earlier one are not rerun
    const int w = effective_waves((int)s->P.n_coef, sizes[c] / 64);
    bool dup = false;
    for (int d = 0; d < c; ++d) dup = dup || seen[d] == w;
    seen[c] = w;
    if (dup) {
      if (ms) ms[c] = NAN;
      continue;
    }
    s->tuned_wg = sizes[c];
    // the faster of two runs: one slow run (a host hiccup) does not decide
    double t = INFINITY;
    for (int rep = 0; rep < 2; ++rep) {
      nebula_run_stats_t st;
      const int rc = nh_session_stage_two(s, engine, nullptr, nullptr, nullptr, &st, err, err_len);
      if (rc != NH_OK) return rc;
      t = std::min(t, st.t_stage_two);
    }
    if (ms) ms[c] = 1e3 * t;
    if (t < best) {
      best = t;
      pick = sizes[c];
    }
  }
  s->tuned_wg = pick;
  if (wg_size) *wg_size = pick;
  return NH_OK;
}

int nh_session_finish(nh_session_t* s, nebula_fit_t* out, double* t_stage_three, char* err,
                      size_t err_len) {
  if (!s || !out) {
    put_err(err, err_len, "null argument");
    return NH_ERR_INVALID_ARGUMENT;
  }
  std::memset(out, 0, sizeof(*out));
  try {
    const auto t0 = Clock::now();
    run_stage_three(s->P, s->R, waves_of(s->o), out);
    if (t_stage_three) *t_stage_three = secs(t0, Clock::now());
  } catch (const ApiError& e) {
    nebula_fit_free(out);
    put_err(err, err_len, e.message);
    return e.code;
  } catch (const std::exception& e) {
    nebula_fit_free(out);
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  return NH_OK;
}

int nh_session_release_plans(nh_session_t* s) {
  if (!s) return NH_ERR_INVALID_ARGUMENT;
  s->R.stage_two_plans.clear();
  return NH_OK;
}

void nh_session_free(nh_session_t* s) { delete s; }

int nebula_sparse_hip_ex(size_t n_genes, size_t n_cells, const uint32_t* indptr,
                         const uint32_t* indices, const double* data, const uint64_t* subject_id,
                         const double* design, size_t n_coef, const double* offset,
                         const nebula_params_t* params, const nebula_hip_options_t* options,
                         nebula_fit_t* out, nebula_run_stats_t* stats, char* err,
                         size_t err_len) {
  if (!out) {
    put_err(err, err_len, "null output");
    return NH_ERR_INVALID_ARGUMENT;
  }
  std::memset(out, 0, sizeof(*out));
  const auto t0 = Clock::now();
  nh_session_t* s = nullptr;
  int rc = nh_session_create(n_genes, n_cells, indptr, indices, data, subject_id, design, n_coef,
                             offset, params, options, &s, err, err_len);
  if (rc != NH_OK) return rc;
  nebula_run_stats_t st;
  std::memset(&st, 0, sizeof(st));
  const int32_t engine = s->o.learned ? NH_ENGINE_LEARNED : s->o.engine;
  rc = nh_session_stage_two(s, engine, nullptr, nullptr, nullptr, &st, err, err_len);
  if (rc == NH_OK) rc = nh_session_finish(s, out, &st.t_stage_three, err, err_len);
  nh_session_free(s);
  st.t_total = secs(t0, Clock::now());
  if (rc == NH_OK && stats) *stats = st;
  return rc;
}

int nebula_sparse_hip(size_t n_genes, size_t n_cells, const uint32_t* indptr,
                      const uint32_t* indices, const double* data, const uint64_t* subject_id,
                      const double* design, size_t n_coef, const double* offset,
                      const nebula_params_t* params, nebula_fit_t* out, char* err, size_t err_len) {
  return nebula_sparse_hip_ex(n_genes, n_cells, indptr, indices, data, subject_id, design, n_coef,
                              offset, params, nullptr, out, nullptr, err, err_len);
}

void nebula_fit_free(nebula_fit_t* f) {
  if (!f) return;
  std::free(f->coefficients);
  std::free(f->covariance);
  std::free(f->se);
  std::free(f->subject_overdispersion);
  std::free(f->cell_overdispersion);
  std::free(f->convergence);
  std::free(f->sigma_at_bound);
  std::free(f->gene_index);
  std::memset(f, 0, sizeof(*f));
}

}  // extern "C"

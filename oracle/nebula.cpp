// The NEBULA driver, f64 on the CPU: a restatement of src/sc/nebula.rs.
//
// nebula_sparse_with set-up  <- nebula.rs:494-669 (validation, subject runs,
//                               offsets, centred design, totals, filter)
// plan_gene (stage one)      <- nebula.rs:812-901, minimise_marginal :1075,
//                               descend_marginal :1132, bound_restart :1190,
//                               newton_marginal :1225
// VarianceObjective          <- nebula.rs:1365-1505 (request, tail, assemble)
// StageTwoSearch (HL)        <- nebula.rs:1568-1760 (BOBYQA pass, order-one
//                               retry, abandon on a non-finite value)
// finish_gene (stage three)  <- nebula.rs:918-1000, cholesky_inverse :2107
// assemble                   <- nebula.rs:2233-2277
// Helpers from src/sc/ptmg.rs: offset_summary :1827, centre_design :1748,
// cell_level_columns :1962, design_cv :1888, cumsum_y :1596,
// positive_indices :1681.  Zero-count tables (zeros.cpp, src/sc/zeros.rs) are
// built as nebula.rs:641 builds them and used by stage one (quasi-Newton on
// the tabled marginal likelihood) and by every penalised fit.
// NEBULA-LN's one-component refit (Nelder-Mead plus polish, nebula.rs:1798-1950)
// is restated in search.hpp.
//
// Test infrastructure: only the tests, the smoke check and the benchmark's CPU
// baseline load this library, as the checker.
#include <algorithm>
#include <array>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "../edge-rs_amd/csrc/bobyqa.hpp"
#include "lbfgsb.hpp"
#include "oracle.hpp"
#include "search.hpp"

namespace oracle {

namespace {

constexpr double kBetaBound = 100.0;
constexpr double kMinCellsPerSubjectLn = 30.0;
constexpr double kHighOrderCutoff = 3.0;
constexpr double kBoundSlack = 1.0 + 1e-4;
constexpr double kKappaFloor = 20.0;
constexpr double kKappaSigmaNumerator = 8.0;
constexpr double kStageOneFtol = 1e-6;
constexpr double kStageOnePgtol = 1e-7;
constexpr int kNewtonMaxIter = 50;
constexpr double kArmijo = 1e-4;
constexpr int kMaxHalvings = 40;
constexpr double kShift = 1e-8;
constexpr int kRestarts = 3;
constexpr double kRestartGain = 1e-6;
constexpr int kLbfgsMaxIter = 300;
constexpr int kLbfgsMaxLineSearch = 40;
constexpr int kOuterFailed = -50;

bool cholesky_lower(std::vector<double>& a, std::size_t n) {
  for (std::size_t i = 0; i < n; ++i)
    for (std::size_t j = 0; j <= i; ++j) {
      double sum = a[i * n + j];
      for (std::size_t k = 0; k < j; ++k) sum -= a[i * n + k] * a[j * n + k];
      if (i == j) {
        if (sum <= 0.0 || std::isnan(sum)) return false;
        a[i * n + j] = std::sqrt(sum);
      } else {
        a[i * n + j] = sum / a[j * n + j];
      }
    }
  return true;
}

bool cholesky_factor(const std::vector<double>& a, std::size_t n, std::vector<double>& l) {
  l = a;
  if (!cholesky_lower(l, n)) return false;
  for (std::size_t i = 0; i < n; ++i)
    if (!std::isfinite(l[i * n + i])) return false;
  return true;
}

bool solve_factored(const std::vector<double>& l, std::size_t n, std::vector<double>& x) {
  for (std::size_t i = 0; i < n; ++i) {
    double s = x[i];
    for (std::size_t k = 0; k < i; ++k) s -= l[i * n + k] * x[k];
    x[i] = s / l[i * n + i];
  }
  for (std::size_t i = n; i-- > 0;) {
    double s = x[i];
    for (std::size_t k = i + 1; k < n; ++k) s -= l[k * n + i] * x[k];
    x[i] = s / l[i * n + i];
  }
  for (double v : x)
    if (!std::isfinite(v)) return false;
  return true;
}

bool cholesky_solve(const std::vector<double>& a, std::size_t n,
                    const std::vector<double>& b, std::vector<double>& x) {
  std::vector<double> l;
  if (!cholesky_factor(a, n, l)) return false;
  x = b;
  return solve_factored(l, n, x);
}

bool cholesky_inverse(const std::vector<double>& a, std::size_t n, std::vector<double>& inv) {
  std::vector<double> l;
  if (!cholesky_factor(a, n, l)) return false;
  inv.assign(n * n, 0.0);
  std::vector<double> col(n);
  for (std::size_t j = 0; j < n; ++j) {
    std::fill(col.begin(), col.end(), 0.0);
    col[j] = 1.0;
    if (!solve_factored(l, n, col)) return false;
    for (std::size_t i = 0; i < n; ++i) inv[i * n + j] = col[i];
  }
  return true;
}

struct Shared {
  const double* design;
  const double* log_offset;
  const std::size_t* fid;
  std::size_t n_cells, n_coef, k, intercept;
  double log_mean_offset, cells_per_subject, cv2;
  std::vector<std::size_t> cell_columns;
  int method;
  Params params;
  bool tables;                 // zero-count tables apply (zeros.rs:219-262)
  const ZeroCells* zeros;      // the tables, when they apply and are used
};

struct GeneCounts {
  std::vector<std::size_t> cells;
  std::vector<double> counts;
  double mean_count = 0;
  std::size_t n_one = 0, n_two = 0;
};

struct Plan {
  unsigned ord;
  double sigma, gamma;
  int convergence;
  std::vector<double> beta_start;
  int refit;  // 0 none, 1 both, 2 subject only
};

std::vector<double> bound_restart(const std::vector<double>& best,
                                  const std::vector<double>& start,
                                  const std::vector<double>& lo,
                                  const std::vector<double>& hi, bool& moved) {
  std::vector<double> x0 = best;
  moved = false;
  for (std::size_t j = 0; j < best.size(); ++j)
    for (double bound : {lo[j], hi[j]})
      if (best[j] == bound && start[j] != bound && bound * start[j] > 0.0) {
        x0[j] = std::sqrt(bound * start[j]);
        moved = true;
      }
  return x0;
}

bool newton_marginal(Ptmg& pt, const std::vector<double>& x0,
                     const std::vector<double>& lo, const std::vector<double>& hi,
                     std::vector<double>& xout, double& fout) {
  const std::size_t n = x0.size();
  std::vector<double> x = x0, g(n), h(n * n);
  double f = pt.value_gradient_hessian(x.data(), g.data(), h.data());
  auto finite = [](double v, const std::vector<double>& gg, const std::vector<double>& hh) {
    if (!std::isfinite(v)) return false;
    for (double a : gg)
      if (!std::isfinite(a)) return false;
    for (double a : hh)
      if (!std::isfinite(a)) return false;
    return true;
  };
  if (!finite(f, g, h)) return false;
  const bool speculative = pt.gene().n_coef <= 8;
  for (int it = 0; it < kNewtonMaxIter; ++it) {
    double projected = 0.0;
    for (std::size_t i = 0; i < n; ++i)
      projected = std::max(projected, std::fabs(std::min(std::max(x[i] - g[i], lo[i]), hi[i]) - x[i]));
    if (projected < kStageOnePgtol) {
      xout = x;
      fout = f;
      return true;
    }
    std::vector<std::size_t> free;
    for (std::size_t i = 0; i < n; ++i)
      if (!((x[i] <= lo[i] && g[i] > 0.0) || (x[i] >= hi[i] && g[i] < 0.0))) free.push_back(i);
    const std::size_t m = free.size();
    std::vector<double> rhs(m);
    double largest = 0.0;
    for (std::size_t p = 0; p < m; ++p) {
      rhs[p] = -g[free[p]];
      largest = std::max(largest, std::fabs(h[free[p] * n + free[p]]));
    }
    double shift = 0.0;
    std::vector<double> step_free;
    for (;;) {
      std::vector<double> block(m * m);
      for (std::size_t p = 0; p < m; ++p) {
        for (std::size_t q = 0; q < m; ++q) block[p * m + q] = h[free[p] * n + free[q]];
        block[p * m + p] += shift;
      }
      if (cholesky_solve(block, m, rhs, step_free)) break;
      shift = shift == 0.0 ? kShift * std::max(largest, 1.0) : shift * 10.0;
      if (!std::isfinite(shift)) return false;
    }
    std::vector<double> d(n, 0.0);
    for (std::size_t p = 0; p < m; ++p) d[free[p]] = step_free[p];
    double alpha = 1.0;
    bool accepted = false;
    std::vector<double> xt(n), gt(n), ht(n * n);
    double ft = 0.0;
    bool have_derivs = false;
    for (int hv = 0; hv < kMaxHalvings; ++hv) {
      for (std::size_t i = 0; i < n; ++i) xt[i] = std::min(std::max(x[i] + alpha * d[i], lo[i]), hi[i]);
      double dec = 0.0;
      for (std::size_t i = 0; i < n; ++i) dec += g[i] * (xt[i] - x[i]);
      bool ok;
      if (speculative) {
        ft = pt.value_gradient_hessian(xt.data(), gt.data(), ht.data());
        ok = finite(ft, gt, ht);
        have_derivs = true;
      } else {
        ft = pt.value_gradient(xt.data(), gt.data());
        ok = std::isfinite(ft);
        have_derivs = false;
      }
      if (ok && ft <= f + kArmijo * dec) {
        accepted = true;
        break;
      }
      alpha *= 0.5;
    }
    if (!accepted) return false;
    if ((f - ft) <= kStageOneFtol) {
      xout = xt;
      fout = ft;
      return true;
    }
    if (!have_derivs) {
      ft = pt.value_gradient_hessian(xt.data(), gt.data(), ht.data());
      if (!finite(ft, gt, ht)) return false;
    }
    x = xt;
    f = ft;
    g = gt;
    h = ht;
  }
  return false;
}

std::pair<std::vector<double>, double> descend_marginal(
    Ptmg& pt, bool dense, const std::vector<double>& x0, double f0,
    const std::vector<double>& lo, const std::vector<double>& hi) {
  if (dense) {
    std::vector<double> x;
    double f;
    if (newton_marginal(pt, x0, lo, hi, x, f) && f < f0) return {x, f};
  }
  auto fg = [&](const std::vector<double>& x, std::vector<double>& g) {
    return pt.value_gradient(x.data(), g.data());
  };
  LbfgsbResult r = lbfgsb_minimise(fg, x0, lo, hi, kStageOneFtol / std::max(std::fabs(f0), 1.0),
                                   kStageOnePgtol, kLbfgsMaxIter, kLbfgsMaxLineSearch);
  if (r.ok && r.f < f0) return {r.x, r.f};
  return {x0, f0};
}

std::pair<std::vector<double>, bool> minimise_marginal(
    Ptmg& pt, bool tables, const std::vector<double>& start,
    const std::vector<double>& lo, const std::vector<double>& hi) {
  const std::size_t n = start.size();
  std::vector<double> x0(n), g(n);
  for (std::size_t j = 0; j < n; ++j) x0[j] = std::min(std::max(start[j], lo[j]), hi[j]);
  double f0 = pt.value_gradient(x0.data(), g.data());
  if (!std::isfinite(f0)) return {x0, true};
  auto bf = descend_marginal(pt, !tables, x0, f0, lo, hi);
  std::vector<double> best = bf.first, last = bf.first, anchor = x0;
  double best_f = bf.second;
  for (int r = 0; r < kRestarts; ++r) {
    bool moved;
    std::vector<double> x1 = bound_restart(last, anchor, lo, hi, moved);
    if (!moved) break;
    double f1 = pt.value_gradient(x1.data(), g.data());
    if (!std::isfinite(f1)) break;
    auto xf = descend_marginal(pt, !tables, x1, f1, lo, hi);
    if (xf.second < best_f - kRestartGain) {
      best = xf.first;
      best_f = xf.second;
    }
    last = xf.first;
    anchor = x1;
  }
  return {best, false};
}

double design_cv(const Shared& sh, const double* beta) {
  const std::size_t n = sh.n_cells, nb = sh.n_coef;
  std::vector<double> mu(n);
  for (std::size_t i = 0; i < n; ++i) {
    double eta = sh.log_offset[i];
    for (std::size_t c : sh.cell_columns) eta += sh.design[i * nb + c] * beta[c];
    mu[i] = std::exp(eta);
  }
  double mean = 0.0;
  for (double v : mu) mean += v;
  mean /= (double)n;
  if (mean <= 0.0) return 0.0;
  double ss = 0.0;
  for (double v : mu) ss += (v - mean) * (v - mean);
  return ss / (double)n / (mean * mean);
}

Plan plan_gene(const Shared& sh, const GeneCounts& gc, const double* totals) {
  const std::size_t nb = sh.n_coef;
  const Params& p = sh.params;
  Plan plan;
  plan.ord = gc.mean_count * sh.cells_per_subject < kHighOrderCutoff ? 3 : 1;
  GeneData gd{sh.design, sh.log_offset, sh.n_cells, nb, gc.counts.data(), gc.cells.data(),
              gc.counts.size(), totals, sh.fid, sh.k, sh.zeros};
  Ptmg pt(gd);
  double lmc = std::log(gc.mean_count);
  std::vector<double> start(nb + 2, 0.0), lo(nb + 2, -kBetaBound), hi(nb + 2, kBetaBound);
  start[sh.intercept] = lmc - sh.log_mean_offset;
  start[nb] = 1.0;
  start[nb + 1] = 1.0;
  lo[nb] = p.min_subject;
  lo[nb + 1] = p.min_cell;
  hi[nb] = p.max_subject;
  hi[nb + 1] = p.max_cell;
  auto res = minimise_marginal(pt, sh.tables, start, lo, hi);
  plan.convergence = res.second ? 0 : 1;
  plan.sigma = res.first[nb];
  plan.gamma = res.first[nb + 1];
  plan.beta_start.assign(nb, 0.0);
  plan.beta_start[sh.intercept] = lmc - sh.log_mean_offset;
  if (sh.method == 1) {
    plan.refit = 1;
  } else {
    double cv2p = sh.cell_columns.empty() ? sh.cv2 : design_cv(sh, res.first.data());
    double gni = sh.cells_per_subject * plan.gamma;
    if (gni < p.cutoff_cell || plan.convergence == 0 || std::isnan(cv2p)) {
      plan.refit = 1;
    } else {
      double kappa_obs = gni / (1.0 + cv2p);
      bool weak = kappa_obs < kKappaFloor ||
                  (kappa_obs < p.kappa && plan.sigma < kKappaSigmaNumerator / kappa_obs);
      plan.refit = weak ? 2 : 0;
    }
  }
  return plan;
}

struct Objective {
  const PmlData* data;
  const std::vector<double>* beta_start;
  std::size_t intercept;
  PmlParams params;
  double n_cells, k;
  const std::vector<double>* counts;
  double n_pos, n_one, n_two;
  // NEBULA-LN's one-component restriction: the cell-level size held fixed
  // (VarianceObjective::fixed_cell, nebula.rs:1388); NaN for both components
  double fixed_cell = std::numeric_limits<double>::quiet_NaN();
  double tail(double cell) const {
    double t = 0.0;
    for (double y : *counts)
      if (y != 1.0 && y != 2.0) t += ln_gamma(y + cell);
    return t;
  }
  double assemble(double subject, double cell, const PmlResult& fit) const {
    if (fit.second_order < -1.0) return std::numeric_limits<double>::infinity();
    auto ab = gamma_prior(std::exp(subject));
    double base = std::isnan(fit.log_likelihood) ? fit.log_likelihood_prev : fit.log_likelihood;
    double ll = base + n_cells * cell * std::log(cell) + k * ab.first * std::log(ab.second) -
                k * ln_gamma(ab.first);
    ll += tail(cell) - (n_pos - n_one - n_two) * ln_gamma(cell) + (n_one + n_two) * std::log(cell) +
          n_two * std::log(cell + 1.0);
    ll += -0.5 * fit.log_det + std::log(1.0 + fit.second_order);
    return std::isfinite(ll) ? -ll : std::numeric_limits<double>::infinity();
  }
  double value(const double* x) const {
    double subject = x[0], cell = std::isnan(fixed_cell) ? x[1] : fixed_cell;
    if (!(std::isfinite(subject) && subject > 0.0 && std::isfinite(cell) && cell > 0.0))
      return std::numeric_limits<double>::infinity();
    std::vector<double> beta = *beta_start;
    beta[intercept] -= subject / 2.0;
    try {
      PmlResult fit = opt_pml(*data, beta.data(), subject, cell, params);
      return assemble(subject, cell, fit);
    } catch (const Error&) {
      return std::numeric_limits<double>::infinity();
    }
  }
};

// StageTwoSearch over both components (nebula.rs:1743-1796): one BOBYQA pass,
// retried from the start at order one when a raised-order pass hit a
// rejected evaluation.  f(x, order) is the objective.
template <class F>
SearchOutcome bobyqa_search(F&& f, unsigned ord, const double start[2], const double lo[2],
                            const double hi[2], double* fbest) {
  SearchOutcome out;
  unsigned order = ord;
  for (;;) {
    double x0[2] = {std::min(std::max(start[0], lo[0]), hi[0]),
                    std::min(std::max(start[1], lo[1]), hi[1])};
    nh::Bobyqa<2> b;
    b.init(2, x0, lo, hi);
    bool found = b.status != nh::BOBYQA_BAD_BOX;
    while (found && b.want) {
      std::vector<double> xv{b.xask[0], b.xask[1]};
      double v = f(xv, order);
      ++out.evaluations;
      if (!std::isfinite(v)) {
        found = false;
        break;
      }
      b.tell(v);
    }
    if (found) {
      out.x = {b.xbest_out[0], b.xbest_out[1]};
      out.failed = b.status == nh::BOBYQA_ROUNDOFF_LIMITED;
      if (fbest) *fbest = b.fbest_out;
    }
    if (found || order == 1) {
      out.found = found;
      out.order = order;
      return out;
    }
    order = 1;
  }
}

// refine_variance (nebula.rs:1972-1994): BOBYQA for both components, the
// simplex search for one (the cell-level size held at obj.fixed_cell).
bool refine_variance(Objective obj, unsigned ord, std::size_t n, const double start[2],
                     const double lo[2], const double hi[2], double x[2],
                     bool& failed, StageTwoInfo* info) {
  auto f = [&](const std::vector<double>& xv, unsigned order) {
    obj.params.ord = order;
    return obj.value(xv.data());
  };
  double fbest = 0.0;
  SearchOutcome r;
  if (n == 2) {
    r = bobyqa_search(f, ord, start, lo, hi, &fbest);
  } else {
    r = simplex_search(f, std::vector<double>{start[0]}, std::vector<double>{lo[0]},
                       std::vector<double>{hi[0]}, ord);
    // the polish moves the incumbent without evaluating it
    fbest = std::numeric_limits<double>::quiet_NaN();
  }
  if (r.found) {
    for (std::size_t j = 0; j < n; ++j) x[j] = r.x[j];
    failed = r.failed;
  }
  if (info) {
    info->evaluations = r.evaluations;
    if (r.found) {
      info->objective = fbest;
      info->ord = (int)r.order;
    }
  }
  return r.found;
}

struct Outcome {
  std::vector<double> beta, cov, se;
  double subject, cell;
  int convergence;
};

Outcome finish_gene(const Shared& sh, const PmlData& pd, const Plan& plan, int refit_state,
                    const double v[2], bool failed) {
  const std::size_t nb = sh.n_coef;
  const Params& p = sh.params;
  double sigma = plan.sigma, gamma = plan.gamma;
  int convergence = plan.convergence;
  if (plan.refit != 0 && refit_state != 0) {
    if (refit_state == 1) {
      sigma = v[0];
      if (plan.refit == 1) gamma = v[1];
      convergence = failed ? kOuterFailed : 1;
    } else {
      convergence = kOuterFailed;
    }
  }
  PmlParams fp;
  fp.reml = p.reml;
  fp.eps = p.eps;
  fp.ord = 1;
  std::vector<double> beta = plan.beta_start;
  beta[sh.intercept] -= sigma / 2.0;
  PmlResult fit = opt_pml(pd, beta.data(), sigma, gamma, fp);
  bool at_bound = sigma == p.max_subject || gamma == p.min_cell;
  int code = check_convergence(fit, fp, at_bound, convergence);
  std::size_t packed = nb * (nb + 1) / 2;
  Outcome o;
  std::vector<double> inv;
  if (code != -25 && cholesky_inverse(fit.information, nb, inv)) {
    o.cov.assign(packed, 0.0);
    for (std::size_t j = 0; j < nb; ++j)
      for (std::size_t i = 0; i <= j; ++i) o.cov[j * (j + 1) / 2 + i] = inv[i * nb + j];
    o.se.resize(nb);
    for (std::size_t j = 0; j < nb; ++j) o.se[j] = std::sqrt(inv[j * nb + j]);
  } else {
    code = -25;
    o.cov.assign(packed, std::numeric_limits<double>::quiet_NaN());
    o.se.assign(nb, std::numeric_limits<double>::quiet_NaN());
  }
  o.beta = fit.beta;
  o.subject = sigma;
  o.cell = 1.0 / gamma;
  o.convergence = code;
  return o;
}


}  // namespace

SearchOutcome stage_two_search(const std::vector<double>& start, const std::vector<double>& lo,
                               const std::vector<double>& hi, unsigned ord,
                               const std::function<double(const std::vector<double>&, unsigned)>& f) {
  if (start.size() == 2) {
    const double s[2] = {start[0], start[1]}, l[2] = {lo[0], lo[1]}, h[2] = {hi[0], hi[1]};
    return bobyqa_search(f, ord, s, l, h, nullptr);
  }
  return simplex_search(f, start, lo, hi, ord);
}

namespace {
// When set, nebula_sparse evaluates the profile objective at these points
// instead of fitting: (kept gene index, subject, cell) -> value.
struct ProfileProbe {
  std::size_t n;
  const std::uint64_t* gene;
  const double* point;
  double* out;
};
thread_local ProfileProbe* g_probe = nullptr;
}  // namespace

void profile_objective(std::size_t n_genes, std::size_t n_cells, const std::uint32_t* indptr,
                       const std::uint32_t* indices, const double* data,
                       const std::uint64_t* subject_id, const double* design,
                       std::size_t n_coef, const double* offset, const Params& params,
                       std::size_t n, const std::uint64_t* gene, const double* point,
                       double* out, int threads) {
  ProfileProbe probe{n, gene, point, out};
  g_probe = &probe;
  try {
    nebula_sparse(n_genes, n_cells, indptr, indices, data, subject_id, design, n_coef, offset,
                  params, threads, nullptr, nullptr);
  } catch (...) {
    g_probe = nullptr;
    throw;
  }
  g_probe = nullptr;
}

Fit nebula_sparse(std::size_t n_genes, std::size_t n_cells, const std::uint32_t* indptr,
                  const std::uint32_t* indices, const double* data,
                  const std::uint64_t* subject_id, const double* design,
                  std::size_t n_coef, const double* offset, const Params& params,
                  int threads, std::vector<StageTwoInfo>* stage_two,
                  double* t_stage_two_out) {
  // NebulaParams::validate (nebula.rs:281-307)
  const double los[2] = {params.min_subject, params.min_cell};
  const double his[2] = {params.max_subject, params.max_cell};
  for (int i = 0; i < 2; ++i) {
    if (!(std::isfinite(los[i]) && los[i] > 0.0)) fail(3, "min." + std::to_string(i));
    if (!(std::isfinite(his[i]) && his[i] > los[i]))
      fail(11, "invalid bounds for index " + std::to_string(i));
  }
  if (!(std::isfinite(params.eps) && params.eps > 0.0)) fail(3, "eps");
  if (params.mincp == 0) fail(3, "mincp");
  if (n_genes == 0 || n_cells == 0) fail(1, "empty counts");
  if (n_cells < 2) fail(2, "NEBULA needs more than one cell in the count matrix.");
  if (n_coef == 0) fail(3, "n_coef");
  if (indptr[0] != 0) fail(6, "indptr must start at zero");
  for (std::size_t g = 0; g < n_genes; ++g) {
    if (indptr[g + 1] < indptr[g]) fail(6, "indptr must be non-decreasing");
    for (std::uint32_t i = indptr[g]; i < indptr[g + 1]; ++i) {
      if (indices[i] >= n_cells) fail(6, "cell index out of range");
      if (i > indptr[g] && indices[i] <= indices[i - 1])
        fail(6, "cell indices must increase within a gene");
    }
  }
  const std::size_t nnz = indptr[n_genes];
  for (std::size_t i = 0; i < nnz; ++i)
    if (!(std::isfinite(data[i]) && data[i] >= 0.0))
      fail(2, "The counts hold a value that is not a non-negative finite number.");

  // subject_boundaries (nebula.rs:2138-2154)
  std::vector<std::size_t> fid{0};
  {
    std::vector<std::uint64_t> seen{subject_id[0]};
    for (std::size_t i = 1; i < n_cells; ++i)
      if (subject_id[i] != subject_id[i - 1]) {
        if (std::find(seen.begin(), seen.end(), subject_id[i]) != seen.end())
          fail(7, "subject " + std::to_string(subject_id[i]) + " is not contiguous");
        seen.push_back(subject_id[i]);
        fid.push_back(i);
      }
    fid.push_back(n_cells);
  }
  const std::size_t k = fid.size() - 1;
  if (k < 2) fail(8, "need at least 2 subjects");

  // offset_summary (ptmg.rs:1827-1865)
  std::vector<double> log_offset(n_cells);
  double mean_offset = 1.0, cv = 0.0;
  if (offset) {
    for (std::size_t i = 0; i < n_cells; ++i)
      if (!(std::isfinite(offset[i]) && offset[i] > 0.0)) fail(3, "offset");
    double s = 0.0;
    for (std::size_t i = 0; i < n_cells; ++i) s += offset[i];
    mean_offset = s / (double)n_cells;
    double ss = 0.0;
    for (std::size_t i = 0; i < n_cells; ++i) ss += (offset[i] - mean_offset) * (offset[i] - mean_offset);
    cv = std::sqrt(ss / (double)n_cells) / mean_offset;
    for (std::size_t i = 0; i < n_cells; ++i) log_offset[i] = std::log(offset[i]);
  }
  const double log_mean_offset = std::log(mean_offset);

  // centre_design (ptmg.rs:1748-1790)
  std::vector<double> centred(n_cells * n_coef, 0.0), sds(n_coef, 0.0);
  for (std::size_t j = 0; j < n_coef; ++j) {
    double mean = 0.0;
    for (std::size_t i = 0; i < n_cells; ++i) mean += design[i * n_coef + j];
    mean /= (double)n_cells;
    double var = 0.0;
    for (std::size_t i = 0; i < n_cells; ++i) {
      double dd = design[i * n_coef + j] - mean;
      centred[i * n_coef + j] = dd;
      var += dd * dd;
    }
    double sd = std::sqrt(var / (double)n_cells);
    if (sd > 0.0) {
      for (std::size_t i = 0; i < n_cells; ++i) centred[i * n_coef + j] /= sd;
      sds[j] = sd;
    } else if (design[j] != 0.0) {
      for (std::size_t i = 0; i < n_cells; ++i) centred[i * n_coef + j] = 1.0;
      sds[j] = sd;
    } else {
      sds[j] = -1.0;
    }
  }
  for (double s : sds)
    if (s < 0.0) fail(2, "Some predictors are a zero vector.");
  std::vector<std::size_t> constant;
  for (std::size_t j = 0; j < n_coef; ++j)
    if (sds[j] == 0.0) constant.push_back(j);
  if (constant.empty()) fail(9, "missing intercept");
  if (constant.size() > 1) fail(2, "More than one predictor has zero variation.");

  Shared sh;
  sh.design = centred.data();
  sh.log_offset = log_offset.data();
  sh.fid = fid.data();
  sh.n_cells = n_cells;
  sh.n_coef = n_coef;
  sh.k = k;
  sh.intercept = constant[0];
  sh.log_mean_offset = log_mean_offset;
  sh.cells_per_subject = (double)n_cells / (double)k;
  sh.cv2 = cv * cv;
  for (std::size_t j = 0; j < n_coef; ++j) {
    bool varies = false;
    for (std::size_t s = 0; s < k && !varies; ++s) {
      double first = centred[fid[s] * n_coef + j];
      for (std::size_t i = fid[s]; i < fid[s + 1]; ++i)
        if (centred[i * n_coef + j] != first) {
          varies = true;
          break;
        }
    }
    if (varies) sh.cell_columns.push_back(j);
  }
  sh.method = sh.cells_per_subject < kMinCellsPerSubjectLn ? 1 : params.method;
  sh.params = params;
  // zero-count tables from the centred design and the log offsets
  // (nebula.rs:641); params.tables = false sweeps every cell instead, the
  // check that both give the same fits
  ZeroCells zc;
  sh.tables = build_zero_cells(centred.data(), log_offset.data(), n_cells, fid.data(), k,
                               n_coef, zc);
  sh.zeros = sh.tables && params.tables ? &zc : nullptr;

  // cumsum_y and the expression filter (nebula.rs:621-636)
  std::vector<std::size_t> subject_of(n_cells);
  for (std::size_t s = 0; s < k; ++s)
    for (std::size_t i = fid[s]; i < fid[s + 1]; ++i) subject_of[i] = s;
  std::vector<double> totals(n_genes * k, 0.0);
  for (std::size_t g = 0; g < n_genes; ++g)
    for (std::uint32_t i = indptr[g]; i < indptr[g + 1]; ++i)
      if (data[i] > 0.0) totals[g * k + subject_of[indices[i]]] += data[i];
  std::vector<std::size_t> kept;
  for (std::size_t g = 0; g < n_genes; ++g) {
    double total = 0.0;
    for (std::size_t s = 0; s < k; ++s) total += totals[g * k + s];
    std::size_t stored = indptr[g + 1] - indptr[g];
    bool pooled = total / (double)n_cells > params.cpc && stored >= params.mincp;
    if (pooled && params.min_subjects > 0) {
      std::size_t expressing = 0;
      for (std::size_t s = 0; s < k; ++s)
        if (totals[g * k + s] / (double)(fid[s + 1] - fid[s]) > params.cpc) ++expressing;
      pooled = expressing >= params.min_subjects;
    }
    if (pooled) kept.push_back(g);
  }
  if (kept.empty()) fail(10, "no genes after filtering");

  if (g_probe) {
    ProfileProbe* pr = g_probe;
    for (std::size_t q = 0; q < pr->n; ++q)
      if (pr->gene[q] >= kept.size()) fail(2, "probe gene out of range");
    std::atomic<std::size_t> next{0};
    auto probe_worker = [&]() {
      for (;;) {
        const std::size_t q = next.fetch_add(1);
        if (q >= pr->n) return;
        const std::size_t g = kept[pr->gene[q]];
        GeneCounts gc;
        double sum = 0.0;
        for (std::uint32_t i = indptr[g]; i < indptr[g + 1]; ++i)
          if (data[i] > 0.0) {
            gc.cells.push_back(indices[i]);
            gc.counts.push_back(data[i]);
            sum += data[i];
            if (data[i] == 1.0) ++gc.n_one;
            else if (data[i] == 2.0) ++gc.n_two;
          }
        gc.mean_count = sum / (double)n_cells;
        const double* tot = &totals[g * k];
        PmlData pd{sh.design, sh.log_offset, n_cells, n_coef, gc.counts.data(), gc.cells.data(),
                   gc.counts.size(), fid.data(), k, tot, sh.zeros};
        std::vector<double> bs(n_coef, 0.0);
        bs[sh.intercept] = std::log(gc.mean_count) - sh.log_mean_offset;
        Objective obj{&pd, &bs, sh.intercept, PmlParams{}, (double)n_cells, (double)k,
                      &gc.counts, (double)gc.counts.size(), (double)gc.n_one, (double)gc.n_two};
        obj.params.eps = params.eps;
        obj.params.ord = gc.mean_count * sh.cells_per_subject < kHighOrderCutoff ? 3 : 1;
        pr->out[q] = obj.value(pr->point + 2 * q);
      }
    };
    int pt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
    pt = (int)std::min<std::size_t>((std::size_t)pt, std::max<std::size_t>(pr->n, 1));
    std::vector<std::thread> pool;
    for (int t = 1; t < pt; ++t) pool.emplace_back(probe_worker);
    probe_worker();
    for (auto& t : pool) t.join();
    return Fit{};
  }

  // Three phases over the kept genes, each spread over the threads: stage
  // one, stage two (timed on the wall clock, every thread busy with it), stage
  // three.  Per gene the arithmetic is nebula.rs:1012-1033's.
  const std::size_t n_kept = kept.size();
  std::vector<Outcome> outcomes(n_kept);
  std::vector<GeneCounts> gcs(n_kept);
  std::vector<Plan> plans(n_kept);
  std::vector<int> refit_state(n_kept, 0);
  std::vector<std::array<double, 2>> found_v(n_kept, {0.0, 0.0});
  std::vector<char> found_failed(n_kept, 0);
  if (stage_two) stage_two->assign(n_kept, StageTwoInfo{});
  std::atomic<bool> errored{false};
  Error first_error{0, ""};
  std::mutex mu;
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  nt = (int)std::min<std::size_t>((std::size_t)nt, n_kept);
  auto run_phase = [&](auto&& body) {
    std::atomic<std::size_t> next{0};
    auto worker = [&]() {
      for (;;) {
        std::size_t idx = next.fetch_add(1);
        if (idx >= n_kept || errored.load()) return;
        try {
          body(idx);
        } catch (const Error& e) {
          std::lock_guard<std::mutex> lock(mu);
          if (!errored.exchange(true)) first_error = e;
          return;
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nt; ++t) pool.emplace_back(worker);
    worker();
    for (auto& t : pool) t.join();
  };
  auto pml_of = [&](std::size_t idx) {
    const GeneCounts& gc = gcs[idx];
    return PmlData{sh.design, sh.log_offset, n_cells, n_coef, gc.counts.data(), gc.cells.data(),
                   gc.counts.size(), fid.data(), k, &totals[kept[idx] * k], sh.zeros};
  };
  // stage one
  run_phase([&](std::size_t idx) {
    const std::size_t g = kept[idx];
    GeneCounts& gc = gcs[idx];
    double sum = 0.0;
    for (std::uint32_t i = indptr[g]; i < indptr[g + 1]; ++i)
      if (data[i] > 0.0) {
        gc.cells.push_back(indices[i]);
        gc.counts.push_back(data[i]);
        sum += data[i];
        if (data[i] == 1.0) ++gc.n_one;
        else if (data[i] == 2.0) ++gc.n_two;
      }
    gc.mean_count = sum / (double)n_cells;
    plans[idx] = plan_gene(sh, gc, &totals[g * k]);
  });
  if (errored) throw first_error;
  // stage two
  auto t0 = std::chrono::steady_clock::now();
  run_phase([&](std::size_t idx) {
    const Plan& plan = plans[idx];
    const GeneCounts& gc = gcs[idx];
    StageTwoInfo info;
    info.start_sigma = plan.sigma;
    info.start_gamma = plan.gamma;
    info.beta_start = plan.beta_start;
    if (plan.refit != 0) {
      PmlData pd = pml_of(idx);
      Objective obj{&pd, &plan.beta_start, sh.intercept, PmlParams{}, (double)n_cells, (double)k,
                    &gc.counts, (double)gc.counts.size(), (double)gc.n_one, (double)gc.n_two};
      obj.params.reml = params.reml;
      obj.params.eps = params.eps;
      double start[2] = {plan.sigma, plan.gamma};
      double lo[2] = {params.min_subject, params.min_cell};
      double hi[2] = {params.max_subject, params.max_cell};
      const std::size_t n = plan.refit == 1 ? 2 : 1;
      if (n == 1) obj.fixed_cell = plan.gamma;
      bool failed = false;
      bool found = refine_variance(obj, plan.ord, n, start, lo, hi, found_v[idx].data(), failed,
                                   &info);
      refit_state[idx] = found ? 1 : 2;
      found_failed[idx] = failed ? 1 : 0;
      info.searched = (int)n;
      info.sigma = found_v[idx][0];
      info.gamma = n == 2 ? found_v[idx][1] : plan.gamma;
    }
    if (stage_two) (*stage_two)[idx] = info;
  });
  if (errored) throw first_error;
  const double stage_two_s =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  // stage three
  run_phase([&](std::size_t idx) {
    PmlData pd = pml_of(idx);
    outcomes[idx] = finish_gene(sh, pd, plans[idx], refit_state[idx], found_v[idx].data(),
                                found_failed[idx] != 0);
  });
  if (errored) throw first_error;
  if (t_stage_two_out) *t_stage_two_out = stage_two_s;

  // assemble (nebula.rs:2233-2277)
  Fit fit;
  fit.n_coef = n_coef;
  fit.gene_index = kept;
  std::vector<double> scale = sds;
  scale[sh.intercept] = 1.0;
  for (const Outcome& o : outcomes) {
    for (std::size_t j = 0; j < n_coef; ++j) {
      fit.coefficients.push_back(o.beta[j] / scale[j]);
      fit.se.push_back(o.se[j] / scale[j]);
    }
    for (std::size_t j = 0; j < n_coef; ++j)
      for (std::size_t i = 0; i <= j; ++i)
        fit.covariance.push_back(o.cov[j * (j + 1) / 2 + i] / (scale[i] * scale[j]));
    fit.subject_overdispersion.push_back(o.subject);
    fit.cell_overdispersion.push_back(o.cell);
    fit.convergence.push_back(o.convergence);
    fit.sigma_at_bound.push_back(o.subject <= params.min_subject * kBoundSlack ? 1 : 0);
  }
  return fit;
}

}  // namespace oracle

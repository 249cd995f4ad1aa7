// Test infrastructure (the oracle): NEBULA-LN's one-component stage-two
// search, f64, closure-driven.  A restatement of
//   NelderMeadStepper        <- src/numeric/optimise.rs:383-670 (scipy's
//                               simplex: 5 % initial steps, 0.00025 for a zero
//                               coordinate, reflect 1 / expand 2 / contract
//                               1/2 / shrink 1/2, stable sort by total order,
//                               collapse at 16 ulps of the best vertex)
//   StageTwoSearch, one component <- src/sc/nebula.rs:1568-1951 (clamped
//                               start; bounded simplex, every point clamped;
//                               one restart off a bound the start was inside,
//                               bound_restart :1190; the clamped minimiser
//                               once more; three quadratic polish stencils at
//                               widths 1e-2, 3e-3, 1e-3 of |x| with floors
//                               1e-5 / 1e-4, least squares through the normal
//                               equations, a Newton step of at most 2 widths;
//                               an order-one retry when a raised-order pass saw
//                               a non-finite value)
//   cholesky_lower           <- src/utils/linalg.rs:36-56, cholesky_factor /
//                               cholesky_solve <- src/sc/nebula.rs:2068-2095
// The evaluation order is the reference's: its ask/tell machine hands out
// one simplex point at a time and a polish stencil's points in index order,
// so evaluating each point as it is asked visits the same points in the same
// order.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <vector>

namespace oracle {

// nebula.rs:100-118,173-179 and optimise.rs:40-56
constexpr double kPolishWidths[3] = {1e-2, 3e-3, 1e-3};
constexpr double kPolishFloor[2] = {1e-5, 1e-4};
constexpr double kPolishMaxStep = 2.0;
constexpr double kVarianceXatol = 1e-7;
constexpr double kVarianceFatol = 1e-7;
constexpr int kVarianceMaxIter = 500;
constexpr double kNmStep = 0.05;
constexpr double kNmZeroStep = 0.00025;
constexpr double kNmCollapsedUlps = 16.0;

// f64::total_cmp as a strict weak order: -NaN < -inf < ... < +inf < +NaN.
inline bool total_less(double a, double b) {
  std::int64_t x, y;
  std::memcpy(&x, &a, 8);
  std::memcpy(&y, &b, 8);
  x ^= (std::int64_t)((std::uint64_t)(x >> 63) >> 1);
  y ^= (std::int64_t)((std::uint64_t)(y >> 63) >> 1);
  return x < y;
}

struct NmResult {
  std::vector<double> x;
  double f;
  int iterations;
  bool converged;
};

// The Nelder-Mead stepper driven by f(x) (optimise.rs:441-670).
template <class F>
NmResult nelder_mead(F&& f, const std::vector<double>& x0, double xatol, double fatol,
                     int max_iter) {
  const std::size_t n = x0.size();
  std::vector<std::vector<double>> simplex(n + 1, x0);
  for (std::size_t i = 0; i < n; ++i) {
    if (simplex[i + 1][i] != 0.0)
      simplex[i + 1][i] *= 1.0 + kNmStep;
    else
      simplex[i + 1][i] = kNmZeroStep;
  }
  std::vector<double> values(n + 1);
  for (std::size_t i = 0; i <= n; ++i) values[i] = f(simplex[i]);
  std::vector<std::size_t> order(n + 1);
  std::iota(order.begin(), order.end(), 0);
  auto sort_order = [&]() {
    std::stable_sort(order.begin(), order.end(),
                     [&](std::size_t a, std::size_t b) { return total_less(values[a], values[b]); });
  };
  int iterations = 0;
  bool converged = false;
  std::vector<double> centroid(n), reflected(n), trial(n);
  for (int next_iter = 0;; ++next_iter) {
    if (next_iter >= max_iter) break;
    iterations = next_iter + 1;
    sort_order();
    const std::size_t best = order[0], worst = order[n], second_worst = order[n - 1];
    double spread_x = 0.0, spread_f = 0.0;
    for (std::size_t q = 1; q <= n; ++q) {
      const std::size_t i = order[q];
      double m = 0.0;
      for (std::size_t j = 0; j < n; ++j)
        m = std::max(m, std::fabs(simplex[i][j] - simplex[best][j]));
      spread_x = std::max(spread_x, m);
      spread_f = std::max(spread_f, std::fabs(values[i] - values[best]));
    }
    double scale = std::numeric_limits<double>::min();
    for (std::size_t j = 0; j < n; ++j) scale = std::max(scale, std::fabs(simplex[best][j]));
    const bool collapsed =
        spread_x <= kNmCollapsedUlps * std::numeric_limits<double>::epsilon() * scale;
    if (collapsed || (spread_x <= xatol && spread_f <= fatol)) {
      converged = true;
      break;
    }
    std::fill(centroid.begin(), centroid.end(), 0.0);
    for (std::size_t q = 0; q < n; ++q)
      for (std::size_t j = 0; j < n; ++j) centroid[j] += simplex[order[q]][j];
    for (std::size_t j = 0; j < n; ++j) centroid[j] /= (double)n;
    for (std::size_t j = 0; j < n; ++j)
      reflected[j] = centroid[j] + (centroid[j] - simplex[worst][j]);
    const double fr = f(reflected);
    if (fr < values[best]) {
      for (std::size_t j = 0; j < n; ++j)
        trial[j] = centroid[j] + 2.0 * (centroid[j] - simplex[worst][j]);
      const double fe = f(trial);
      if (fe < fr) {
        simplex[worst] = trial;
        values[worst] = fe;
      } else {
        simplex[worst] = reflected;
        values[worst] = fr;
      }
    } else if (fr < values[second_worst]) {
      simplex[worst] = reflected;
      values[worst] = fr;
    } else {
      const bool inside = fr >= values[worst];
      for (std::size_t j = 0; j < n; ++j)
        trial[j] = inside ? centroid[j] - 0.5 * (centroid[j] - simplex[worst][j])
                          : centroid[j] + 0.5 * (centroid[j] - simplex[worst][j]);
      const double fc = f(trial);
      const bool accept = inside ? fc < values[worst] : fc <= fr;
      if (accept) {
        simplex[worst] = trial;
        values[worst] = fc;
      } else {
        for (std::size_t q = 1; q <= n; ++q) {
          const std::size_t i = order[q];
          for (std::size_t j = 0; j < n; ++j) {
            const double anchor = simplex[best][j];
            simplex[i][j] = anchor + 0.5 * (simplex[i][j] - anchor);
          }
          values[i] = f(simplex[i]);
        }
      }
    }
  }
  sort_order();
  return NmResult{simplex[order[0]], values[order[0]], iterations, converged};
}

inline bool chol_lower(std::vector<double>& a, std::size_t n) {
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

// cholesky_solve with cholesky_factor's extra check of finite pivots.
inline bool chol_solve(std::vector<double> a, std::size_t n, std::vector<double>& x) {
  if (!chol_lower(a, n)) return false;
  for (std::size_t i = 0; i < n; ++i)
    if (!std::isfinite(a[i * n + i])) return false;
  for (std::size_t i = 0; i < n; ++i) {
    double s = x[i];
    for (std::size_t k = 0; k < i; ++k) s -= a[i * n + k] * x[k];
    x[i] = s / a[i * n + i];
  }
  for (std::size_t i = n; i-- > 0;) {
    double s = x[i];
    for (std::size_t k = i + 1; k < n; ++k) s -= a[k * n + i] * x[k];
    x[i] = s / a[i * n + i];
  }
  for (double v : x)
    if (!std::isfinite(v)) return false;
  return true;
}

inline bool restart_point(const std::vector<double>& best, const std::vector<double>& start,
                          const std::vector<double>& lo, const std::vector<double>& hi,
                          std::vector<double>& x0) {
  x0 = best;
  bool moved = false;
  for (std::size_t j = 0; j < best.size(); ++j)
    for (double bound : {lo[j], hi[j]})
      if (best[j] == bound && start[j] != bound && bound * start[j] > 0.0) {
        x0[j] = std::sqrt(bound * start[j]);
        moved = true;
      }
  return moved;
}

struct SearchOutcome {
  bool found = false;
  bool failed = false;
  std::vector<double> x;
  unsigned order = 1;
  int evaluations = 0;
};

// One quadratic polish step (nebula.rs:1890-1950): returns false when the
// model is unusable; otherwise writes the new incumbent.
inline bool fit_polish(const std::vector<double>& x, const std::vector<double>& step,
                       const std::vector<std::vector<double>>& offsets,
                       const std::vector<double>& values, const std::vector<double>& lo,
                       const std::vector<double>& hi, std::vector<double>& out) {
  const std::size_t n = x.size();
  const std::size_t nt = 1 + 2 * n + n * (n - 1) / 2;
  std::vector<double> normal(nt * nt, 0.0), rhs(nt, 0.0), row(nt);
  for (std::size_t pt = 0; pt < values.size(); ++pt) {
    const std::vector<double>& u = offsets[pt];
    std::size_t c = 0;
    row[c++] = 1.0;
    for (std::size_t j = 0; j < n; ++j) row[c++] = u[j];
    for (std::size_t j = 0; j < n; ++j) row[c++] = 0.5 * u[j] * u[j];
    for (std::size_t i = 0; i < n; ++i)
      for (std::size_t j = i + 1; j < n; ++j) row[c++] = u[i] * u[j];
    for (std::size_t a = 0; a < nt; ++a) {
      rhs[a] += row[a] * values[pt];
      for (std::size_t b = 0; b < nt; ++b) normal[a * nt + b] += row[a] * row[b];
    }
  }
  std::vector<double> coef = rhs;
  if (!chol_solve(normal, nt, coef)) return false;
  std::vector<double> hess(n * n, 0.0), delta(n);
  for (std::size_t j = 0; j < n; ++j) hess[j * n + j] = coef[1 + n + j];
  std::size_t cross = 1 + 2 * n;
  for (std::size_t i = 0; i < n; ++i)
    for (std::size_t j = i + 1; j < n; ++j) {
      hess[i * n + j] = coef[cross];
      hess[j * n + i] = coef[cross];
      ++cross;
    }
  for (std::size_t j = 0; j < n; ++j) delta[j] = -coef[1 + j];
  if (!chol_solve(hess, n, delta)) return false;
  for (double d : delta)
    if (!std::isfinite(d) || std::fabs(d) > kPolishMaxStep) return false;
  out.resize(n);
  for (std::size_t j = 0; j < n; ++j)
    out[j] = std::min(std::max(x[j] + delta[j] * step[j], lo[j]), hi[j]);
  return true;
}

// The one-component (or general simplex) stage-two search.  f(x, order)
// returns the objective (+inf for a rejected fit).
template <class F>
SearchOutcome simplex_search(F&& f, const std::vector<double>& start,
                             const std::vector<double>& lo, const std::vector<double>& hi,
                             unsigned ord) {
  const std::size_t n = start.size();
  auto clamp = [&](const std::vector<double>& x) {
    std::vector<double> c(n);
    for (std::size_t j = 0; j < n; ++j) c[j] = std::min(std::max(x[j], lo[j]), hi[j]);
    return c;
  };
  SearchOutcome out;
  unsigned order = ord;
  for (;;) {
    bool invalid = false;
    auto eval = [&](const std::vector<double>& x) {
      const double v = f(x, order);
      ++out.evaluations;
      if (!std::isfinite(v)) invalid = true;
      return v;
    };
    bool found = false;
    std::vector<double> best;
    const std::vector<double> x0 = clamp(start);
    if (std::isfinite(eval(x0))) {
      auto clamped_eval = [&](const std::vector<double>& x) { return eval(clamp(x)); };
      NmResult r = nelder_mead(clamped_eval, x0, kVarianceXatol, kVarianceFatol, kVarianceMaxIter);
      best = clamp(r.x);
      std::vector<double> x1;
      if (restart_point(best, x0, lo, hi, x1)) {
        const std::vector<double> first = best;
        const double f_first = r.f;
        r = nelder_mead(clamped_eval, x1, kVarianceXatol, kVarianceFatol, kVarianceMaxIter);
        best = clamp(r.x);
        if (f_first <= r.f) best = first;
      }
      if (std::isfinite(eval(best))) {
        found = true;
        for (double width : kPolishWidths) {
          std::vector<double> step(n);
          std::vector<std::array<double, 3>> nodes(n);
          for (std::size_t j = 0; j < n; ++j) {
            step[j] = std::max(width * std::fabs(best[j]), kPolishFloor[j]);
            if (best[j] - step[j] < lo[j])
              nodes[j] = {0.0, 1.0, 2.0};
            else if (best[j] + step[j] > hi[j])
              nodes[j] = {-2.0, -1.0, 0.0};
            else
              nodes[j] = {-1.0, 0.0, 1.0};
          }
          std::size_t n_points = 1;
          for (std::size_t j = 0; j < n; ++j) n_points *= 3;
          std::vector<std::vector<double>> offsets(n_points, std::vector<double>(n));
          std::vector<double> values(n_points);
          bool all_finite = true;
          for (std::size_t pt = 0; pt < n_points; ++pt) {
            std::size_t rest = pt;
            std::vector<double> ask(n);
            for (std::size_t j = 0; j < n; ++j) {
              const double node = nodes[j][rest % 3];
              rest /= 3;
              const double p = std::min(std::max(best[j] + node * step[j], lo[j]), hi[j]);
              offsets[pt][j] = (p - best[j]) / step[j];
              ask[j] = best[j] + offsets[pt][j] * step[j];
            }
            values[pt] = eval(ask);
            if (!std::isfinite(values[pt])) all_finite = false;
          }
          std::vector<double> next;
          if (all_finite && fit_polish(best, step, offsets, values, lo, hi, next)) best = next;
        }
      }
    }
    if ((found && !invalid) || order == 1) {
      out.found = found;
      out.x = best;
      out.order = order;
      return out;
    }
    order = 1;
  }
}

}  // namespace oracle

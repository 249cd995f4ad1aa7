// Test infrastructure (the oracle): box-constrained limited-memory BFGS, the
// stage-one quasi-Newton.  A restatement of edge-rs's L-BFGS-B
// (src/numeric/lbfgsb.rs), itself Byrd, Lu, Nocedal and Zhu (SIAM J. Sci.
// Comput. 16(5), 1995) with a strong-Wolfe line search (Nocedal and Wright,
// Algorithms 3.5 and 3.6) in place of More and Thuente's dcsrch:
//
//   Memory (s, y pairs, theta)    <- lbfgsb.rs:130-315 (admit when
//                                   s'y > 2.2e-16 y'y, theta = y'y / s'y; the
//                                   middle matrix [[-D, L'], [L, theta S'S]]
//                                   solved by partial-pivot elimination)
//   solve_dense                   <- :317-361
//   cauchy_point                  <- :400-495 (generalised Cauchy point,
//                                   Algorithm CP)
//   subspace_minimise             <- :528-639 (direct primal method, step
//                                   truncated to the box)
//   max_feasible_step, line_search, zoom <- :641-874 (steps capped at 1e3,
//                                   expansion by 2, bisection zoom)
//   minimise                      <- :876-1020 (pgtol on the projected
//                                   gradient, ftol on the relative decrease,
//                                   memory reset on a stalled line search)
//
// Every sum is a left fold in the reference's order, so for the same
// objective the iterates are the reference's to rounding.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

#include "oracle.hpp"

namespace oracle {

namespace lbfgsb_detail {

constexpr double kWolfeC1 = 1e-4;
constexpr double kWolfeC2 = 0.9;
constexpr double kCurvatureEps = 2.2e-16;
constexpr double kMaxStep = 1e3;
constexpr double kEps = std::numeric_limits<double>::epsilon();
constexpr double kInf = std::numeric_limits<double>::infinity();

using Vec = std::vector<double>;

inline double dot(const Vec& a, const Vec& b) {
  double s = 0.0;
  for (std::size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

inline double clampd(double v, double lo, double hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Gaussian elimination with partial pivoting; false when a pivot is below
// 1e-300 (lbfgsb.rs:317-361).
inline bool solve_dense(Vec m, Vec& x, std::size_t dim) {
  for (std::size_t col = 0; col < dim; ++col) {
    std::size_t pivot = col;
    double best = std::fabs(m[col * dim + col]);
    for (std::size_t row = col + 1; row < dim; ++row) {
      const double v = std::fabs(m[row * dim + col]);
      if (v > best) {
        best = v;
        pivot = row;
      }
    }
    if (best < 1e-300) return false;
    if (pivot != col) {
      for (std::size_t j = 0; j < dim; ++j) std::swap(m[col * dim + j], m[pivot * dim + j]);
      std::swap(x[col], x[pivot]);
    }
    const double diag = m[col * dim + col];
    for (std::size_t row = col + 1; row < dim; ++row) {
      const double factor = m[row * dim + col] / diag;
      if (factor == 0.0) continue;
      for (std::size_t j = col; j < dim; ++j) m[row * dim + j] -= factor * m[col * dim + j];
      x[row] -= factor * x[col];
    }
  }
  for (std::size_t col = dim; col-- > 0;) {
    double acc = x[col];
    for (std::size_t j = col + 1; j < dim; ++j) acc -= m[col * dim + j] * x[j];
    x[col] = acc / m[col * dim + col];
  }
  return true;
}

struct Memory {
  std::vector<Vec> s, y;  // oldest first
  double theta = 1.0;
  std::size_t limit;
  explicit Memory(std::size_t l) : limit(l) {}
  std::size_t k() const { return s.size(); }
  void reset() {
    s.clear();
    y.clear();
    theta = 1.0;
  }
  bool push(Vec sv, Vec yv) {
    const double sy = dot(sv, yv);
    const double yy = dot(yv, yv);
    if (sy <= kCurvatureEps * yy || yy <= 0.0) return false;
    if (s.size() == limit) {
      s.erase(s.begin());
      y.erase(y.begin());
    }
    theta = yy / sy;
    s.push_back(std::move(sv));
    y.push_back(std::move(yv));
    return true;
  }
  // W = [Y, theta S]: entry (i, j)
  double w(std::size_t i, std::size_t j) const {
    const std::size_t kk = k();
    return j < kk ? y[j][i] : theta * s[j - kk][i];
  }
  Vec wt_dot(const Vec& v) const {
    const std::size_t kk = k();
    Vec out(2 * kk, 0.0);
    for (std::size_t j = 0; j < 2 * kk; ++j) {
      double acc = 0.0;
      for (std::size_t i = 0; i < v.size(); ++i) acc += w(i, j) * v[i];
      out[j] = acc;
    }
    return out;
  }
  Vec middle() const {
    const std::size_t kk = k(), dim = 2 * kk;
    Vec m(dim * dim, 0.0);
    for (std::size_t i = 0; i < kk; ++i) m[i * dim + i] = -dot(s[i], y[i]);
    for (std::size_t i = 0; i < kk; ++i)
      for (std::size_t j = 0; j < kk; ++j)
        if (i > j) {
          const double sy = dot(s[i], y[j]);
          m[(kk + i) * dim + j] = sy;
          m[j * dim + (kk + i)] = sy;
        }
    for (std::size_t i = 0; i < kk; ++i)
      for (std::size_t j = 0; j < kk; ++j) m[(kk + i) * dim + (kk + j)] = theta * dot(s[i], s[j]);
    return m;
  }
  bool apply_m(const Vec& v, Vec& out) const {
    const std::size_t dim = 2 * k();
    out = v;
    if (dim == 0) return true;
    return solve_dense(middle(), out, dim);
  }
};

struct CauchyPoint {
  Vec x;
  std::vector<bool> free;
  Vec c;
};

inline bool cauchy_point(const Vec& x, const Vec& g, const Vec& lower, const Vec& upper,
                         const Memory& mem, CauchyPoint& cp) {
  const std::size_t n = x.size(), k = mem.k();
  Vec t(n, kInf), d(n, 0.0);
  for (std::size_t i = 0; i < n; ++i) {
    if (g[i] < 0.0 && std::isfinite(upper[i]))
      t[i] = (x[i] - upper[i]) / g[i];
    else if (g[i] > 0.0 && std::isfinite(lower[i]))
      t[i] = (x[i] - lower[i]) / g[i];
    d[i] = t[i] == 0.0 ? 0.0 : -g[i];
  }
  Vec xcp = x;
  std::vector<bool> free(n);
  for (std::size_t i = 0; i < n; ++i) free[i] = t[i] > 0.0;
  std::vector<std::size_t> order;
  for (std::size_t i = 0; i < n; ++i)
    if (t[i] > 0.0 && std::isfinite(t[i])) order.push_back(i);
  std::stable_sort(order.begin(), order.end(),
                   [&](std::size_t a, std::size_t b) { return total_less(t[a], t[b]); });
  Vec c(2 * k, 0.0);
  Vec p = mem.wt_dot(d);
  double dd = 0.0;
  for (double v : d) dd += v * v;
  double f_prime = -dd;
  Vec mp;
  if (!mem.apply_m(p, mp)) return false;
  double f_double = -mem.theta * f_prime - dot(p, mp);
  const double f_double_initial = std::max(f_double, kEps);
  f_double = std::max(f_double, kEps);
  double dt_min = -f_prime / f_double;
  double t_old = 0.0;
  std::size_t cursor = 0;
  while (cursor < order.size()) {
    const std::size_t b = order[cursor];
    const double t_b = t[b];
    const double dt = t_b - t_old;
    if (dt_min < dt) break;
    xcp[b] = d[b] > 0.0 ? upper[b] : lower[b];
    const double zb = xcp[b] - x[b];
    free[b] = false;
    for (std::size_t j = 0; j < c.size(); ++j) c[j] += dt * p[j];
    const double gb = g[b];
    Vec wb(2 * k);
    for (std::size_t j = 0; j < 2 * k; ++j) wb[j] = mem.w(b, j);
    Vec mc, mwb;
    if (!mem.apply_m(c, mc) || !mem.apply_m(p, mp) || !mem.apply_m(wb, mwb)) return false;
    f_prime += dt * f_double + gb * gb + mem.theta * gb * zb - gb * dot(wb, mc);
    f_double += -mem.theta * gb * gb - 2.0 * gb * dot(wb, mp) - gb * gb * dot(wb, mwb);
    f_double = std::max(f_double, kEps * f_double_initial);
    for (std::size_t j = 0; j < p.size(); ++j) p[j] += gb * wb[j];
    d[b] = 0.0;
    dt_min = -f_prime / f_double;
    t_old = t_b;
    ++cursor;
  }
  dt_min = std::max(dt_min, 0.0);
  const double t_final = t_old + dt_min;
  for (std::size_t i = 0; i < n; ++i)
    if (free[i] && t[i] >= t_final) xcp[i] = x[i] + t_final * d[i];
  for (std::size_t j = 0; j < c.size(); ++j) c[j] += dt_min * p[j];
  cp.x = std::move(xcp);
  cp.free = std::move(free);
  cp.c = std::move(c);
  return true;
}

inline bool subspace_minimise(const Vec& x, const Vec& g, const Vec& lower, const Vec& upper,
                              const CauchyPoint& cp, const Memory& mem, Vec& out) {
  const std::size_t n = x.size(), k = mem.k();
  if (k == 0) {
    out = cp.x;
    return true;
  }
  std::vector<std::size_t> free;
  for (std::size_t i = 0; i < n; ++i)
    if (cp.free[i]) free.push_back(i);
  if (free.empty()) {
    out = cp.x;
    return true;
  }
  Vec mc;
  if (!mem.apply_m(cp.c, mc)) return false;
  const std::size_t dim = 2 * k;
  Vec r(free.size());
  for (std::size_t slot = 0; slot < free.size(); ++slot) {
    const std::size_t i = free[slot];
    double wmc = 0.0;
    for (std::size_t j = 0; j < dim; ++j) wmc += mem.w(i, j) * mc[j];
    r[slot] = g[i] + mem.theta * (cp.x[i] - x[i]) - wmc;
  }
  Vec v(dim, 0.0);
  for (std::size_t j = 0; j < dim; ++j) {
    double acc = 0.0;
    for (std::size_t slot = 0; slot < free.size(); ++slot) acc += mem.w(free[slot], j) * r[slot];
    v[j] = acc;
  }
  Vec mv;
  if (!mem.apply_m(v, mv)) return false;
  v = mv;
  Vec wtw(dim * dim, 0.0);
  for (std::size_t a = 0; a < dim; ++a)
    for (std::size_t b = 0; b < dim; ++b) {
      double acc = 0.0;
      for (std::size_t i : free) acc += mem.w(i, a) * mem.w(i, b);
      wtw[a * dim + b] = acc;
    }
  Vec n_mat(dim * dim, 0.0), col(dim), m_col;
  for (std::size_t b = 0; b < dim; ++b) {
    for (std::size_t a = 0; a < dim; ++a) col[a] = wtw[a * dim + b];
    if (!mem.apply_m(col, m_col)) return false;
    for (std::size_t a = 0; a < dim; ++a) n_mat[a * dim + b] = -m_col[a] / mem.theta;
  }
  for (std::size_t a = 0; a < dim; ++a) n_mat[a * dim + a] += 1.0;
  if (!solve_dense(n_mat, v, dim)) return false;
  Vec d(free.size());
  for (std::size_t slot = 0; slot < free.size(); ++slot) {
    const std::size_t i = free[slot];
    double acc = 0.0;
    for (std::size_t j = 0; j < dim; ++j) acc += mem.w(i, j) * v[j];
    d[slot] = -(r[slot] / mem.theta + acc / (mem.theta * mem.theta));
  }
  double alpha = 1.0;
  for (std::size_t slot = 0; slot < free.size(); ++slot) {
    const std::size_t i = free[slot];
    if (d[slot] > 0.0 && std::isfinite(upper[i]))
      alpha = std::min(alpha, (upper[i] - cp.x[i]) / d[slot]);
    else if (d[slot] < 0.0 && std::isfinite(lower[i]))
      alpha = std::min(alpha, (lower[i] - cp.x[i]) / d[slot]);
  }
  alpha = clampd(alpha, 0.0, 1.0);
  out = cp.x;
  for (std::size_t slot = 0; slot < free.size(); ++slot) {
    const std::size_t i = free[slot];
    out[i] = clampd(cp.x[i] + alpha * d[slot], lower[i], upper[i]);
  }
  return true;
}

inline double max_feasible_step(const Vec& x, const Vec& dir, const Vec& lower, const Vec& upper) {
  double step = kInf;
  for (std::size_t i = 0; i < x.size(); ++i) {
    if (dir[i] > 0.0 && std::isfinite(upper[i]))
      step = std::min(step, (upper[i] - x[i]) / dir[i]);
    else if (dir[i] < 0.0 && std::isfinite(lower[i]))
      step = std::min(step, (lower[i] - x[i]) / dir[i]);
  }
  return std::max(step, 0.0);
}

struct Accepted {
  Vec x;
  double f;
  Vec g;
  std::size_t evals;
};

template <class F>
bool zoom(F& fg, const Vec& x, const Vec& dir, const Vec& lower, const Vec& upper, double f0,
          double slope0, double alpha_lo, double f_lo, double alpha_hi, std::size_t budget,
          std::size_t evals_so_far, Accepted& out) {
  const std::size_t n = x.size();
  Vec grad(n, 0.0), point(n);
  std::size_t evals = evals_so_far;
  for (std::size_t it = 0; it < budget; ++it) {
    const double alpha = 0.5 * (alpha_lo + alpha_hi);
    if (std::fabs(alpha_hi - alpha_lo) < 1e-16) break;
    for (std::size_t i = 0; i < n; ++i) point[i] = clampd(x[i] + alpha * dir[i], lower[i], upper[i]);
    const double f_new = fg(point, grad);
    ++evals;
    if (!std::isfinite(f_new) || f_new > f0 + kWolfeC1 * alpha * slope0 || f_new >= f_lo) {
      alpha_hi = alpha;
      continue;
    }
    const double slope = dot(grad, dir);
    if (std::fabs(slope) <= -kWolfeC2 * slope0) {
      out = Accepted{point, f_new, grad, evals};
      return true;
    }
    if (slope * (alpha_hi - alpha_lo) >= 0.0) alpha_hi = alpha_lo;
    alpha_lo = alpha;
    f_lo = f_new;
  }
  return false;
}

template <class F>
bool line_search(F& fg, const Vec& x, const Vec& dir, const Vec& lower, const Vec& upper,
                 double f0, const Vec& g0, std::size_t max_evals, Accepted& out) {
  const std::size_t n = x.size();
  const double slope0 = dot(g0, dir);
  if (slope0 >= 0.0) return false;
  const double step_max = std::min(max_feasible_step(x, dir, lower, upper), kMaxStep);
  if (step_max <= 0.0) return false;
  std::size_t evals = 0;
  Vec grad(n, 0.0), point(n);
  double alpha_prev = 0.0, f_prev = f0;
  double alpha = std::min(1.0, step_max);
  bool have_best = false;
  Accepted best;
  for (std::size_t iteration = 0; iteration < max_evals; ++iteration) {
    for (std::size_t i = 0; i < n; ++i) point[i] = clampd(x[i] + alpha * dir[i], lower[i], upper[i]);
    const double f_new = fg(point, grad);
    ++evals;
    if (!std::isfinite(f_new)) {
      alpha *= 0.5;
      if (alpha < 1e-20) break;
      continue;
    }
    const bool sufficient = f_new <= f0 + kWolfeC1 * alpha * slope0;
    if (!sufficient || (iteration > 0 && f_new >= f_prev))
      return zoom(fg, x, dir, lower, upper, f0, slope0, alpha_prev, f_prev, alpha,
                  max_evals - evals, evals, out);
    const double slope = dot(grad, dir);
    if (std::fabs(slope) <= -kWolfeC2 * slope0) {
      out = Accepted{point, f_new, grad, evals};
      return true;
    }
    if (slope >= 0.0)
      return zoom(fg, x, dir, lower, upper, f0, slope0, alpha, f_new, alpha_prev,
                  max_evals - evals, evals, out);
    if (alpha >= step_max) {
      out = Accepted{point, f_new, grad, evals};
      return true;
    }
    best = Accepted{point, f_new, grad, 0};
    have_best = true;
    alpha_prev = alpha;
    f_prev = f_new;
    alpha = std::min(alpha * 2.0, step_max);
  }
  if (!have_best) return false;
  best.evals = evals;
  out = best;
  return true;
}

inline double projected_gradient_norm(const Vec& x, const Vec& g, const Vec& lower,
                                      const Vec& upper) {
  double norm = 0.0;
  for (std::size_t i = 0; i < x.size(); ++i)
    norm = std::max(norm, std::fabs(clampd(x[i] - g[i], lower[i], upper[i]) - x[i]));
  return norm;
}

}  // namespace lbfgsb_detail

// minimise (lbfgsb.rs:876-1020) with memory 10; fg(x, g) returns f and writes
// the gradient.  ok is false only for inputs the reference rejects (lengths,
// an inverted bound); the point is inside the box either way.
template <class F>
LbfgsbResult lbfgsb_minimise(F&& fg, const std::vector<double>& x0,
                             const std::vector<double>& lower,
                             const std::vector<double>& upper, double ftol,
                             double pgtol, std::size_t max_iter,
                             std::size_t max_line_search) {
  using namespace lbfgsb_detail;
  const std::size_t n = x0.size();
  if (n == 0 || lower.size() != n || upper.size() != n) return LbfgsbResult{x0, NAN, false};
  for (std::size_t i = 0; i < n; ++i)
    if (lower[i] > upper[i]) return LbfgsbResult{x0, NAN, false};
  Vec x(n), grad(n, 0.0);
  for (std::size_t i = 0; i < n; ++i) x[i] = clampd(x0[i], lower[i], upper[i]);
  double f = fg(x, grad);
  Memory mem(10);
  for (std::size_t iter = 0; iter < max_iter; ++iter) {
    if (projected_gradient_norm(x, grad, lower, upper) <= pgtol) break;
    CauchyPoint cp;
    Vec target;
    const bool have = cauchy_point(x, grad, lower, upper, mem, cp) &&
                      subspace_minimise(x, grad, lower, upper, cp, mem, target);
    Vec dir(n);
    for (std::size_t i = 0; i < n; ++i)
      dir[i] = have ? target[i] - x[i] : clampd(x[i] - grad[i], lower[i], upper[i]) - x[i];
    bool all_zero = true;
    for (double v : dir) all_zero = all_zero && v == 0.0;
    if (all_zero) break;
    Accepted acc;
    if (!line_search(fg, x, dir, lower, upper, f, grad, max_line_search, acc)) {
      if (mem.k() > 0) {
        mem.reset();
        continue;
      }
      break;
    }
    Vec s(n), y(n);
    for (std::size_t i = 0; i < n; ++i) {
      s[i] = acc.x[i] - x[i];
      y[i] = acc.g[i] - grad[i];
    }
    mem.push(std::move(s), std::move(y));
    const double reduction = f - acc.f;
    const double scale = std::max(std::max(std::fabs(f), std::fabs(acc.f)), 1.0);
    x = acc.x;
    f = acc.f;
    grad = acc.g;
    if (reduction <= ftol * scale) break;
  }
  return LbfgsbResult{x, f, true};
}

}  // namespace oracle

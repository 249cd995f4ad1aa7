// Bounded limited-memory quasi-Newton for the oracle's stage one.
//
// edge-rs runs its own L-BFGS-B (src/numeric/lbfgsb.rs:876, Byrd-Lu-Nocedal-
// Zhu 1995 with a strong-Wolfe line search) as stage one's fallback and as the
// only stage-one optimiser where zero-count tables apply
// (src/sc/nebula.rs:1132-1176).  This is a projected L-BFGS of the same family,
// written from the published method (Bertsekas' projected Newton framework
// with the L-BFGS two-loop recursion on the free variables, Armijo search
// along the projection arc) rather than the compact-representation Cauchy
// point.  Its iterates follow no reference trace (not pinned at the iterate
// level); what the tests pin is the optimum it reaches, which is what stage
// one hands on: the reference's known-answer problems for its L-BFGS-B
// (src/numeric/lbfgsb.rs:1040-1296) at their tolerances, scipy's L-BFGS-B on
// boxed quadratics, and the R goldens through stage one.  On a badly scaled
// box it stops within pgtol short of the corner the Cauchy point reaches.  Stopping rules match the reference's parameters:
// relative objective decrease under ftol, projected-gradient inf-norm under
// pgtol, max_iter iterations, max_line_search evaluations per search.
#pragma once

#include <algorithm>
#include <cmath>
#include <deque>
#include <vector>

#include "oracle.hpp"

namespace oracle {

template <class F>
LbfgsbResult lbfgsb_minimise(F&& fg, const std::vector<double>& x0,
                             const std::vector<double>& lower,
                             const std::vector<double>& upper, double ftol,
                             double pgtol, std::size_t max_iter,
                             std::size_t max_line_search) {
  const std::size_t n = x0.size();
  const std::size_t memory = 10;
  std::vector<double> x(n), g(n), xn(n), gn(n), d(n), q(n);
  for (std::size_t i = 0; i < n; ++i) x[i] = std::min(std::max(x0[i], lower[i]), upper[i]);
  double f = fg(x, g);
  LbfgsbResult res{x, f, std::isfinite(f)};
  if (!std::isfinite(f)) return res;
  std::deque<std::vector<double>> S, Y;
  std::deque<double> R;
  auto projected_norm = [&](const std::vector<double>& xx, const std::vector<double>& gg) {
    double m = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
      double p = std::min(std::max(xx[i] - gg[i], lower[i]), upper[i]) - xx[i];
      m = std::max(m, std::fabs(p));
    }
    return m;
  };
  for (std::size_t it = 0; it < max_iter; ++it) {
    if (projected_norm(x, g) < pgtol) break;
    // Free variables: not held on a bound by an outward gradient.
    std::vector<char> free(n, 1);
    for (std::size_t i = 0; i < n; ++i)
      if ((x[i] <= lower[i] && g[i] > 0.0) || (x[i] >= upper[i] && g[i] < 0.0)) free[i] = 0;
    // Two-loop recursion restricted to the free variables.
    for (std::size_t i = 0; i < n; ++i) q[i] = free[i] ? g[i] : 0.0;
    std::vector<double> a(S.size());
    for (std::size_t m = S.size(); m-- > 0;) {
      double sq = 0.0;
      for (std::size_t i = 0; i < n; ++i) sq += free[i] ? S[m][i] * q[i] : 0.0;
      a[m] = R[m] * sq;
      for (std::size_t i = 0; i < n; ++i)
        if (free[i]) q[i] -= a[m] * Y[m][i];
    }
    double gamma0 = 1.0;
    if (!S.empty()) {
      double sy = 0.0, yy = 0.0;
      for (std::size_t i = 0; i < n; ++i) {
        sy += S.back()[i] * Y.back()[i];
        yy += Y.back()[i] * Y.back()[i];
      }
      if (yy > 0.0) gamma0 = sy / yy;
    } else {
      double gn2 = 0.0;
      for (std::size_t i = 0; i < n; ++i) gn2 += free[i] ? g[i] * g[i] : 0.0;
      gamma0 = gn2 > 0.0 ? std::min(1.0, 1.0 / std::sqrt(gn2)) : 1.0;
    }
    for (std::size_t i = 0; i < n; ++i) q[i] *= gamma0;
    for (std::size_t m = 0; m < S.size(); ++m) {
      double yq = 0.0;
      for (std::size_t i = 0; i < n; ++i) yq += free[i] ? Y[m][i] * q[i] : 0.0;
      double b = R[m] * yq;
      for (std::size_t i = 0; i < n; ++i)
        if (free[i]) q[i] += S[m][i] * (a[m] - b);
    }
    double slope = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
      d[i] = free[i] ? -q[i] : 0.0;
      slope += g[i] * d[i];
    }
    if (!(slope < 0.0)) {
      S.clear();
      Y.clear();
      R.clear();
      for (std::size_t i = 0; i < n; ++i) d[i] = free[i] ? -g[i] : 0.0;
    }
    // Armijo along the projection arc.
    double step = 1.0;
    bool accepted = false;
    double fn = f;
    for (std::size_t ls = 0; ls < max_line_search; ++ls) {
      for (std::size_t i = 0; i < n; ++i)
        xn[i] = std::min(std::max(x[i] + step * d[i], lower[i]), upper[i]);
      double dec = 0.0;
      for (std::size_t i = 0; i < n; ++i) dec += g[i] * (xn[i] - x[i]);
      fn = fg(xn, gn);
      if (std::isfinite(fn) && fn <= f + 1e-4 * dec) {
        accepted = true;
        break;
      }
      step *= 0.5;
    }
    if (!accepted) break;
    std::vector<double> s(n), y(n);
    double sy = 0.0, yy = 0.0;
    for (std::size_t i = 0; i < n; ++i) {
      s[i] = xn[i] - x[i];
      y[i] = gn[i] - g[i];
      sy += s[i] * y[i];
      yy += y[i] * y[i];
    }
    if (sy > 2.2e-16 * yy) {
      S.push_back(s);
      Y.push_back(y);
      R.push_back(1.0 / sy);
      if (S.size() > memory) {
        S.pop_front();
        Y.pop_front();
        R.pop_front();
      }
    }
    double fold = f;
    x = xn;
    g = gn;
    f = fn;
    if (fold - f <= ftol * std::max(std::max(std::fabs(fold), std::fabs(f)), 1.0)) break;
  }
  res.x = x;
  res.f = f;
  res.ok = true;
  return res;
}

}  // namespace oracle

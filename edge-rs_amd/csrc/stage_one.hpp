// Stage one (edge-rs plan_gene / minimise_marginal, src/sc/nebula.rs:812-1325)
// with every pass over cells on the device.
//
// Each gene's search is a C++20 coroutine that suspends whenever it needs the
// marginal likelihood at a point; the driver gathers every suspended gene's
// point, evaluates them all in one launch of the ptmg kernel, turns the
// device sums into value, gradient and Hessian on the host
// (assemble_ptmg: ptmg.rs's gradient_parts, finish_gradient and Hessian
// assembly) and resumes the searches.  Search logic per gene:
//   minimise_marginal  nebula.rs:1075-1116 (bound restarts, STAGE_ONE_RESTARTS)
//   descend_marginal   nebula.rs:1132-1176 (Newton, else the quasi-Newton)
//   newton_marginal    nebula.rs:1225-1325 (projected Newton, Levenberg shift,
//                      Armijo halvings, speculative Hessians)
//   quasi-Newton       projected L-BFGS with an Armijo arc search (the same
//                      method as oracle/lbfgsb.hpp; edge-rs's own is L-BFGS-B,
//                      src/numeric/lbfgsb.rs:876)
#pragma once

#include <algorithm>
#include <cmath>
#include <coroutine>
#include <deque>
#include <utility>
#include <vector>

#include "host_common.hpp"

namespace nh {

struct EvalReq {
  std::vector<double> x;
  bool hess = false;
};
struct EvalRes {
  double f = INFINITY;
  std::vector<double> g, h;
};

struct S1Task {
  struct promise_type {
    EvalReq req;
    EvalRes res;
    bool waiting = false;
    S1Task get_return_object() {
      return S1Task{std::coroutine_handle<promise_type>::from_promise(*this)};
    }
    std::suspend_never initial_suspend() noexcept { return {}; }
    std::suspend_always final_suspend() noexcept { return {}; }
    void return_void() {}
    void unhandled_exception() { std::terminate(); }
  };
  std::coroutine_handle<promise_type> h;
  explicit S1Task(std::coroutine_handle<promise_type> hh) : h(hh) {}
  S1Task(S1Task&& o) noexcept : h(o.h) { o.h = nullptr; }
  S1Task& operator=(S1Task&& o) noexcept {
    if (h) h.destroy();
    h = o.h;
    o.h = nullptr;
    return *this;
  }
  S1Task(const S1Task&) = delete;
  ~S1Task() {
    if (h) h.destroy();
  }
  bool waiting() const { return h && !h.done() && h.promise().waiting; }
};

struct Eval {
  EvalReq req;
  S1Task::promise_type* p = nullptr;
  bool await_ready() const noexcept { return false; }
  void await_suspend(std::coroutine_handle<S1Task::promise_type> hh) noexcept {
    p = &hh.promise();
    p->req = std::move(req);
    p->waiting = true;
  }
  EvalRes await_resume() noexcept {
    p->waiting = false;
    return std::move(p->res);
  }
};

struct S1Result {
  std::vector<double> x;
  bool failed = false;
};

// Cholesky solve of a small SPD system; false when not positive definite or
// not finite (src/utils/linalg.rs:36-104, nebula.rs:2063-2080).
inline bool chol_solve(const std::vector<double>& a, size_t n, const std::vector<double>& b,
                       std::vector<double>& x) {
  std::vector<double> l = a;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j <= i; ++j) {
      double sum = l[i * n + j];
      for (size_t q = 0; q < j; ++q) sum -= l[i * n + q] * l[j * n + q];
      if (i == j) {
        if (sum <= 0.0 || std::isnan(sum)) return false;
        l[i * n + j] = std::sqrt(sum);
      } else {
        l[i * n + j] = sum / l[j * n + j];
      }
    }
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(l[i * n + i])) return false;
  x = b;
  for (size_t i = 0; i < n; ++i) {
    double s = x[i];
    for (size_t q = 0; q < i; ++q) s -= l[i * n + q] * x[q];
    x[i] = s / l[i * n + i];
  }
  for (size_t i = n; i-- > 0;) {
    double s = x[i];
    for (size_t q = i + 1; q < n; ++q) s -= l[q * n + i] * x[q];
    x[i] = s / l[i * n + i];
  }
  for (double v : x)
    if (!std::isfinite(v)) return false;
  return true;
}

inline bool all_finite(double f, const std::vector<double>& g, const std::vector<double>& h) {
  if (!std::isfinite(f)) return false;
  for (double v : g)
    if (!std::isfinite(v)) return false;
  for (double v : h)
    if (!std::isfinite(v)) return false;
  return true;
}

// One gene's stage-one search.  `tables` selects the quasi-Newton only, as
// edge-rs does where zero-count tables apply.
inline S1Task stage_one_search(std::vector<double> start, std::vector<double> lo,
                               std::vector<double> hi, bool tables, size_t n_coef,
                               S1Result* out) {
  constexpr double kFtol = 1e-6, kPgtol = 1e-7, kArmijo = 1e-4, kShift = 1e-8, kGain = 1e-6;
  constexpr int kNewtonIter = 50, kHalvings = 40, kRestarts = 3, kLbfgsIter = 300, kLs = 40;
  const size_t n = start.size();
  std::vector<double> x0(n);
  for (size_t j = 0; j < n; ++j) x0[j] = std::min(std::max(start[j], lo[j]), hi[j]);
  EvalRes r0 = co_await Eval{EvalReq{x0, false}};
  const double f0 = r0.f;
  if (!std::isfinite(f0)) {
    out->x = x0;
    out->failed = true;
    co_return;
  }
  std::vector<double> best, last, anchor;
  double best_f = 0.0;
  for (int attempt = 0; attempt <= kRestarts; ++attempt) {
    std::vector<double> xs;
    double fs;
    if (attempt == 0) {
      xs = x0;
      fs = f0;
    } else {
      // bound_restart (nebula.rs:1190-1211)
      xs = last;
      bool moved = false;
      for (size_t j = 0; j < n; ++j)
        for (double bound : {lo[j], hi[j]})
          if (last[j] == bound && anchor[j] != bound && bound * anchor[j] > 0.0) {
            xs[j] = std::sqrt(bound * anchor[j]);
            moved = true;
          }
      if (!moved) break;
      EvalRes r1 = co_await Eval{EvalReq{xs, false}};
      fs = r1.f;
      if (!std::isfinite(fs)) break;
    }
    // ---- descend_marginal from (xs, fs) ----
    std::vector<double> xd = xs;
    double fd = fs;
    bool have = false;
    if (!tables) {
      // newton_marginal
      std::vector<double> x = xs;
      EvalRes e = co_await Eval{EvalReq{x, true}};
      double f = e.f;
      std::vector<double> g = e.g, h = e.h;
      bool ok = all_finite(f, g, h);
      const bool speculative = n_coef <= 8;
      bool converged = false;
      for (int it = 0; ok && it < kNewtonIter; ++it) {
        double projected = 0.0;
        for (size_t i = 0; i < n; ++i)
          projected = std::max(projected, std::fabs(std::min(std::max(x[i] - g[i], lo[i]), hi[i]) - x[i]));
        if (projected < kPgtol) {
          converged = true;
          break;
        }
        std::vector<size_t> fr;
        for (size_t i = 0; i < n; ++i)
          if (!((x[i] <= lo[i] && g[i] > 0.0) || (x[i] >= hi[i] && g[i] < 0.0))) fr.push_back(i);
        const size_t m = fr.size();
        std::vector<double> rhs(m), step;
        double largest = 0.0;
        for (size_t p = 0; p < m; ++p) {
          rhs[p] = -g[fr[p]];
          largest = std::max(largest, std::fabs(h[fr[p] * n + fr[p]]));
        }
        double shift = 0.0;
        bool solved = false;
        for (;;) {
          std::vector<double> blk(m * m);
          for (size_t p = 0; p < m; ++p) {
            for (size_t q = 0; q < m; ++q) blk[p * m + q] = h[fr[p] * n + fr[q]];
            blk[p * m + p] += shift;
          }
          if (chol_solve(blk, m, rhs, step)) {
            solved = true;
            break;
          }
          shift = shift == 0.0 ? kShift * std::max(largest, 1.0) : shift * 10.0;
          if (!std::isfinite(shift)) break;
        }
        if (!solved) {
          ok = false;
          break;
        }
        std::vector<double> d(n, 0.0);
        for (size_t p = 0; p < m; ++p) d[fr[p]] = step[p];
        double a = 1.0;
        bool accepted = false;
        std::vector<double> xt(n), gt, ht;
        double ft = 0.0;
        bool derivs = false;
        for (int hv = 0; hv < kHalvings; ++hv) {
          for (size_t i = 0; i < n; ++i) xt[i] = std::min(std::max(x[i] + a * d[i], lo[i]), hi[i]);
          double dec = 0.0;
          for (size_t i = 0; i < n; ++i) dec += g[i] * (xt[i] - x[i]);
          EvalRes et = co_await Eval{EvalReq{xt, speculative}};
          ft = et.f;
          bool good = speculative ? all_finite(ft, et.g, et.h) : std::isfinite(ft);
          if (good && ft <= f + kArmijo * dec) {
            accepted = true;
            gt = std::move(et.g);
            ht = std::move(et.h);
            derivs = speculative;
            break;
          }
          a *= 0.5;
        }
        if (!accepted) {
          ok = false;
          break;
        }
        if ((f - ft) <= kFtol) {
          x = xt;
          f = ft;
          converged = true;
          break;
        }
        if (!derivs) {
          EvalRes eh = co_await Eval{EvalReq{xt, true}};
          ft = eh.f;
          gt = std::move(eh.g);
          ht = std::move(eh.h);
          if (!all_finite(ft, gt, ht)) {
            ok = false;
            break;
          }
        }
        x = xt;
        f = ft;
        g = gt;
        h = ht;
      }
      if (ok && converged && f < fs) {
        xd = x;
        fd = f;
        have = true;
      }
    }
    if (!have) {
      // projected L-BFGS from (xs, fs)
      const double ftol = kFtol / std::max(std::fabs(fs), 1.0);
      std::vector<double> x = xs, g;
      double f;
      {
        EvalRes e = co_await Eval{EvalReq{x, false}};
        f = e.f;
        g = std::move(e.g);
      }
      std::deque<std::vector<double>> S, Y;
      std::deque<double> R;
      if (std::isfinite(f)) {
        for (int it = 0; it < kLbfgsIter; ++it) {
          double pn = 0.0;
          for (size_t i = 0; i < n; ++i)
            pn = std::max(pn, std::fabs(std::min(std::max(x[i] - g[i], lo[i]), hi[i]) - x[i]));
          if (pn < kPgtol) break;
          std::vector<char> fr(n, 1);
          for (size_t i = 0; i < n; ++i)
            if ((x[i] <= lo[i] && g[i] > 0.0) || (x[i] >= hi[i] && g[i] < 0.0)) fr[i] = 0;
          std::vector<double> q(n), al(S.size()), d(n);
          for (size_t i = 0; i < n; ++i) q[i] = fr[i] ? g[i] : 0.0;
          for (size_t mm = S.size(); mm-- > 0;) {
            double sq = 0.0;
            for (size_t i = 0; i < n; ++i) sq += fr[i] ? S[mm][i] * q[i] : 0.0;
            al[mm] = R[mm] * sq;
            for (size_t i = 0; i < n; ++i)
              if (fr[i]) q[i] -= al[mm] * Y[mm][i];
          }
          double g0s = 1.0;
          if (!S.empty()) {
            double sy = 0.0, yy = 0.0;
            for (size_t i = 0; i < n; ++i) {
              sy += S.back()[i] * Y.back()[i];
              yy += Y.back()[i] * Y.back()[i];
            }
            if (yy > 0.0) g0s = sy / yy;
          } else {
            double gn2 = 0.0;
            for (size_t i = 0; i < n; ++i) gn2 += fr[i] ? g[i] * g[i] : 0.0;
            g0s = gn2 > 0.0 ? std::min(1.0, 1.0 / std::sqrt(gn2)) : 1.0;
          }
          for (size_t i = 0; i < n; ++i) q[i] *= g0s;
          for (size_t mm = 0; mm < S.size(); ++mm) {
            double yq = 0.0;
            for (size_t i = 0; i < n; ++i) yq += fr[i] ? Y[mm][i] * q[i] : 0.0;
            double b = R[mm] * yq;
            for (size_t i = 0; i < n; ++i)
              if (fr[i]) q[i] += S[mm][i] * (al[mm] - b);
          }
          double slope = 0.0;
          for (size_t i = 0; i < n; ++i) {
            d[i] = fr[i] ? -q[i] : 0.0;
            slope += g[i] * d[i];
          }
          if (!(slope < 0.0)) {
            S.clear();
            Y.clear();
            R.clear();
            for (size_t i = 0; i < n; ++i) d[i] = fr[i] ? -g[i] : 0.0;
          }
          double stp = 1.0, fn = f;
          std::vector<double> xn(n), gn;
          bool acc = false;
          for (int ls = 0; ls < kLs; ++ls) {
            for (size_t i = 0; i < n; ++i) xn[i] = std::min(std::max(x[i] + stp * d[i], lo[i]), hi[i]);
            double dec = 0.0;
            for (size_t i = 0; i < n; ++i) dec += g[i] * (xn[i] - x[i]);
            EvalRes et = co_await Eval{EvalReq{xn, false}};
            fn = et.f;
            if (std::isfinite(fn) && fn <= f + 1e-4 * dec) {
              gn = std::move(et.g);
              acc = true;
              break;
            }
            stp *= 0.5;
          }
          if (!acc) break;
          std::vector<double> s(n), y(n);
          double sy = 0.0, yy = 0.0;
          for (size_t i = 0; i < n; ++i) {
            s[i] = xn[i] - x[i];
            y[i] = gn[i] - g[i];
            sy += s[i] * y[i];
            yy += y[i] * y[i];
          }
          if (sy > 2.2e-16 * yy) {
            S.push_back(s);
            Y.push_back(y);
            R.push_back(1.0 / sy);
            if (S.size() > 10) {
              S.pop_front();
              Y.pop_front();
              R.pop_front();
            }
          }
          const double fold = f;
          x = xn;
          g = gn;
          f = fn;
          if (fold - f <= ftol * std::max(std::max(std::fabs(fold), std::fabs(f)), 1.0)) break;
        }
        if (f < fs) {
          xd = x;
          fd = f;
        }
      }
    }
    // ---- end of descend ----
    if (attempt == 0) {
      best = xd;
      best_f = fd;
      last = xd;
      anchor = x0;
    } else {
      if (fd < best_f - kGain) {
        best = xd;
        best_f = fd;
      }
      last = xd;
      anchor = xs;
    }
  }
  out->x = best;
  out->failed = false;
}

}  // namespace nh

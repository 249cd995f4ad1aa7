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
//   quasi-Newton       L-BFGS-B (src/numeric/lbfgsb.rs:876-1020: compact
//                      representation, generalised Cauchy point, subspace
//                      minimisation, strong-Wolfe line search with bisection
//                      zoom, memory reset on a stalled search), the middle
//                      matrix factored once per iteration and replayed on each
//                      right-hand side (the same operations as eliminating it
//                      afresh for each, lbfgsb.rs:317-361)
#pragma once

#include <algorithm>
#include <cmath>
#include <coroutine>
#include <cstdint>
#include <cstring>
#include <deque>
#include <limits>
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

// ---------------------------------------------------------------------------
// L-BFGS-B's pieces that need no objective values (lbfgsb.rs:130-639).
struct QnMemory {
  std::vector<std::vector<double>> s, y;  // oldest first, at most kLimit pairs
  double theta = 1.0;
  static constexpr size_t kLimit = 10;
  // Gaussian elimination of the middle matrix [[-D, L'], [L, theta S'S]]
  // with partial pivoting: U in lu's upper triangle, the row factors below,
  // the pivot of each column; usable when ok
  std::vector<double> lu;
  std::vector<size_t> piv;
  bool ok = true;

  size_t k() const { return s.size(); }
  double w(size_t i, size_t j) const { return j < k() ? y[j][i] : theta * s[j - k()][i]; }
  static double dot(const std::vector<double>& a, const std::vector<double>& b) {
    double t = 0.0;
    for (size_t i = 0; i < a.size(); ++i) t += a[i] * b[i];
    return t;
  }
  void reset() {
    s.clear();
    y.clear();
    theta = 1.0;
    factor();
  }
  void push(std::vector<double> sv, std::vector<double> yv) {
    const double sy = dot(sv, yv), yy = dot(yv, yv);
    if (sy <= 2.2e-16 * yy || yy <= 0.0) return;
    if (s.size() == kLimit) {
      s.erase(s.begin());
      y.erase(y.begin());
    }
    theta = yy / sy;
    s.push_back(std::move(sv));
    y.push_back(std::move(yv));
    factor();
  }
  void factor() {
    const size_t kk = k(), dim = 2 * kk;
    lu.assign(dim * dim, 0.0);
    piv.assign(dim, 0);
    ok = true;
    for (size_t i = 0; i < kk; ++i) lu[i * dim + i] = -dot(s[i], y[i]);
    for (size_t i = 0; i < kk; ++i)
      for (size_t j = 0; j < i; ++j) {
        const double sy = dot(s[i], y[j]);
        lu[(kk + i) * dim + j] = sy;
        lu[j * dim + (kk + i)] = sy;
      }
    for (size_t i = 0; i < kk; ++i)
      for (size_t j = 0; j < kk; ++j) lu[(kk + i) * dim + (kk + j)] = theta * dot(s[i], s[j]);
    for (size_t col = 0; col < dim; ++col) {
      size_t pv = col;
      double best = std::fabs(lu[col * dim + col]);
      for (size_t row = col + 1; row < dim; ++row)
        if (std::fabs(lu[row * dim + col]) > best) {
          best = std::fabs(lu[row * dim + col]);
          pv = row;
        }
      piv[col] = pv;
      if (best < 1e-300) {
        ok = false;
        return;
      }
      if (pv != col)
        for (size_t j = 0; j < dim; ++j) std::swap(lu[col * dim + j], lu[pv * dim + j]);
      const double diag = lu[col * dim + col];
      for (size_t row = col + 1; row < dim; ++row) {
        const double f = lu[row * dim + col] / diag;
        if (f != 0.0)
          for (size_t j = col + 1; j < dim; ++j) lu[row * dim + j] -= f * lu[col * dim + j];
        lu[row * dim + col] = f;  // kept for the right-hand sides
      }
    }
  }
  // v <- M v (the middle matrix's inverse applied by elimination).  The
  // factors moved with the row swaps, so the swaps go first; each entry then
  // takes the same updates, in the same column order, as elimination with
  // the right-hand side alongside.
  bool apply_m(std::vector<double>& v) const {
    const size_t dim = 2 * k();
    if (dim == 0) return true;
    if (!ok) return false;
    for (size_t col = 0; col < dim; ++col)
      if (piv[col] != col) std::swap(v[col], v[piv[col]]);
    for (size_t col = 0; col < dim; ++col)
      for (size_t row = col + 1; row < dim; ++row) {
        const double f = lu[row * dim + col];
        if (f != 0.0) v[row] -= f * v[col];
      }
    for (size_t col = dim; col-- > 0;) {
      double acc = v[col];
      for (size_t j = col + 1; j < dim; ++j) acc -= lu[col * dim + j] * v[j];
      v[col] = acc / lu[col * dim + col];
    }
    return true;
  }
};

// Partial-pivot elimination of a dense system (lbfgsb.rs:317-361).
inline bool qn_solve_dense(std::vector<double> m, std::vector<double>& x, size_t dim) {
  for (size_t col = 0; col < dim; ++col) {
    size_t pv = col;
    double best = std::fabs(m[col * dim + col]);
    for (size_t row = col + 1; row < dim; ++row)
      if (std::fabs(m[row * dim + col]) > best) {
        best = std::fabs(m[row * dim + col]);
        pv = row;
      }
    if (best < 1e-300) return false;
    if (pv != col) {
      for (size_t j = 0; j < dim; ++j) std::swap(m[col * dim + j], m[pv * dim + j]);
      std::swap(x[col], x[pv]);
    }
    const double diag = m[col * dim + col];
    for (size_t row = col + 1; row < dim; ++row) {
      const double f = m[row * dim + col] / diag;
      if (f == 0.0) continue;
      for (size_t j = col; j < dim; ++j) m[row * dim + j] -= f * m[col * dim + j];
      x[row] -= f * x[col];
    }
  }
  for (size_t col = dim; col-- > 0;) {
    double acc = x[col];
    for (size_t j = col + 1; j < dim; ++j) acc -= m[col * dim + j] * x[j];
    x[col] = acc / m[col * dim + col];
  }
  return true;
}

inline double qn_clamp(double v, double a, double b) { return v < a ? a : (v > b ? b : v); }

// f64::total_cmp's strict order
inline bool qn_total_less(double a, double b) {
  int64_t x, y;
  std::memcpy(&x, &a, 8);
  std::memcpy(&y, &b, 8);
  x ^= (int64_t)((uint64_t)(x >> 63) >> 1);
  y ^= (int64_t)((uint64_t)(y >> 63) >> 1);
  return x < y;
}

// Generalised Cauchy point then subspace minimisation (Algorithm CP and the
// direct primal method, lbfgsb.rs:400-639); false when a small system is
// singular (the caller then takes the projected gradient step).
inline bool qn_target(const std::vector<double>& x, const std::vector<double>& g,
                      const std::vector<double>& lo, const std::vector<double>& hi,
                      const QnMemory& mem, std::vector<double>& out) {
  const size_t n = x.size(), k = mem.k(), dim = 2 * k;
  const double inf = std::numeric_limits<double>::infinity();
  const double eps = std::numeric_limits<double>::epsilon();
  std::vector<double> t(n, inf), d(n, 0.0);
  for (size_t i = 0; i < n; ++i) {
    if (g[i] < 0.0 && std::isfinite(hi[i]))
      t[i] = (x[i] - hi[i]) / g[i];
    else if (g[i] > 0.0 && std::isfinite(lo[i]))
      t[i] = (x[i] - lo[i]) / g[i];
    d[i] = t[i] == 0.0 ? 0.0 : -g[i];
  }
  std::vector<double> xcp = x;
  std::vector<char> fr(n);
  for (size_t i = 0; i < n; ++i) fr[i] = t[i] > 0.0;
  std::vector<size_t> order;
  for (size_t i = 0; i < n; ++i)
    if (t[i] > 0.0 && std::isfinite(t[i])) order.push_back(i);
  std::stable_sort(order.begin(), order.end(),
                   [&](size_t a, size_t b) { return qn_total_less(t[a], t[b]); });
  std::vector<double> c(dim, 0.0), p(dim, 0.0), mp, mc, wb(dim), mwb;
  for (size_t j = 0; j < dim; ++j) {
    double acc = 0.0;
    for (size_t i = 0; i < n; ++i) acc += mem.w(i, j) * d[i];
    p[j] = acc;
  }
  double dd = 0.0;
  for (double v : d) dd += v * v;
  double f1 = -dd;
  mp = p;
  if (!mem.apply_m(mp)) return false;
  double f2 = -mem.theta * f1 - QnMemory::dot(p, mp);
  const double f2_initial = std::max(f2, eps);
  f2 = std::max(f2, eps);
  double dt_min = -f1 / f2, t_old = 0.0;
  for (size_t cur = 0; cur < order.size(); ++cur) {
    const size_t b = order[cur];
    const double dt = t[b] - t_old;
    if (dt_min < dt) break;
    xcp[b] = d[b] > 0.0 ? hi[b] : lo[b];
    const double zb = xcp[b] - x[b];
    fr[b] = 0;
    for (size_t j = 0; j < dim; ++j) c[j] += dt * p[j];
    const double gb = g[b];
    for (size_t j = 0; j < dim; ++j) wb[j] = mem.w(b, j);
    mc = c;
    mp = p;
    mwb = wb;
    if (!mem.apply_m(mc) || !mem.apply_m(mp) || !mem.apply_m(mwb)) return false;
    f1 += dt * f2 + gb * gb + mem.theta * gb * zb - gb * QnMemory::dot(wb, mc);
    f2 += -mem.theta * gb * gb - 2.0 * gb * QnMemory::dot(wb, mp) - gb * gb * QnMemory::dot(wb, mwb);
    f2 = std::max(f2, eps * f2_initial);
    for (size_t j = 0; j < dim; ++j) p[j] += gb * wb[j];
    d[b] = 0.0;
    dt_min = -f1 / f2;
    t_old = t[b];
  }
  dt_min = std::max(dt_min, 0.0);
  const double t_final = t_old + dt_min;
  for (size_t i = 0; i < n; ++i)
    if (fr[i] && t[i] >= t_final) xcp[i] = x[i] + t_final * d[i];
  for (size_t j = 0; j < dim; ++j) c[j] += dt_min * p[j];
  // subspace minimisation over the free variables
  out = xcp;
  std::vector<size_t> free;
  for (size_t i = 0; i < n; ++i)
    if (fr[i]) free.push_back(i);
  if (k == 0 || free.empty()) return true;
  mc = c;
  if (!mem.apply_m(mc)) return false;
  const size_t nf = free.size();
  std::vector<double> r(nf), v(dim, 0.0);
  for (size_t q = 0; q < nf; ++q) {
    const size_t i = free[q];
    double wmc = 0.0;
    for (size_t j = 0; j < dim; ++j) wmc += mem.w(i, j) * mc[j];
    r[q] = g[i] + mem.theta * (xcp[i] - x[i]) - wmc;
  }
  for (size_t j = 0; j < dim; ++j) {
    double acc = 0.0;
    for (size_t q = 0; q < nf; ++q) acc += mem.w(free[q], j) * r[q];
    v[j] = acc;
  }
  if (!mem.apply_m(v)) return false;
  std::vector<double> nm(dim * dim, 0.0), col(dim);
  for (size_t b = 0; b < dim; ++b) {
    for (size_t a = 0; a < dim; ++a) {
      double acc = 0.0;
      for (size_t i : free) acc += mem.w(i, a) * mem.w(i, b);
      col[a] = acc;
    }
    if (!mem.apply_m(col)) return false;
    for (size_t a = 0; a < dim; ++a) nm[a * dim + b] = -col[a] / mem.theta;
  }
  for (size_t a = 0; a < dim; ++a) nm[a * dim + a] += 1.0;
  if (!qn_solve_dense(nm, v, dim)) return false;
  std::vector<double> dz(nf);
  for (size_t q = 0; q < nf; ++q) {
    double acc = 0.0;
    for (size_t j = 0; j < dim; ++j) acc += mem.w(free[q], j) * v[j];
    dz[q] = -(r[q] / mem.theta + acc / (mem.theta * mem.theta));
  }
  double alpha = 1.0;
  for (size_t q = 0; q < nf; ++q) {
    const size_t i = free[q];
    if (dz[q] > 0.0 && std::isfinite(hi[i]))
      alpha = std::min(alpha, (hi[i] - xcp[i]) / dz[q]);
    else if (dz[q] < 0.0 && std::isfinite(lo[i]))
      alpha = std::min(alpha, (lo[i] - xcp[i]) / dz[q]);
  }
  alpha = qn_clamp(alpha, 0.0, 1.0);
  for (size_t q = 0; q < nf; ++q) {
    const size_t i = free[q];
    out[i] = qn_clamp(xcp[i] + alpha * dz[q], lo[i], hi[i]);
  }
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
  // the Newton descent's first step needs the start's Hessian: ask for it
  // with the start's value (the same sums, so the same bits) rather than in
  // a second round
  EvalRes r0 = co_await Eval{EvalReq{x0, !tables}};
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
      EvalRes e;
      if (attempt == 0)
        e = r0;  // the start, evaluated with its Hessian
      else
        e = co_await Eval{EvalReq{x, true}};
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
      // L-BFGS-B from (xs, fs): minimise (lbfgsb.rs:876-1020) with memory 10,
      // ftol relative to the start's objective (nebula.rs:1150-1156)
      const double ftol = kFtol / std::max(std::fabs(fs), 1.0);
      std::vector<double> x = xs, g;
      double f;
      {
        EvalRes e = co_await Eval{EvalReq{x, false}};
        f = e.f;
        g = std::move(e.g);
      }
      QnMemory mem;
      mem.factor();
      for (int it = 0; it < kLbfgsIter; ++it) {
        double pn = 0.0;
        for (size_t i = 0; i < n; ++i)
          pn = std::max(pn, std::fabs(qn_clamp(x[i] - g[i], lo[i], hi[i]) - x[i]));
        if (pn <= kPgtol) break;
        std::vector<double> target, dir(n);
        const bool have_t = qn_target(x, g, lo, hi, mem, target);
        bool zero = true;
        for (size_t i = 0; i < n; ++i) {
          dir[i] = have_t ? target[i] - x[i] : qn_clamp(x[i] - g[i], lo[i], hi[i]) - x[i];
          zero = zero && dir[i] == 0.0;
        }
        if (zero) break;
        // strong-Wolfe line search with bisection zoom (lbfgsb.rs:675-874)
        double slope0 = 0.0;
        for (size_t i = 0; i < n; ++i) slope0 += g[i] * dir[i];
        double step_max = std::numeric_limits<double>::infinity();
        for (size_t i = 0; i < n; ++i) {
          if (dir[i] > 0.0 && std::isfinite(hi[i]))
            step_max = std::min(step_max, (hi[i] - x[i]) / dir[i]);
          else if (dir[i] < 0.0 && std::isfinite(lo[i]))
            step_max = std::min(step_max, (lo[i] - x[i]) / dir[i]);
        }
        step_max = std::min(std::max(step_max, 0.0), 1e3);
        bool found = false;
        std::vector<double> xa, ga, pt(n);
        double fa = 0.0;
        if (slope0 < 0.0 && step_max > 0.0) {
          int evals = 0;
          double a_prev = 0.0, f_prev = f, a = std::min(1.0, step_max);
          bool have_best = false, zoom = false;
          double z_lo = 0.0, zf_lo = 0.0, z_hi = 0.0;
          std::vector<double> xb, gb;
          double fb = 0.0;
          for (int ls = 0; ls < kLs; ++ls) {
            for (size_t i = 0; i < n; ++i) pt[i] = qn_clamp(x[i] + a * dir[i], lo[i], hi[i]);
            EvalRes et = co_await Eval{EvalReq{pt, false}};
            ++evals;
            const double fn = et.f;
            if (!std::isfinite(fn)) {
              a *= 0.5;
              if (a < 1e-20) break;
              continue;
            }
            if (!(fn <= f + 1e-4 * a * slope0) || (ls > 0 && fn >= f_prev)) {
              zoom = true;
              z_lo = a_prev;
              zf_lo = f_prev;
              z_hi = a;
              break;
            }
            double sl = 0.0;
            for (size_t i = 0; i < n; ++i) sl += et.g[i] * dir[i];
            const bool curvature = std::fabs(sl) <= -0.9 * slope0;
            if (!curvature && sl >= 0.0) {
              zoom = true;
              z_lo = a;
              zf_lo = fn;
              z_hi = a_prev;
              break;
            }
            // the curvature condition, or the box edge with the objective
            // still descending
            if (curvature || a >= step_max) {
              found = true;
              xa = pt;
              fa = fn;
              ga = std::move(et.g);
              break;
            }
            have_best = true;
            xb = pt;
            fb = fn;
            gb = std::move(et.g);
            a_prev = a;
            f_prev = fn;
            a = std::min(a * 2.0, step_max);
          }
          if (zoom) {
            for (int zz = evals; zz < kLs; ++zz) {
              const double am = 0.5 * (z_lo + z_hi);
              if (std::fabs(z_hi - z_lo) < 1e-16) break;
              for (size_t i = 0; i < n; ++i) pt[i] = qn_clamp(x[i] + am * dir[i], lo[i], hi[i]);
              EvalRes et = co_await Eval{EvalReq{pt, false}};
              const double fn = et.f;
              if (!std::isfinite(fn) || fn > f + 1e-4 * am * slope0 || fn >= zf_lo) {
                z_hi = am;
                continue;
              }
              double sl = 0.0;
              for (size_t i = 0; i < n; ++i) sl += et.g[i] * dir[i];
              if (std::fabs(sl) <= -0.9 * slope0) {
                found = true;
                xa = pt;
                fa = fn;
                ga = std::move(et.g);
                break;
              }
              if (sl * (z_hi - z_lo) >= 0.0) z_hi = z_lo;
              z_lo = am;
              zf_lo = fn;
            }
          } else if (!found && have_best) {
            found = true;
            xa = xb;
            fa = fb;
            ga = gb;
          }
        }
        if (!found) {
          if (mem.k() > 0) {
            mem.reset();
            continue;
          }
          break;
        }
        std::vector<double> sv(n), yv(n);
        for (size_t i = 0; i < n; ++i) {
          sv[i] = xa[i] - x[i];
          yv[i] = ga[i] - g[i];
        }
        mem.push(std::move(sv), std::move(yv));
        const double reduction = f - fa;
        const double scale = std::max(std::max(std::fabs(f), std::fabs(fa)), 1.0);
        x = xa;
        f = fa;
        g = ga;
        if (reduction <= ftol * scale) break;
      }
      if (f < fs) {
        xd = x;
        fd = f;
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

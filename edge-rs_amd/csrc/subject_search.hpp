// NEBULA-LN's one-component stage-two search: the subject-level variance
// alone, the cell-level size held at stage one's value, as a
// reverse-communication state machine that runs on the host (plain engine)
// and on one lane of a device workgroup (persistent kernel).
//
// Semantics of edge-rs 0.4.0's StageTwoSearch with one component
// (src/sc/nebula.rs:1568-1951): evaluate the clamped start; a bounded
// Nelder-Mead simplex (NelderMeadStepper, src/numeric/optimise.rs:383-670, at
// n = 1: initial vertices x0 and 1.05 x0, reflect / expand / contract /
// shrink, vertices ordered by f64::total_cmp with ties kept, stop when both
// spreads are within 1e-7 or the vertices are 16 ulps apart, at most 500
// iterations), every asked point clamped into the box; one restart of the
// simplex when it ends on a bound the start was strictly inside, from the
// geometric mean of that bound and the start (bound_restart, :1190), keeping
// the better of the two; the clamped minimiser evaluated once more; three
// quadratic polish stencils of three points at widths 1e-2, 3e-3, 1e-3 of
// |x| (at least 1e-5), shifted off a bound, fitted by least squares through
// the normal equations and stepped by Newton when the step is finite and at
// most two widths; a non-finite value marks the pass, and a marked pass at a
// raised Laplace order is repeated from the start at order one; a
// non-finite start or check value ends the pass with nothing found.
//
// State is a fixed-size plain struct (no allocation), so it sits in LDS or
// host memory alike.  Contraction is off: the host and device runs, and the
// oracle's restatement, must take the same steps for the same values.
#pragma once

#ifndef NH_HD
#if defined(__HIPCC__)
#define NH_HD __host__ __device__
#else
#define NH_HD
#endif
#endif

#include <cmath>
#include <cstdint>
#include <cstring>

namespace nh {

#if defined(__clang__)
#define NH_SS_STRICT _Pragma("clang fp contract(off)")
#else
#define NH_SS_STRICT
#endif

struct SubjectSearch {
  // constants (src/sc/nebula.rs:100-118,173-179; src/numeric/optimise.rs:40-56)
  static constexpr double kXatol = 1e-7, kFatol = 1e-7;
  static constexpr int kMaxIter = 500;
  static constexpr double kStep = 0.05, kZeroStep = 0.00025, kCollapsedUlps = 16.0;
  static constexpr double kPolishFloor = 1e-5, kPolishMaxStep = 2.0;
  static constexpr int kWidths = 3;

  enum Stage : int { kStart, kSimplex, kCheck, kPolish, kDone };
  enum Move : int { kInit0, kInit1, kReflect, kExpand, kContract, kShrink, kFinished };

  // ---- interface ----
  int n_ask;       // points waiting for values (0 once done)
  double ask[3];   // those points
  int order;       // Laplace order they are to be evaluated at
  bool found;      // once done: whether a minimiser was found
  double x_out;    // that minimiser

  // ---- search ----
  double lo, hi, start;
  int stage;
  bool invalid, restarted, pinned;
  double pinned_x, pinned_f, best;
  int width;
  double step, offs[3];

  // ---- simplex (two vertices) ----
  double vx[2], vf[2];
  int rank[2];  // vertex indices by value, best first
  int move, next_iter;
  double centroid, reflected, f_reflected, pending;
  bool inside;

  NH_HD static double clampd(double v, double a, double b) { return v < a ? a : (v > b ? b : v); }

  // f64::total_cmp's order: a strictly below b.
  NH_HD static bool below(double a, double b) {
    int64_t x, y;
    std::memcpy(&x, &a, 8);
    std::memcpy(&y, &b, 8);
    x ^= (int64_t)((uint64_t)(x >> 63) >> 1);
    y ^= (int64_t)((uint64_t)(y >> 63) >> 1);
    return x < y;
  }

  NH_HD void init(double x0, double lower, double upper, int ord) {
    lo = lower;
    hi = upper;
    start = x0;
    order = ord;
    found = false;
    x_out = 0.0;
    begin_pass();
  }

  NH_HD bool done() const { return stage == kDone; }

  // One value per asked point, in order.
  NH_HD void tell(const double* v) {
    NH_SS_STRICT
    for (int i = 0; i < n_ask; ++i)
      if (!__builtin_isfinite(v[i])) invalid = true;
    switch (stage) {
      case kStart:
        if (!__builtin_isfinite(v[0])) {
          end_pass(false);
          return;
        }
        simplex_new(clampd(start, lo, hi));
        stage = kSimplex;
        ask_simplex();
        return;
      case kSimplex:
        simplex_tell(v[0]);
        ask_simplex();
        return;
      case kCheck:
        if (!__builtin_isfinite(v[0])) {
          end_pass(false);
          return;
        }
        width = 0;
        begin_polish();
        return;
      case kPolish: {
        bool all = true;
        for (int i = 0; i < 3; ++i) all = all && __builtin_isfinite(v[i]);
        double next;
        if (all && polish_step(v, next)) best = next;
        ++width;
        begin_polish();
        return;
      }
      default:
        return;
    }
  }

 private:
  NH_HD void begin_pass() {
    invalid = false;
    restarted = false;
    pinned = false;
    stage = kStart;
    n_ask = 1;
    ask[0] = clampd(start, lo, hi);
  }

  NH_HD void end_pass(bool got) {
    if ((got && !invalid) || order == 1) {
      found = got;
      if (got) x_out = best;
      stage = kDone;
      n_ask = 0;
      return;
    }
    order = 1;
    begin_pass();
  }

  // ---- Nelder-Mead at n = 1 ----
  NH_HD void simplex_new(double x0) {
    NH_SS_STRICT
    vx[0] = x0;
    vx[1] = x0 != 0.0 ? x0 * (1.0 + kStep) : kZeroStep;
    vf[0] = vf[1] = 0.0;
    rank[0] = 0;
    rank[1] = 1;
    next_iter = 0;
    move = kInit0;
    pending = vx[0];
  }

  NH_HD void sort_rank() {
    // stable: swap only when the second is strictly below the first
    if (below(vf[rank[1]], vf[rank[0]])) {
      const int t = rank[0];
      rank[0] = rank[1];
      rank[1] = t;
    }
  }

  NH_HD void simplex_finish() {
    sort_rank();
    move = kFinished;
  }

  NH_HD void simplex_iterate() {
    NH_SS_STRICT
    if (next_iter >= kMaxIter) {
      simplex_finish();
      return;
    }
    ++next_iter;
    sort_rank();
    const int b = rank[0], w = rank[1];
    // folded from 0 with f64::max, which drops a NaN (inf - inf)
    const double spread_x = fmax(0.0, fabs(vx[w] - vx[b]));
    const double spread_f = fmax(0.0, fabs(vf[w] - vf[b]));
    const double scale = fmax(2.2250738585072014e-308, fabs(vx[b]));
    const bool collapsed = spread_x <= kCollapsedUlps * 2.220446049250313e-16 * scale;
    if (collapsed || (spread_x <= kXatol && spread_f <= kFatol)) {
      simplex_finish();
      return;
    }
    centroid = (0.0 + vx[b]) / 1.0;
    reflected = centroid + (centroid - vx[w]);
    pending = reflected;
    move = kReflect;
  }

  NH_HD void simplex_tell(double value) {
    NH_SS_STRICT
    const int b = rank[0], w = rank[1];  // the current iteration's ranks
    switch (move) {
      case kInit0:
        vf[0] = value;
        move = kInit1;
        pending = vx[1];
        return;
      case kInit1:
        vf[1] = value;
        simplex_iterate();
        return;
      case kReflect:
        f_reflected = value;
        if (value < vf[b]) {
          pending = centroid + 2.0 * (centroid - vx[w]);
          move = kExpand;
        } else if (value < vf[b]) {
          // the second-worst vertex is the best one at n = 1
          vx[w] = reflected;
          vf[w] = value;
          simplex_iterate();
        } else {
          inside = value >= vf[w];
          pending = inside ? centroid - 0.5 * (centroid - vx[w]) : centroid + 0.5 * (centroid - vx[w]);
          move = kContract;
        }
        return;
      case kExpand:
        if (value < f_reflected) {
          vx[w] = pending;
          vf[w] = value;
        } else {
          vx[w] = reflected;
          vf[w] = f_reflected;
        }
        simplex_iterate();
        return;
      case kContract: {
        const bool accept = inside ? value < vf[w] : value <= f_reflected;
        if (accept) {
          vx[w] = pending;
          vf[w] = value;
          simplex_iterate();
        } else {
          // shrink the one other vertex halfway to the best
          vx[w] = vx[b] + 0.5 * (vx[w] - vx[b]);
          pending = vx[w];
          move = kShrink;
        }
        return;
      }
      case kShrink:
        vf[w] = value;
        simplex_iterate();
        return;
      default:
        return;
    }
  }

  // After a simplex value: ask its next point, or move to the check.
  NH_HD void ask_simplex() {
    if (move != kFinished) {
      n_ask = 1;
      ask[0] = clampd(pending, lo, hi);
      return;
    }
    double b = clampd(vx[rank[0]], lo, hi);
    const double fb = vf[rank[0]];
    const double x0 = clampd(start, lo, hi);
    if (!restarted) {
      double moved = b;
      bool any = false;
      const double bounds[2] = {lo, hi};
      for (int i = 0; i < 2; ++i)
        if (b == bounds[i] && x0 != bounds[i] && bounds[i] * x0 > 0.0) {
          moved = sqrt(bounds[i] * x0);
          any = true;
        }
      if (any) {
        restarted = true;
        pinned = true;
        pinned_x = b;
        pinned_f = fb;
        simplex_new(moved);
        ask_simplex();
        return;
      }
    }
    if (pinned) {
      pinned = false;
      if (pinned_f <= fb) b = pinned_x;
    }
    best = b;
    stage = kCheck;
    n_ask = 1;
    ask[0] = best;
  }

  // ---- polish ----
  NH_HD void begin_polish() {
    NH_SS_STRICT
    if (width == kWidths) {
      end_pass(true);
      return;
    }
    const double widths[kWidths] = {1e-2, 3e-3, 1e-3};
    step = fmax(widths[width] * fabs(best), kPolishFloor);
    double nodes[3] = {-1.0, 0.0, 1.0};
    if (best - step < lo) {
      nodes[0] = 0.0;
      nodes[1] = 1.0;
      nodes[2] = 2.0;
    } else if (best + step > hi) {
      nodes[0] = -2.0;
      nodes[1] = -1.0;
      nodes[2] = 0.0;
    }
    for (int i = 0; i < 3; ++i) {
      const double p = clampd(best + nodes[i] * step, lo, hi);
      offs[i] = (p - best) / step;
      ask[i] = best + offs[i] * step;
    }
    n_ask = 3;
    stage = kPolish;
  }

  // Cholesky of an n x n row-major matrix in place, the reference's loop
  // order; false on a pivot that is not strictly positive or not finite.
  NH_HD static bool chol(double* a, int n) {
    NH_SS_STRICT
    for (int i = 0; i < n; ++i)
      for (int j = 0; j <= i; ++j) {
        double sum = a[i * n + j];
        for (int k = 0; k < j; ++k) sum -= a[i * n + k] * a[j * n + k];
        if (i == j) {
          if (sum <= 0.0 || __builtin_isnan(sum)) return false;
          a[i * n + j] = sqrt(sum);
        } else {
          a[i * n + j] = sum / a[j * n + j];
        }
      }
    for (int i = 0; i < n; ++i)
      if (!__builtin_isfinite(a[i * n + i])) return false;
    return true;
  }
  NH_HD static bool chol_solve(double* a, int n, double* x) {
    NH_SS_STRICT
    if (!chol(a, n)) return false;
    for (int i = 0; i < n; ++i) {
      double s = x[i];
      for (int k = 0; k < i; ++k) s -= a[i * n + k] * x[k];
      x[i] = s / a[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = x[i];
      for (int k = i + 1; k < n; ++k) s -= a[k * n + i] * x[k];
      x[i] = s / a[i * n + i];
    }
    for (int i = 0; i < n; ++i)
      if (!__builtin_isfinite(x[i])) return false;
    return true;
  }

  // Least squares quadratic c0 + c1 u + c2 u^2 / 2 through the stencil, then
  // one Newton step in units of the stencil width.
  NH_HD bool polish_step(const double* v, double& next) const {
    NH_SS_STRICT
    double nrm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, rhs[3] = {0, 0, 0};
    for (int pt = 0; pt < 3; ++pt) {
      const double u = offs[pt];
      const double row[3] = {1.0, u, 0.5 * u * u};
      for (int a = 0; a < 3; ++a) {
        rhs[a] += row[a] * v[pt];
        for (int b = 0; b < 3; ++b) nrm[a * 3 + b] += row[a] * row[b];
      }
    }
    if (!chol_solve(nrm, 3, rhs)) return false;
    double h = rhs[2], d = -rhs[1];
    if (!chol_solve(&h, 1, &d)) return false;
    if (!__builtin_isfinite(d) || fabs(d) > kPolishMaxStep) return false;
    next = clampd(best + d * step, lo, hi);
    return true;
  }
};

}  // namespace nh

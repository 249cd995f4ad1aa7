// One gene's stage-two search control: which point to evaluate next and what
// to do with its value.  The same code runs on the host (plain engine, one
// per search) and on lane 0 of a persistent-kernel workgroup, between the
// workgroup's penalised fits; a device test kernel drives it with synthetic
// objectives.
//
// Semantics of edge-rs 0.4.0's StageTwoSearch (src/sc/nebula.rs:1568-1951):
// * both components (NEBULA-HL, and LN's genes that refit both): BOBYQA from
//   the clamped stage-one point; a non-finite value ends the pass, and a pass
//   at a raised Laplace order that ends without a minimiser is repeated from
//   the start at order one; NLopt's RoundoffLimited keeps the point and marks
//   the search failed (nebula's -50);
// * the subject variance alone (LN's weak genes): SubjectSearch, the cell
//   size held at stage one's value.
// Learned start (no counterpart in edge-rs), both components only: a first
// BOBYQA pass from the network's prediction (WARM), then the objective at the
// cold start point (CHECK); the warm result is kept when it is at least as
// good, otherwise the cold search runs as it would have without the network.
//
// Inner fits start from the gene's first finite fit (src/gpu/stage_two.rs:
// 36-43); the warm search moves that start to its best point so far.  tell()
// reports whether the fit just evaluated becomes that start.
#pragma once

#include <cmath>

#include "bobyqa.hpp"
#include "subject_search.hpp"

namespace nh {

enum SearchKind : int { kSearchSubject = 1, kSearchBoth = 2 };
enum SearchPhase : int { PH_COLD = 0, PH_WARM = 1, PH_CHECK = 2 };

struct SearchControl {
  // ---- set by begin ----
  int kind, ord0;
  double lo[2], hi[2];
  double cold[2], warm[2];
  double warm_step, warm_xtol;
  // ---- state ----
  int phase, order;
  int active, found, failed, evals, rejected, from_warm, have_warm;
  int ask_i;
  double anchor_f;            // objective at the inner fits' current start
  double wx[2], wf;           // the warm pass's result
  int wfailed;
  double vals[3];
  Bobyqa<2, 2, true> bq;  // on the device a SearchControl lives in LDS
  SubjectSearch ss;

  NH_HD static double clampd(double v, double a, double b) { return v < a ? a : (v > b ? b : v); }

  // start: stage one's (sigma, gamma); pred: the network's (log sigma^2,
  // log phi), or null for a cold search.
  NH_HD void begin(int search_kind, int ord, const double* start, const float* pred,
                   const double* lower, const double* upper, double step_rel, double xtol_rel) {
    kind = search_kind;
    ord0 = ord;
    order = ord;
    for (int i = 0; i < 2; ++i) {
      lo[i] = lower[i];
      hi[i] = upper[i];
      cold[i] = clampd(start[i], lo[i], hi[i]);
    }
    warm_step = step_rel;
    warm_xtol = xtol_rel;
    have_warm = 0;
    evals = 0;
    found = 0;
    failed = 0;
    rejected = 0;
    from_warm = 0;
    ask_i = 0;
    anchor_f = INFINITY;
    active = 0;
    if (kind == kSearchSubject) {
      // the cell size is held at stage one's value as is
      cold[1] = start[1];
      phase = PH_COLD;
      ss.init(start[0], lo[0], hi[0], ord);
      active = ss.done() ? 0 : 1;
      return;
    }
    if (pred) {
      warm[0] = fmin(fmax(exp((double)pred[0]), lo[0]), hi[0]);
      warm[1] = fmin(fmax(exp(-(double)pred[1]), lo[1]), hi[1]);
      begin_pass(PH_WARM);
    }
    if (!active) begin_pass(PH_COLD);
  }

  // The point to evaluate next and its Laplace order.
  NH_HD void point(double& subject, double& cell) const {
    if (kind == kSearchSubject) {
      subject = ss.ask[ask_i];
      cell = cold[1];
    } else if (phase == PH_CHECK) {
      subject = cold[0];
      cell = cold[1];
    } else {
      subject = bq.xask[0];
      cell = bq.xask[1];
    }
  }
  NH_HD int point_order() const { return kind == kSearchSubject ? ss.order : order; }

  // The value at point(); returns 1 when this fit becomes the inner fits'
  // start.
  NH_HD int tell(double v) {
    evals += 1;
    const int ok = __builtin_isfinite(v) && (!have_warm || (phase == PH_WARM && v < anchor_f));
    if (ok) {
      anchor_f = v;
      have_warm = 1;
    }
    if (kind == kSearchSubject) {
      vals[ask_i++] = v;
      if (ask_i == ss.n_ask) {
        ask_i = 0;
        ss.tell(vals);
        if (ss.done()) {
          active = 0;
          found = ss.found ? 1 : 0;
          failed = 0;
        }
      }
      return ok;
    }
    if (phase == PH_CHECK) {
      // keep the warm result unless the cold start point does better
      if (__builtin_isfinite(v) && wf <= v) {
        active = 0;
        found = 1;
        failed = wfailed;
        from_warm = 1;
      } else {
        fall_back_cold();
      }
      return ok;
    }
    bool pass_over = false, got = false;
    if (!__builtin_isfinite(v)) {
      pass_over = true;
    } else {
      bq.tell(v);
      if (!bq.want) {
        pass_over = true;
        got = true;
      }
    }
    if (!pass_over) return ok;
    if (got || order == 1) {
      if (phase == PH_WARM && got) {
        wx[0] = bq.xbest_out[0];
        wx[1] = bq.xbest_out[1];
        wf = bq.fbest_out;
        wfailed = bq.status == BOBYQA_ROUNDOFF_LIMITED ? 1 : 0;
        phase = PH_CHECK;
      } else if (phase == PH_WARM) {
        fall_back_cold();
      } else {
        active = 0;
        found = got ? 1 : 0;
        failed = bq.status == BOBYQA_ROUNDOFF_LIMITED ? 1 : 0;
      }
    } else {
      order = 1;
      begin_pass(phase);
      if (!active && phase == PH_WARM) {
        rejected = 1;
        order = ord0;
        begin_pass(PH_COLD);
      }
    }
    return ok;
  }

  // The search's outcome, once !active.
  NH_HD void result(double* x, double& f, int& ord_out) const {
    if (kind == kSearchSubject) {
      x[0] = ss.x_out;
      x[1] = cold[1];
      f = NAN;  // the polish moves the incumbent without evaluating it
      ord_out = ss.order;
      return;
    }
    x[0] = from_warm ? wx[0] : bq.xbest_out[0];
    x[1] = from_warm ? wx[1] : bq.xbest_out[1];
    f = from_warm ? wf : bq.fbest_out;
    ord_out = order;
  }

 private:
  NH_HD void fall_back_cold() {
    rejected = 1;
    order = ord0;
    have_warm = 0;
    begin_pass(PH_COLD);
  }

  // A BOBYQA pass of the given phase at the current order.
  NH_HD void begin_pass(int ph) {
    phase = ph;
    if (ph == PH_WARM) {
      // The initial step is a fraction of the predicted value, widened to
      // twice the distance to stage one's estimate where the two disagree
      // (a prediction far from stage one is less certain), at most to a
      // tenth of the predicted value: wider first steps cost more BOBYQA
      // stages than they save where the network is right and stage one is
      // not.  The final precision rhoend = warm_xtol * warm_step * x stays
      // the same whatever the step: xtol is scaled down when it is widened.
      double step[2];
      for (int i = 0; i < 2; ++i) {
        const double base = warm_step * warm[i];
        step[i] = fmax(base, fmin(2.0 * fabs(warm[i] - cold[i]), 0.1 * warm[i]));
        step[i] = fmin(fmax(step[i], 1e-3 * lo[i]), 0.45 * (hi[i] - lo[i]));
      }
      const double xtol = warm_xtol * fmin(1.0, warm_step * warm[0] / step[0]);
      bq.init_with_step(2, warm, lo, hi, step, xtol, 1000);
    } else {
      bq.init(2, cold, lo, hi);
    }
    active = (bq.status != BOBYQA_BAD_BOX && bq.want) ? 1 : 0;
  }
};

}  // namespace nh

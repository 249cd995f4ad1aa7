// Powell's BOBYQA (bound-constrained derivative-free minimisation) with the
// set-up NLopt 2.7.1 applies before it, as a reverse-communication stepper.
//
// Follows src/numeric/bobyqa.rs of edge-rs 0.4.0 (itself NLopt's
// src/algs/bobyqa/bobyqa.c): default initial step (bobyqa.rs:344-372), the
// rescaling to equal steps (:374-386), rhobeg from the step and
// rhoend = xtol_rel * rhobeg (:401-402), the box adjustment of x0 (:407-438),
// prelim (:636-767), bobyqb labels 20..720 (:785-1538), rescue (:1540-1932),
// trsbox (:2007), altmov (:2438) and update (:2722).  For identical objective
// values it asks for the same points in the same order as nloptr::bobyqa; the
// nloptr traces pin that bit for bit.
//
// The same code runs on the host (plain search) and on the device (one lane of
// a workgroup drives a gene's whole stage-two search inside the persistent
// kernel).  All state lives in fixed-size members, so a stepper is a plain
// struct that can sit in registers/scratch or global memory.  Floating-point
// contraction is switched off below: an FMA where the C has a multiply and an
// add would change the trace.
//
// Control flow is a resumable routine: `advance()` runs until the next point
// is needed (`want == true`, point in `xask`) or the search ends.  Labels keep
// Powell's numbers so each block can be compared with the Fortran/C.
#pragma once

#ifndef NH_HD
#if defined(__HIPCC__)
#define NH_HD __host__ __device__
#else
#define NH_HD
#endif
#endif

#include <cmath>

namespace nh {

enum BobyqaStatus : int {
  BOBYQA_RUNNING = 0,
  BOBYQA_SUCCESS = 1,          // NLopt SUCCESS
  BOBYQA_XTOL_REACHED = 4,     // NLopt XTOL_REACHED
  BOBYQA_MAXEVAL_REACHED = 5,  // NLopt MAXEVAL_REACHED
  BOBYQA_ROUNDOFF_LIMITED = -4,
  BOBYQA_BAD_BOX = -2,         // NLopt refuses the box (bobyqa.rs:413-419)
};

// Every routine below opens with NH_STRICT_FP: no a*b+c is fused, on the host
// or the device, so the arithmetic is the C's operation for operation.
#if defined(__clang__)
#define NH_STRICT_FP _Pragma("clang fp contract(off)")
#else
#define NH_STRICT_FP
#endif

// Local n, npt and ndim, hiding the members of the same names: compile-time
// constants when the dimension is fixed.  On the device, an IN_LDS object's
// address is declared to be in workgroup memory.
#if defined(__HIP_DEVICE_COMPILE__)
#define NH_BQ_WHERE \
  if constexpr (IN_LDS) __builtin_assume(__builtin_amdgcn_is_shared((const void*)this));
#else
#define NH_BQ_WHERE
#endif
#define NH_BQ_DIMS                                                   \
  NH_BQ_WHERE                                                        \
  [[maybe_unused]] const int n = kFix ? NFIX : this->n;              \
  [[maybe_unused]] const int npt = kFix ? 2 * NFIX + 1 : this->npt;  \
  [[maybe_unused]] const int ndim = kFix ? 3 * NFIX + 1 : this->ndim;

NH_HD inline bool nh_finite(double v) { return __builtin_isfinite(v); }
NH_HD inline bool nh_inf(double v) { return __builtin_isinf(v); }

// NFIX > 0 fixes the dimension at compile time (n = NFIX, npt = 2n + 1,
// ndim = npt + n; init refuses any other n): every loop over the coordinates
// and the interpolation points then has a constant trip count, so the device
// compiler unrolls it and addresses Powell's arrays at constant offsets
// instead of running loop counters and 64-bit address arithmetic on one lane.
// The arithmetic and its order are the same as with a run-time n.
//
// IN_LDS: on the device the object lives in workgroup memory (LDS), so its
// members are read and written with ds_ instructions at 32-bit addresses
// rather than through flat (generic) pointers.
template <int NMAX, int NFIX = 0, bool IN_LDS = false>
struct Bobyqa {
  static_assert(NFIX == 0 || NFIX == NMAX, "a fixed dimension is NMAX");
  static constexpr bool kFix = NFIX > 0;
  static constexpr int NPTMAX = 2 * NMAX + 1;
  static constexpr int NDIMMAX = NPTMAX + NMAX;
  static constexpr int NHMAX = NMAX * (NMAX + 1) / 2;
  static constexpr int WMAX = 3 * NDIMMAX + NMAX + NPTMAX;

  // ---- public reverse-communication interface ----
  bool want;             // a point is waiting for its value
  double xask[NMAX];     // that point, on the caller's scale
  int status;            // BobyqaStatus
  int nevals;            // evaluations requested so far
  double xbest_out[NMAX];
  double fbest_out;

  // ---- problem ----
  int n, npt, ndim, maxeval;
  double s[NMAX + 1];    // NLopt rescaling, one-based
  double xl[NMAX + 1], xu[NMAX + 1];
  double rhobeg, rhoend;

  // ---- Powell's arrays (one-based; matrices column-major with helpers) ----
  double x[NMAX + 1];
  double xbase[NMAX + 1], xopt[NMAX + 1], gopt[NMAX + 1], sl[NMAX + 1],
      su[NMAX + 1], xnew[NMAX + 1], xalt[NMAX + 1], d[NMAX + 1];
  double xpt[NPTMAX * NMAX];
  double fval[NPTMAX + 1];
  double hq[NHMAX + 1];
  double pq[NPTMAX + 1];
  double bmat[NDIMMAX * NMAX];
  double zmat[NPTMAX * (NPTMAX - NMAX - 1) + 1];
  double vlag[NDIMMAX + 1];
  double w[WMAX + 1];

  // ---- bobyqb scalars kept across pauses ----
  int kopt, knew, kbase, nresc, nfsav, itest;
  long ntrits;
  double diffa, diffb, diffc, rho, delta, xoptsq, dsq, crvmin, dnorm, adelt,
      distsq, alpha, cauchy, beta, denom, vquad, ratio, fopt, fsave;
  double f;  // last value told
  int resume;

  // prelim
  int nf, nfm, nfx, ipt, jpt;
  double stepa, stepb, fbeg;
  // rescue
  int kpt, rip, riq;
  double rvquad, fbase, rxp, rxq;

  NH_HD int ix(int k, int j) const {
    NH_BQ_DIMS
    return (k - 1) + (j - 1) * npt;
  }
  NH_HD int ib(int k, int j) const {
    NH_BQ_DIMS
    return (k - 1) + (j - 1) * ndim;
  }
  NH_HD int iz(int k, int j) const {
    NH_BQ_DIMS
    return (k - 1) + (j - 1) * npt;
  }

  // ---------------------------------------------------------------------
  // Set-up: nlopt_set_default_initial_step, nlopt_compute_rescaling and the
  // head of bobyqa() (bobyqa.rs:318-514).  Leaves the first point asked for.
  // ---------------------------------------------------------------------
  NH_HD void init(int n_, const double* x0, const double* lower,
                  const double* upper, double xtol_rel = 1e-6,
                  int maxeval_ = 1000) {
    NH_STRICT_FP
    double dx[NMAX];
    for (int i = 0; i < n_; ++i) {
      double lb = lower[i], ub = upper[i], xi = x0[i];
      double step = INFINITY;
      if (nh_finite(ub) && nh_finite(lb) && (ub - lb) * 0.25 < step && ub > lb)
        step = (ub - lb) * 0.25;
      if (nh_finite(ub) && ub - xi < step && ub > xi) step = (ub - xi) * 0.75;
      if (nh_finite(lb) && xi - lb < step && xi > lb) step = (xi - lb) * 0.75;
      if (nh_inf(step)) {
        if (nh_finite(ub) && std::fabs(ub - xi) < std::fabs(step)) step = (ub - xi) * 1.1;
        if (nh_finite(lb) && std::fabs(xi - lb) < std::fabs(step)) step = (xi - lb) * 1.1;
      }
      if (nh_inf(step) || step == 0.0 || (step != 0.0 && std::fabs(step) < 2.2250738585072014e-308))
        step = xi;
      if (nh_inf(step) || step == 0.0) step = 1.0;
      dx[i] = step;
    }
    init_with_step(n_, x0, lower, upper, dx, xtol_rel, maxeval_);
  }

  // The same set-up with the initial step given (NLopt's
  // nlopt_set_initial_step instead of its default); rhoend = xtol_rel * rhobeg.
  NH_HD void init_with_step(int n_, const double* x0, const double* lower,
                            const double* upper, const double* dx, double xtol_rel,
                            int maxeval_) {
    NH_STRICT_FP
    n = n_;
    npt = 2 * n + 1;
    ndim = npt + n;
    maxeval = maxeval_;
    nevals = 0;
    want = false;
    status = BOBYQA_RUNNING;
    fbest_out = INFINITY;
    if (kFix && n_ != NFIX) {
      status = BOBYQA_BAD_BOX;
      return;
    }
    for (int i = 0; i <= n; ++i) s[i] = 1.0;
    bool differ = false;
    for (int i = 1; i < n; ++i) differ = differ || (dx[i] != dx[i - 1]);
    if (n > 1 && differ)
      for (int i = 1; i < n; ++i) s[i + 1] = dx[i] / dx[0];
    for (int i = 1; i <= n; ++i) {
      if (s[i] == 0.0 || !nh_finite(s[i])) {
        status = BOBYQA_BAD_BOX;
        return;
      }
    }
    for (int j = 1; j <= n; ++j) {
      x[j] = x0[j - 1] / s[j];
      xl[j] = lower[j - 1] / s[j];
      xu[j] = upper[j - 1] / s[j];
      if (xl[j] > xu[j]) {
        double t = xl[j];
        xl[j] = xu[j];
        xu[j] = t;
      }
    }
    rhobeg = std::fabs(dx[0] / s[1]);
    rhoend = xtol_rel * rhobeg;
    for (int j = 1; j <= n; ++j) {
      double temp = xu[j] - xl[j];
      if (temp < rhobeg + rhobeg) {
        status = BOBYQA_BAD_BOX;
        return;
      }
      sl[j] = xl[j] - x[j];
      su[j] = xu[j] - x[j];
      if (sl[j] >= -rhobeg) {
        if (sl[j] >= 0.0) {
          x[j] = xl[j];
          sl[j] = 0.0;
          su[j] = temp;
        } else {
          x[j] = xl[j] + rhobeg;
          sl[j] = -rhobeg;
          double a = xu[j] - x[j];
          su[j] = a > rhobeg ? a : rhobeg;
        }
      } else if (su[j] <= rhobeg) {
        if (su[j] <= 0.0) {
          x[j] = xu[j];
          sl[j] = -temp;
          su[j] = 0.0;
        } else {
          x[j] = xu[j] - rhobeg;
          double a = xl[j] - x[j];
          sl[j] = a < -rhobeg ? a : -rhobeg;
          su[j] = rhobeg;
        }
      }
    }
    for (int i = 0; i < NPTMAX * NMAX; ++i) xpt[i] = 0.0;
    for (int i = 0; i < NDIMMAX * NMAX; ++i) bmat[i] = 0.0;
    for (int i = 0; i < NPTMAX * (NPTMAX - NMAX - 1) + 1; ++i) zmat[i] = 0.0;
    for (int i = 0; i <= NHMAX; ++i) hq[i] = 0.0;
    for (int i = 0; i <= NPTMAX; ++i) { pq[i] = 0.0; fval[i] = 0.0; }
    for (int i = 0; i <= NDIMMAX; ++i) vlag[i] = 0.0;
    for (int i = 0; i <= WMAX; ++i) w[i] = 0.0;
    for (int i = 0; i <= NMAX; ++i) {
      xbase[i] = xopt[i] = gopt[i] = xnew[i] = xalt[i] = d[i] = 0.0;
    }
    for (int j = 1; j <= n; ++j) xbase[j] = x[j];
    kopt = 1; knew = 0; kbase = 1; nresc = 0; nfsav = 0; itest = 0; ntrits = 0;
    diffa = diffb = diffc = 0.0;
    rho = delta = rhobeg;
    xoptsq = dsq = crvmin = dnorm = adelt = distsq = alpha = cauchy = beta = 0.0;
    denom = vquad = ratio = fopt = fsave = f = 0.0;
    nf = 0; nfm = 0; nfx = 0; ipt = jpt = 0;
    stepa = stepb = fbeg = 0.0;
    kpt = rip = riq = 0;
    rvquad = fbase = rxp = rxq = 0.0;
    resume = 0;
    advance();
  }

  NH_HD bool done() const { return !want; }

  // Value at the point last asked for.
  NH_HD void tell(double value) {
    NH_STRICT_FP
    f = value;
    want = false;
    advance();
  }

 private:
  NH_HD bool stop_evals() const { return maxeval > 0 && nevals >= maxeval; }

  NH_HD void ask_point(int where) {
    NH_STRICT_FP
    NH_BQ_DIMS
    for (int j = 1; j <= n; ++j) xask[j - 1] = x[j] * s[j];
    ++nevals;
    resume = where;
    want = true;
  }

  NH_HD void finish_at_best() {
    NH_STRICT_FP
    NH_BQ_DIMS
    for (int i = 1; i <= n; ++i) xbest_out[i - 1] = x[i] * s[i];
  }

  enum : int { R_START = 0, R_PRELIM = 1, R_360 = 2, R_RESCUE = 3, R_DONE = 4 };

 public:
  // The resumable body of prelim + bobyqb + rescue.
  NH_HD void advance() {
    NH_STRICT_FP
    NH_BQ_DIMS
    switch (resume) {
      case R_START: goto L_prelim_next;
      case R_PRELIM: goto L_prelim_post;
      case R_360: goto L_360_post;
      case R_RESCUE: goto L_rescue_post;
      default: return;
    }

    // ===================== prelim =====================
  L_prelim_next:
    nfm = nf;
    nfx = nf - n;
    nf = nf + 1;
    if (nfm <= 2 * n) {
      if (nfm >= 1 && nfm <= n) {
        stepa = rhobeg;
        if (su[nfm] == 0.0) stepa = -stepa;
        xpt[ix(nf, nfm)] = stepa;
      } else if (nfm > n) {
        stepa = xpt[ix(nf - n, nfx)];
        stepb = -rhobeg;
        if (sl[nfx] == 0.0) stepb = fmin(2.0 * rhobeg, su[nfx]);
        if (su[nfx] == 0.0) stepb = fmax(-2.0 * rhobeg, sl[nfx]);
        xpt[ix(nf, nfx)] = stepb;
      }
    } else {
      int itemp = (nfm - (n + 1)) / n;
      int jj = nfm - itemp * n - n;
      int ii = jj + itemp;
      if (ii > n) {
        itemp = jj;
        jj = ii - n;
        ii = itemp;
      }
      ipt = ii;
      jpt = jj;
      xpt[ix(nf, ipt)] = xpt[ix(ipt + 1, ipt)];
      xpt[ix(nf, jpt)] = xpt[ix(jpt + 1, jpt)];
    }
    for (int j = 1; j <= n; ++j) {
      double p = xpt[ix(nf, j)];
      x[j] = fmin(fmax(xl[j], xbase[j] + p), xu[j]);
      if (p == sl[j]) x[j] = xl[j];
      if (p == su[j]) x[j] = xu[j];
    }
    ask_point(R_PRELIM);
    return;

  L_prelim_post: {
    const double rhosq = rhobeg * rhobeg;
    const double recip = 1.0 / rhosq;
    fval[nf] = f;
    if (nf == 1) {
      fbeg = f;
      kopt = 1;
    } else if (f < fval[kopt]) {
      kopt = nf;
    }
    if (nf <= 2 * n + 1) {
      if (nf >= 2 && nf <= n + 1) {
        gopt[nfm] = (f - fbeg) / stepa;
        if (npt < nf + n) {
          bmat[ib(1, nfm)] = -1.0 / stepa;
          bmat[ib(nf, nfm)] = 1.0 / stepa;
          bmat[ib(npt + nfm, nfm)] = -0.5 * rhosq;
        }
      } else if (nf >= n + 2) {
        int ih = nfx * (nfx + 1) / 2;
        double temp = (f - fbeg) / stepb;
        double diff = stepb - stepa;
        hq[ih] = 2.0 * (temp - gopt[nfx]) / diff;
        gopt[nfx] = (gopt[nfx] * stepb - temp * stepa) / diff;
        if (stepa * stepb < 0.0 && f < fval[nf - n]) {
          fval[nf] = fval[nf - n];
          fval[nf - n] = f;
          if (kopt == nf) kopt = nf - n;
          xpt[ix(nf - n, nfx)] = stepb;
          xpt[ix(nf, nfx)] = stepa;
        }
        bmat[ib(1, nfx)] = -(stepa + stepb) / (stepa * stepb);
        bmat[ib(nf, nfx)] = -0.5 / xpt[ix(nf - n, nfx)];
        bmat[ib(nf - n, nfx)] = -bmat[ib(1, nfx)] - bmat[ib(nf, nfx)];
        zmat[iz(1, nfx)] = sqrt(2.0) / (stepa * stepb);
        zmat[iz(nf, nfx)] = sqrt(0.5) / rhosq;
        zmat[iz(nf - n, nfx)] = -zmat[iz(1, nfx)] - zmat[iz(nf, nfx)];
      }
    } else {
      int ih = ipt * (ipt - 1) / 2 + jpt;
      zmat[iz(1, nfx)] = recip;
      zmat[iz(nf, nfx)] = recip;
      zmat[iz(ipt + 1, nfx)] = -recip;
      zmat[iz(jpt + 1, nfx)] = -recip;
      double temp = xpt[ix(nf, ipt)] * xpt[ix(nf, jpt)];
      hq[ih] = (fbeg - fval[ipt + 1] - fval[jpt + 1] + f) / temp;
    }
  }
    if (stop_evals()) {
      status = BOBYQA_MAXEVAL_REACHED;
      goto L_after_prelim;
    }
    if (nf < npt) goto L_prelim_next;

  L_after_prelim:
    xoptsq = 0.0;
    for (int i = 1; i <= n; ++i) {
      xopt[i] = xpt[ix(kopt, i)];
      xoptsq += xopt[i] * xopt[i];
    }
    fsave = fval[1];
    if (status != BOBYQA_RUNNING) goto L720;

    // ===================== bobyqb =====================
    kbase = 1;
    rho = rhobeg;
    delta = rho;
    nresc = nevals;
    ntrits = 0;
    diffa = 0.0;
    diffb = 0.0;
    itest = 0;
    nfsav = nevals;

  L20:
    if (kopt != kbase) {
      int ih = 0;
      for (int j = 1; j <= n; ++j) {
        for (int i = 1; i <= j; ++i) {
          ++ih;
          if (i < j) gopt[j] += hq[ih] * xopt[i];
          gopt[i] += hq[ih] * xopt[j];
        }
      }
      if (nevals > npt) {
        for (int k = 1; k <= npt; ++k) {
          double temp = 0.0;
          for (int j = 1; j <= n; ++j) temp += xpt[ix(k, j)] * xopt[j];
          temp *= pq[k];
          for (int i = 1; i <= n; ++i) gopt[i] += temp * xpt[ix(k, i)];
        }
      }
    }

  L60:
    trsbox_step();
    dnorm = fmin(delta, sqrt(dsq));
    if (dnorm < 0.5 * rho) {
      ntrits = -1;
      distsq = (10.0 * rho) * (10.0 * rho);
      if (nevals <= nfsav + 2) goto L650;
      {
        double errbig = fmax(fmax(diffa, diffb), diffc);
        double frhosq = rho * 0.125 * rho;
        if (crvmin > 0.0 && errbig > frhosq * crvmin) goto L650;
        double bdtol = errbig / rho;
        for (int j = 1; j <= n; ++j) {
          double bdtest = bdtol;
          if (xnew[j] == sl[j]) bdtest = w[j];
          if (xnew[j] == su[j]) bdtest = -w[j];
          if (bdtest < bdtol) {
            double curv = hq[(j + j * j) / 2];
            for (int k = 1; k <= npt; ++k) {
              double v = xpt[ix(k, j)];
              curv += pq[k] * (v * v);
            }
            bdtest += 0.5 * curv * rho;
            if (bdtest < bdtol) goto L650;
          }
        }
      }
      goto L680;
    }
    ++ntrits;

  L90:
    if (dsq <= xoptsq * 0.001) shift_base();
    if (ntrits == 0) goto L210;
    goto L230;

  L190:
    nfsav = nevals;
    kbase = kopt;
    if (!rescue_prepare()) goto L_rescue_end;
    kpt = 0;
  L_rescue_next:
    ++kpt;
    if (kpt > npt) goto L_rescue_end;
    if (w[pid(kpt)] == 0.0) goto L_rescue_next;
    if (stop_evals()) {
      status = BOBYQA_MAXEVAL_REACHED;
      goto L_rescue_end;
    }
    rescue_point();
    ask_point(R_RESCUE);
    return;
  L_rescue_post:
    fval[kpt] = f;
    if (f < fval[kopt]) kopt = kpt;
    if (stop_evals()) {
      status = BOBYQA_MAXEVAL_REACHED;
      goto L_rescue_end;
    }
    rescue_update();
    goto L_rescue_next;
  L_rescue_end:
    xoptsq = 0.0;
    if (kopt != kbase) {
      for (int i = 1; i <= n; ++i) {
        xopt[i] = xpt[ix(kopt, i)];
        xoptsq += xopt[i] * xopt[i];
      }
    }
    if (status != BOBYQA_RUNNING) goto L720;
    nresc = nevals;
    if (nfsav < nevals) {
      nfsav = nevals;
      goto L20;
    }
    if (ntrits > 0) goto L60;

  L210:
    altmov_step();
    for (int i = 1; i <= n; ++i) d[i] = xnew[i] - xopt[i];

  L230: {
    const int nptm = npt - n - 1;
    for (int k = 1; k <= npt; ++k) {
      double suma = 0.0, sumb = 0.0, sum = 0.0;
      for (int j = 1; j <= n; ++j) {
        suma += xpt[ix(k, j)] * d[j];
        sumb += xpt[ix(k, j)] * xopt[j];
        sum += bmat[ib(k, j)] * d[j];
      }
      w[k] = suma * (0.5 * suma + sumb);
      vlag[k] = sum;
      w[npt + k] = suma;
    }
    beta = 0.0;
    for (int jj = 1; jj <= nptm; ++jj) {
      double sum = 0.0;
      for (int k = 1; k <= npt; ++k) sum += zmat[iz(k, jj)] * w[k];
      beta -= sum * sum;
      for (int k = 1; k <= npt; ++k) vlag[k] += sum * zmat[iz(k, jj)];
    }
    dsq = 0.0;
    double bsum = 0.0, dx = 0.0;
    for (int j = 1; j <= n; ++j) {
      dsq += d[j] * d[j];
      double sum = 0.0;
      for (int k = 1; k <= npt; ++k) sum += w[k] * bmat[ib(k, j)];
      bsum += sum * d[j];
      int jp = npt + j;
      for (int i = 1; i <= n; ++i) sum += bmat[ib(jp, i)] * d[i];
      vlag[jp] = sum;
      bsum += sum * d[j];
      dx += d[j] * xopt[j];
    }
    beta = dx * dx + dsq * (xoptsq + dx + dx + 0.5 * dsq) + beta - bsum;
    vlag[kopt] += 1.0;

    if (ntrits == 0) {
      double vk = vlag[knew];
      denom = vk * vk + alpha * beta;
      if (denom < cauchy && cauchy > 0.0) {
        for (int i = 1; i <= n; ++i) {
          xnew[i] = xalt[i];
          d[i] = xnew[i] - xopt[i];
        }
        cauchy = 0.0;
        goto L230;
      }
      if (denom <= 0.5 * (vk * vk)) {
        if (nevals > nresc) goto L190;
        status = BOBYQA_ROUNDOFF_LIMITED;
        goto L720;
      }
    } else {
      double delsq = delta * delta;
      double scaden = 0.0, biglsq = 0.0;
      knew = 0;
      for (int k = 1; k <= npt; ++k) {
        if (k == kopt) continue;
        double hdiag = 0.0;
        for (int jj = 1; jj <= nptm; ++jj) {
          double z = zmat[iz(k, jj)];
          hdiag += z * z;
        }
        double den = beta * hdiag + vlag[k] * vlag[k];
        double dist = 0.0;
        for (int j = 1; j <= n; ++j) {
          double v = xpt[ix(k, j)] - xopt[j];
          dist += v * v;
        }
        double r = dist / delsq;
        double temp = fmax(1.0, r * r);
        if (temp * den > scaden) {
          scaden = temp * den;
          knew = k;
          denom = den;
        }
        biglsq = fmax(biglsq, temp * (vlag[k] * vlag[k]));
      }
      if (scaden <= 0.5 * biglsq) {
        if (nevals > nresc) goto L190;
        status = BOBYQA_ROUNDOFF_LIMITED;
        goto L720;
      }
    }
  }

  L360:
    for (int i = 1; i <= n; ++i) {
      x[i] = fmin(fmax(xl[i], xbase[i] + xnew[i]), xu[i]);
      if (xnew[i] == sl[i]) x[i] = xl[i];
      if (xnew[i] == su[i]) x[i] = xu[i];
    }
    if (stop_evals()) {
      status = BOBYQA_MAXEVAL_REACHED;
      goto L720;
    }
    ask_point(R_360);
    return;

  L_360_post:
    if (ntrits == -1) {
      fsave = f;
      status = BOBYQA_XTOL_REACHED;
      if (fsave < fval[kopt]) {
        fbest_out = f;
        for (int i = 1; i <= n; ++i) xbest_out[i - 1] = x[i] * s[i];
        resume = R_DONE;
        return;
      }
      goto L720;
    }
    if (!model_update()) goto L720;
    if (ntrits == 0) goto L60;
    if (f <= fopt + 0.1 * vquad) goto L60;
    {
      double a = 2.0 * delta, b = 10.0 * rho;
      distsq = fmax(a * a, b * b);
    }

  L650:
    knew = 0;
    for (int k = 1; k <= npt; ++k) {
      double sum = 0.0;
      for (int j = 1; j <= n; ++j) {
        double v = xpt[ix(k, j)] - xopt[j];
        sum += v * v;
      }
      if (sum > distsq) {
        knew = k;
        distsq = sum;
      }
    }
    if (knew > 0) {
      double dist = sqrt(distsq);
      if (ntrits == -1) {
        delta = fmin(0.1 * delta, 0.5 * dist);
        if (delta <= rho * 1.5) delta = rho;
      }
      ntrits = 0;
      adelt = fmax(fmin(0.1 * dist, delta), rho);
      dsq = adelt * adelt;
      goto L90;
    }
    if (ntrits == -1) goto L680;
    if (ratio > 0.0) goto L60;
    if (fmax(delta, dnorm) > rho) goto L60;

  L680:
    if (rho > rhoend) {
      delta = 0.5 * rho;
      ratio = rho / rhoend;
      if (ratio <= 16.0)
        rho = rhoend;
      else if (ratio <= 250.0)
        rho = sqrt(ratio) * rhoend;
      else
        rho *= 0.1;
      delta = fmax(delta, rho);
      ntrits = 0;
      nfsav = nevals;
      goto L60;
    }
    if (ntrits == -1) goto L360;

  L720:
    for (int i = 1; i <= n; ++i) {
      x[i] = fmin(fmax(xl[i], xbase[i] + xopt[i]), xu[i]);
      if (xopt[i] == sl[i]) x[i] = xl[i];
      if (xopt[i] == su[i]) x[i] = xu[i];
    }
    fbest_out = fval[kopt];
    for (int i = 1; i <= n; ++i) xbest_out[i - 1] = x[i] * s[i];
    if (status == BOBYQA_RUNNING) status = BOBYQA_SUCCESS;
    resume = R_DONE;
    return;
  }

 private:
  // Label 90 body: move xbase to xopt (bobyqa.rs:904-980).
  NH_HD void shift_base() {
    NH_STRICT_FP
    NH_BQ_DIMS
    const int nptm = npt - n - 1;
    double fracsq = xoptsq * 0.25;
    double sumpq = 0.0;
    for (int k = 1; k <= npt; ++k) {
      sumpq += pq[k];
      double sum = -0.5 * xoptsq;
      for (int i = 1; i <= n; ++i) sum += xpt[ix(k, i)] * xopt[i];
      w[npt + k] = sum;
      double temp = fracsq - 0.5 * sum;
      for (int i = 1; i <= n; ++i) {
        w[i] = bmat[ib(k, i)];
        vlag[i] = sum * xpt[ix(k, i)] + temp * xopt[i];
        int ip = npt + i;
        for (int j = 1; j <= i; ++j)
          bmat[ib(ip, j)] = bmat[ib(ip, j)] + w[i] * vlag[j] + vlag[i] * w[j];
      }
    }
    for (int jj = 1; jj <= nptm; ++jj) {
      double sumz = 0.0, sumw = 0.0;
      for (int k = 1; k <= npt; ++k) {
        double z = zmat[iz(k, jj)];
        sumz += z;
        vlag[k] = w[npt + k] * z;
        sumw += vlag[k];
      }
      for (int j = 1; j <= n; ++j) {
        double sum = (fracsq * sumz - 0.5 * sumw) * xopt[j];
        for (int k = 1; k <= npt; ++k) sum += vlag[k] * xpt[ix(k, j)];
        w[j] = sum;
        for (int k = 1; k <= npt; ++k) bmat[ib(k, j)] += sum * zmat[iz(k, jj)];
      }
      for (int i = 1; i <= n; ++i) {
        int ip = i + npt;
        double temp = w[i];
        for (int j = 1; j <= i; ++j) bmat[ib(ip, j)] += temp * w[j];
      }
    }
    int ih = 0;
    for (int j = 1; j <= n; ++j) {
      w[j] = -0.5 * sumpq * xopt[j];
      for (int k = 1; k <= npt; ++k) {
        w[j] += pq[k] * xpt[ix(k, j)];
        xpt[ix(k, j)] -= xopt[j];
      }
      for (int i = 1; i <= j; ++i) {
        ++ih;
        hq[ih] = hq[ih] + w[i] * xopt[j] + xopt[i] * w[j];
        bmat[ib(npt + i, j)] = bmat[ib(npt + j, i)];
      }
    }
    for (int i = 1; i <= n; ++i) {
      xbase[i] += xopt[i];
      xnew[i] -= xopt[i];
      sl[i] -= xopt[i];
      su[i] -= xopt[i];
      xopt[i] = 0.0;
    }
    xoptsq = 0.0;
  }

  // After the value at xnew: model error, radius, knew, and the update of the
  // quadratic model (bobyqa.rs:1201-1431).  Returns false when the search must
  // stop (status set).
  NH_HD bool model_update() {
    NH_STRICT_FP
    NH_BQ_DIMS
    const int nptm = npt - n - 1;
    const int nh = n * (n + 1) / 2;
    fopt = fval[kopt];
    double vq = 0.0;
    int ih = 0;
    for (int j = 1; j <= n; ++j) {
      vq += d[j] * gopt[j];
      for (int i = 1; i <= j; ++i) {
        ++ih;
        double temp = d[i] * d[j];
        if (i == j) temp *= 0.5;
        vq += hq[ih] * temp;
      }
    }
    for (int k = 1; k <= npt; ++k) {
      double v = w[npt + k];
      vq += 0.5 * pq[k] * (v * v);
    }
    vquad = vq;
    double diff = f - fopt - vquad;
    diffc = diffb;
    diffb = diffa;
    diffa = fabs(diff);
    if (dnorm > rho) nfsav = nevals;

    if (ntrits > 0) {
      if (vquad >= 0.0) {
        status = BOBYQA_ROUNDOFF_LIMITED;
        return false;
      }
      ratio = (f - fopt) / vquad;
      if (ratio <= 0.1)
        delta = fmin(0.5 * delta, dnorm);
      else if (ratio <= 0.7)
        delta = fmax(0.5 * delta, dnorm);
      else
        delta = fmax(0.5 * delta, dnorm + dnorm);
      if (delta <= rho * 1.5) delta = rho;

      if (f < fopt) {
        int ksav = knew;
        double densav = denom;
        double delsq = delta * delta;
        double scaden = 0.0, biglsq = 0.0;
        knew = 0;
        for (int k = 1; k <= npt; ++k) {
          double hdiag = 0.0;
          for (int jj = 1; jj <= nptm; ++jj) {
            double z = zmat[iz(k, jj)];
            hdiag += z * z;
          }
          double den = beta * hdiag + vlag[k] * vlag[k];
          double dist = 0.0;
          for (int j = 1; j <= n; ++j) {
            double v = xpt[ix(k, j)] - xnew[j];
            dist += v * v;
          }
          double r = dist / delsq;
          double temp = fmax(1.0, r * r);
          if (temp * den > scaden) {
            scaden = temp * den;
            knew = k;
            denom = den;
          }
          biglsq = fmax(biglsq, temp * (vlag[k] * vlag[k]));
        }
        if (scaden <= 0.5 * biglsq) {
          knew = ksav;
          denom = densav;
        }
      }
    }

    update_hz(knew);
    ih = 0;
    double pqold = pq[knew];
    pq[knew] = 0.0;
    for (int i = 1; i <= n; ++i) {
      double temp = pqold * xpt[ix(knew, i)];
      for (int j = 1; j <= i; ++j) {
        ++ih;
        hq[ih] += temp * xpt[ix(knew, j)];
      }
    }
    for (int jj = 1; jj <= nptm; ++jj) {
      double temp = diff * zmat[iz(knew, jj)];
      for (int k = 1; k <= npt; ++k) pq[k] += temp * zmat[iz(k, jj)];
    }
    fval[knew] = f;
    for (int i = 1; i <= n; ++i) {
      xpt[ix(knew, i)] = xnew[i];
      w[i] = bmat[ib(knew, i)];
    }
    for (int k = 1; k <= npt; ++k) {
      double suma = 0.0;
      for (int jj = 1; jj <= nptm; ++jj) suma += zmat[iz(knew, jj)] * zmat[iz(k, jj)];
      if (nh_inf(suma)) {
        status = BOBYQA_ROUNDOFF_LIMITED;
        return false;
      }
      double sumb = 0.0;
      for (int j = 1; j <= n; ++j) sumb += xpt[ix(k, j)] * xopt[j];
      double temp = suma * sumb;
      for (int i = 1; i <= n; ++i) w[i] += temp * xpt[ix(k, i)];
    }
    for (int i = 1; i <= n; ++i) gopt[i] += diff * w[i];

    if (f < fopt) {
      kopt = knew;
      xoptsq = 0.0;
      ih = 0;
      for (int j = 1; j <= n; ++j) {
        xopt[j] = xnew[j];
        xoptsq += xopt[j] * xopt[j];
        for (int i = 1; i <= j; ++i) {
          ++ih;
          if (i < j) gopt[j] += hq[ih] * d[i];
          gopt[i] += hq[ih] * d[j];
        }
      }
      for (int k = 1; k <= npt; ++k) {
        double temp = 0.0;
        for (int j = 1; j <= n; ++j) temp += xpt[ix(k, j)] * d[j];
        temp *= pq[k];
        for (int i = 1; i <= n; ++i) gopt[i] += temp * xpt[ix(k, i)];
      }
    }

    if (ntrits > 0) {
      for (int k = 1; k <= npt; ++k) {
        vlag[k] = fval[k] - fval[kopt];
        w[k] = 0.0;
      }
      for (int j = 1; j <= nptm; ++j) {
        double sum = 0.0;
        for (int k = 1; k <= npt; ++k) sum += zmat[iz(k, j)] * vlag[k];
        for (int k = 1; k <= npt; ++k) w[k] += sum * zmat[iz(k, j)];
      }
      for (int k = 1; k <= npt; ++k) {
        double sum = 0.0;
        for (int j = 1; j <= n; ++j) sum += xpt[ix(k, j)] * xopt[j];
        w[k + npt] = w[k];
        w[k] *= sum;
      }
      double gqsq = 0.0, gisq = 0.0;
      for (int i = 1; i <= n; ++i) {
        double sum = 0.0;
        for (int k = 1; k <= npt; ++k)
          sum = sum + bmat[ib(k, i)] * vlag[k] + xpt[ix(k, i)] * w[k];
        if (xopt[i] == sl[i]) {
          double a = fmin(0.0, gopt[i]);
          gqsq += a * a;
          double b = fmin(0.0, sum);
          gisq += b * b;
        } else if (xopt[i] == su[i]) {
          double a = fmax(0.0, gopt[i]);
          gqsq += a * a;
          double b = fmax(0.0, sum);
          gisq += b * b;
        } else {
          gqsq += gopt[i] * gopt[i];
          gisq += sum * sum;
        }
        vlag[npt + i] = sum;
      }
      ++itest;
      if (gqsq < 10.0 * gisq) itest = 0;
      if (itest >= 3) {
        int top = npt > nh ? npt : nh;
        for (int i = 1; i <= top; ++i) {
          if (i <= n) gopt[i] = vlag[npt + i];
          if (i <= npt) pq[i] = w[npt + i];
          if (i <= nh) hq[i] = 0.0;
          itest = 0;
        }
      }
    }
    return true;
  }

  // ---- rescue (bobyqa.rs:1540-1932), split around its evaluations ----
  NH_HD int pa(int r, int j) const { return 1 + (r - 1) + 2 * (j - 1); }
  NH_HD int pid(int k) const {
    NH_BQ_DIMS
    return 2 * n + k;
  }
  NH_HD int rw(int i) const {
    NH_BQ_DIMS
    return ndim + n + i;
  }
  NH_HD void pid_parts(int k, int& ip, int& iq) const {
    NH_STRICT_FP
    NH_BQ_DIMS
    double np = (double)(n + 1);
    double v = w[pid(k)];
    long a = (long)v;
    long b = (long)(np * v - (double)a * np);
    ip = a > 0 ? (int)a : 0;
    iq = b > 0 ? (int)b : 0;
  }

  // Returns false when every point was replaced without provisional values
  // (Powell's NREM = 0 exit to label 350), true to go on to label 260.
  NH_HD bool rescue_prepare() {
    NH_STRICT_FP
    NH_BQ_DIMS
    const int np = n + 1;
    const int nptm = npt - np;
    const double sfrac = 0.5 / (double)np;
    double sumpq = 0.0, winc = 0.0;
    for (int k = 1; k <= npt; ++k) {
      double dist = 0.0;
      for (int j = 1; j <= n; ++j) {
        xpt[ix(k, j)] -= xopt[j];
        dist += xpt[ix(k, j)] * xpt[ix(k, j)];
      }
      sumpq += pq[k];
      w[rw(ndim + k)] = dist;
      winc = fmax(winc, dist);
      for (int j = 1; j <= nptm; ++j) zmat[iz(k, j)] = 0.0;
    }
    int ih = 0;
    for (int j = 1; j <= n; ++j) {
      w[rw(j)] = 0.5 * sumpq * xopt[j];
      for (int k = 1; k <= npt; ++k) w[rw(j)] += pq[k] * xpt[ix(k, j)];
      for (int i = 1; i <= j; ++i) {
        ++ih;
        hq[ih] = hq[ih] + w[rw(i)] * xopt[j] + w[rw(j)] * xopt[i];
      }
    }
    for (int j = 1; j <= n; ++j) {
      xbase[j] += xopt[j];
      sl[j] -= xopt[j];
      su[j] -= xopt[j];
      xopt[j] = 0.0;
      w[pa(1, j)] = fmin(delta, su[j]);
      w[pa(2, j)] = fmax(-delta, sl[j]);
      if (w[pa(1, j)] + w[pa(2, j)] < 0.0) {
        double t = w[pa(1, j)];
        w[pa(1, j)] = w[pa(2, j)];
        w[pa(2, j)] = t;
      }
      if (fabs(w[pa(2, j)]) < 0.5 * fabs(w[pa(1, j)])) w[pa(2, j)] = 0.5 * w[pa(1, j)];
      for (int i = 1; i <= ndim; ++i) bmat[ib(i, j)] = 0.0;
    }
    fbase = fval[kopt];
    w[pid(1)] = sfrac;
    for (int j = 1; j <= n; ++j) {
      int jp = j + 1;
      int jpn = jp + n;
      w[pid(jp)] = (double)j + sfrac;
      double a1 = w[pa(1, j)], a2 = w[pa(2, j)];
      if (jpn <= npt) {
        w[pid(jpn)] = (double)j / (double)np + sfrac;
        double temp = 1.0 / (a1 - a2);
        bmat[ib(jp, j)] = -temp + 1.0 / a1;
        bmat[ib(jpn, j)] = temp + 1.0 / a2;
        bmat[ib(1, j)] = -bmat[ib(jp, j)] - bmat[ib(jpn, j)];
        zmat[iz(1, j)] = sqrt(2.0) / fabs(a1 * a2);
        zmat[iz(jp, j)] = zmat[iz(1, j)] * a2 * temp;
        zmat[iz(jpn, j)] = -zmat[iz(1, j)] * a1 * temp;
      } else {
        bmat[ib(1, j)] = -1.0 / a1;
        bmat[ib(jp, j)] = 1.0 / a1;
        bmat[ib(j + npt, j)] = -0.5 * (a1 * a1);
      }
    }
    if (npt >= n + np) {
      for (int k = 2 * np; k <= npt; ++k) {
        int iw = (int)(((double)(k - np) - 0.5) / (double)n);
        int ip = k - np - iw * n;
        int iq = ip + iw;
        if (iq > n) iq -= n;
        w[pid(k)] = (double)ip + (double)iq / (double)np + sfrac;
        double temp = 1.0 / (w[pa(1, ip)] * w[pa(1, iq)]);
        zmat[iz(1, k - np)] = temp;
        zmat[iz(ip + 1, k - np)] = -temp;
        zmat[iz(iq + 1, k - np)] = -temp;
        zmat[iz(k, k - np)] = temp;
      }
    }
    int nrem = npt;
    int kold = 1;
    int kn = kopt;
    double rbeta = 0.0, rdenom = 0.0;
    for (;;) {  // label 80
      for (int j = 1; j <= n; ++j) {
        double t = bmat[ib(kold, j)];
        bmat[ib(kold, j)] = bmat[ib(kn, j)];
        bmat[ib(kn, j)] = t;
      }
      for (int j = 1; j <= nptm; ++j) {
        double t = zmat[iz(kold, j)];
        zmat[iz(kold, j)] = zmat[iz(kn, j)];
        zmat[iz(kn, j)] = t;
      }
      w[pid(kold)] = w[pid(kn)];
      w[pid(kn)] = 0.0;
      w[rw(ndim + kn)] = 0.0;
      --nrem;
      if (kn != kopt) {
        double t = vlag[kold];
        vlag[kold] = vlag[kn];
        vlag[kn] = t;
        update_with(rbeta, rdenom, kn, rw(1));
        if (nrem == 0) return false;
        for (int k = 1; k <= npt; ++k) w[rw(ndim + k)] = fabs(w[rw(ndim + k)]);
      }
      for (;;) {  // label 120
        double dsqmin = 0.0;
        for (int k = 1; k <= npt; ++k) {
          double v = w[rw(ndim + k)];
          if (v > 0.0 && (dsqmin == 0.0 || v < dsqmin)) {
            kn = k;
            dsqmin = v;
          }
        }
        if (dsqmin == 0.0) return true;  // on to label 260
        for (int j = 1; j <= n; ++j) w[rw(npt + j)] = xpt[ix(kn, j)];
        for (int k = 1; k <= npt; ++k) {
          double sum = 0.0;
          if (k == kopt) {
          } else if (w[pid(k)] == 0.0) {
            for (int j = 1; j <= n; ++j) sum += w[rw(npt + j)] * xpt[ix(k, j)];
          } else {
            int ip, iq;
            pid_parts(k, ip, iq);
            if (ip > 0) sum = w[rw(npt + ip)] * w[pa(1, ip)];
            if (iq > 0) {
              int iw = ip == 0 ? 2 : 1;
              sum += w[rw(npt + iq)] * w[pa(iw, iq)];
            }
          }
          w[rw(k)] = 0.5 * sum * sum;
        }
        for (int k = 1; k <= npt; ++k) {
          double sum = 0.0;
          for (int j = 1; j <= n; ++j) sum += bmat[ib(k, j)] * w[rw(npt + j)];
          vlag[k] = sum;
        }
        rbeta = 0.0;
        for (int j = 1; j <= nptm; ++j) {
          double sum = 0.0;
          for (int k = 1; k <= npt; ++k) sum += zmat[iz(k, j)] * w[rw(k)];
          rbeta -= sum * sum;
          for (int k = 1; k <= npt; ++k) vlag[k] += sum * zmat[iz(k, j)];
        }
        double bsum = 0.0, dist = 0.0;
        for (int j = 1; j <= n; ++j) {
          double sum = 0.0;
          for (int k = 1; k <= npt; ++k) sum += bmat[ib(k, j)] * w[rw(k)];
          int jp = j + npt;
          bsum += sum * w[rw(jp)];
          for (int ip = npt + 1; ip <= ndim; ++ip) sum += bmat[ib(ip, j)] * w[rw(ip)];
          bsum += sum * w[rw(jp)];
          vlag[jp] = sum;
          double v = xpt[ix(kn, j)];
          dist += v * v;
        }
        rbeta = 0.5 * dist * dist + rbeta - bsum;
        vlag[kopt] += 1.0;
        rdenom = 0.0;
        double vlmxsq = 0.0;
        for (int k = 1; k <= npt; ++k) {
          if (w[pid(k)] != 0.0) {
            double hdiag = 0.0;
            for (int j = 1; j <= nptm; ++j) {
              double z = zmat[iz(k, j)];
              hdiag += z * z;
            }
            double den = rbeta * hdiag + vlag[k] * vlag[k];
            if (den > rdenom) {
              kold = k;
              rdenom = den;
            }
          }
          vlmxsq = fmax(vlmxsq, vlag[k] * vlag[k]);
        }
        if (rdenom <= vlmxsq * 0.01) {
          w[rw(ndim + kn)] = -w[rw(ndim + kn)] - winc;
          continue;
        }
        break;  // back to label 80
      }
    }
  }

  // Label 260: place provisional point kpt and its model value.
  NH_HD void rescue_point() {
    NH_STRICT_FP
    NH_BQ_DIMS
    int ih = 0;
    for (int j = 1; j <= n; ++j) {
      w[rw(j)] = xpt[ix(kpt, j)];
      xpt[ix(kpt, j)] = 0.0;
      double temp = pq[kpt] * w[rw(j)];
      for (int i = 1; i <= j; ++i) {
        ++ih;
        hq[ih] += temp * w[rw(i)];
      }
    }
    pq[kpt] = 0.0;
    int ip, iq;
    pid_parts(kpt, ip, iq);
    rip = ip;
    riq = iq;
    if (ip > 0) {
      rxp = w[pa(1, ip)];
      xpt[ix(kpt, ip)] = rxp;
    }
    if (iq > 0) {
      rxq = w[pa(1, iq)];
      if (ip == 0) rxq = w[pa(2, iq)];
      xpt[ix(kpt, iq)] = rxq;
    }
    double vq = fbase;
    int ihp = 0;
    if (ip > 0) {
      ihp = (ip + ip * ip) / 2;
      vq += rxp * (gopt[ip] + 0.5 * rxp * hq[ihp]);
    }
    if (iq > 0) {
      int ihq = (iq + iq * iq) / 2;
      vq += rxq * (gopt[iq] + 0.5 * rxq * hq[ihq]);
      if (ip > 0) {
        int iw = (ihp > ihq ? ihp : ihq) - (ip > iq ? ip - iq : iq - ip);
        vq += rxp * rxq * hq[iw];
      }
    }
    for (int k = 1; k <= npt; ++k) {
      double temp = 0.0;
      if (ip > 0) temp += rxp * xpt[ix(k, ip)];
      if (iq > 0) temp += rxq * xpt[ix(k, iq)];
      vq += 0.5 * pq[k] * temp * temp;
    }
    rvquad = vq;
    for (int i = 1; i <= n; ++i) {
      double p = xpt[ix(kpt, i)];
      x[i] = fmin(fmax(xl[i], xbase[i] + p), xu[i]);
      if (p == sl[i]) x[i] = xl[i];
      if (p == su[i]) x[i] = xu[i];
    }
  }

  NH_HD void rescue_update() {
    NH_STRICT_FP
    NH_BQ_DIMS
    const int nptm = npt - n - 1;
    double diff = f - rvquad;
    for (int i = 1; i <= n; ++i) gopt[i] += diff * bmat[ib(kpt, i)];
    for (int k = 1; k <= npt; ++k) {
      double sum = 0.0;
      for (int j = 1; j <= nptm; ++j) sum += zmat[iz(k, j)] * zmat[iz(kpt, j)];
      double temp = diff * sum;
      if (w[pid(k)] == 0.0) {
        pq[k] += temp;
      } else {
        int ip, iq;
        pid_parts(k, ip, iq);
        int ihq = (iq * iq + iq) / 2;
        if (ip == 0) {
          double a = w[pa(2, iq)];
          hq[ihq] += temp * (a * a);
        } else {
          int ihp = (ip * ip + ip) / 2;
          double a = w[pa(1, ip)];
          hq[ihp] += temp * (a * a);
          if (iq > 0) {
            double b = w[pa(1, iq)];
            hq[ihq] += temp * (b * b);
            int iw = (ihp > ihq ? ihp : ihq) - (iq > ip ? iq - ip : ip - iq);
            hq[iw] += temp * a * b;
          }
        }
      }
    }
    w[pid(kpt)] = 0.0;
  }

  // ---- UPDATE (bobyqa.rs:2722-2787); work space starts at w[wbase] ----
  NH_HD void update_hz(int kn) { update_with(beta, denom, kn, 1); }
  NH_HD void update_with(double ubeta, double udenom, int kn, int wbase) {
    NH_STRICT_FP
    NH_BQ_DIMS
    double* wk = w + wbase;
    double* vl = vlag + 1;
    const int nptm = npt - n - 1;
    double ztest = 0.0;
    for (int k = 1; k <= npt; ++k)
      for (int j = 1; j <= nptm; ++j) ztest = fmax(ztest, fabs(zmat[iz(k, j)]));
    ztest *= 1e-20;
    for (int j = 2; j <= nptm; ++j) {
      if (fabs(zmat[iz(kn, j)]) > ztest) {
        double a = zmat[iz(kn, 1)], b = zmat[iz(kn, j)];
        double temp = sqrt(a * a + b * b);
        double tempa = a / temp, tempb = b / temp;
        for (int i = 1; i <= npt; ++i) {
          double t = tempa * zmat[iz(i, 1)] + tempb * zmat[iz(i, j)];
          zmat[iz(i, j)] = tempa * zmat[iz(i, j)] - tempb * zmat[iz(i, 1)];
          zmat[iz(i, 1)] = t;
        }
      }
      zmat[iz(kn, j)] = 0.0;
    }
    for (int i = 1; i <= npt; ++i) wk[i - 1] = zmat[iz(kn, 1)] * zmat[iz(i, 1)];
    double ualpha = wk[kn - 1];
    double tau = vl[kn - 1];
    vl[kn - 1] -= 1.0;
    double temp = sqrt(udenom);
    double tempb = zmat[iz(kn, 1)] / temp;
    double tempa = tau / temp;
    for (int i = 1; i <= npt; ++i) zmat[iz(i, 1)] = tempa * zmat[iz(i, 1)] - tempb * vl[i - 1];
    for (int j = 1; j <= n; ++j) {
      int jp = npt + j;
      wk[jp - 1] = bmat[ib(kn, j)];
      double ta = (ualpha * vl[jp - 1] - tau * wk[jp - 1]) / udenom;
      double tb = (-ubeta * wk[jp - 1] - tau * vl[jp - 1]) / udenom;
      for (int i = 1; i <= jp; ++i) {
        bmat[ib(i, j)] = bmat[ib(i, j)] + ta * vl[i - 1] + tb * wk[i - 1];
        if (i > npt) bmat[ib(jp, i - npt)] = bmat[ib(i, j)];
      }
    }
  }

  // ---- TRSBOX (bobyqa.rs:2007-2432) with Powell's labels ----
  NH_HD void trsbox_step() {
    NH_STRICT_FP
    NH_BQ_DIMS
    double* gnew = w + 1;
    double* xbdi = w + 1 + n;
    double* sv = w + 1 + 2 * n;
    double* hs = w + 1 + 3 * n;
    double* hred = w + 1 + 4 * n;
    // one-based views
    auto XO = [&](int i) -> double { return xopt[i]; };
    long iterc = 0;
    int nact = 0;
    for (int i = 1; i <= n; ++i) {
      xbdi[i - 1] = 0.0;
      if (XO(i) <= sl[i]) {
        if (gopt[i] >= 0.0) xbdi[i - 1] = -1.0;
      } else if (XO(i) >= su[i] && gopt[i] <= 0.0) {
        xbdi[i - 1] = 1.0;
      }
      if (xbdi[i - 1] != 0.0) ++nact;
      d[i] = 0.0;
      gnew[i - 1] = gopt[i];
    }
    double delsq = delta * delta;
    double qred = 0.0;
    crvmin = -1.0;
    double tbeta = 0.0, stepsq = 0.0, gredsq = 0.0, ggsav = 0.0, dredsq = 0.0,
           dredg = 0.0, sredg = 0.0, angbd = 0.0, xsav = 0.0, shs = 0.0, sdec = 0.0;
    long itermax = 0, itcsav = 0;
    int iact = 0;

  T20:
    tbeta = 0.0;
  T30:
    stepsq = 0.0;
    for (int i = 1; i <= n; ++i) {
      if (xbdi[i - 1] != 0.0)
        sv[i - 1] = 0.0;
      else if (tbeta == 0.0)
        sv[i - 1] = -gnew[i - 1];
      else
        sv[i - 1] = tbeta * sv[i - 1] - gnew[i - 1];
      stepsq += sv[i - 1] * sv[i - 1];
    }
    if (stepsq == 0.0) goto T190;
    if (tbeta == 0.0) {
      gredsq = stepsq;
      itermax = iterc + n - nact;
    }
    if (gredsq * delsq <= qred * 1e-4 * qred) goto T190;
    goto T210;

  T50: {
    double resid = delsq, ds = 0.0;
    shs = 0.0;
    for (int i = 1; i <= n; ++i) {
      if (xbdi[i - 1] == 0.0) {
        resid -= d[i] * d[i];
        ds += sv[i - 1] * d[i];
        shs += sv[i - 1] * hs[i - 1];
      }
    }
    if (resid <= 0.0) goto T90;
    double temp = sqrt(stepsq * resid + ds * ds);
    double blen = ds < 0.0 ? (temp - ds) / stepsq : resid / (temp + ds);
    double stplen = blen;
    if (shs > 0.0) stplen = fmin(blen, gredsq / shs);
    iact = 0;
    for (int i = 1; i <= n; ++i) {
      if (sv[i - 1] != 0.0) {
        double xsum = XO(i) + d[i];
        double t = sv[i - 1] > 0.0 ? (su[i] - xsum) / sv[i - 1] : (sl[i] - xsum) / sv[i - 1];
        if (t < stplen) {
          stplen = t;
          iact = i;
        }
      }
    }
    sdec = 0.0;
    if (stplen > 0.0) {
      ++iterc;
      double t = shs / stepsq;
      if (iact == 0 && t > 0.0) {
        crvmin = fmin(crvmin, t);
        if (crvmin == -1.0) crvmin = t;
      }
      ggsav = gredsq;
      gredsq = 0.0;
      for (int i = 1; i <= n; ++i) {
        gnew[i - 1] += stplen * hs[i - 1];
        if (xbdi[i - 1] == 0.0) gredsq += gnew[i - 1] * gnew[i - 1];
        d[i] += stplen * sv[i - 1];
      }
      sdec = fmax(stplen * (ggsav - 0.5 * stplen * shs), 0.0);
      qred += sdec;
    }
    if (iact > 0) {
      ++nact;
      xbdi[iact - 1] = 1.0;
      if (sv[iact - 1] < 0.0) xbdi[iact - 1] = -1.0;
      delsq -= d[iact] * d[iact];
      if (delsq <= 0.0) goto T90;
      goto T20;
    }
    if (stplen < blen) {
      if (iterc == itermax) goto T190;
      if (sdec <= qred * 0.01) goto T190;
      tbeta = gredsq / ggsav;
      goto T30;
    }
  }
  T90:
    crvmin = 0.0;
  T100:
    if (nact + 1 >= n) goto T190;
    dredsq = 0.0;
    dredg = 0.0;
    gredsq = 0.0;
    for (int i = 1; i <= n; ++i) {
      if (xbdi[i - 1] == 0.0) {
        dredsq += d[i] * d[i];
        dredg += d[i] * gnew[i - 1];
        gredsq += gnew[i - 1] * gnew[i - 1];
        sv[i - 1] = d[i];
      } else {
        sv[i - 1] = 0.0;
      }
    }
    itcsav = iterc;
    goto T210;

  T120: {
    ++iterc;
    double temp = gredsq * dredsq - dredg * dredg;
    if (temp <= qred * 1e-4 * qred) goto T190;
    temp = sqrt(temp);
    for (int i = 1; i <= n; ++i) {
      if (xbdi[i - 1] == 0.0)
        sv[i - 1] = (dredg * d[i] - dredsq * gnew[i - 1]) / temp;
      else
        sv[i - 1] = 0.0;
    }
    sredg = -temp;
    angbd = 1.0;
    iact = 0;
    for (int i = 1; i <= n; ++i) {
      if (xbdi[i - 1] == 0.0) {
        double tempa = XO(i) + d[i] - sl[i];
        double tempb = su[i] - XO(i) - d[i];
        if (tempa <= 0.0) {
          ++nact;
          xbdi[i - 1] = -1.0;
          goto T100;
        } else if (tempb <= 0.0) {
          ++nact;
          xbdi[i - 1] = 1.0;
          goto T100;
        }
        double ssq = d[i] * d[i] + sv[i - 1] * sv[i - 1];
        double a = XO(i) - sl[i];
        double t = ssq - a * a;
        if (t > 0.0) {
          t = sqrt(t) - sv[i - 1];
          if (angbd * t > tempa) {
            angbd = tempa / t;
            iact = i;
            xsav = -1.0;
          }
        }
        double b = su[i] - XO(i);
        t = ssq - b * b;
        if (t > 0.0) {
          t = sqrt(t) + sv[i - 1];
          if (angbd * t > tempb) {
            angbd = tempb / t;
            iact = i;
            xsav = 1.0;
          }
        }
      }
    }
    goto T210;
  }

  T150: {
    shs = 0.0;
    double dhs = 0.0, dhd = 0.0;
    for (int i = 1; i <= n; ++i) {
      if (xbdi[i - 1] == 0.0) {
        shs += sv[i - 1] * hs[i - 1];
        dhs += d[i] * hs[i - 1];
        dhd += d[i] * hred[i - 1];
      }
    }
    double redmax = 0.0, redsav = 0.0, rdprev = 0.0, rdnext = 0.0;
    long isav = 0;
    long iu = (long)(angbd * 17.0 + 3.1);
    for (long i = 1; i <= iu; ++i) {
      double angt = angbd * (double)i / (double)iu;
      double sth = (angt + angt) / (1.0 + angt * angt);
      double temp = shs + angt * (angt * dhd - dhs - dhs);
      double rednew = sth * (angt * dredg - sredg - 0.5 * sth * temp);
      if (rednew > redmax) {
        redmax = rednew;
        isav = i;
        rdprev = redsav;
      } else if (i == isav + 1) {
        rdnext = rednew;
      }
      redsav = rednew;
    }
    if (isav == 0) goto T190;
    double angt = 0.0;
    if (isav < iu) {
      double temp = (rdnext - rdprev) / (redmax + redmax - rdprev - rdnext);
      angt = angbd * ((double)isav + 0.5 * temp) / (double)iu;
    }
    double cth = (1.0 - angt * angt) / (1.0 + angt * angt);
    double sth = (angt + angt) / (1.0 + angt * angt);
    double temp = shs + angt * (angt * dhd - dhs - dhs);
    sdec = sth * (angt * dredg - sredg - 0.5 * sth * temp);
    if (sdec <= 0.0) goto T190;
    dredg = 0.0;
    gredsq = 0.0;
    for (int i = 1; i <= n; ++i) {
      gnew[i - 1] = gnew[i - 1] + (cth - 1.0) * hred[i - 1] + sth * hs[i - 1];
      if (xbdi[i - 1] == 0.0) {
        d[i] = cth * d[i] + sth * sv[i - 1];
        dredg += d[i] * gnew[i - 1];
        gredsq += gnew[i - 1] * gnew[i - 1];
      }
      hred[i - 1] = cth * hred[i - 1] + sth * hs[i - 1];
    }
    qred += sdec;
    if (iact > 0 && isav == iu) {
      ++nact;
      xbdi[iact - 1] = xsav;
      goto T100;
    }
    if (sdec > qred * 0.01) goto T120;
    goto T190;
  }

  T190:
    dsq = 0.0;
    for (int i = 1; i <= n; ++i) {
      xnew[i] = fmax(fmin(XO(i) + d[i], su[i]), sl[i]);
      if (xbdi[i - 1] == -1.0) xnew[i] = sl[i];
      if (xbdi[i - 1] == 1.0) xnew[i] = su[i];
      d[i] = xnew[i] - XO(i);
      dsq += d[i] * d[i];
    }
    return;

  T210: {
    int ih = 0;
    for (int j = 1; j <= n; ++j) {
      hs[j - 1] = 0.0;
      for (int i = 1; i <= j; ++i) {
        ++ih;
        if (i < j) hs[j - 1] += hq[ih] * sv[i - 1];
        hs[i - 1] += hq[ih] * sv[j - 1];
      }
    }
    for (int k = 1; k <= npt; ++k) {
      if (pq[k] != 0.0) {
        double temp = 0.0;
        for (int j = 1; j <= n; ++j) temp += xpt[ix(k, j)] * sv[j - 1];
        temp *= pq[k];
        for (int i = 1; i <= n; ++i) hs[i - 1] += temp * xpt[ix(k, i)];
      }
    }
    if (crvmin != 0.0) goto T50;
    if (iterc > itcsav) goto T150;
    for (int i = 0; i < n; ++i) hred[i] = hs[i];
    goto T120;
  }
  }

  // ---- ALTMOV (bobyqa.rs:2438-2708) ----
  NH_HD void altmov_step() {
    NH_STRICT_FP
    NH_BQ_DIMS
    double* glag = w + 1;
    double* hcol = w + 1 + n;
    double* wk = w + ndim + 1;
    const double cnst = 1.0 + sqrt(2.0);
    for (int k = 1; k <= npt; ++k) hcol[k - 1] = 0.0;
    for (int j = 1; j <= npt - n - 1; ++j) {
      double temp = zmat[iz(knew, j)];
      for (int k = 1; k <= npt; ++k) hcol[k - 1] += temp * zmat[iz(k, j)];
    }
    alpha = hcol[knew - 1];
    double ha = 0.5 * alpha;
    for (int i = 1; i <= n; ++i) glag[i - 1] = bmat[ib(knew, i)];
    for (int k = 1; k <= npt; ++k) {
      double temp = 0.0;
      for (int j = 1; j <= n; ++j) temp += xpt[ix(k, j)] * xopt[j];
      temp *= hcol[k - 1];
      for (int i = 1; i <= n; ++i) glag[i - 1] += temp * xpt[ix(k, i)];
    }
    double presav = 0.0, stpsav = 0.0;
    int ksav = 0;
    long ibdsav = 0;
    for (int k = 1; k <= npt; ++k) {
      if (k == kopt) continue;
      double dderiv = 0.0, dist = 0.0;
      for (int i = 1; i <= n; ++i) {
        double temp = xpt[ix(k, i)] - xopt[i];
        dderiv += glag[i - 1] * temp;
        dist += temp * temp;
      }
      double subd = adelt / sqrt(dist);
      double slbd = -subd;
      long ilbd = 0, iubd = 0;
      double sumin = fmin(1.0, subd);
      for (int i = 1; i <= n; ++i) {
        double temp = xpt[ix(k, i)] - xopt[i];
        if (temp > 0.0) {
          if (slbd * temp < sl[i] - xopt[i]) {
            slbd = (sl[i] - xopt[i]) / temp;
            ilbd = -i;
          }
          if (subd * temp > su[i] - xopt[i]) {
            subd = fmax(sumin, (su[i] - xopt[i]) / temp);
            iubd = i;
          }
        } else if (temp < 0.0) {
          if (slbd * temp > su[i] - xopt[i]) {
            slbd = (su[i] - xopt[i]) / temp;
            ilbd = i;
          }
          if (subd * temp < sl[i] - xopt[i]) {
            subd = fmax(sumin, (sl[i] - xopt[i]) / temp);
            iubd = -i;
          }
        }
      }
      double step, vl;
      long isbd;
      if (k == knew) {
        double diff = dderiv - 1.0;
        step = slbd;
        vl = slbd * (dderiv - slbd * diff);
        isbd = ilbd;
        double temp = subd * (dderiv - subd * diff);
        if (fabs(temp) > fabs(vl)) {
          step = subd;
          vl = temp;
          isbd = iubd;
        }
        double tempd = 0.5 * dderiv;
        double tempa = tempd - diff * slbd;
        double tempb = tempd - diff * subd;
        if (tempa * tempb < 0.0) {
          temp = tempd * tempd / diff;
          if (fabs(temp) > fabs(vl)) {
            step = tempd / diff;
            vl = temp;
            isbd = 0;
          }
        }
      } else {
        step = slbd;
        vl = slbd * (1.0 - slbd);
        isbd = ilbd;
        double temp = subd * (1.0 - subd);
        if (fabs(temp) > fabs(vl)) {
          step = subd;
          vl = temp;
          isbd = iubd;
        }
        if (subd > 0.5 && fabs(vl) < 0.25) {
          step = 0.5;
          vl = 0.25;
          isbd = 0;
        }
        vl *= dderiv;
      }
      double temp = step * (1.0 - step) * dist;
      double predsq = vl * vl * (vl * vl + ha * temp * temp);
      if (predsq > presav) {
        presav = predsq;
        ksav = k;
        stpsav = step;
        ibdsav = isbd;
      }
    }
    // With no positive prediction the C reads row zero; stpsav is zero then,
    // so any row gives xnew = xopt (bobyqa.rs:2624-2627).
    if (ksav < 1) ksav = 1;
    for (int i = 1; i <= n; ++i) {
      double temp = xopt[i] + stpsav * (xpt[ix(ksav, i)] - xopt[i]);
      xnew[i] = fmax(sl[i], fmin(su[i], temp));
    }
    if (ibdsav < 0) xnew[-ibdsav] = sl[-ibdsav];
    if (ibdsav > 0) xnew[ibdsav] = su[ibdsav];

    double bigstp = adelt + adelt;
    int iflag = 0;
    double csave = 0.0;
    for (;;) {
      double wfixsq = 0.0, ggfree = 0.0;
      for (int i = 1; i <= n; ++i) {
        wk[i - 1] = 0.0;
        double tempa = fmin(xopt[i] - sl[i], glag[i - 1]);
        double tempb = fmax(xopt[i] - su[i], glag[i - 1]);
        if (tempa > 0.0 || tempb < 0.0) {
          wk[i - 1] = bigstp;
          ggfree += glag[i - 1] * glag[i - 1];
        }
      }
      if (ggfree == 0.0) {
        cauchy = 0.0;
        return;
      }
      double step = 0.0;
      for (;;) {
        double temp = adelt * adelt - wfixsq;
        if (temp > 0.0) {
          double wsqsav = wfixsq;
          step = sqrt(temp / ggfree);
          ggfree = 0.0;
          for (int i = 1; i <= n; ++i) {
            if (wk[i - 1] == bigstp) {
              double t = xopt[i] - step * glag[i - 1];
              if (t <= sl[i]) {
                wk[i - 1] = sl[i] - xopt[i];
                wfixsq += wk[i - 1] * wk[i - 1];
              } else if (t >= su[i]) {
                wk[i - 1] = su[i] - xopt[i];
                wfixsq += wk[i - 1] * wk[i - 1];
              } else {
                ggfree += glag[i - 1] * glag[i - 1];
              }
            }
          }
          if (wfixsq > wsqsav && ggfree > 0.0) continue;
        }
        break;
      }
      double gw = 0.0;
      for (int i = 1; i <= n; ++i) {
        if (wk[i - 1] == bigstp) {
          wk[i - 1] = -step * glag[i - 1];
          xalt[i] = fmax(sl[i], fmin(su[i], xopt[i] + wk[i - 1]));
        } else if (wk[i - 1] == 0.0) {
          xalt[i] = xopt[i];
        } else if (glag[i - 1] > 0.0) {
          xalt[i] = sl[i];
        } else {
          xalt[i] = su[i];
        }
        gw += glag[i - 1] * wk[i - 1];
      }
      double curv = 0.0;
      for (int k = 1; k <= npt; ++k) {
        double temp = 0.0;
        for (int j = 1; j <= n; ++j) temp += xpt[ix(k, j)] * wk[j - 1];
        curv += hcol[k - 1] * temp * temp;
      }
      if (iflag == 1) curv = -curv;
      if (curv > -gw && curv < -cnst * gw) {
        double scale = -gw / curv;
        for (int i = 1; i <= n; ++i) {
          double temp = xopt[i] + scale * wk[i - 1];
          xalt[i] = fmax(sl[i], fmin(su[i], temp));
        }
        double a = 0.5 * gw * scale;
        cauchy = a * a;
      } else {
        double a = gw + 0.5 * curv;
        cauchy = a * a;
      }
      if (iflag == 0) {
        for (int i = 1; i <= n; ++i) {
          glag[i - 1] = -glag[i - 1];
          wk[n + i - 1] = xalt[i];
        }
        csave = cauchy;
        iflag = 1;
        continue;
      }
      if (csave > cauchy) {
        for (int i = 1; i <= n; ++i) xalt[i] = wk[n + i - 1];
        cauchy = csave;
      }
      return;
    }
  }
};

}  // namespace nh



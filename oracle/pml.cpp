// The penalised NB inner fit, f64: a restatement of src/sc/pml.rs.
//
// optimise        <- pml.rs:1614-1800 (Newton over (beta, log w) through the
//                    Schur complement, per-coordinate backtracking, codes)
// sweep           <- pml.rs:1419-1590 (dense pass at a count of zero, or the
//                    zero-count tables' groups plus the loose cells, then the
//                    positive counts; exact because every term is linear in y)
// laplace_corr.   <- pml.rs:1828-1919
// ldlt_solve      <- pml.rs:1940-2021 (Eigen-style pivoted LDL')
// determinant     <- pml.rs:2036-2071, min_eigenvalue <- pml.rs:2084-2133
// check_conv      <- pml.rs:1111-1139
// Sums associate as the reference does where that is cheap; agreement with
// nebula's values is to rounding (1e-11 on the quoted fixtures).
#include <algorithm>
#include <cmath>
#include <limits>

#include "oracle.hpp"

namespace oracle {

void fail(int code, const std::string& message) { throw Error{code, message}; }

namespace {

constexpr double kGradientTolerance = 0.01;  // pml.rs:67
constexpr double kStepCutoff = 40.0;         // pml.rs:74
constexpr double kEigenCutoff = 1e-8;        // pml.rs:80
constexpr int kJacobiSweeps = 30;            // pml.rs:86

struct Penalty {
  bool gamma;  // true: gamma prior (NBGMM), false: log-normal (NBLMM)
  double alpha, lambda;
  double accumulate(double ll, const std::vector<double>& log_w,
                    const std::vector<double>& w) const {
    if (gamma) {
      double sl = 0.0, sw = 0.0;
      for (double v : log_w) sl += v;
      for (double v : w) sw += v;
      return ll + (alpha * sl - lambda * sw);
    }
    double quad = 0.0;
    for (double v : log_w) quad += v * v;
    return ll - quad / alpha / 2.0 - (double)log_w.size() / 2.0 * std::log(alpha);
  }
  double gradient(double log_w, double w) const {
    return gamma ? alpha - lambda * w : -log_w / alpha;
  }
  double curvature(double w) const { return gamma ? lambda * w : 1.0 / alpha; }
};

struct Sweep {
  double ll = 0.0;
  std::vector<double> w, db, dw, vw, vwb, vb;
  void size(std::size_t nb, std::size_t k) {
    w.assign(k, 0.0);
    db.assign(nb, 0.0);
    dw.assign(k, 0.0);
    vw.assign(k, 0.0);
    vwb.assign(k * nb, 0.0);
    vb.assign(nb * nb, 0.0);
  }
};

std::vector<std::size_t> positive_runs(const PmlData& d) {
  std::vector<std::size_t> run(d.k + 1);
  for (std::size_t s = 0; s <= d.k; ++s) {
    std::size_t b = d.subject_start[s];
    run[s] = std::lower_bound(d.cells, d.cells + d.nnz, b) - d.cells;
  }
  return run;
}

// One pass per subject at (beta, log_w): penalised log-likelihood, gradient
// and curvature blocks.  NB > 0 fixes the width at compile time.
template <int NB>
void sweep_width(const PmlData& d, const double* beta, const double* log_w,
                 double gamma, const Penalty& pen, std::vector<double>& stored,
                 const std::vector<std::size_t>& run, Sweep& out) {
  const std::size_t nb = NB > 0 ? (std::size_t)NB : d.nb;
  const std::size_t k = d.k;
  std::vector<double> heap(NB > 0 ? 0 : 2 * nb + nb * nb);
  double local[2 * 8 + 64];
  double* sums = NB > 0 ? local : heap.data();
  std::fill(out.db.begin(), out.db.end(), 0.0);
  std::fill(out.vb.begin(), out.vb.end(), 0.0);
  double ll = 0.0;
  const double ln_gamma = std::log(gamma);
  std::vector<double> lw(log_w, log_w + k);
  for (std::size_t s = 0; s < k; ++s) {
    const std::size_t c0 = d.subject_start[s], c1 = d.subject_start[s + 1];
    const double lws = log_w[s];
    const double ws = std::exp(lws);
    std::fill(sums, sums + 2 * nb + nb * nb, 0.0);
    double* score = sums;
    double* first = sums + nb;
    double* second = sums + 2 * nb;
    double resid = 0.0, weight = 0.0, sum_log = 0.0, linear = 0.0, wlog = 0.0;
    auto fold = [&](std::size_t r, double dd, double pp) {
      const double* row = d.design + r * nb;
      resid += dd;
      weight += pp;
      for (std::size_t a = 0; a < nb; ++a) {
        score[a] += row[a] * dd;
        double pa = pp * row[a];
        first[a] += pa;
        for (std::size_t b = a; b < nb; ++b) second[a * nb + b] += pa * row[b];
      }
    };
    if (!d.zeros) {
      for (std::size_t r = c0; r < c1; ++r) {
        const double* row = d.design + r * nb;
        double eta = d.offset[r];
        for (std::size_t j = 0; j < nb; ++j) eta += row[j] * beta[j];
        stored[r] = std::exp(eta + lws);
      }
      for (std::size_t r = c0; r < c1; ++r) {
        double e = stored[r];
        double t = e + gamma;
        sum_log += std::log(t);
        double inv = 1.0 / t;
        double u = e * inv;
        fold(r, -gamma * u, gamma * u * inv);
      }
    } else {
      // per group at a count of zero: sum ln t = n ln gamma + L, sum u = A,
      // sum gamma u / t = B at ln tau = log w + x'beta - ln gamma
      // (pml.rs:1510-1536); then the loose cells one by one
      const ZeroCells& z = *d.zeros;
      for (std::size_t gi = z.subject_groups[s]; gi < z.subject_groups[s + 1]; ++gi) {
        const ZeroGroup& g = z.groups[gi];
        const double* row = d.design + g.row * nb;
        double log_tau = lws - ln_gamma;
        for (std::size_t j = 0; j < nb; ++j) log_tau += row[j] * beta[j];
        double lab[3];
        g.eval(0, 3, log_tau, lab);
        sum_log += g.n * ln_gamma + lab[0];
        fold(g.row, -gamma * lab[1], lab[2]);
      }
      for (std::size_t li = z.subject_loose[s]; li < z.subject_loose[s + 1]; ++li) {
        const std::size_t r = z.loose[li];
        const double* row = d.design + r * nb;
        double eta = d.offset[r];
        for (std::size_t j = 0; j < nb; ++j) eta += row[j] * beta[j];
        const double e = std::exp(eta + lws);
        stored[r] = e;
        const double t = e + gamma;
        sum_log += std::log(t);
        const double inv = 1.0 / t;
        const double u = e * inv;
        fold(r, -gamma * u, gamma * u * inv);
      }
    }
    for (std::size_t i = run[s]; i < run[s + 1]; ++i) {
      std::size_t c = d.cells[i];
      double y = d.counts[i];
      const double* row = d.design + c * nb;
      double eta = d.offset[c];
      for (std::size_t j = 0; j < nb; ++j) eta += row[j] * beta[j];
      double e;
      if (d.zeros) {
        e = std::exp(eta + lws);
        stored[c] = e;
      } else {
        e = stored[c];
      }
      double t = e + gamma;
      linear += eta * y;
      wlog += y * std::log(t);
      double inv = 1.0 / t;
      double u = e * inv;
      fold(c, y * (1.0 - u), y * u * inv);
    }
    ll += linear + lws * d.subject_total[s] - gamma * sum_log - wlog;
    out.w[s] = ws;
    out.dw[s] = resid + pen.gradient(lws, ws);
    out.vw[s] = gamma * weight + pen.curvature(ws);
    for (std::size_t a = 0; a < nb; ++a) {
      out.db[a] += score[a];
      out.vwb[s * nb + a] = gamma * first[a];
      for (std::size_t b = a; b < nb; ++b) out.vb[a * nb + b] += gamma * second[a * nb + b];
    }
  }
  for (std::size_t a = 0; a < nb; ++a)
    for (std::size_t b = a + 1; b < nb; ++b) out.vb[b * nb + a] = out.vb[a * nb + b];
  out.ll = pen.accumulate(ll, lw, out.w);
}

void sweep(const PmlData& d, const double* beta, const double* log_w,
           double gamma, const Penalty& pen, std::vector<double>& stored,
           const std::vector<std::size_t>& run, Sweep& out) {
  switch (d.nb) {
    case 1: return sweep_width<1>(d, beta, log_w, gamma, pen, stored, run, out);
    case 2: return sweep_width<2>(d, beta, log_w, gamma, pen, stored, run, out);
    case 3: return sweep_width<3>(d, beta, log_w, gamma, pen, stored, run, out);
    case 4: return sweep_width<4>(d, beta, log_w, gamma, pen, stored, run, out);
    case 5: return sweep_width<5>(d, beta, log_w, gamma, pen, stored, run, out);
    case 6: return sweep_width<6>(d, beta, log_w, gamma, pen, stored, run, out);
    case 7: return sweep_width<7>(d, beta, log_w, gamma, pen, stored, run, out);
    case 8: return sweep_width<8>(d, beta, log_w, gamma, pen, stored, run, out);
    default: return sweep_width<0>(d, beta, log_w, gamma, pen, stored, run, out);
  }
}

void validate(const PmlData& d) {
  if (d.n_cells == 0) fail(3, "n_cells must be positive");
  if (d.k == 0) fail(3, "n_subjects must be positive");
  if (d.nb == 0) fail(4, "design has no columns");
  if (d.subject_start[0] != 0 || d.subject_start[d.k] != d.n_cells)
    fail(2, "subject_start must run from 0 to n_cells");
  for (std::size_t s = 0; s < d.k; ++s)
    if (d.subject_start[s] >= d.subject_start[s + 1]) fail(7, "empty subject block");
  for (std::size_t i = 0; i < d.nnz; ++i) {
    if (d.cells[i] >= d.n_cells) fail(2, "cell index out of range");
    if (i > 0 && d.cells[i] <= d.cells[i - 1]) fail(2, "cell indices must increase");
  }
}

// Higher Laplace terms (pml.rs:1828-1919).
double laplace_correction(const PmlData& d, const std::vector<double>& stored,
                          const std::vector<std::size_t>& run, const double* beta,
                          const double* log_w, const std::vector<double>& w, double gamma,
                          double lambda, unsigned ord) {
  const bool high = ord > 2;
  const double ln_gamma = std::log(gamma);
  auto terms = [&](double e, double out[4]) {
    double t = e + gamma;
    double p2 = e / (t * t);
    double p3 = p2 / t;
    double p4 = p3 / t;
    double p5 = p4 / t;
    double e2 = e * e;
    out[0] = p2;
    out[1] = p3 * (gamma - e);
    out[2] = p4 * (gamma * gamma + e2 - (4.0 * gamma) * e);
    out[3] = p5 * (gamma * gamma * gamma - (11.0 * gamma * gamma) * e +
                   (11.0 * gamma) * e2 - e2 * e);
  };
  double third_acc = 0.0, fourth_acc = 0.0, fifth_acc = 0.0;
  for (std::size_t s = 0; s < d.k; ++s) {
    double sums[4] = {0, 0, 0, 0}, f[4];
    if (!d.zeros) {
      for (std::size_t r = d.subject_start[s]; r < d.subject_start[s + 1]; ++r) {
        terms(stored[r], f);
        for (int j = 0; j < 4; ++j) sums[j] += gamma * f[j];
      }
    } else {
      // loose cells, then each group's B..E (gamma times each factor)
      const ZeroCells& z = *d.zeros;
      for (std::size_t li = z.subject_loose[s]; li < z.subject_loose[s + 1]; ++li) {
        terms(stored[z.loose[li]], f);
        for (int j = 0; j < 4; ++j) sums[j] += gamma * f[j];
      }
      for (std::size_t gi = z.subject_groups[s]; gi < z.subject_groups[s + 1]; ++gi) {
        const ZeroGroup& g = z.groups[gi];
        const double* row = d.design + g.row * d.nb;
        double log_tau = log_w[s] - ln_gamma;
        for (std::size_t j = 0; j < d.nb; ++j) log_tau += row[j] * beta[j];
        g.eval(2, 4, log_tau, f);
        for (int j = 0; j < 4; ++j) sums[j] += f[j];
      }
    }
    for (std::size_t i = run[s]; i < run[s + 1]; ++i) {
      terms(stored[d.cells[i]], f);
      for (int j = 0; j < 4; ++j) sums[j] += d.counts[i] * f[j];
    }
    double prior = lambda * w[s];
    double vw = gamma * sums[0] + prior;
    double third = gamma * sums[1] + prior;
    third_acc += third * third / (vw * vw * vw);
    if (high) {
      double fourth = gamma * sums[2] + prior;
      double fifth = gamma * sums[3] + prior;
      double vs = vw * vw;
      fourth_acc += fourth / vs;
      fifth_acc += fifth * third / (vs * vs);
    }
  }
  double second = 5.0 * third_acc / 24.0;
  if (high) {
    second -= fourth_acc / 8.0;
    second += 7.0 * fifth_acc / 48.0;
  }
  return second;
}

PmlResult optimise(const PmlData& d, const double* beta_init,
                   const double* log_w_init, const Penalty& pen, double gamma,
                   const PmlParams& params) {
  validate(d);
  const std::size_t nb = d.nb, k = d.k;
  if (!(std::isfinite(gamma) && gamma > 0.0)) fail(12, "invalid cell dispersion");
  std::vector<double> beta(beta_init, beta_init + nb);
  std::vector<double> log_w(k, 0.0);
  if (log_w_init) log_w.assign(log_w_init, log_w_init + k);
  std::vector<double> stored(d.n_cells, 0.0);
  auto run = positive_runs(d);
  Sweep cur, trial;
  cur.size(nb, k);
  trial.size(nb, k);
  sweep(d, beta.data(), log_w.data(), gamma, pen, stored, run, cur);

  std::vector<double> vw_step(k), vb2(nb * nb), factor(nb * nb), step_beta(nb),
      step_log_w(k), dwvw(k), damp_beta(nb), damp_log_w(k), new_beta(nb),
      new_log_w(k);
  double ll = cur.ll, ll_prev = 0.0, likdif = 0.0;
  std::size_t step = 0, backtracks = 0;
  while (step == 0 || (likdif > params.eps && step < params.max_iter)) {
    ++step;
    std::fill(damp_beta.begin(), damp_beta.end(), 1.0);
    std::fill(damp_log_w.begin(), damp_log_w.end(), 1.0);
    vw_step = cur.vw;
    for (std::size_t s = 0; s < k; ++s) dwvw[s] = cur.dw[s] / cur.vw[s];
    for (std::size_t a = 0; a < nb; ++a)
      for (std::size_t b = a; b < nb; ++b) {
        double acc = 0.0;
        for (std::size_t s = 0; s < k; ++s) acc += cur.vwb[s * nb + a] * cur.vwb[s * nb + b] / cur.vw[s];
        double v = cur.vb[a * nb + b] - acc;
        vb2[a * nb + b] = v;
        vb2[b * nb + a] = v;
      }
    for (std::size_t j = 0; j < nb; ++j) {
      double acc = 0.0;
      for (std::size_t s = 0; s < k; ++s) acc += cur.vwb[s * nb + j] * dwvw[s];
      step_beta[j] = cur.db[j] - acc;
    }
    factor = vb2;
    ldlt_solve(factor.data(), step_beta.data(), nb);
    for (std::size_t s = 0; s < k; ++s) {
      double acc = 0.0;
      for (std::size_t j = 0; j < nb; ++j) acc += cur.vwb[s * nb + j] * step_beta[j];
      step_log_w[s] = dwvw[s] - acc / cur.vw[s];
    }
    for (std::size_t j = 0; j < nb; ++j) new_beta[j] = beta[j] + step_beta[j];
    for (std::size_t s = 0; s < k; ++s) new_log_w[s] = log_w[s] + step_log_w[s];

    ll_prev = ll;
    sweep(d, new_beta.data(), new_log_w.data(), gamma, pen, stored, run, trial);
    ll = trial.ll;
    likdif = ll - ll_prev;
    backtracks = 0;
    double min_step = kStepCutoff;
    while (likdif < 0.0 || std::isinf(ll)) {
      ++backtracks;
      min_step /= 2.0;
      if (backtracks > params.max_backtrack) {
        likdif = 0.0;
        ll = ll_prev;
        double gb = 0.0, gw = 0.0;
        for (double v : cur.db) gb = std::max(gb, std::fabs(v));
        for (double v : cur.dw) gw = std::max(gw, std::fabs(v));
        if (gb > kGradientTolerance || gw > kGradientTolerance) ++backtracks;
        break;
      }
      for (std::size_t j = 0; j < nb; ++j) {
        double dd = step_beta[j];
        if (dd < kStepCutoff && dd > -kStepCutoff) {
          damp_beta[j] /= 2.0;
          new_beta[j] = beta[j] + dd * damp_beta[j];
        } else if (dd > 0.0) {
          new_beta[j] = beta[j] + min_step;
        } else {
          new_beta[j] = beta[j] - min_step;
        }
      }
      for (std::size_t s = 0; s < k; ++s) {
        double dd = step_log_w[s];
        if (dd < kStepCutoff && dd > -kStepCutoff) {
          damp_log_w[s] /= 2.0;
          new_log_w[s] = log_w[s] + dd * damp_log_w[s];
        } else if (dd > 0.0) {
          new_log_w[s] = log_w[s] + min_step;
        } else {
          new_log_w[s] = log_w[s] - min_step;
        }
      }
      sweep(d, new_beta.data(), new_log_w.data(), gamma, pen, stored, run, trial);
      ll = trial.ll;
      likdif = ll - ll_prev;
    }
    beta = new_beta;
    log_w = new_log_w;
    std::swap(cur, trial);
  }

  double log_det = 0.0;
  for (std::size_t s = 0; s < k; ++s) log_det += std::log(std::fabs(vw_step[s]));
  if (params.reml) {
    factor = vb2;
    log_det += std::log(std::fabs(determinant(factor.data(), nb)));
  }
  double second = 0.0;
  if (pen.gamma && params.ord > 1)
    second = laplace_correction(d, stored, run, beta.data(), log_w.data(), cur.w, gamma,
                                pen.lambda, params.ord);

  PmlResult r;
  r.beta = beta;
  r.log_w = log_w;
  r.information = vb2;
  r.log_likelihood = ll;
  r.log_likelihood_prev = ll_prev;
  r.log_det = log_det;
  r.second_order = second;
  r.iterations = step;
  r.backtracks = backtracks;
  r.convergence = check_convergence(r, params, false, 1);
  return r;
}

Penalty gamma_penalty(double subject) {
  double exps = std::exp(subject);
  if (!(std::isfinite(exps) && exps > 1.0))
    fail(2, "PmlVariance::subject must be finite and strictly positive for opt_pml");
  auto ab = gamma_prior(exps);
  return Penalty{true, ab.first, ab.second};
}

}  // namespace

std::pair<double, double> gamma_prior(double exps) {
  return {1.0 / (exps - 1.0), 1.0 / (std::sqrt(exps) * (exps - 1.0))};
}

PmlResult opt_pml(const PmlData& d, const double* beta_init, double subject,
                  double cell, const PmlParams& p, const double* log_w_init) {
  return optimise(d, beta_init, log_w_init, gamma_penalty(subject), cell, p);
}

PmlResult opt_pml_nbm(const PmlData& d, const double* beta_init,
                      double subject, double cell, const PmlParams& p) {
  if (!(std::isfinite(subject) && subject > 0.0)) fail(2, "subject must be positive");
  return optimise(d, beta_init, nullptr, Penalty{false, subject, 0.0}, cell, p);
}

void pml_log_likelihood_gradient(const PmlData& d, const double* beta,
                                 const double* log_w, double subject,
                                 double cell, double* objective,
                                 double* gradient) {
  validate(d);
  // nebula's pml_ll_der_eigen reads subject as both shape and rate
  // (pml.rs:586-596).
  Penalty pen{true, subject, subject};
  std::vector<double> stored(d.n_cells);
  auto run = positive_runs(d);
  Sweep s;
  s.size(d.nb, d.k);
  sweep(d, beta, log_w, cell, pen, stored, run, s);
  *objective = -s.ll;
  for (std::size_t j = 0; j < d.nb; ++j) gradient[j] = -s.db[j];
  for (std::size_t j = 0; j < d.k; ++j) gradient[d.nb + j] = -s.dw[j];
}

int check_convergence(const PmlResult& r, const PmlParams& p,
                      bool variance_at_bound, int initial) {
  int code = initial;
  if (code == 1) {
    if (variance_at_bound)
      code = -60;
    else if (std::isnan(r.log_likelihood))
      code = -30;
    else if (r.iterations == p.max_iter)
      code = -20;
    else if (r.backtracks == p.max_backtrack + 1)
      code = -10;
    else if (r.backtracks == p.max_backtrack + 2)
      code = -40;
    else
      code = 1;
  }
  std::size_t nb = r.beta.size();
  if (nb > 1 && min_eigenvalue(r.information.data(), nb) < kEigenCutoff) code = -25;
  return code;
}

void ldlt_solve(double* a, double* b, std::size_t n) {
  std::vector<std::size_t> perm(n);
  std::vector<double> tmp(n);
  for (std::size_t st = 0; st < n; ++st) {
    std::size_t piv = st;
    double best = std::fabs(a[st * n + st]);
    for (std::size_t i = st + 1; i < n; ++i) {
      double v = std::fabs(a[i * n + i]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
    perm[st] = piv;
    if (piv != st) {
      for (std::size_t j = 0; j < st; ++j) std::swap(a[st * n + j], a[piv * n + j]);
      for (std::size_t i = piv + 1; i < n; ++i) std::swap(a[i * n + st], a[i * n + piv]);
      std::swap(a[st * n + st], a[piv * n + piv]);
      for (std::size_t i = st + 1; i < piv; ++i) std::swap(a[i * n + st], a[piv * n + i]);
    }
    if (st > 0) {
      for (std::size_t j = 0; j < st; ++j) tmp[j] = a[j * n + j] * a[st * n + j];
      double acc = 0.0;
      for (std::size_t j = 0; j < st; ++j) acc += a[st * n + j] * tmp[j];
      a[st * n + st] -= acc;
      for (std::size_t i = st + 1; i < n; ++i) {
        double ac = 0.0;
        for (std::size_t j = 0; j < st; ++j) ac += a[i * n + j] * tmp[j];
        a[i * n + st] -= ac;
      }
    }
    double dd = a[st * n + st];
    if (dd != 0.0)
      for (std::size_t i = st + 1; i < n; ++i) a[i * n + st] /= dd;
  }
  for (std::size_t st = 0; st < n; ++st) std::swap(b[st], b[perm[st]]);
  for (std::size_t i = 0; i < n; ++i) {
    double acc = b[i];
    for (std::size_t j = 0; j < i; ++j) acc -= a[i * n + j] * b[j];
    b[i] = acc;
  }
  for (std::size_t i = 0; i < n; ++i) {
    double dd = a[i * n + i];
    b[i] = std::fabs(dd) > std::numeric_limits<double>::min() ? b[i] / dd : 0.0;
  }
  for (std::size_t ii = n; ii-- > 0;) {
    double acc = b[ii];
    for (std::size_t j = ii + 1; j < n; ++j) acc -= a[j * n + ii] * b[j];
    b[ii] = acc;
  }
  for (std::size_t st = n; st-- > 0;) std::swap(b[st], b[perm[st]]);
}

double determinant(double* a, std::size_t n) {
  double det = 1.0;
  for (std::size_t st = 0; st < n; ++st) {
    std::size_t piv = st;
    double best = std::fabs(a[st * n + st]);
    for (std::size_t i = st + 1; i < n; ++i) {
      double v = std::fabs(a[i * n + st]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
    if (piv != st) {
      for (std::size_t j = 0; j < n; ++j) std::swap(a[st * n + j], a[piv * n + j]);
      det = -det;
    }
    double dd = a[st * n + st];
    det *= dd;
    if (dd == 0.0) return 0.0;
    for (std::size_t i = st + 1; i < n; ++i) {
      double f = a[i * n + st] / dd;
      a[i * n + st] = f;
      for (std::size_t j = st + 1; j < n; ++j) a[i * n + j] -= f * a[st * n + j];
    }
  }
  return det;
}

double min_eigenvalue(const double* m, std::size_t n) {
  if (n == 0) return std::numeric_limits<double>::quiet_NaN();
  if (n == 1) return m[0];
  std::vector<double> a(m, m + n * n);
  double frob = 0.0;
  for (double v : a) frob += v * v;
  const double eps = std::numeric_limits<double>::epsilon();
  double tol = std::max(frob * eps * eps, std::numeric_limits<double>::min());
  for (int sweep_i = 0; sweep_i < kJacobiSweeps; ++sweep_i) {
    double off = 0.0;
    for (std::size_t p = 0; p < n; ++p)
      for (std::size_t q = p + 1; q < n; ++q) off += a[p * n + q] * a[p * n + q];
    if (off <= tol) break;
    for (std::size_t p = 0; p < n; ++p)
      for (std::size_t q = p + 1; q < n; ++q) {
        double apq = a[p * n + q];
        if (apq == 0.0) continue;
        double theta = (a[q * n + q] - a[p * n + p]) / (2.0 * apq);
        double sgn = std::signbit(theta) ? -1.0 : 1.0;
        double t = sgn / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        double c = 1.0 / std::sqrt(t * t + 1.0);
        double s = t * c;
        for (std::size_t r = 0; r < n; ++r) {
          double arp = a[r * n + p], arq = a[r * n + q];
          a[r * n + p] = c * arp - s * arq;
          a[r * n + q] = s * arp + c * arq;
        }
        for (std::size_t r = 0; r < n; ++r) {
          double apr = a[p * n + r], aqr = a[q * n + r];
          a[p * n + r] = c * apr - s * aqr;
          a[q * n + r] = s * apr + c * aqr;
        }
      }
  }
  double mn = std::numeric_limits<double>::infinity();
  for (std::size_t i = 0; i < n; ++i) mn = std::min(mn, a[i * n + i]);
  return mn;
}

}  // namespace oracle

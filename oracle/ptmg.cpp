// Stage one's marginal Poisson-gamma likelihood in (beta, sigma, phi), f64:
// a restatement of src/sc/ptmg.rs (nebula's ptmg_ll_der_hes_eigen plus the
// gamma-function corrections of R/ptmg.R).
//
// Terms          <- ptmg.rs:288-348  (reparametrised scalars, literal exp - 1)
// gamma_terms    <- ptmg.rs:404-478
// value+gradient <- ptmg.rs:661-896 (linear_predictor, Subjects, evaluate,
//                   gradient_parts, finish_gradient)
// tabled value+gradient <- ptmg.rs:1462-1580 (zero counts from the tables)
// Hessian        <- ptmg.rs:999-1252 (evaluate_hessian) with the per-subject
//                   sums of hessian_subject (:1310-1444)
// Pinned by the nebula 1.5.8 values quoted in ptmg.rs:2163-2445 and by finite
// differences (tests/test_oracle_ptmg.py).
#include <algorithm>
#include <cmath>
#include <vector>

#include "oracle.hpp"

namespace oracle {

namespace {

struct Terms {
  double exps_s, alpha, lambda, gamma, log_lambda, log_gamma, alpha_pr,
      lambda_pr, alpha_dpr, lambda_dpr;
  Terms(double sigma, double phi) {
    double exps = std::exp(sigma);
    double m1 = exps - 1.0;
    exps_s = std::sqrt(exps);
    alpha = 1.0 / m1;
    lambda = alpha / exps_s;
    double sq = m1 * m1;
    gamma = phi;
    log_lambda = std::log(lambda);
    log_gamma = std::log(phi);
    alpha_pr = -exps / sq;
    lambda_pr = (1.0 - 3.0 * exps) / (2.0 * exps_s * sq);
    alpha_dpr = 2.0 * exps * exps / (sq * m1) - exps / sq;
    lambda_dpr = -3.0 * exps / (2.0 * exps_s * sq) +
                 (3.0 * exps - 1.0) / (4.0 * exps_s * sq) +
                 (3.0 * exps - 1.0) * exps_s / (sq * m1);
  }
};

struct GammaTerms {
  double sv = 0, sd = 0, st = 0, cv = 0, cd = 0, ct = 0;
};

GammaTerms gamma_terms(const GeneData& g, double alpha, double gamma, int order) {
  GammaTerms o;
  double npos = 0, ln_sum = 0, di_sum = 0, tri_sum = 0;
  for (std::size_t s = 0; s < g.k; ++s) {
    double t = g.subject_totals[s];
    if (t > 0.0) {
      npos += 1.0;
      ln_sum += ln_gamma(t + alpha);
      if (order >= 1) di_sum += digamma(t + alpha);
      if (order >= 2) tri_sum += trigamma(t + alpha);
    }
  }
  o.sv = ln_sum - npos * ln_gamma(alpha);
  if (order >= 1) o.sd = di_sum - npos * digamma(alpha);
  if (order >= 2) o.st = tri_sum - npos * trigamma(alpha);
  double n1 = 0, n2 = 0, nb = 0, lb = 0, db = 0, tb = 0;
  for (std::size_t i = 0; i < g.nnz; ++i) {
    double y = g.counts[i];
    if (y == 1.0) {
      n1 += 1.0;
    } else if (y == 2.0) {
      n2 += 1.0;
    } else {
      nb += 1.0;
      lb += ln_gamma(y + gamma);
      if (order >= 1) db += digamma(y + gamma);
      if (order >= 2) tb += trigamma(y + gamma);
    }
  }
  double ns = n1 + n2;
  o.cv = lb - nb * ln_gamma(gamma) + ns * std::log(gamma) + n2 * std::log(gamma + 1.0);
  if (order >= 1) o.cd = db - nb * digamma(gamma) + ns / gamma + n2 / (gamma + 1.0);
  if (order >= 2)
    o.ct = tb - nb * trigamma(gamma) - ns / (gamma * gamma) -
           n2 / ((gamma + 1.0) * (gamma + 1.0));
  return o;
}

struct Subjects {
  std::vector<double> mlog, ym, ymm, imu;
};

Subjects subjects(const GeneData& g, const Terms& t, const std::vector<double>& csx,
                  double& total) {
  Subjects o;
  o.mlog.resize(g.k);
  o.ym.resize(g.k);
  o.ymm.resize(g.k);
  o.imu.resize(g.k);
  for (std::size_t s = 0; s < g.k; ++s) {
    double ystar = g.subject_totals[s] + t.alpha;
    double mu = csx[s] + t.lambda;
    double lm = std::log(mu);
    o.mlog[s] = lm;
    o.ym[s] = ystar / mu;
    o.ymm[s] = ystar / mu / mu;
    o.imu[s] = 1.0 / mu;
    total -= ystar * lm;
  }
  total += (double)g.k * t.alpha * t.log_lambda;
  total += (double)g.n_cells * t.gamma * t.log_gamma;
  return o;
}

struct GradParts {
  std::vector<double> grad, ymm_d, hbl0;
  double dbim = 0, sum_ymm_d = 0;
};

GradParts gradient_parts(const GeneData& g, const Terms& t, const Subjects& sj,
                         const std::vector<double>& xf, const std::vector<double>& d41,
                         const std::vector<double>& d42c, double slpey,
                         double gstar_sum) {
  const std::size_t nb = g.n_coef, k = g.k;
  const double n = (double)g.n_cells;
  GradParts o;
  o.ymm_d.resize(k);
  for (std::size_t s = 0; s < k; ++s) o.ymm_d[s] = sj.ymm[s] * d42c[s];
  o.grad.assign(nb + 2, 0.0);
  for (std::size_t j = 0; j < nb; ++j) {
    double acc = 0.0;
    for (std::size_t s = 0; s < k; ++s)
      acc += xf[s * nb + j] * o.ymm_d[s] - d41[s * nb + j] * sj.ym[s];
    o.grad[j] = acc;
  }
  for (std::size_t i = 0; i < g.nnz; ++i) {
    const double* row = g.design + g.cells[i] * nb;
    for (std::size_t j = 0; j < nb; ++j) o.grad[j] += row[j] * g.counts[i];
  }
  double smlog = 0, sym = 0;
  for (std::size_t s = 0; s < k; ++s) {
    smlog += sj.mlog[s];
    sym += sj.ym[s];
  }
  double ldm = t.log_lambda * (double)k - smlog;
  double adlmy = t.exps_s * (double)k - sym;
  o.hbl0.resize(k);
  for (std::size_t s = 0; s < k; ++s) o.hbl0[s] = d42c[s] * sj.imu[s];
  for (std::size_t s = 0; s < k; ++s) o.dbim += o.hbl0[s];
  for (std::size_t s = 0; s < k; ++s) o.sum_ymm_d += o.ymm_d[s];
  o.grad[nb] = -t.alpha_pr * o.dbim + t.lambda_pr * o.sum_ymm_d + t.alpha_pr * ldm +
               t.lambda_pr * adlmy;
  o.grad[nb + 1] = t.log_gamma * n + n - slpey - gstar_sum;
  return o;
}

}  // namespace

Ptmg::Ptmg(const GeneData& g) : g_(g) {
  y_cell_.assign(g.n_cells, 0.0);
  for (std::size_t i = 0; i < g.nnz; ++i) y_cell_[g.cells[i]] = g.counts[i];
  extb_.assign(g.n_cells, 0.0);
  cumsumxtb_.assign(g.k, 0.0);
}

double Ptmg::linear_predictor(const double* beta) {
  const std::size_t nb = g_.n_coef;
  double total = 0.0;
  for (std::size_t s = 0; s < g_.k; ++s) {
    double csx = 0.0;
    for (std::size_t i = g_.fid[s]; i < g_.fid[s + 1]; ++i) {
      const double* row = g_.design + i * nb;
      double eta = g_.log_offset[i];
      for (std::size_t j = 0; j < nb; ++j) eta += row[j] * beta[j];
      double y = y_cell_[i];
      if (y != 0.0) total += eta * y;
      double e = std::exp(eta);
      extb_[i] = e;
      csx += e;
    }
    cumsumxtb_[s] = csx;
  }
  return total;
}

double Ptmg::value(const double* params) {
  std::vector<double> grad(g_.n_coef + 2);
  return value_gradient(params, grad.data());
}

// Value and gradient with the zero counts summed from the run's tables
// (ptmg.rs:1462-1580): per group, sum ym e is ym exp(x'beta) sum O exactly and
// the log terms are L and A at s = ln(ym) + x'beta - ln(gamma); loose and
// positive cells one by one.
double Ptmg::value_gradient_tabled(const double* params, double* grad_out) {
  const std::size_t nb = g_.n_coef, k = g_.k;
  const ZeroCells& z = *g_.zeros;
  const double* beta = params;
  Terms t(params[nb], params[nb + 1]);
  const double gamma = t.gamma, ln_gamma = t.log_gamma;
  auto row_of = [&](std::size_t r) { return g_.design + r * nb; };
  auto dot = [&](const double* row) {
    double acc = 0.0;
    for (std::size_t j = 0; j < nb; ++j) acc += row[j] * beta[j];
    return acc;
  };
  // positive entries of each subject: [run[s], run[s + 1])
  std::vector<std::size_t> run(k + 1);
  for (std::size_t s = 0; s <= k; ++s)
    run[s] = std::lower_bound(g_.cells, g_.cells + g_.nnz, g_.fid[s]) - g_.cells;
  double total = 0.0;
  std::vector<double> group_xb(z.groups.size(), 0.0), xf(k * nb, 0.0);
  for (std::size_t s = 0; s < k; ++s) {
    double* xs = &xf[s * nb];
    double csx = 0.0;
    for (std::size_t gi = z.subject_groups[s]; gi < z.subject_groups[s + 1]; ++gi) {
      const ZeroGroup& g = z.groups[gi];
      const double* row = row_of(g.row);
      const double xb = dot(row);
      group_xb[gi] = xb;
      const double e = std::exp(xb) * g.sum_o;
      csx += e;
      for (std::size_t j = 0; j < nb; ++j) xs[j] += row[j] * e;
    }
    for (std::size_t li = z.subject_loose[s]; li < z.subject_loose[s + 1]; ++li) {
      const std::size_t r = z.loose[li];
      const double* row = row_of(r);
      const double e = std::exp(g_.log_offset[r] + dot(row));
      extb_[r] = e;
      csx += e;
      for (std::size_t j = 0; j < nb; ++j) xs[j] += row[j] * e;
    }
    for (std::size_t i = run[s]; i < run[s + 1]; ++i) {
      const std::size_t c = g_.cells[i];
      const double eta = g_.log_offset[c] + dot(row_of(c));
      total += eta * g_.counts[i];
      extb_[c] = std::exp(eta);
    }
    cumsumxtb_[s] = csx;
  }
  Subjects sj = subjects(g_, t, cumsumxtb_, total);
  double slpey = 0.0, gstar_sum = 0.0;
  std::vector<double> d41(k * nb, 0.0), d42c(k, 0.0);
  for (std::size_t s = 0; s < k; ++s) {
    const double ym = sj.ym[s];
    const double ln_ym = std::log(ym);
    double* ds = &d41[s * nb];
    double acc = 0.0;
    total += ym * cumsumxtb_[s];
    for (std::size_t gi = z.subject_groups[s]; gi < z.subject_groups[s + 1]; ++gi) {
      const ZeroGroup& g = z.groups[gi];
      const double* row = row_of(g.row);
      double la[2];
      g.eval(0, 2, ln_ym + group_xb[gi] - ln_gamma, la);
      const double sum_log_w = g.n * ln_gamma + la[0];
      total -= gamma * sum_log_w;
      slpey += sum_log_w;
      gstar_sum += g.n - la[1];
      const double ge = gamma / ym * la[1];  // sum gamma e / w
      acc += ge;
      for (std::size_t j = 0; j < nb; ++j) ds[j] += row[j] * ge;
    }
    auto cell = [&](std::size_t r, double y, bool dense) {
      const double e = extb_[r];
      const double w = ym * e + gamma;
      const double lw = std::log(w);
      const double gw = dense ? gamma / w : y / w;
      if (dense) {
        total -= gamma * lw;
        slpey += lw;
      } else {
        total -= y * lw;
      }
      gstar_sum += gw;
      const double ge = gw * e;
      acc += ge;
      const double* row = row_of(r);
      for (std::size_t j = 0; j < nb; ++j) ds[j] += row[j] * ge;
    };
    for (std::size_t li = z.subject_loose[s]; li < z.subject_loose[s + 1]; ++li)
      cell(z.loose[li], 0.0, true);
    for (std::size_t i = run[s]; i < run[s + 1]; ++i) cell(g_.cells[i], g_.counts[i], false);
    d42c[s] = acc - cumsumxtb_[s];
  }
  GradParts gp = gradient_parts(g_, t, sj, xf, d41, d42c, slpey, gstar_sum);
  GammaTerms ex = gamma_terms(g_, t.alpha, gamma, 1);
  for (std::size_t j = 0; j < nb + 2; ++j) grad_out[j] = -gp.grad[j];
  grad_out[nb] -= t.alpha_pr * ex.sd;
  grad_out[nb + 1] -= ex.cd;
  return -total - ex.sv - ex.cv;
}

double Ptmg::value_gradient(const double* params, double* grad_out) {
  if (g_.zeros) return value_gradient_tabled(params, grad_out);
  const std::size_t nb = g_.n_coef, k = g_.k;
  Terms t(params[nb], params[nb + 1]);
  const double gamma = t.gamma;
  double total = linear_predictor(params);
  Subjects sj = subjects(g_, t, cumsumxtb_, total);
  std::vector<double> xf(k * nb, 0.0), d41(k * nb, 0.0), d42c(k, 0.0);
  double slpey = 0.0, gstar_sum = 0.0;
  for (std::size_t s = 0; s < k; ++s) {
    double ym = sj.ym[s];
    double acc = 0.0;
    for (std::size_t i = g_.fid[s]; i < g_.fid[s + 1]; ++i) {
      double e = extb_[i];
      double we = ym * e;
      double w = we + gamma;
      double lw = std::log(w);
      double iw = 1.0 / w;
      total += we;
      total -= (gamma + y_cell_[i]) * lw;
      slpey += lw;
      double gg = (gamma + y_cell_[i]) * iw;
      gstar_sum += gg;
      const double* row = g_.design + i * nb;
      for (std::size_t j = 0; j < nb; ++j) {
        double v = row[j] * e;
        xf[s * nb + j] += v;
        d41[s * nb + j] += gg * v;
      }
      acc += gg * e;
    }
    d42c[s] = acc - cumsumxtb_[s];
  }
  GradParts gp = gradient_parts(g_, t, sj, xf, d41, d42c, slpey, gstar_sum);
  GammaTerms ex = gamma_terms(g_, t.alpha, gamma, 1);
  for (std::size_t j = 0; j < nb + 2; ++j) grad_out[j] = -gp.grad[j];
  grad_out[nb] -= t.alpha_pr * ex.sd;
  grad_out[nb + 1] -= ex.cd;
  return -total - ex.sv - ex.cv;
}

double Ptmg::value_gradient_hessian(const double* params, double* grad_out,
                                    double* hess) {
  const std::size_t nb = g_.n_coef, k = g_.k, nn = nb * nb, np = nb + 2;
  Terms t(params[nb], params[nb + 1]);
  const double gamma = t.gamma;
  double total = linear_predictor(params);
  Subjects sj = subjects(g_, t, cumsumxtb_, total);
  std::vector<double> xf(k * nb, 0.0), d41(k * nb, 0.0), t1(k * nb, 0.0),
      t2(k * nb, 0.0), bp(k * nn, 0.0), bg(k * nn, 0.0), bgte(k * nn, 0.0),
      hb2f(k), b2f(k), d42c(k);
  double slpey = 0, gstar_sum = 0, sum_tempa = 0, sum_gt = 0;
  for (std::size_t s = 0; s < k; ++s) {
    double ym = sj.ym[s];
    double acc = 0, hb2 = 0, b2 = 0;
    double* xs = &xf[s * nb];
    double* ds = &d41[s * nb];
    double* t1s = &t1[s * nb];
    double* t2s = &t2[s * nb];
    double* ps = &bp[s * nn];
    double* gs = &bg[s * nn];
    double* es = &bgte[s * nn];
    for (std::size_t i = g_.fid[s]; i < g_.fid[s + 1]; ++i) {
      double e = extb_[i], y = y_cell_[i];
      double we = ym * e;
      double w = we + gamma;
      double lw = std::log(w);
      double iw = 1.0 / w;
      total += we;
      total -= (gamma + y) * lw;
      slpey += lw;
      double gg = (gamma + y) * iw;
      gstar_sum += gg;
      double gt = gg * iw;
      double gte = gt * e;
      sum_tempa += iw;
      sum_gt += gt;
      acc += gg * e;
      hb2 += gte - e * iw;
      b2 += gte * e;
      double f1 = gt - iw;
      const double* row = g_.design + i * nb;
      for (std::size_t a = 0; a < nb; ++a) {
        double xe = row[a] * e;
        xs[a] += xe;
        ds[a] += gg * xe;
        t1s[a] += xe * f1;
        t2s[a] += xe * gte;
        for (std::size_t b = a; b < nb; ++b) {
          double c = xe * row[b];
          ps[a * nb + b] += c;
          gs[a * nb + b] += c * gg;
          es[a * nb + b] += c * gte;
        }
      }
    }
    hb2f[s] = hb2;
    b2f[s] = b2;
    d42c[s] = acc - cumsumxtb_[s];
  }
  GradParts gp = gradient_parts(g_, t, sj, xf, d41, d42c, slpey, gstar_sum);
  double value = -total;
  for (std::size_t j = 0; j < np; ++j) grad_out[j] = -gp.grad[j];

  std::vector<double> h(np * np, 0.0);
  const double lr = t.lambda_pr / t.lambda;
  double hs = (t.alpha_dpr * t.log_lambda + 2.0 * t.alpha_pr * lr - t.alpha * lr * lr +
               t.alpha / t.lambda * t.lambda_dpr) *
              (double)k;
  double s_hbl0_imu = 0, s_ymmd_imu = 0, s_imu = 0, s_mlog = 0, s_ym = 0, s_ymm = 0;
  for (std::size_t s = 0; s < k; ++s) {
    s_hbl0_imu += gp.hbl0[s] * sj.imu[s];
    s_ymmd_imu += gp.ymm_d[s] * sj.imu[s];
  }
  hs -= gp.dbim * t.alpha_dpr - 2.0 * t.alpha_pr * t.lambda_pr * s_hbl0_imu -
        t.lambda_dpr * gp.sum_ymm_d + 2.0 * t.lambda_pr * t.lambda_pr * s_ymmd_imu;
  for (std::size_t s = 0; s < k; ++s) {
    s_imu += sj.imu[s];
    s_mlog += sj.mlog[s];
    s_ym += sj.ym[s];
    s_ymm += sj.ymm[s];
  }
  hs -= 2.0 * t.alpha_pr * t.lambda_pr * s_imu + t.alpha_dpr * s_mlog +
        t.lambda_dpr * s_ym - t.lambda_pr * t.lambda_pr * s_ymm;

  std::vector<double> hsb(nb), hgb(nb), apm(k);
  for (std::size_t j = 0; j < nb; ++j) {
    double d41_imu = 0, d41_ymm = 0, xh_imu = 0, xh_ymm = 0;
    for (std::size_t s = 0; s < k; ++s) {
      d41_imu += d41[s * nb + j] * sj.imu[s];
      d41_ymm += d41[s * nb + j] * sj.ymm[s];
      xh_imu += xf[s * nb + j] * gp.hbl0[s] * sj.imu[s];
      xh_ymm += xf[s * nb + j] * gp.hbl0[s] * sj.ymm[s];
    }
    hsb[j] = -t.alpha_pr * d41_imu + t.lambda_pr * d41_ymm -
             (-t.alpha_pr * xh_imu + 2.0 * t.lambda_pr * xh_ymm);
  }
  for (std::size_t s = 0; s < k; ++s) apm[s] = t.alpha_pr * sj.imu[s] - t.lambda_pr * sj.ymm[s];
  double hg = (double)g_.n_cells / gamma - 2.0 * sum_tempa + sum_gt;
  for (std::size_t j = 0; j < nb; ++j) {
    double a = 0, b = 0;
    for (std::size_t s = 0; s < k; ++s) {
      a += t1[s * nb + j] * sj.ym[s];
      b += xf[s * nb + j] * hb2f[s] * sj.ymm[s];
    }
    hgb[j] = a - b;
  }
  double hsg = 0.0;
  for (std::size_t s = 0; s < k; ++s) hsg += apm[s] * hb2f[s];
  for (std::size_t j = 0; j < nb; ++j) {
    double a = 0, b = 0;
    for (std::size_t s = 0; s < k; ++s) {
      a += t2[s * nb + j] * apm[s] * sj.ym[s];
      b += xf[s * nb + j] * apm[s] * b2f[s] * sj.ymm[s];
    }
    hsb[j] += a - b;
  }
  {
    double add = 0.0;
    for (std::size_t s = 0; s < k; ++s) add += b2f[s] * apm[s] * apm[s];
    hs += add;
  }
  std::vector<double> ymu2(k), yyi(k), b2s(k), ymi(k);
  for (std::size_t s = 0; s < k; ++s) {
    ymu2[s] = sj.ym[s] * sj.ym[s];
    yyi[s] = ymu2[s] * sj.imu[s];
    b2s[s] = b2f[s] * yyi[s] * sj.imu[s];
    ymi[s] = gp.ymm_d[s] * sj.imu[s];
  }
  for (std::size_t p = 0; p < nb; ++p)
    for (std::size_t q = p; q < nb; ++q) {
      double acc = 0.0;
      for (std::size_t s = 0; s < k; ++s) {
        std::size_t at = s * nn + p * nb + q;
        acc -= bg[at] * sj.ym[s];
        acc += bgte[at] * ymu2[s];
        acc += gp.ymm_d[s] * bp[at];
      }
      for (std::size_t s = 0; s < k; ++s) {
        double xp = xf[s * nb + p], xq = xf[s * nb + q];
        acc += (d41[s * nb + p] * xq + d41[s * nb + q] * xp) * sj.ymm[s];
        acc -= yyi[s] * (t2[s * nb + p] * xq + t2[s * nb + q] * xp);
        double cross = xp * xq;
        acc -= cross * sj.ymm[s];
        acc += b2s[s] * cross;
        acc -= 2.0 * cross * ymi[s];
      }
      h[p * np + q] = acc;
      h[q * np + p] = acc;
    }
  h[nb * np + nb] = hs;
  h[(nb + 1) * np + nb + 1] = hg;
  h[nb * np + nb + 1] = hsg;
  h[(nb + 1) * np + nb] = hsg;
  for (std::size_t j = 0; j < nb; ++j) {
    h[nb * np + j] = hsb[j];
    h[j * np + nb] = hsb[j];
    h[(nb + 1) * np + j] = hgb[j];
    h[j * np + nb + 1] = hgb[j];
  }
  for (double& v : h) v = -v;
  GammaTerms ex = gamma_terms(g_, t.alpha, gamma, 2);
  value = value - ex.sv - ex.cv;
  grad_out[nb] -= t.alpha_pr * ex.sd;
  grad_out[nb + 1] -= ex.cd;
  h[nb * np + nb] -= t.alpha_dpr * ex.sd + t.alpha_pr * t.alpha_pr * ex.st;
  h[(nb + 1) * np + nb + 1] -= ex.ct;
  for (std::size_t i = 0; i < np * np; ++i) hess[i] = h[i];
  return value;
}

}  // namespace oracle

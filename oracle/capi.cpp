// C entry points of the oracle library, for the ctypes checker in
// oracle/oracle.py.  Test infrastructure only (see oracle.hpp).
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/nebula_hip.h"
#include "lbfgsb.hpp"
#include "oracle.hpp"

using namespace oracle;

namespace {
void put_err(char* err, std::size_t len, const std::string& m) {
  if (err && len) {
    std::strncpy(err, m.c_str(), len - 1);
    err[len - 1] = 0;
  }
}
template <class T>
T* dup(const std::vector<T>& v) {
  T* p = (T*)std::malloc(sizeof(T) * (v.empty() ? 1 : v.size()));
  if (!v.empty()) std::memcpy(p, v.data(), sizeof(T) * v.size());
  return p;
}
PmlData make_pml(const double* design, const double* offset, std::size_t n_cells,
                 std::size_t nb, const double* counts, const std::uint64_t* cells,
                 std::size_t nnz, const std::uint64_t* start, std::size_t k,
                 const double* totals, std::vector<std::size_t>& c,
                 std::vector<std::size_t>& s) {
  c.assign(cells, cells + nnz);
  s.assign(start, start + k + 1);
  return PmlData{design, offset, n_cells, nb, counts, c.data(), nnz, s.data(), k, totals};
}
}  // namespace

extern "C" {

int oracle_nebula_sparse(std::size_t n_genes, std::size_t n_cells, const std::uint32_t* indptr,
                         const std::uint32_t* indices, const double* data,
                         const std::uint64_t* subject_id, const double* design,
                         std::size_t n_coef, const double* offset, const nebula_params_t* p,
                         int threads, nebula_fit_t* out, double* stage_two_info,
                         double* t_stage_two, int tables, char* err, std::size_t err_len) {
  Params prm;
  prm.tables = tables != 0;
  if (p) {
    prm.min_subject = p->min_subject;
    prm.min_cell = p->min_cell;
    prm.max_subject = p->max_subject;
    prm.max_cell = p->max_cell;
    prm.method = p->method;
    prm.reml = p->reml != 0;
    prm.cutoff_cell = p->cutoff_cell;
    prm.kappa = p->kappa;
    prm.cpc = p->cpc;
    prm.mincp = p->mincp;
    prm.min_subjects = p->min_subjects;
    prm.eps = p->eps;
  }
  try {
    std::vector<StageTwoInfo> info;
    Fit f = nebula_sparse(n_genes, n_cells, indptr, indices, data, subject_id, design, n_coef,
                          offset, prm, threads, stage_two_info ? &info : nullptr, t_stage_two);
    out->n_genes = f.gene_index.size();
    out->n_coef = f.n_coef;
    out->coefficients = dup(f.coefficients);
    out->covariance = dup(f.covariance);
    out->se = dup(f.se);
    out->subject_overdispersion = dup(f.subject_overdispersion);
    out->cell_overdispersion = dup(f.cell_overdispersion);
    std::vector<std::int32_t> conv(f.convergence.begin(), f.convergence.end());
    out->convergence = dup(conv);
    out->sigma_at_bound = dup(f.sigma_at_bound);
    std::vector<std::uint64_t> gi(f.gene_index.begin(), f.gene_index.end());
    out->gene_index = dup(gi);
    // Per kept gene: searched, sigma, gamma, objective, evaluations, ord,
    // start sigma, start gamma (8 doubles).
    if (stage_two_info)
      for (std::size_t i = 0; i < info.size(); ++i) {
        double* r = stage_two_info + 8 * i;
        r[0] = info[i].searched;
        r[1] = info[i].sigma;
        r[2] = info[i].gamma;
        r[3] = info[i].objective;
        r[4] = info[i].evaluations;
        r[5] = info[i].ord;
        r[6] = info[i].start_sigma;
        r[7] = info[i].start_gamma;
      }
    return 0;
  } catch (const Error& e) {
    put_err(err, err_len, e.message);
    return e.code;
  }
}

int oracle_profile_objective(std::size_t n_genes, std::size_t n_cells,
                             const std::uint32_t* indptr, const std::uint32_t* indices,
                             const double* data, const std::uint64_t* subject_id,
                             const double* design, std::size_t n_coef, const double* offset,
                             const nebula_params_t* p, std::size_t n, const std::uint64_t* gene,
                             const double* point, double* out, int threads, int tables,
                             char* err, std::size_t err_len) {
  Params prm;
  prm.tables = tables != 0;
  if (p) {
    prm.min_subject = p->min_subject;
    prm.min_cell = p->min_cell;
    prm.max_subject = p->max_subject;
    prm.max_cell = p->max_cell;
    prm.method = p->method;
    prm.cutoff_cell = p->cutoff_cell;
    prm.kappa = p->kappa;
    prm.cpc = p->cpc;
    prm.mincp = p->mincp;
    prm.min_subjects = p->min_subjects;
    prm.eps = p->eps;
  }
  try {
    profile_objective(n_genes, n_cells, indptr, indices, data, subject_id, design, n_coef, offset,
                      prm, n, gene, point, out, threads);
    return 0;
  } catch (const Error& e) {
    put_err(err, err_len, e.message);
    return e.code;
  }
}

// StageTwoSearch on a caller's objective f(x, n, order, call, ctx), the
// reference's run_search harness (src/sc/nebula.rs:3516-3542): n = 2 runs
// BOBYQA, n = 1 the simplex search.  Returns 1 when a minimiser was found
// (written to x_out), 0 when the search was abandoned.
typedef double (*oracle_search_fn)(const double* x, int n, int order, int call, void* ctx);
int oracle_stage_two_search(int n, const double* start, const double* lo, const double* hi,
                            int ord, oracle_search_fn f, void* ctx, double* x_out,
                            int* failed, int* order_out, int* evaluations) {
  int calls = 0;
  std::vector<double> s(start, start + n), l(lo, lo + n), h(hi, hi + n);
  SearchOutcome r = stage_two_search(s, l, h, (unsigned)ord,
                                     [&](const std::vector<double>& x, unsigned order) {
                                       ++calls;
                                       return f(x.data(), n, (int)order, calls, ctx);
                                     });
  if (r.found)
    for (int j = 0; j < n; ++j) x_out[j] = r.x[j];
  *failed = r.failed ? 1 : 0;
  *order_out = (int)r.order;
  *evaluations = r.evaluations;
  return r.found ? 1 : 0;
}

void oracle_fit_free(nebula_fit_t* f) {
  std::free(f->coefficients);
  std::free(f->covariance);
  std::free(f->se);
  std::free(f->subject_overdispersion);
  std::free(f->cell_overdispersion);
  std::free(f->convergence);
  std::free(f->sigma_at_bound);
  std::free(f->gene_index);
  std::memset(f, 0, sizeof(*f));
}

// opt_pml / opt_pml_nbm.  scalars: ll, ll_prev, log_det, second;
// ints: convergence, iterations, backtracks.
int oracle_opt_pml(const double* design, const double* offset, std::size_t n_cells,
                   std::size_t nb, const double* counts, const std::uint64_t* cells,
                   std::size_t nnz, const std::uint64_t* start, std::size_t k,
                   const double* totals, const double* beta_init, const double* log_w_init,
                   double subject, double cell, int reml, double eps, int ord, int max_iter,
                   int nbm, int tabled, double* beta, double* log_w, double* information,
                   double* scalars, int* ints, char* err, std::size_t err_len) {
  std::vector<std::size_t> c, s;
  PmlData d = make_pml(design, offset, n_cells, nb, counts, cells, nnz, start, k, totals, c, s);
  ZeroCells zc;
  if (tabled) {
    if (!build_zero_cells(design, offset, n_cells, s.data(), k, nb, zc)) {
      put_err(err, err_len, "zero-count tables do not apply to this design");
      return 3;
    }
    d.zeros = &zc;
  }
  PmlParams p;
  p.reml = reml != 0;
  p.eps = eps;
  p.ord = (unsigned)ord;
  if (max_iter > 0) p.max_iter = (std::size_t)max_iter;
  try {
    PmlResult r = nbm ? opt_pml_nbm(d, beta_init, subject, cell, p)
                      : opt_pml(d, beta_init, subject, cell, p, log_w_init);
    std::memcpy(beta, r.beta.data(), nb * sizeof(double));
    std::memcpy(log_w, r.log_w.data(), k * sizeof(double));
    std::memcpy(information, r.information.data(), nb * nb * sizeof(double));
    scalars[0] = r.log_likelihood;
    scalars[1] = r.log_likelihood_prev;
    scalars[2] = r.log_det;
    scalars[3] = r.second_order;
    ints[0] = r.convergence;
    ints[1] = (int)r.iterations;
    ints[2] = (int)r.backtracks;
    return 0;
  } catch (const Error& e) {
    put_err(err, err_len, e.message);
    return e.code;
  }
}

int oracle_pml_ll_grad(const double* design, const double* offset, std::size_t n_cells,
                       std::size_t nb, const double* counts, const std::uint64_t* cells,
                       std::size_t nnz, const std::uint64_t* start, std::size_t k,
                       const double* totals, const double* beta, const double* log_w,
                       double subject, double cell, double* objective, double* gradient) {
  std::vector<std::size_t> c, s;
  PmlData d = make_pml(design, offset, n_cells, nb, counts, cells, nnz, start, k, totals, c, s);
  try {
    pml_log_likelihood_gradient(d, beta, log_w, subject, cell, objective, gradient);
    return 0;
  } catch (const Error& e) {
    return e.code;
  }
}

// ptmg value / gradient / Hessian at params ([beta, sigma, phi]).
int oracle_ptmg(const double* design, const double* log_offset, std::size_t n_cells,
                std::size_t n_coef, const double* counts, const std::uint64_t* cells,
                std::size_t nnz, const double* totals, const std::uint64_t* fid, std::size_t k,
                const double* params, int tabled, double* value, double* grad, double* hess) {
  std::vector<std::size_t> c(cells, cells + nnz), f(fid, fid + k + 1);
  GeneData g{design, log_offset, n_cells, n_coef, counts, c.data(), nnz, totals, f.data(), k};
  ZeroCells zc;
  if (tabled) {
    if (!build_zero_cells(design, log_offset, n_cells, f.data(), k, n_coef, zc)) return 3;
    g.zeros = &zc;
  }
  Ptmg pt(g);
  if (hess)
    *value = pt.value_gradient_hessian(params, grad, hess);
  else
    *value = pt.value_gradient(params, grad);
  return 0;
}

// One zero-count group built from log offsets log_o[n]: its table evaluated
// at ns points (m sums from `from`, row-major ns x m), the compensated direct
// sums there (ns x 6) when `direct` is given, and (s_lo, s_hi, n_base,
// n_panels) in info[4].
void oracle_zero_group(const double* log_o, std::size_t n, const double* s, std::size_t ns,
                       int from, int m, double* out, double* direct, double* info) {
  std::vector<double> lo(log_o, log_o + n);
  ZeroGroup g = build_zero_group(0, lo);
  for (std::size_t i = 0; i < ns; ++i) g.eval(from, m, s[i], out + i * (std::size_t)m);
  if (direct) {
    std::vector<std::pair<double, double>> distinct;
    for (double l : lo) distinct.push_back({l, 1.0});
    for (std::size_t i = 0; i < ns; ++i) zero_direct_sums(distinct, s[i], direct + i * 6);
  }
  if (info) {
    info[0] = g.s_lo;
    info[1] = g.s_hi;
    info[2] = (double)g.n_base;
    info[3] = (double)(g.panels.size() / (6 * 17));
  }
}

// ZeroCells::build on a design: 1 when built, 0 when the tables do not apply.
// Outputs (when built): subject_groups[k + 1], subject_loose[k + 1],
// loose[n_cells] (the first subject_loose[k] entries), group_row and group_n
// (n_groups entries, at most n_cells).
int oracle_zero_cells(const double* design, const double* log_offset, std::size_t n_cells,
                      const std::uint64_t* fid, std::size_t k, std::size_t n_coef,
                      std::uint64_t* subject_groups, std::uint64_t* subject_loose,
                      std::uint64_t* loose, std::uint64_t* group_row, double* group_n) {
  std::vector<std::size_t> f(fid, fid + k + 1);
  ZeroCells z;
  if (!build_zero_cells(design, log_offset, n_cells, f.data(), k, n_coef, z)) return 0;
  for (std::size_t s = 0; s <= k; ++s) {
    subject_groups[s] = z.subject_groups[s];
    subject_loose[s] = z.subject_loose[s];
  }
  for (std::size_t i = 0; i < z.loose.size(); ++i) loose[i] = z.loose[i];
  for (std::size_t i = 0; i < z.groups.size(); ++i) {
    group_row[i] = z.groups[i].row;
    group_n[i] = z.groups[i].n;
  }
  return 1;
}

void oracle_special(int which, const double* x, double* y, std::size_t n) {
  for (std::size_t i = 0; i < n; ++i)
    y[i] = which == 0 ? ln_gamma(x[i]) : which == 1 ? digamma(x[i]) : trigamma(x[i]);
}

double oracle_min_eigenvalue(const double* m, std::size_t n) { return min_eigenvalue(m, n); }
void oracle_ldlt_solve(double* a, double* b, std::size_t n) { ldlt_solve(a, b, n); }
double oracle_determinant(double* a, std::size_t n) { return determinant(a, n); }

typedef double (*fg_fn)(const double* x, double* g, void* ctx);
int oracle_lbfgsb(std::size_t n, const double* x0, const double* lo, const double* hi,
                  double ftol, double pgtol, int max_iter, fg_fn f, void* ctx, double* x_out,
                  double* f_out) {
  std::vector<double> x(x0, x0 + n), l(lo, lo + n), h(hi, hi + n);
  auto fg = [&](const std::vector<double>& xx, std::vector<double>& gg) {
    return f(xx.data(), gg.data(), ctx);
  };
  LbfgsbResult r = lbfgsb_minimise(fg, x, l, h, ftol, pgtol, (std::size_t)max_iter, 40);
  std::memcpy(x_out, r.x.data(), n * sizeof(double));
  *f_out = r.f;
  return r.ok ? 0 : 1;
}

}  // extern "C"

// CPU f64 oracle for NEBULA-HL stage two: a restatement of edge-rs 0.4.0's
// `nebula_sparse` (src/sc/nebula.rs), its penalised inner fit (src/sc/pml.rs),
// its stage-one marginal likelihood (src/sc/ptmg.rs) and the numeric pieces
// they call.  TEST INFRASTRUCTURE: only the tests, the smoke check and the
// benchmark's CPU-baseline leg load it, as the checker the HIP path is held
// to.  The product library never links it.
//
// Pinning (each function's check is named in tests/test_oracle_*.py):
//   opt_pml, opt_pml_nbm, pml_log_likelihood_gradient  <- nebula 1.5.8 values
//       quoted in src/sc/pml.rs:2151-2700 (1e-11 relative)
//   ptmg value / gradient / Hessian and the helpers     <- src/sc/ptmg.rs:2140-2898
//   BOBYQA (shared header edge-rs_amd/csrc/bobyqa.hpp)  <- nloptr traces, exact
//   one-component simplex search (search.hpp)          <- the reference's
//       state-machine cases (src/sc/nebula.rs:3516-3619), scipy's Nelder-Mead
//       on the same simplex, and the NEBULA-LN R goldens (sc, sc_blocks,
//       sc_cat_ln)
//   nebula_sparse end to end                            <- R nebula goldens
//       tests/data/nebula_*.csv and tests/data/e2e/sc_*
//   zero-count tables (zeros.cpp)                       <- their own compensated
//       direct sums over the whole range (1e-13, zeros.rs:512-535), numpy and
//       finite differences for the sums, zeros.rs:543-557's grouping case, and
//       the tabled fit / marginal likelihood against the dense ones
//       (pml.rs:2926-2960)
//   digamma / trigamma / lgamma                         <- scipy.special
//   quasi-Newton (stage-one fallback)                   <- the reference's
//       L-BFGS-B known-answer problems (src/numeric/lbfgsb.rs:1040-1296) and
//       scipy at the optimum; iterates follow no reference trace.
#pragma once

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "search.hpp"

namespace oracle {

// EdgeErrors carried as (code, message); codes are include/nebula_hip.h's.
struct Error {
  int code;
  std::string message;
};
[[noreturn]] void fail(int code, const std::string& message);

// ---- numeric/gamma.rs ----
double ln_gamma(double x);
double digamma(double x);
double trigamma(double x);

// ---- sc/zeros.rs ----
// One subject's cells sharing a design row: the six sums over them at a
// count of zero, L = sum ln(1 + v) and its five derivatives in s = ln(tau)
// (v = tau O), as piecewise Chebyshev interpolants.
struct ZeroGroup {
  std::size_t row = 0;  // a member cell (the first)
  double n = 0, sum_o = 0, sum_inv_o = 0, sum_log_o = 0, s_lo = 0, s_hi = 0, lo = 0;
  std::size_t n_base = 0;
  std::vector<std::pair<std::uint32_t, std::uint8_t>> index;  // first panel, depth
  std::vector<double> panels;  // per panel: 6 forms x 17 coefficients
  // m sums from sum `from` (0 L .. 5 E); C, D, E need B in the same call
  void eval(int from, int m, double s, double* out) const;
};

struct ZeroCells {
  std::vector<ZeroGroup> groups;
  std::vector<std::size_t> subject_groups;  // k + 1
  std::vector<std::size_t> loose;           // cells in no group, ascending per subject
  std::vector<std::size_t> subject_loose;   // k + 1
};

ZeroGroup build_zero_group(std::size_t row, const std::vector<double>& log_o);
bool build_zero_cells(const double* design, const double* log_offset, std::size_t n_cells,
                      const std::size_t* fid, std::size_t k, std::size_t n_coef, ZeroCells& z);
void zero_direct_sums(const std::vector<std::pair<double, double>>& distinct, double s,
                      double* out);

// ---- sc/pml.rs ----
struct PmlData {
  const double* design;         // n_cells * nb, row-major
  const double* offset;         // log offset, n_cells
  std::size_t n_cells;
  std::size_t nb;
  const double* counts;         // positive counts
  const std::size_t* cells;     // their cells, increasing
  std::size_t nnz;
  const std::size_t* subject_start;  // k + 1
  std::size_t k;
  const double* subject_total;  // k
  const ZeroCells* zeros = nullptr;  // tables built from this design, or none
};

struct PmlParams {
  bool reml = false;
  double eps = 1e-6;
  unsigned ord = 1;
  std::size_t max_iter = 50;
  std::size_t max_backtrack = 10;
};

struct PmlResult {
  std::vector<double> beta, log_w, information;
  double log_likelihood = 0, log_likelihood_prev = 0, log_det = 0,
         second_order = 0;
  int convergence = 1;
  std::size_t iterations = 0, backtracks = 0;
};

std::pair<double, double> gamma_prior(double exps);
PmlResult opt_pml(const PmlData& d, const double* beta_init, double subject,
                  double cell, const PmlParams& p,
                  const double* log_w_init = nullptr);
PmlResult opt_pml_nbm(const PmlData& d, const double* beta_init,
                      double subject, double cell, const PmlParams& p);
void pml_log_likelihood_gradient(const PmlData& d, const double* beta,
                                 const double* log_w, double subject,
                                 double cell, double* objective,
                                 double* gradient);
int check_convergence(const PmlResult& r, const PmlParams& p,
                      bool variance_at_bound, int initial);
void ldlt_solve(double* a, double* b, std::size_t n);
double determinant(double* a, std::size_t n);
double min_eigenvalue(const double* m, std::size_t n);

// ---- sc/ptmg.rs ----
struct GeneData {
  const double* design;
  const double* log_offset;
  std::size_t n_cells, n_coef;
  const double* counts;
  const std::size_t* cells;
  std::size_t nnz;
  const double* subject_totals;
  const std::size_t* fid;
  std::size_t k;
  const ZeroCells* zeros = nullptr;
};

class Ptmg {
 public:
  explicit Ptmg(const GeneData& g);
  double value(const double* params);
  double value_gradient(const double* params, double* grad);
  double value_gradient_hessian(const double* params, double* grad,
                                double* hess);
  const GeneData& gene() const { return g_; }

 private:
  double linear_predictor(const double* beta);
  double value_gradient_tabled(const double* params, double* grad);
  GeneData g_;
  std::vector<double> y_cell_, extb_, cumsumxtb_;
};

// ---- numeric/lbfgsb.rs (own compact bounded L-BFGS) ----
struct LbfgsbResult {
  std::vector<double> x;
  double f;
  bool ok;
};
template <class F>
LbfgsbResult lbfgsb_minimise(F&& fg, const std::vector<double>& x0,
                             const std::vector<double>& lower,
                             const std::vector<double>& upper, double ftol,
                             double pgtol, std::size_t max_iter,
                             std::size_t max_line_search);

// ---- sc/nebula.rs driver ----
struct Params {
  double min_subject = 1e-4, min_cell = 1e-4, max_subject = 10.0,
         max_cell = 1000.0;
  int method = 0;  // 0 LN, 1 HL
  bool reml = false;
  double cutoff_cell = 20.0, kappa = 800.0, cpc = 0.005;
  std::size_t mincp = 5, min_subjects = 0;
  double eps = 1e-6;
  bool tables = true;  // false: sweep every cell even where tables apply
};

struct Fit {
  std::size_t n_coef = 0;
  std::vector<double> coefficients, covariance, se, subject_overdispersion,
      cell_overdispersion;
  std::vector<int> convergence;
  std::vector<std::uint8_t> sigma_at_bound;
  std::vector<std::size_t> gene_index;
};

// Per-gene stage-two trace for the parity tests: the search's minimiser, the
// profile objective there and the evaluation count.
struct StageTwoInfo {
  double sigma = 0, gamma = 0, objective = 0;
  int evaluations = 0;
  int ord = 1;
  int searched = 0;
  double start_sigma = 0, start_gamma = 0;
  std::vector<double> beta_start;
};

Fit nebula_sparse(std::size_t n_genes, std::size_t n_cells,
                  const std::uint32_t* indptr, const std::uint32_t* indices,
                  const double* data, const std::uint64_t* subject_id,
                  const double* design, std::size_t n_coef,
                  const double* offset, const Params& params, int threads,
                  std::vector<StageTwoInfo>* stage_two,
                  double* t_stage_two_out);

// StageTwoSearch driven by an objective f(x, order) (nebula.rs:1568-1951):
// BOBYQA for two components, the simplex search of search.hpp for one.
SearchOutcome stage_two_search(const std::vector<double>& start, const std::vector<double>& lo,
                               const std::vector<double>& hi, unsigned ord,
                               const std::function<double(const std::vector<double>&, unsigned)>& f);

// The profile objective (VarianceObjective::value, nebula.rs:1489) at n
// points (subject, cell) for kept genes gene[q], after the same set-up as
// nebula_sparse.
void profile_objective(std::size_t n_genes, std::size_t n_cells, const std::uint32_t* indptr,
                       const std::uint32_t* indices, const double* data,
                       const std::uint64_t* subject_id, const double* design,
                       std::size_t n_coef, const double* offset, const Params& params,
                       std::size_t n, const std::uint64_t* gene, const double* point,
                       double* out, int threads = 0);

}  // namespace oracle

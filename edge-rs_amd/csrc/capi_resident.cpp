// C ABI: request-level batch API, BOBYQA entry points and defaults.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/nebula_hip.h"
#include "bobyqa.hpp"
#include "driver.hpp"
#include "persistent.hpp"
#include "resident.hpp"
#include "search_control.hpp"
#include "stage_one.hpp"

struct nh_resident {
  nh::Resident r;
};

namespace nh {
// The handle's resident batch, for the entry points defined in other files.
Resident* resident_of(nh_resident_t* h) { return &h->r; }
}  // namespace nh

namespace {
void put_err(char* err, size_t len, const std::string& m) {
  if (err && len) {
    std::strncpy(err, m.c_str(), len - 1);
    err[len - 1] = 0;
  }
}
}  // namespace

extern "C" {

void nebula_params_default(nebula_params_t* p) {
  p->min_subject = 1e-4;
  p->min_cell = 1e-4;
  p->max_subject = 10.0;
  p->max_cell = 1000.0;
  p->method = NH_METHOD_LN;
  p->reml = 0;
  p->cutoff_cell = 20.0;
  p->kappa = 800.0;
  p->cpc = 0.005;
  p->mincp = 5;
  p->min_subjects = 0;
  p->eps = 1e-6;
}

void nebula_hip_options_default(nebula_hip_options_t* o) {
  o->engine = NH_ENGINE_BATCHED;
  o->learned = 0;
  o->mlp_weights = nullptr;
  o->wg_size = 0;
  o->device = 0;
  o->host_threads = 0;
  o->learned_step = 0.0;
  o->learned_xtol = 0.0;
  o->zero_tables = 1;
}

int nh_resident_upload(size_t n_cells, size_t n_coef, size_t n_subjects, const double* design,
                       const double* log_offset, const uint64_t* subject_start, size_t n_genes,
                       const uint64_t* gene_ptr, const uint32_t* cells, const double* counts,
                       int32_t device, nh_resident_t** out, char* err, size_t err_len) {
  if (!out || !design || !log_offset || !subject_start || !gene_ptr) {
    put_err(err, err_len, "null argument");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (n_coef < 1 || n_coef > 8) {
    put_err(err, err_len, "the device fit takes 1 to 8 design columns");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (n_subjects < 1 || subject_start[0] != 0 || subject_start[n_subjects] != n_cells) {
    put_err(err, err_len, "subject_start must run from 0 to n_cells");
    return NH_ERR_INVALID_ARGUMENT;
  }
  for (size_t s = 0; s < n_subjects; ++s)
    if (subject_start[s] >= subject_start[s + 1]) {
      put_err(err, err_len, "empty subject block");
      return NH_ERR_SUBJECTS_NOT_CONTIGUOUS;
    }
  if (gene_ptr[0] != 0) {
    put_err(err, err_len, "gene_ptr must start at zero");
    return NH_ERR_MALFORMED_SPARSE;
  }
  for (size_t g = 0; g < n_genes; ++g)
    if (gene_ptr[g + 1] < gene_ptr[g]) {
      put_err(err, err_len, "gene_ptr must be non-decreasing");
      return NH_ERR_MALFORMED_SPARSE;
    }
  if (gene_ptr[n_genes] > 0 && (!cells || !counts)) {
    put_err(err, err_len, "null cells or counts with stored entries");
    return NH_ERR_INVALID_ARGUMENT;
  }
  for (size_t g = 0; g < n_genes; ++g)
    for (uint64_t i = gene_ptr[g]; i < gene_ptr[g + 1]; ++i)
      if (cells[i] >= n_cells || (i > gene_ptr[g] && cells[i] <= cells[i - 1])) {
        put_err(err, err_len, "cells must be in range and increasing within a gene");
        return NH_ERR_MALFORMED_SPARSE;
      }
  nh_resident* h = new (std::nothrow) nh_resident;
  if (!h) {
    put_err(err, err_len, "out of host memory");
    return NH_ERR_GPU;
  }
  try {
    h->r.upload(n_cells, n_coef, n_subjects, design, log_offset, subject_start, n_genes, gene_ptr,
                cells, counts, device);
  } catch (const std::invalid_argument& e) {
    delete h;
    put_err(err, err_len, e.what());
    return NH_ERR_INVALID_ARGUMENT;
  } catch (const std::exception& e) {
    delete h;
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  *out = h;
  return NH_OK;
}

}  // extern "C"

namespace {

// Checks the requests and converts them to the device layout; returns NH_OK
// or an error code with the message in err.
int build_requests(const nh::Resident& r, const nh_request_t* requests, size_t n_req,
                   std::vector<nh::DevRequest>& req, std::vector<double>& bi,
                   std::vector<double>& lwi, bool& any_lw, char* err, size_t err_len) {
  const size_t nb = r.nb, k = r.k;
  if (n_req && !requests) {
    put_err(err, err_len, "null requests");
    return NH_ERR_INVALID_ARGUMENT;
  }
  req.assign(n_req, nh::DevRequest{});
  bi.assign(n_req * nb, 0.0);
  any_lw = false;
  for (size_t q = 0; q < n_req; ++q) any_lw = any_lw || requests[q].log_w_init;
  lwi.assign(any_lw ? n_req * k : 0, 0.0);
  for (size_t q = 0; q < n_req; ++q) {
    const nh_request_t& x = requests[q];
    if (x.gene >= r.n_genes || x.ord < 1 || x.ord > 3 || !x.beta_init) {
      put_err(err, err_len, "invalid request " + std::to_string(q));
      return NH_ERR_INVALID_ARGUMENT;
    }
    const double exps = std::exp(x.subject);
    if (!(std::isfinite(exps) && exps > 1.0)) {
      put_err(err, err_len, "subject must be finite and strictly positive");
      return NH_ERR_INVALID_ARGUMENT;
    }
    if (!(std::isfinite(x.cell) && x.cell > 0.0)) {
      put_err(err, err_len, "invalid cell dispersion");
      return NH_ERR_INVALID_DISPERSION;
    }
    req[q] = nh::DevRequest{x.gene, x.ord, x.subject, x.cell};
    for (size_t j = 0; j < nb; ++j) bi[q * nb + j] = x.beta_init[j];
    if (any_lw && x.log_w_init)
      for (size_t s = 0; s < k; ++s) lwi[q * k + s] = x.log_w_init[s];
  }
  return NH_OK;
}

}  // namespace

extern "C" {

int nh_resident_submit(nh_resident_t* h, const nh_request_t* requests, size_t n_req, double eps,
                       int32_t mode, char* err, size_t err_len) {
  if (!h) {
    put_err(err, err_len, "null handle");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (mode != NH_SOLVE_F32_FINISH && mode != NH_SOLVE_F64) {
    put_err(err, err_len, "unknown solve mode");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (h->r.pending_launched) {
    put_err(err, err_len, "a launch is already in flight on this handle; collect it first");
    return NH_ERR_INVALID_ARGUMENT;
  }
  std::vector<nh::DevRequest> req;
  std::vector<double> bi, lwi;
  bool any_lw = false;
  const int rc = build_requests(h->r, requests, n_req, req, bi, lwi, any_lw, err, err_len);
  if (rc != NH_OK) return rc;
  try {
    h->r.submit(req, bi, any_lw ? &lwi : nullptr, eps, mode, 8);
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  return NH_OK;
}

int nh_resident_collect(nh_resident_t* h, nh_reply_t* replies, size_t n_req, char* err,
                        size_t err_len) {
  if (!h) {
    put_err(err, err_len, "null handle");
    return NH_ERR_INVALID_ARGUMENT;
  }
  nh::Resident& r = h->r;
  if (!r.pending_launched) {
    put_err(err, err_len, "no launch in flight on this handle");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (n_req != r.in_flight() || (n_req && !replies)) {
    put_err(err, err_len, "replies must match the submitted request count (" +
                              std::to_string(r.in_flight()) + ")");
    return NH_ERR_INVALID_ARGUMENT;
  }
  const size_t nb = r.nb, k = r.k;
  std::vector<nh::DevReply> out;
  std::vector<double> ob, olw, oinfo;
  try {
    r.collect(out, ob, olw, &oinfo);
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  for (size_t q = 0; q < n_req; ++q) {
    nh_reply_t& o = replies[q];
    o.log_likelihood = out[q].ll;
    o.log_likelihood_prev = out[q].ll_prev;
    o.log_det = out[q].log_det;
    o.second_order = out[q].second;
    o.iterations = out[q].iterations;
    o.backtracks = out[q].backtracks;
    o.finish_steps = out[q].finish_steps;
    o.convergence = nh::fit_convergence(out[q], &oinfo[q * nb * nb], nb, false, NH_CONV_SUCCESS);
    if (o.beta) std::memcpy(o.beta, &ob[q * nb], nb * sizeof(double));
    if (o.log_w) std::memcpy(o.log_w, &olw[q * k], k * sizeof(double));
    if (o.information) std::memcpy(o.information, &oinfo[q * nb * nb], nb * nb * sizeof(double));
  }
  return NH_OK;
}

int nh_resident_solve(nh_resident_t* h, const nh_request_t* requests, size_t n_req, double eps,
                      int32_t mode, nh_reply_t* replies, char* err, size_t err_len) {
  const int rc = nh_resident_submit(h, requests, n_req, eps, mode, err, err_len);
  if (rc != NH_OK) return rc;
  return nh_resident_collect(h, replies, n_req, err, err_len);
}

void nh_resident_free(nh_resident_t* h) { delete h; }

int nh_bobyqa(int n, const double* x0, const double* lower, const double* upper, double xtol_rel,
              int maxeval, nh_objective_fn f, void* ctx, double* x_out, double* f_out,
              int* evals_out) {
  if (n < 1 || n > 3) return nh::BOBYQA_BAD_BOX;
  nh::Bobyqa<3> b;
  b.init(n, x0, lower, upper, xtol_rel, maxeval);
  if (b.status == nh::BOBYQA_BAD_BOX) return b.status;
  while (b.want) b.tell(f(b.xask, ctx));
  for (int i = 0; i < n; ++i) x_out[i] = b.xbest_out[i];
  *f_out = b.fbest_out;
  *evals_out = b.nevals;
  return b.status;
}

int nh_bobyqa_device_trace(int problem, int n, const double* x0, const double* lower,
                           const double* upper, double* trace, int max_trace, int* evals_out,
                           char* err, size_t err_len) {
  if (n < 1 || n > 3 || max_trace < 0) {
    put_err(err, err_len, "invalid arguments");
    return -100 - NH_ERR_INVALID_ARGUMENT;
  }
  try {
    nh::DevBuf<double> dx0, dlo, dhi, dtr;
    nh::DevBuf<int> dres;
    hipStream_t st = nullptr;
    dx0.upload(x0, n, st);
    dlo.upload(lower, n, st);
    dhi.upload(upper, n, st);
    dtr.alloc((size_t)max_trace * (n + 1));
    dres.alloc(2);
    NH_HIP(nh::launch_bobyqa_trace(problem, n, dx0.p, dlo.p, dhi.p, dtr.p, max_trace, dres.p, st));
    int res[2];
    NH_HIP(hipMemcpy(res, dres.p, sizeof(res), hipMemcpyDeviceToHost));
    int rows = res[1] < max_trace ? res[1] : max_trace;
    if (rows > 0)
      NH_HIP(hipMemcpy(trace, dtr.p, sizeof(double) * (size_t)rows * (n + 1),
                       hipMemcpyDeviceToHost));
    *evals_out = res[1];
    return res[0];
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return -100 - NH_ERR_GPU;
  }
}

int nh_stage_one_search(int n, const double* start, const double* lower, const double* upper,
                        nh_value_grad_fn f, void* ctx, double* x_out, int* failed,
                        int* evals_out) {
  if (n < 1 || n > 16 || !start || !lower || !upper || !f || !x_out || !failed || !evals_out)
    return NH_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < n; ++i)
    if (!(lower[i] <= upper[i])) return NH_ERR_INVALID_BOUNDS;
  std::vector<double> st(start, start + n), lo(lower, lower + n), hi(upper, upper + n);
  nh::S1Result res;
  nh::S1Task task = nh::stage_one_search(st, lo, hi, true, 0, &res);
  int evals = 0;
  while (task.waiting()) {
    auto& pr = task.h.promise();
    nh::EvalRes er;
    er.g.assign(n, 0.0);
    er.f = f(pr.req.x.data(), er.g.data(), ctx);
    ++evals;
    pr.res = std::move(er);
    task.h.resume();
  }
  for (int i = 0; i < n; ++i) x_out[i] = res.x[i];
  *failed = res.failed ? 1 : 0;
  *evals_out = evals;
  return NH_OK;
}

int nh_stage_two_search(int kind, const double* start, const double* lower, const double* upper,
                        int ord, nh_search_fn f, void* ctx, double* x_out, int* found,
                        int* failed, int* order_out, int* evals_out) {
  if ((kind != nh::kSearchSubject && kind != nh::kSearchBoth) || (ord != 1 && ord != 3) ||
      !start || !lower || !upper || !f || !x_out || !found || !failed || !order_out || !evals_out)
    return NH_ERR_INVALID_ARGUMENT;
  for (int i = 0; i < kind; ++i)
    if (!(lower[i] < upper[i])) return NH_ERR_INVALID_BOUNDS;
  const double lo[2] = {lower[0], kind == nh::kSearchBoth ? lower[1] : 0.0};
  const double hi[2] = {upper[0], kind == nh::kSearchBoth ? upper[1] : 1.0};
  const double st[2] = {start[0], start[1]};
  nh::SearchControl c;
  c.begin(kind, ord, st, nullptr, lo, hi, 0.0, 0.0);
  while (c.active) {
    double sub, cell;
    c.point(sub, cell);
    c.tell(f(sub, cell, c.point_order(), ctx));
  }
  double fv;
  c.result(x_out, fv, *order_out);
  *found = c.found;
  *failed = c.failed;
  *evals_out = c.evals;
  return NH_OK;
}

int nh_stage_two_search_device(int problem, int kind, int ord, const double* start,
                               const double* lower, const double* upper, int inj_from,
                               int inj_to, int inj_order, int inj_kind, int32_t device,
                               double* trace, int max_trace, double* x_out, int* found,
                               int* failed, int* order_out, int* evals_out, char* err,
                               size_t err_len) {
  if (problem < 0 || problem > 4 || (kind != nh::kSearchSubject && kind != nh::kSearchBoth) ||
      (ord != 1 && ord != 3) || inj_kind < 0 || inj_kind > 2 || max_trace < 0 || !start ||
      !lower || !upper || (max_trace && !trace) || !x_out || !found || !failed || !order_out ||
      !evals_out) {
    put_err(err, err_len, "invalid arguments");
    return NH_ERR_INVALID_ARGUMENT;
  }
  for (int i = 0; i < kind; ++i)
    if (!(lower[i] < upper[i])) {
      put_err(err, err_len, "invalid bounds");
      return NH_ERR_INVALID_BOUNDS;
    }
  const double lo[2] = {lower[0], kind == nh::kSearchBoth ? lower[1] : 0.0};
  const double hi[2] = {upper[0], kind == nh::kSearchBoth ? upper[1] : 1.0};
  const double st2[2] = {start[0], start[1]};
  try {
    NH_HIP(hipSetDevice(device));
    nh::DevBuf<double> dst, dlo, dhi, dtr;
    nh::DevBuf<nh::DevSearchOut> dout;
    hipStream_t st = nullptr;
    dst.upload(st2, 2, st);
    dlo.upload(lo, 2, st);
    dhi.upload(hi, 2, st);
    dtr.alloc((size_t)std::max(max_trace, 1) * 4);
    dout.alloc(1);
    NH_HIP(nh::launch_search_trace(problem, kind, ord, dst.p, dlo.p, dhi.p, inj_from, inj_to,
                                   inj_order, inj_kind, dtr.p, max_trace, dout.p, st));
    nh::DevSearchOut o;
    NH_HIP(hipMemcpy(&o, dout.p, sizeof(o), hipMemcpyDeviceToHost));
    const int rows = std::min(o.evals, max_trace);
    if (rows > 0)
      NH_HIP(hipMemcpy(trace, dtr.p, sizeof(double) * (size_t)rows * 4, hipMemcpyDeviceToHost));
    x_out[0] = o.x[0];
    x_out[1] = o.x[1];
    *found = o.found;
    *failed = o.failed;
    *order_out = o.ord;
    *evals_out = o.evals;
    return NH_OK;
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
}

int nh_selftest_copy(size_t mib, int32_t reps, int32_t device, char* err, size_t err_len) {
  if (mib == 0 || mib > 4096 || reps < 1 || reps > 100) {
    put_err(err, err_len, "invalid arguments");
    return NH_ERR_INVALID_ARGUMENT;
  }
  try {
    NH_HIP(hipSetDevice(device));
    const size_t words = mib * (size_t)(1 << 20) / 16;
    nh::DevBuf<uint4> src, dst;
    src.alloc(words);
    dst.alloc(words);
    NH_HIP(hipMemset(src.p, 0x5a, words * 16));
    for (int r = 0; r < reps; ++r) NH_HIP(nh::launch_copy_words(src.p, dst.p, words, nullptr));
    NH_HIP(hipDeviceSynchronize());
    uint4 last;
    NH_HIP(hipMemcpy(&last, dst.p + (words - 1), sizeof(last), hipMemcpyDeviceToHost));
    if (last.x != 0x5a5a5a5au) throw std::runtime_error("copy self-test read back a wrong value");
    return NH_OK;
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
}

int nh_selftest_f64_math(const double* x, size_t n, int32_t which, int32_t device,
                         double* out, char* err, size_t err_len) {
  if ((n && (!x || !out)) || which < 0 || which > 3) {
    put_err(err, err_len, "invalid arguments");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (n == 0) return NH_OK;
  try {
    NH_HIP(hipSetDevice(device));
    nh::DevBuf<double> dx, dout;
    hipStream_t st = nullptr;
    dx.upload(x, n, st);
    dout.alloc(n);
    NH_HIP(nh::launch_f64_math(dx.p, n, which, dout.p, st));
    NH_HIP(hipMemcpy(out, dout.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    return NH_OK;
  } catch (const std::exception& e) {
    put_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
}

}  // extern "C"

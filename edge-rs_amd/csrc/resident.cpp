// Host side of the resident batch: upload once, solve many request batches.
#include <stdexcept>

#include "driver.hpp"
#include "resident.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace nh {

Resident::~Resident() {
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

void Resident::upload(size_t nc, size_t nb_, size_t k_, const double* design,
                      const double* log_offset, const uint64_t* subject_start, size_t ng,
                      const uint64_t* gene_ptr, const uint32_t* cells, const double* counts,
                      int dev) {
  if (nb_ < 1 || nb_ > 8) throw std::invalid_argument("the device fit takes 1 to 8 design columns");
  if (nc >= 0xffffffffull) throw std::invalid_argument("too many cells for u32 cell indices");
  device = dev;
  NH_HIP(hipSetDevice(device));
  if (!stream) NH_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  if (!ev0) NH_HIP(hipEventCreate(&ev0));
  if (!ev1) NH_HIP(hipEventCreate(&ev1));
  n_cells = nc;
  nb = nb_;
  k = k_;
  n_genes = ng;
  // the intercept: the first column that is 1 in every cell (nebula's centred
  // design always has one); the fit keeps its sums implicitly
  icpt = nb;
  for (size_t j = 0; j < nb && icpt == nb; ++j) {
    bool ones = true;
    for (size_t i = 0; i < n_cells && ones; ++i) ones = design[i * nb + j] == 1.0;
    if (ones) icpt = j;
  }
  if (icpt == nb)
    throw std::invalid_argument("the design needs an intercept column (1 in every cell)");
  rw = packed_row_width((uint32_t)nb);
  rw64 = packed_row_width64((uint32_t)nb);
  // the sweeps address a cell's f64 row at a 32-bit byte offset from the
  // array's base
  if (n_cells * rw64 * sizeof(double) > 0xffffffffull)
    throw std::invalid_argument("too many cells for 32-bit row offsets at this design width");
  std::vector<float> h32(n_cells * rw, 0.0f);
  std::vector<double> h64(n_cells * rw64, 0.0);
  for (size_t i = 0; i < n_cells; ++i) {
    size_t q = 0;
    for (size_t j = 0; j < nb; ++j) {
      if (j == icpt) continue;
      h32[i * rw + q] = (float)design[i * nb + j];
      h64[i * rw64 + q] = design[i * nb + j];
      ++q;
    }
    h32[i * rw + q] = (float)log_offset[i];
    h64[i * rw64 + q] = log_offset[i];
  }
  row32.upload(h32.data(), h32.size(), stream);
  row64.upload(h64.data(), h64.size(), stream);
  x64.upload(design, n_cells * nb, stream);
  off64.upload(log_offset, n_cells, stream);
  std::vector<uint32_t> hf(k + 1);
  for (size_t s = 0; s <= k; ++s) hf[s] = (uint32_t)subject_start[s];
  fid.upload(hf.data(), hf.size(), stream);
  const size_t nnz = gene_ptr[n_genes];
  gptr.upload(gene_ptr, n_genes + 1, stream);
  gcell.upload(cells, nnz, stream);
  // f32 counts and subject totals per gene, genes in parallel
  RawVec<float> hy(nnz);
  std::vector<double> tot(n_genes * k, 0.0);
  std::vector<uint8_t> inexact(n_genes, 0);
  // (per-gene flags and sums kept in registers: neighbouring genes run on
  // different threads and would share cache lines)
  parallel_for(n_genes, 0, [&](size_t g) {
    size_t s = 0;
    bool inex = false;
    double t = 0.0;
    for (uint64_t i = gene_ptr[g]; i < gene_ptr[g + 1]; ++i) {
      const float f = (float)counts[i];
      hy[i] = f;
      inex |= (double)f != counts[i];
      if (cells[i] >= hf[s + 1]) {
        tot[g * k + s] = t;
        t = 0.0;
        while (cells[i] >= hf[s + 1]) ++s;
      }
      t += counts[i];
    }
    if (gene_ptr[g + 1] > gene_ptr[g]) tot[g * k + s] = t;
    inexact[g] = inex;
  }, 16);
  gy32.upload(hy.data(), nnz, stream);
  // the f64 counts: widened on the device when every count is an f32 value
  // (whole counts below 2^24 are), else sent as they are
  bool exact = true;
  for (uint8_t b : inexact) exact = exact && !b;
  if (exact) {
    gy64.alloc(nnz);
    NH_HIP(launch_widen(gy32.p, nnz, gy64.p, stream));
  } else {
    gy64.upload(counts, nnz, stream);
  }
  gtot.upload(tot.data(), tot.size(), stream);
  counters.alloc(1);
  reset_counters();
  NH_HIP(hipStreamSynchronize(stream));
}

DeviceData Resident::view() const {
  DeviceData d;
  d.row32 = row32.p;
  d.row64 = row64.p;
  d.x64 = x64.p;
  d.off64 = off64.p;
  d.fid = fid.p;
  d.gptr = gptr.p;
  d.gcell = gcell.p;
  d.gy32 = gy32.p;
  d.gy64 = gy64.p;
  d.gtot = gtot.p;
  d.n_cells = (uint32_t)n_cells;
  d.nb = (uint32_t)nb;
  d.k = (uint32_t)k;
  d.icpt = (uint32_t)icpt;
  d.rw = rw;
  d.rw64 = rw64;
  d.tabled = tabled ? 1u : 0u;
  d.zgroup = zgroup.p;
  d.zsg = zsg.p;
  d.zloose = zloose.p;
  d.zsl = zsl.p;
  d.zindex = zindex.p;
  d.zcoef = zcoef.p;
  return d;
}

void Resident::set_tables(const ZeroTables& z) {
  if (z.subject_groups.size() != k + 1 || z.subject_loose.size() != k + 1)
    throw std::invalid_argument("zero-count tables built for another subject count");
  NH_HIP(hipSetDevice(device));
  zgroup.upload(z.groups.data(), z.groups.size(), stream);
  zsg.upload(z.subject_groups.data(), z.subject_groups.size(), stream);
  zloose.upload(z.loose.data(), z.loose.size(), stream);
  zsl.upload(z.subject_loose.data(), z.subject_loose.size(), stream);
  zindex.upload(z.index.data(), z.index.size(), stream);
  zcoef.upload(z.coef.data(), z.coef.size(), stream);
  NH_HIP(hipStreamSynchronize(stream));
  tabled = true;
}

void Resident::reset_counters() {
  // stream-ordered before the next launch; never captured into a graph
  NH_HIP(hipMemsetAsync(counters.p, 0, sizeof(FitCounters), stream));
}

FitCounters Resident::read_counters() {
  FitCounters c;
  NH_HIP(hipMemcpyAsync(&c, counters.p, sizeof(c), hipMemcpyDeviceToHost, stream));
  NH_HIP(hipStreamSynchronize(stream));
  return c;
}

void Resident::solve(const std::vector<DevRequest>& r, const std::vector<double>& bi,
                     const std::vector<double>* lwi, double eps, int mode, int waves,
                     std::vector<DevReply>& out, std::vector<double>& ob,
                     std::vector<double>& olw, std::vector<double>* oinfo,
                     bool decrement_stop) {
  submit(r, bi, lwi, eps, mode, waves, decrement_stop);
  collect(out, ob, olw, oinfo);
}

void Resident::submit(const std::vector<DevRequest>& r, const std::vector<double>& bi,
                      const std::vector<double>* lwi, double eps, int mode, int waves,
                      bool decrement_stop) {
  if (pending_launched) throw std::logic_error("a launch is already in flight on this handle");
  const size_t n = r.size();
  for (const DevRequest& q : r)
    if (q.gene >= n_genes) throw std::invalid_argument("request names a gene that is not resident");
  if (n == 0) {
    pending = 0;
    pending_launched = true;
    return;
  }
  h_req = r;
  h_beta = bi;
  if (lwi) h_lw = *lwi;
  NH_HIP(hipSetDevice(device));
  req.upload(h_req.data(), n, stream);
  beta_init.upload(h_beta.data(), n * nb, stream);
  if (lwi) lw_init.upload(h_lw.data(), n * k, stream);
  reply.alloc(n);
  out_beta.alloc(n * nb);
  out_lw.alloc(n * k);
  out_info.alloc(n * nb * nb);
  const size_t stride = fit_work_doubles((int)nb, (int)k, waves);
  const bool in_lds = stride * sizeof(double) <= kWorkLdsBytes;
  if (!in_lds) scratch.alloc(stride * n);
  BatchLaunch L;
  L.dd = view();
  L.req = req.p;
  L.beta_init = beta_init.p;
  L.lw_init = lwi ? lw_init.p : nullptr;
  L.out = reply.p;
  L.out_beta = out_beta.p;
  L.out_lw = out_lw.p;
  L.out_info = out_info.p;
  L.scratch = in_lds ? nullptr : scratch.p;
  L.stride = stride;
  L.work_in_lds = in_lds ? 1 : 0;
  L.knobs.eps = eps;
  L.knobs.noise_rel = 1e-8;
  L.knobs.max_iter = kFitMaxIter;
  L.knobs.max_backtrack = kFitMaxBacktrack;
  L.knobs.decrement_stop = decrement_stop ? 1 : 0;
  L.mode = mode;
  L.waves = waves;
  L.n_req = (int)n;
  L.counters = counters.p;
  NH_HIP(hipEventRecord(ev0, stream));
  NH_HIP(launch_batch(L, stream));
  NH_HIP(hipEventRecord(ev1, stream));
  pending = n;
  pending_launched = true;
}

void Resident::collect(std::vector<DevReply>& out, std::vector<double>& ob,
                       std::vector<double>& olw, std::vector<double>* oinfo) {
  if (!pending_launched) throw std::logic_error("no launch in flight on this handle");
  const size_t n = pending;
  pending = 0;
  pending_launched = false;
  out.resize(n);
  ob.resize(n * nb);
  olw.resize(n * k);
  if (oinfo) oinfo->resize(n * nb * nb);
  if (n == 0) return;
  NH_HIP(hipSetDevice(device));
  NH_HIP(hipMemcpyAsync(out.data(), reply.p, n * sizeof(DevReply), hipMemcpyDeviceToHost, stream));
  NH_HIP(hipMemcpyAsync(ob.data(), out_beta.p, n * nb * sizeof(double), hipMemcpyDeviceToHost,
                        stream));
  NH_HIP(hipMemcpyAsync(olw.data(), out_lw.p, n * k * sizeof(double), hipMemcpyDeviceToHost,
                        stream));
  if (oinfo)
    NH_HIP(hipMemcpyAsync(oinfo->data(), out_info.p, n * nb * nb * sizeof(double),
                          hipMemcpyDeviceToHost, stream));
  NH_HIP(hipStreamSynchronize(stream));
  float ms = 0.0f;
  NH_HIP(hipEventElapsedTime(&ms, ev0, ev1));
  kernel_ms += ms;
  ++launches;
}

}  // namespace nh

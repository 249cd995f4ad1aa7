// nh_resident_solve without allocation: the same batch of penalised fits as
// nh_resident_solve, with every device buffer beyond the resident data -- the
// requests and start points, the replies and result arrays, the fit's
// global-memory work area and its sweep counters -- carved out of one
// file-scope device array instead of hipMalloc'd.  For callers that cannot
// allocate (serving loops, code under stream capture).
//
// The array is per device: each GPU the library runs on holds its own copy
// for the life of the process, at its own address, which is looked up once
// per device with hipGetSymbolAddress and used from then on (the array's name
// is a host-side stand-in and is never passed to a kernel).  Every call on a
// device shares that device's copy: calls that overlap on one GPU race, and a
// call leaves the last batch's data behind (what must start at zero, the
// counters, is cleared by a kernel at each call).  Host staging goes through
// fixed-size stack chunks, so the host heap is not touched either.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/nebula_hip.h"
#include "driver.hpp"
#include "kernels.hpp"
#include "resident.hpp"

namespace nh {

constexpr size_t kNoAllocDoubles = size_t(4) << 20;  // 32 MiB per device
static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of this device's copy of `scratch` (device current).
double* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::invalid_argument("device out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw HipError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = (double*)p;
  }
  return g_scratch[device];
}

__global__ void zero_doubles_kernel(double* p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

size_t doubles_for(size_t bytes) { return (bytes + 7) / 8; }
// offsets kept 256-byte aligned
size_t align32(size_t d) { return (d + 31) & ~size_t(31); }

constexpr size_t kChunk = 64;  // requests staged through the stack at a time

}  // namespace

// The scratch layout of one call, in doubles; total > kNoAllocDoubles is
// rejected before any device work.
struct NoAllocLayout {
  size_t req, beta, lw, reply, obeta, olw, oinfo, work, counters, total;
  size_t stride;
  bool in_lds;
  NoAllocLayout(size_t n, size_t nb, size_t k, bool any_lw) {
    stride = fit_work_doubles((int)nb, (int)k, 8);
    in_lds = stride * sizeof(double) <= kWorkLdsBytes;
    size_t at = 0;
    auto take = [&](size_t d) {
      const size_t o = at;
      at = align32(at + d);
      return o;
    };
    req = take(doubles_for(n * sizeof(DevRequest)));
    beta = take(n * nb);
    lw = take(any_lw ? n * k : 0);
    reply = take(doubles_for(n * sizeof(DevReply)));
    obeta = take(n * nb);
    olw = take(n * k);
    oinfo = take(n * nb * nb);
    work = take(in_lds ? 0 : stride * n);
    counters = take(doubles_for(sizeof(FitCounters)));
    total = at;
  }
};

}  // namespace nh

extern "C" int nh_resident_solve_noalloc(nh_resident_t* h, const nh_request_t* requests,
                                         size_t n_req, double eps, int32_t mode,
                                         nh_reply_t* replies, char* err, size_t err_len);

namespace nh {
Resident* resident_of(nh_resident_t* h);
}

namespace {
void noalloc_err(char* err, size_t len, const std::string& m) {
  if (err && len) {
    std::strncpy(err, m.c_str(), len - 1);
    err[len - 1] = 0;
  }
}
}  // namespace

extern "C" int nh_resident_solve_noalloc(nh_resident_t* h, const nh_request_t* requests,
                                         size_t n_req, double eps, int32_t mode,
                                         nh_reply_t* replies, char* err, size_t err_len) {
  using namespace nh;
  if (!h) {
    noalloc_err(err, err_len, "null handle");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (mode != NH_SOLVE_F32_FINISH && mode != NH_SOLVE_F64) {
    noalloc_err(err, err_len, "unknown solve mode");
    return NH_ERR_INVALID_ARGUMENT;
  }
  Resident& r = *resident_of(h);
  if (r.pending_launched) {
    noalloc_err(err, err_len, "a launch is already in flight on this handle; collect it first");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (n_req == 0) return NH_OK;
  if (!requests || !replies) {
    noalloc_err(err, err_len, "null requests or replies");
    return NH_ERR_INVALID_ARGUMENT;
  }
  const size_t nb = r.nb, k = r.k;
  bool any_lw = false;
  for (size_t q = 0; q < n_req; ++q) {
    const nh_request_t& x = requests[q];
    if (x.gene >= r.n_genes || x.ord < 1 || x.ord > 3 || !x.beta_init) {
      noalloc_err(err, err_len, "invalid request " + std::to_string(q));
      return NH_ERR_INVALID_ARGUMENT;
    }
    const double exps = std::exp(x.subject);
    if (!(std::isfinite(exps) && exps > 1.0)) {
      noalloc_err(err, err_len, "subject must be finite and strictly positive");
      return NH_ERR_INVALID_ARGUMENT;
    }
    if (!(std::isfinite(x.cell) && x.cell > 0.0)) {
      noalloc_err(err, err_len, "invalid cell dispersion");
      return NH_ERR_INVALID_DISPERSION;
    }
    any_lw = any_lw || x.log_w_init;
  }
  const NoAllocLayout lay(n_req, nb, k, any_lw);
  if (lay.total > kNoAllocDoubles) {
    noalloc_err(err, err_len,
                "batch needs " + std::to_string(lay.total * 8) + " bytes of device scratch; " +
                    "this entry point holds " + std::to_string(kNoAllocDoubles * 8) +
                    " (use nh_resident_solve, or split the batch)");
    return NH_ERR_INVALID_ARGUMENT;
  }
  try {
    NH_HIP(hipSetDevice(r.device));
    double* base = scratch_on(r.device);
    hipStream_t st = r.stream;
    // requests and start points, staged through the stack
    {
      DevRequest dreq[kChunk];
      double dbeta[kChunk * 8];
      for (size_t c0 = 0; c0 < n_req; c0 += kChunk) {
        const size_t c = std::min(kChunk, n_req - c0);
        for (size_t q = 0; q < c; ++q) {
          const nh_request_t& x = requests[c0 + q];
          dreq[q] = DevRequest{x.gene, x.ord, x.subject, x.cell};
          std::memcpy(&dbeta[q * nb], x.beta_init, nb * sizeof(double));
        }
        NH_HIP(hipMemcpyAsync((char*)(base + lay.req) + c0 * sizeof(DevRequest), dreq,
                              c * sizeof(DevRequest), hipMemcpyHostToDevice, st));
        NH_HIP(hipMemcpyAsync(base + lay.beta + c0 * nb, dbeta, c * nb * sizeof(double),
                              hipMemcpyHostToDevice, st));
        NH_HIP(hipStreamSynchronize(st));  // the stack chunk is reused next
      }
      if (any_lw) {
        hipLaunchKernelGGL(zero_doubles_kernel, dim3(256), dim3(256), 0, st, base + lay.lw,
                           n_req * k);
        NH_HIP(hipGetLastError());
        for (size_t q = 0; q < n_req; ++q)
          if (requests[q].log_w_init)
            NH_HIP(hipMemcpyAsync(base + lay.lw + q * k, requests[q].log_w_init,
                                  k * sizeof(double), hipMemcpyHostToDevice, st));
      }
    }
    hipLaunchKernelGGL(zero_doubles_kernel, dim3(1), dim3(64), 0, st, base + lay.counters,
                       doubles_for(sizeof(FitCounters)));
    NH_HIP(hipGetLastError());
    BatchLaunch L;
    L.dd = r.view();
    L.req = (const DevRequest*)(base + lay.req);
    L.beta_init = base + lay.beta;
    L.lw_init = any_lw ? base + lay.lw : nullptr;
    L.out = (DevReply*)(base + lay.reply);
    L.out_beta = base + lay.obeta;
    L.out_lw = base + lay.olw;
    L.out_info = base + lay.oinfo;
    L.scratch = lay.in_lds ? nullptr : base + lay.work;
    L.stride = lay.stride;
    L.work_in_lds = lay.in_lds ? 1 : 0;
    L.knobs.eps = eps;
    L.knobs.noise_rel = 1e-8;
    L.knobs.max_iter = kFitMaxIter;
    L.knobs.max_backtrack = kFitMaxBacktrack;
    L.knobs.decrement_stop = 0;
    L.mode = mode;
    L.waves = 8;
    L.n_req = (int)n_req;
    L.counters = (FitCounters*)(base + lay.counters);
    NH_HIP(launch_batch(L, st));
    // replies, staged through the stack; result arrays straight into the
    // caller's views
    DevReply drep[kChunk];
    double dinfo[kChunk * 64];
    for (size_t c0 = 0; c0 < n_req; c0 += kChunk) {
      const size_t c = std::min(kChunk, n_req - c0);
      NH_HIP(hipMemcpyAsync(drep, (char*)(base + lay.reply) + c0 * sizeof(DevReply),
                            c * sizeof(DevReply), hipMemcpyDeviceToHost, st));
      NH_HIP(hipMemcpyAsync(dinfo, base + lay.oinfo + c0 * nb * nb, c * nb * nb * sizeof(double),
                            hipMemcpyDeviceToHost, st));
      for (size_t q = 0; q < c; ++q) {
        nh_reply_t& o = replies[c0 + q];
        if (o.beta)
          NH_HIP(hipMemcpyAsync(o.beta, base + lay.obeta + (c0 + q) * nb, nb * sizeof(double),
                                hipMemcpyDeviceToHost, st));
        if (o.log_w)
          NH_HIP(hipMemcpyAsync(o.log_w, base + lay.olw + (c0 + q) * k, k * sizeof(double),
                                hipMemcpyDeviceToHost, st));
      }
      NH_HIP(hipStreamSynchronize(st));
      for (size_t q = 0; q < c; ++q) {
        nh_reply_t& o = replies[c0 + q];
        const DevReply& d = drep[q];
        o.log_likelihood = d.ll;
        o.log_likelihood_prev = d.ll_prev;
        o.log_det = d.log_det;
        o.second_order = d.second;
        o.iterations = d.iterations;
        o.backtracks = d.backtracks;
        o.finish_steps = d.finish_steps;
        o.convergence = fit_convergence(d, &dinfo[q * nb * nb], nb, false, NH_CONV_SUCCESS);
        if (o.information)
          std::memcpy(o.information, &dinfo[q * nb * nb], nb * nb * sizeof(double));
      }
    }
  } catch (const std::exception& e) {
    noalloc_err(err, err_len, e.what());
    return NH_ERR_GPU;
  }
  return NH_OK;
}

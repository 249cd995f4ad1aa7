// Device-resident copy of one run's shared inputs and per-gene counts, and
// batched penalised fits on it (edge-rs ResidentBatch, src/gpu/nebula_gpu.rs:195-648).
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "kernels.hpp"
#include "zero_tables.hpp"

namespace nh {

struct HipError : std::runtime_error {
  explicit HipError(const std::string& m) : std::runtime_error(m) {}
};

#define NH_HIP(call)                                                                 \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::nh::HipError(std::string(#call) + ": " + hipGetErrorString(e_));       \
  } while (0)

// A device allocation that frees itself.
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    NH_HIP(hipMalloc((void**)&p, sizeof(T) * (count ? count : 1)));
    n = count;
  }
  void upload(const T* h, size_t count, hipStream_t st) {
    alloc(count);
    if (count) NH_HIP(hipMemcpyAsync(p, h, sizeof(T) * count, hipMemcpyHostToDevice, st));
  }
};

// Page-locked host buffer (grow-only), mapped into the device's address
// space: kernels read and write it in place through `dev`, so a small
// exchange with the host needs no copy on the stream.  Coherent memory
// (device reads see every host write) suits inputs; non-coherent memory is
// written through the device's L2 and is visible to the host once the
// writing kernel has completed, which suits outputs.
template <class T>
struct PinnedBuf {
  T* p = nullptr;    // host address
  T* dev = nullptr;  // the same memory as the device addresses it
  size_t n = 0;
  bool coherent = true;
  PinnedBuf() = default;
  explicit PinnedBuf(bool coherent_) : coherent(coherent_) {}
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (p) (void)hipHostFree(p);
    p = dev = nullptr;
    n = 0;
    const unsigned flags =
        hipHostMallocMapped | (coherent ? hipHostMallocCoherent : hipHostMallocNonCoherent);
    NH_HIP(hipHostMalloc((void**)&p, sizeof(T) * (count ? count : 1), flags));
    NH_HIP(hipHostGetDevicePointer((void**)&dev, p, 0));
    n = count;
  }
};

struct StageTwoPlan;

struct Resident {
  int device = 0;
  hipStream_t stream = nullptr;
  size_t n_cells = 0, nb = 0, k = 0, n_genes = 0, icpt = 0;
  uint32_t rw = 1;    // values per packed f32 cell row
  uint32_t rw64 = 1;  // values per packed f64 cell row
  DevBuf<float> row32, gy32;
  DevBuf<double> row64;
  DevBuf<double> x64, off64, gy64, gtot;
  DevBuf<uint32_t> fid, gcell;
  DevBuf<uint64_t> gptr;
  DevBuf<FitCounters> counters;
  // zero-count tables, when set_tables() has put them in place
  bool tabled = false;
  DevBuf<ZGroup> zgroup;
  DevBuf<uint32_t> zsg, zloose, zsl, zindex;
  DevBuf<double> zcoef;
  // per-solve staging
  DevBuf<DevRequest> req;
  DevBuf<DevReply> reply;
  DevBuf<double> beta_init, lw_init, out_beta, out_lw, out_info, scratch;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  // the launch in flight: request count and host copies of its inputs
  size_t pending = 0;
  bool pending_launched = false;
  std::vector<DevRequest> h_req;
  std::vector<double> h_beta, h_lw;
  double kernel_ms = 0.0;
  unsigned long long launches = 0;
  // batched stage-two runs built so far (one per launch shape), replayed
  std::vector<std::shared_ptr<StageTwoPlan>> stage_two_plans;

  ~Resident();
  void upload(size_t n_cells, size_t nb, size_t k, const double* design, const double* log_offset,
              const uint64_t* subject_start, size_t n_genes, const uint64_t* gene_ptr,
              const uint32_t* cells, const double* counts, int device);
  DeviceData view() const;
  // Uploads a run's zero-count tables; every later fit's sweeps read them
  // (one lookup per group, the loose cells and the positive counts).
  void set_tables(const ZeroTables& z);
  // Runs n requests, replies filled on the host (submit + collect).
  // decrement_stop: mode 0 fits stop on the Newton decrement (stage-two
  // evaluations).
  void solve(const std::vector<DevRequest>& r, const std::vector<double>& beta_init,
             const std::vector<double>* lw_init, double eps, int mode, int waves,
             std::vector<DevReply>& out, std::vector<double>& out_beta,
             std::vector<double>& out_lw, std::vector<double>* out_info,
             bool decrement_stop = false);
  // Uploads n requests and launches their fits on the handle's stream without
  // waiting; one launch may be in flight (std::logic_error otherwise).  The
  // request arrays are copied, so the caller's may go at once.
  void submit(const std::vector<DevRequest>& r, const std::vector<double>& beta_init,
              const std::vector<double>* lw_init, double eps, int mode, int waves,
              bool decrement_stop = false);
  // Waits for the launch in flight and copies its replies out.
  void collect(std::vector<DevReply>& out, std::vector<double>& out_beta,
               std::vector<double>& out_lw, std::vector<double>* out_info);
  size_t in_flight() const { return pending; }
  FitCounters read_counters();
  void reset_counters();
};

}  // namespace nh

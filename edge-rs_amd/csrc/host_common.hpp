// Host-side pieces of the driver shared by the stage-two engines.
#pragma once

#include <sys/mman.h>

#include <atomic>
#include <cmath>
#include <cstdlib>
#include <new>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>

namespace nh {

// EdgeErrors as (code, message), codes from include/nebula_hip.h.
struct ApiError {
  int code;
  std::string message;
};
[[noreturn]] inline void api_fail(int code, const std::string& m) { throw ApiError{code, m}; }

// Host threads a loop uses when the caller asks for 0 (the default): one per
// core, at most 16 -- the set-up loops are memory-bound well before that.
inline int default_threads() {
  return (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
}

// Runs f(i) for i in [0, n) on up to `threads` host threads (0: the
// default), each thread taking at least `grain` items so that short loops do
// not pay for thread start-up.
inline void parallel_for(size_t n, int threads, const std::function<void(size_t)>& f,
                         size_t grain = 1) {
  if (n == 0) return;
  int nt = threads > 0 ? threads : default_threads();
  const size_t cap = n / std::max<size_t>(1, grain);
  if ((size_t)nt > cap) nt = (int)cap;
  if (nt <= 1) {
    for (size_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> pool;
  for (int t = 0; t < nt; ++t)
    pool.emplace_back([&, t]() {
      for (size_t i = (size_t)t; i < n; i += (size_t)nt) f(i);
    });
  for (auto& th : pool) th.join();
}

// std::allocator that leaves trivially constructible elements uninitialised
// on resize, so a large array is first touched (and its pages faulted in) by
// the threads that fill it rather than by one zeroing pass; arrays of 32 MiB
// and more are 2 MiB aligned and asked for transparent huge pages, so that
// filling them takes one fault per 2 MiB instead of one per 4 KiB.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  static constexpr size_t kHuge = size_t(2) << 20, kHugeMin = size_t(32) << 20;
  template <class U>
  struct rebind {
    using other = NoInitAlloc<U>;
  };
  NoInitAlloc() = default;
  template <class U>
  NoInitAlloc(const NoInitAlloc<U>&) noexcept {}
  T* allocate(size_t n) {
    const size_t bytes = n * sizeof(T);
    if (bytes < kHugeMin) return std::allocator<T>::allocate(n);
    const size_t rounded = (bytes + kHuge - 1) / kHuge * kHuge;
    void* p = std::aligned_alloc(kHuge, rounded);
    if (!p) throw std::bad_alloc();
    (void)madvise(p, rounded, MADV_HUGEPAGE);  // advisory: ignored where unsupported
    return static_cast<T*>(p);
  }
  void deallocate(T* p, size_t n) noexcept {
    if (n * sizeof(T) < kHugeMin)
      std::allocator<T>::deallocate(p, n);
    else
      std::free(p);
  }
  template <class U>
  void construct(U* p) noexcept {
    ::new ((void*)p) U;
  }
  template <class U, class... A>
  void construct(U* p, A&&... a) {
    ::new ((void*)p) U(std::forward<A>(a)...);
  }
};
template <class T>
using RawVec = std::vector<T, NoInitAlloc<T>>;

// A fixed set of helper threads for loops run many times over: run(n, f)
// calls f(i) for every i in [0, n) on the helpers and the calling thread and
// returns when all are done.  f must not throw.
class WorkPool {
 public:
  explicit WorkPool(int helpers) {
    for (int t = 0; t < helpers; ++t) threads_.emplace_back([this] { loop(); });
  }
  WorkPool(const WorkPool&) = delete;
  WorkPool& operator=(const WorkPool&) = delete;
  ~WorkPool() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    wake_.notify_all();
    for (auto& t : threads_) t.join();
  }
  void run(size_t n, const std::function<void(size_t)>& f) {
    if (threads_.empty() || n < 2) {
      for (size_t i = 0; i < n; ++i) f(i);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      f_ = &f;
      n_ = n;
      next_.store(0);
      busy_ = threads_.size();
      ++generation_;
    }
    wake_.notify_all();
    drain();
    std::unique_lock<std::mutex> lk(m_);
    idle_.wait(lk, [&] { return busy_ == 0; });
  }

 private:
  void drain() {
    for (size_t i; (i = next_.fetch_add(1)) < n_;) (*f_)(i);
  }
  void loop() {
    uint64_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(m_);
      wake_.wait(lk, [&] { return stop_ || generation_ != seen; });
      if (stop_) return;
      seen = generation_;
      lk.unlock();
      drain();
      lk.lock();
      // run() waits for every helper, so none can miss a generation
      if (--busy_ == 0) idle_.notify_one();
    }
  }
  std::vector<std::thread> threads_;
  std::mutex m_;
  std::condition_variable wake_, idle_;
  const std::function<void(size_t)>* f_ = nullptr;
  size_t n_ = 0, busy_ = 0;
  std::atomic<size_t> next_{0};
  uint64_t generation_ = 0;
  bool stop_ = false;
};

// gamma prior of the subject random effect (src/sc/pml.rs:427-429)
inline void gamma_prior(double exps, double& alpha, double& lambda) {
  alpha = 1.0 / (exps - 1.0);
  lambda = 1.0 / (std::sqrt(exps) * (exps - 1.0));
}

// Distinct positive counts other than 1 and 2 with multiplicities
// (src/gpu/stage_two.rs:746-766).
struct CountTail {
  std::vector<double> value, mult;
  double n_pos = 0, n_one = 0, n_two = 0;
  double tail(double cell) const {
    double t = 0.0;
    for (size_t i = 0; i < value.size(); ++i) t += mult[i] * std::lgamma(value[i] + cell);
    return t;
  }
};

// nebula's profile objective -loglik (VarianceObjective::assemble,
// src/sc/nebula.rs:1452-1477).
inline double profile_objective(double subject, double cell, double ll, double ll_prev,
                                double log_det, double second, double n_cells, double k,
                                const CountTail& ct) {
  if (second < -1.0) return INFINITY;
  double alpha, lambda;
  gamma_prior(std::exp(subject), alpha, lambda);
  double base = std::isnan(ll) ? ll_prev : ll;
  double v = base + n_cells * cell * std::log(cell) + k * alpha * std::log(lambda) -
             k * std::lgamma(alpha);
  v += ct.tail(cell) - (ct.n_pos - ct.n_one - ct.n_two) * std::lgamma(cell) +
       (ct.n_one + ct.n_two) * std::log(cell) + ct.n_two * std::log(cell + 1.0);
  v += -0.5 * log_det + std::log(1.0 + second);
  return std::isfinite(v) ? -v : INFINITY;
}

}  // namespace nh

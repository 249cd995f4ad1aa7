// Host-side pieces of the driver shared by the stage-two engines.
#pragma once

#include <cmath>
#include <cstdint>
#include <functional>
#include <string>
#include <thread>
#include <vector>

namespace nh {

// EdgeErrors as (code, message), codes from include/nebula_hip.h.
struct ApiError {
  int code;
  std::string message;
};
[[noreturn]] inline void api_fail(int code, const std::string& m) { throw ApiError{code, m}; }

// Runs f(i) for i in [0, n) on up to `threads` host threads.
inline void parallel_for(size_t n, int threads, const std::function<void(size_t)>& f) {
  if (n == 0) return;
  int nt = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  if ((size_t)nt > n) nt = (int)n;
  if (nt <= 1 || n < 16) {
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

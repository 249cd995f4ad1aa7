// Wald tests on a NEBULA fit: glm_sc_test (edge-rs src/sc/test.rs:98-201),
// chisq_sf at 1 df (src/numeric/dist.rs:737-746) and Benjamini-Hochberg
// (src/numeric/stats.rs:40-61).  Host code: O(genes * coefficients^2).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/nebula_hip.h"

namespace {

void put(char* err, size_t len, const std::string& m) {
  if (err && len) {
    std::strncpy(err, m.c_str(), len - 1);
    err[len - 1] = 0;
  }
}

// V[i][j] of a packed upper-triangular covariance (column-major triangle).
double packed_get(const double* v, size_t i, size_t j) {
  const size_t lo = i <= j ? i : j, hi = i <= j ? j : i;
  return v[hi * (hi + 1) / 2 + lo];
}

// P(chi^2_1 > x) = erfc(sqrt(x / 2)); 1 for x <= 0, 0 for x = inf, NaN stays NaN.
double chisq1_sf(double x) {
  if (x <= 0.0) return 1.0;
  if (std::isinf(x)) return 0.0;
  return std::erfc(std::sqrt(0.5 * x));
}

// IEEE total order as a signed integer key (NaN with the sign bit clear
// above +inf), the order edge-rs sorts p-values in.
int64_t total_key(double v) {
  int64_t b;
  std::memcpy(&b, &v, sizeof b);
  return b ^ (int64_t)((uint64_t)(b >> 63) >> 1);  // negatives: all bits but the sign flipped
}

}  // namespace

extern "C" int nh_glm_sc_test(const double* coefficients, size_t coefficients_len,
                              const double* covariance, size_t covariance_len, size_t n_genes,
                              size_t n_coef, size_t coef, const double* contrast,
                              double* log_fc, double* se, double* z, double* p_value,
                              double* fdr, char* err, size_t err_len) {
  if (n_genes == 0 || n_coef == 0) {
    put(err, err_len, "fit dimensions must be positive");
    return NH_ERR_MUST_BE_POSITIVE;
  }
  if (coefficients_len != n_genes * n_coef) {
    put(err, err_len, "coefficients: expected " + std::to_string(n_genes * n_coef) + ", got " +
                          std::to_string(coefficients_len));
    return NH_ERR_LENGTH_MISMATCH;
  }
  const size_t packed = n_coef * (n_coef + 1) / 2;
  if (covariance_len != n_genes * packed) {
    put(err, err_len, "covariance: expected " + std::to_string(n_genes * packed) + ", got " +
                          std::to_string(covariance_len));
    return NH_ERR_LENGTH_MISMATCH;
  }
  if (!coefficients || !covariance || !log_fc || !se || !z || !p_value || !fdr) {
    put(err, err_len, "null argument");
    return NH_ERR_INVALID_ARGUMENT;
  }
  std::vector<double> w(n_coef, 0.0);
  if (contrast) {
    bool any = false;
    for (size_t j = 0; j < n_coef; ++j) {
      w[j] = contrast[j];
      any = any || w[j] != 0.0;
    }
    if (!any) {
      put(err, err_len, "contrast is entirely zero");
      return NH_ERR_INVALID_ARGUMENT;
    }
  } else {
    if (coef >= n_coef) {
      put(err, err_len, "coefficient " + std::to_string(coef) + " out of range for " +
                            std::to_string(n_coef) + " coefficients");
      return NH_ERR_COEF_OUT_OF_RANGE;
    }
    w[coef] = 1.0;
  }
  for (size_t g = 0; g < n_genes; ++g) {
    const double* b = coefficients + g * n_coef;
    const double* v = covariance + g * packed;
    double effect = 0.0;
    for (size_t j = 0; j < n_coef; ++j) effect += b[j] * w[j];
    double var = 0.0;  // c' V c over the non-zero weights, off-diagonal terms included
    for (size_t i = 0; i < n_coef; ++i) {
      if (w[i] == 0.0) continue;
      for (size_t j = 0; j < n_coef; ++j)
        if (w[j] != 0.0) var += w[i] * w[j] * packed_get(v, i, j);
    }
    const double e = var > 0.0 ? std::sqrt(var) : 0.0;
    const double stat = e > 0.0 ? effect / e : 0.0;
    log_fc[g] = effect;
    se[g] = e;
    z[g] = stat;
    p_value[g] = e > 0.0 ? chisq1_sf(stat * stat) : 1.0;
  }
  // Benjamini-Hochberg: p-values in descending total order (stable), running
  // minimum of n / rank * p, capped at 1 (a NaN never lowers the minimum).
  std::vector<size_t> order(n_genes);
  std::iota(order.begin(), order.end(), (size_t)0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
    return total_key(p_value[b]) < total_key(p_value[a]);
  });
  const double n = (double)n_genes;
  double running = INFINITY;
  for (size_t j = 0; j < n_genes; ++j) {
    const size_t idx = order[j];
    const double rank = (double)(n_genes - j);
    running = std::fmin(running, n / rank * p_value[idx]);
    fdr[idx] = std::fmin(running, 1.0);
  }
  return NH_OK;
}

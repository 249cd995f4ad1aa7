// Log-gamma, digamma and trigamma in f64 for the oracle.
//
// edge-rs takes ln_gamma from statrs 0.18.0 (Lanczos; src/numeric/gamma.rs:213)
// and writes digamma/trigamma as upward recurrence to x >= 20 followed by the
// asymptotic Bernoulli series (src/numeric/gamma.rs:229-300).  Here lgamma is
// glibc's (agrees with Lanczos to ~1e-15 relative) and the two polygammas are
// the same published method: Abramowitz & Stegun 6.3.5 / 6.4.6 recurrences,
// then 6.3.18 / 6.4.12 series.  Pinned against scipy.special in
// tests/test_oracle_gamma.py.
#include <cmath>
#include <limits>

#include "oracle.hpp"

namespace oracle {

double ln_gamma(double x) { return std::lgamma(x); }

namespace {
constexpr double kShift = 20.0;
}  // namespace

double digamma(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return std::numeric_limits<double>::quiet_NaN();
  // psi(x) = psi(x + m) - sum_{j<m} 1/(x+j)
  double acc = 0.0;
  while (x < kShift) {
    acc += 1.0 / x;
    x += 1.0;
  }
  // psi(z) ~ ln z - 1/(2z) - sum_k B_2k / (2k z^2k)
  const double z2 = 1.0 / (x * x);
  // B2/2, B4/4, ..., B16/16
  const double c[8] = {1.0 / 12.0,  -1.0 / 120.0,     1.0 / 252.0, -1.0 / 240.0,
                       1.0 / 132.0, -691.0 / 32760.0, 1.0 / 12.0,  -3617.0 / 8160.0};
  double series = 0.0;
  for (int k = 7; k >= 0; --k) series = c[k] + z2 * series;
  series *= z2;
  return std::log(x) - 0.5 / x - series - acc;
}

double trigamma(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return std::numeric_limits<double>::quiet_NaN();
  double acc = 0.0;
  while (x < kShift + 1.0) {
    acc += 1.0 / (x * x);
    x += 1.0;
  }
  // psi'(z) ~ 1/z + 1/(2z^2) + sum_k B_2k / z^(2k+1)
  const double z2 = 1.0 / (x * x);
  const double b[8] = {1.0 / 6.0,  -1.0 / 30.0,      1.0 / 42.0, -1.0 / 30.0,
                       5.0 / 66.0, -691.0 / 2730.0, 7.0 / 6.0,  -3617.0 / 510.0};
  double series = 0.0;
  for (int k = 7; k >= 0; --k) series = b[k] + z2 * series;
  series *= z2 / x;
  return 1.0 / x + 0.5 * z2 + series + acc;
}

}  // namespace oracle

// Empirical-Bayes shrinkage of the per-gene cell-level overdispersions
// (src/sc/shrink.rs:118, shrink_sc_dispersion), host code on NebulaFit's
// output: phi = 1 / dispersion is treated as a sample variance on
// df_residual degrees of freedom and moderated by limma's squeezeVar
// (src/limma/squeeze_var.rs:286) in its equal-df family -- a robust fit
// (Winsorised moments of the log variances about a lowess trend, the prior
// degrees of freedom matched by quadrature of the F density, outliers
// down-weighted per gene), a spline-trended moment fit, or a plain one -- with
// shrink.rs's fallback ladder (trended robust, robust, plain, unshrunk).
//
// Special functions are this file's own: the F and chi-square densities in
// Loader's saddle-point form (stirlerr + bd0, Loader 2000), the regularized
// incomplete beta and gamma by their continued fractions and series
// (Abramowitz & Stegun 26.5.8, 6.5.29/31) with those densities as front
// factors, quantiles by safeguarded Newton; lowess is Cleveland's (1979)
// algorithm with R's stats::lowess window and delta rules; the natural
// cubic spline trend uses the truncated power basis (Hastie, Tibshirani &
// Friedman 5.4-5.5), whose least-squares fitted values do not depend on the
// basis chosen for the spline space.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/nebula_hip.h"

namespace nh {
namespace eb {

namespace {

constexpr double kInf = std::numeric_limits<double>::infinity();
constexpr double kEps = 2.220446049250313e-16;
constexpr double kLnSqrt2Pi = 0.918938533204672741780329736406;
constexpr double kLn2Pi = 1.837877066409345483560659472811;

struct Failure : std::runtime_error {
  int code;
  Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(const std::string& m, int code = NH_ERR_INVALID_ARGUMENT) {
  throw Failure(code, m);
}

// ---- polygamma ----------------------------------------------------------

// Stirling's series remainder lgamma(x) - ((x - 1/2) ln x - x + ln sqrt(2 pi)),
// x >= 10 (terms through B14; the next is < 3e-17 there).
double stirling_tail(double x) {
  const double r = 1.0 / (x * x);
  return (1.0 / 12 -
          r * (1.0 / 360 -
               r * (1.0 / 1260 -
                    r * (1.0 / 1680 - r * (1.0 / 1188 - r * (691.0 / 360360 - r / 156)))))) /
         x;
}

// ln x - digamma(x), x > 0 (recurrence up to 20, then its asymptotic series).
double log_minus_digamma(double x) {
  double shift = 0.0, z = x;
  while (z < 20.0) {
    shift += 1.0 / z;
    z += 1.0;
  }
  const double r = 1.0 / (z * z);
  const double tail =
      0.5 / z + r * (1.0 / 12 - r * (1.0 / 120 - r * (1.0 / 252 - r * (1.0 / 240 - r * (1.0 / 132 - r * (691.0 / 32760 - r / 12))))));
  // ln x - psi(x) = ln x - psi(z) + shift = (ln x - ln z) + (ln z - psi(z)) + shift
  return std::log(x / z) + tail + shift;
}

double trigamma(double x) {
  double acc = 0.0, z = x;
  while (z < 20.0) {
    acc += 1.0 / (z * z);
    z += 1.0;
  }
  const double r = 1.0 / (z * z);
  const double series =
      1.0 / z + 0.5 * r +
      r / z * (1.0 / 6 - r * (1.0 / 30 - r * (1.0 / 42 - r * (1.0 / 30 - r * (5.0 / 66 - r * (691.0 / 2730 - r * 7.0 / 6))))));
  return acc + series;
}

double tetragamma(double x) {
  double acc = 0.0, z = x;
  while (z < 20.0) {
    acc -= 2.0 / (z * z * z);
    z += 1.0;
  }
  const double r = 1.0 / (z * z);
  const double series =
      -r - r / z - 0.5 * r * r +
      r * r * r * (1.0 / 6 - r * (1.0 / 6 - r * (3.0 / 10 - r * (5.0 / 6 - r * (691.0 / 210 - r * 35.0 / 2)))));
  return acc + series;
}

// y with trigamma(y) = x (limma's Newton iteration from 1/2 + 1/x).
double trigamma_inverse(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return std::numeric_limits<double>::quiet_NaN();
  if (x > 1e7) return 1.0 / std::sqrt(x);
  if (x < 1e-6) return 1.0 / x;
  double y = 0.5 + 1.0 / x;
  for (int it = 0; it < 50; ++it) {
    const double tri = trigamma(y);
    const double dif = tri * (1.0 - tri / x) / tetragamma(y);
    y += dif;
    if (y <= 0.0) y = x;
    if (std::fabs(dif / y) < 1e-10) break;
  }
  return y;
}

// ---- Loader's saddle-point pieces ---------------------------------------

// lgamma(n + 1) - (n + 1/2) ln n + n - ln sqrt(2 pi), n > 0
double stirlerr(double n) {
  if (n >= 10.0) return stirling_tail(n);
  return std::lgamma(n + 1.0) - (n + 0.5) * std::log(n) + n - kLnSqrt2Pi;
}

// x ln(x / np) + np - x without cancellation near x = np
double bd0(double x, double np) {
  if (std::fabs(x - np) < 0.1 * (x + np)) {
    double v = (x - np) / (x + np);
    double s = (x - np) * v;
    double ej = 2.0 * x * v;
    const double v2 = v * v;
    for (int j = 1; j < 1000; ++j) {
      ej *= v2;
      const double s1 = s + ej / (2 * j + 1);
      if (s1 == s) return s1;
      s = s1;
    }
    return s;
  }
  return x * std::log(x / np) + np - x;
}

// C(n, x) p^x q^(n-x) for real 0 <= x <= n (p + q = 1, both given)
double binom_raw(double x, double n, double p, double q) {
  if (p == 0.0) return x == 0.0 ? 1.0 : 0.0;
  if (q == 0.0) return x == n ? 1.0 : 0.0;
  if (x == 0.0) {
    if (n == 0.0) return 1.0;
    return std::exp(p < 0.1 ? -bd0(n, n * q) - n * p : n * std::log(q));
  }
  if (x == n) return std::exp(q < 0.1 ? -bd0(n, n * p) - n * q : n * std::log(p));
  const double lc = stirlerr(n) - stirlerr(x) - stirlerr(n - x) - bd0(x, n * p) - bd0(n - x, n * q);
  const double lf = kLn2Pi + std::log(x) + std::log1p(-x / n);
  return std::exp(lc - 0.5 * lf);
}

// x^k e^-lambda / Gamma(k + 1) for real k >= 0
double poisson_raw(double k, double lambda) {
  if (lambda == 0.0) return k == 0.0 ? 1.0 : 0.0;
  if (k == 0.0) return std::exp(-lambda);
  return std::exp(-stirlerr(k) - bd0(k, lambda)) / std::sqrt(2.0 * M_PI * k);
}

// ---- densities ----------------------------------------------------------

// density of chi^2_k at y
double chisq_density(double y, double k) {
  if (y < 0.0) return 0.0;
  const double a = 0.5 * k;
  if (y == 0.0) return a < 1.0 ? kInf : (a == 1.0 ? 0.5 : 0.0);
  if (a < 1.0) return poisson_raw(a, 0.5 * y) * a / y;
  return 0.5 * poisson_raw(a - 1.0, 0.5 * y);
}

// density of F(m, n) at x (n may be infinite: chi^2_m / m)
double f_density(double x, double m, double n) {
  if (x < 0.0) return 0.0;
  if (std::isinf(n)) return m * chisq_density(m * x, m);
  if (x == 0.0) return m > 2.0 ? 0.0 : (m == 2.0 ? 1.0 : kInf);
  const double f = 1.0 / (n + x * m);
  const double q = n * f, p = x * m * f;
  if (m >= 2.0) return m * q / 2.0 * binom_raw((m - 2.0) / 2.0, (m + n - 2.0) / 2.0, p, q);
  return m * m * q / (2.0 * p * (m + n)) * binom_raw(m / 2.0, (m + n) / 2.0, p, q);
}

// ---- incomplete beta and gamma ------------------------------------------

// Continued fraction of I_x(a, b) / front (A&S 26.5.8, modified Lentz),
// convergent for x < (a + 1) / (a + b + 2).
double beta_cf(double a, double b, double x) {
  const double tiny = 1e-300;
  const double qab = a + b, qap = a + 1.0, qam = a - 1.0;
  double c = 1.0, d = 1.0 - qab * x / qap;
  if (std::fabs(d) < tiny) d = tiny;
  d = 1.0 / d;
  double h = d;
  for (int m = 1; m < 100000; ++m) {
    const double m2 = 2.0 * m;
    double aa = m * (b - m) * x / ((qam + m2) * (a + m2));
    d = 1.0 + aa * d;
    if (std::fabs(d) < tiny) d = tiny;
    c = 1.0 + aa / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    h *= d * c;
    aa = -(a + m) * (qab + m) * x / ((a + m2) * (qap + m2));
    d = 1.0 + aa * d;
    if (std::fabs(d) < tiny) d = tiny;
    c = 1.0 + aa / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < 1e-16) break;
  }
  return h;
}

// x^a y^b / (a B(a, b)) with y = 1 - x, in the binomial saddle-point form:
// x^a y^b / B(a, b) = a b / (a + b) C(a + b, a) x^a y^b.
double beta_front(double a, double b, double x, double y) {
  return b / (a + b) * binom_raw(a, a + b, x, y);
}

// Both tails of the regularized incomplete beta at (x, y = 1 - x): lower =
// I_x(a, b), upper = 1 - I_x(a, b), the smaller computed directly.
void ibeta(double a, double b, double x, double y, double& lower, double& upper) {
  if (x <= 0.0) {
    lower = 0.0;
    upper = 1.0;
    return;
  }
  if (y <= 0.0) {
    lower = 1.0;
    upper = 0.0;
    return;
  }
  if (x < (a + 1.0) / (a + b + 2.0)) {
    lower = beta_front(a, b, x, y) * beta_cf(a, b, x);
    upper = 1.0 - lower;
  } else {
    upper = beta_front(b, a, y, x) * beta_cf(b, a, y);
    lower = 1.0 - upper;
  }
}

// Both tails of the regularized incomplete gamma: lower = P(a, x), upper = Q.
void igamma(double a, double x, double& lower, double& upper) {
  if (x <= 0.0) {
    lower = 0.0;
    upper = 1.0;
    return;
  }
  const double front = poisson_raw(a, x);  // x^a e^-x / Gamma(a + 1)
  if (x < a + 1.0) {
    double term = 1.0, sum = 1.0, ap = a;
    for (int n = 1; n < 100000; ++n) {
      ap += 1.0;
      term *= x / ap;
      sum += term;
      if (std::fabs(term) < std::fabs(sum) * 1e-17) break;
    }
    lower = front * sum;
    upper = 1.0 - lower;
  } else {
    // Q = x^a e^-x / Gamma(a) * CF (A&S 6.5.31, modified Lentz)
    const double tiny = 1e-300;
    double b = x + 1.0 - a, c = 1.0 / tiny, d = 1.0 / b, h = d;
    for (int i = 1; i < 100000; ++i) {
      const double an = -i * (i - a);
      b += 2.0;
      d = an * d + b;
      if (std::fabs(d) < tiny) d = tiny;
      c = b + an / c;
      if (std::fabs(c) < tiny) c = tiny;
      d = 1.0 / d;
      const double del = d * c;
      h *= del;
      if (std::fabs(del - 1.0) < 1e-16) break;
    }
    upper = a * front * h;
    lower = 1.0 - upper;
  }
}

// P(F(m, n) > x)
double f_sf(double x, double m, double n) {
  if (x <= 0.0) return 1.0;
  if (std::isinf(x)) return 0.0;
  if (std::isinf(n)) {
    double lo, up;
    igamma(0.5 * m, 0.5 * m * x, lo, up);
    return up;
  }
  const double den = n + m * x;
  double lo, up;
  ibeta(0.5 * m, 0.5 * n, m * x / den, n / den, lo, up);
  return up;
}

// P(chi^2_k > x)
double chisq_sf(double x, double k) {
  double lo, up;
  igamma(0.5 * k, 0.5 * x, lo, up);
  return up;
}

// Quantile of F(m, n) at lower-tail probability p (n may be infinite):
// bracket, then Newton on the smaller tail, bisecting when a step leaves
// the bracket.
double f_quantile(double p, double m, double n) {
  if (!(p > 0.0 && p < 1.0)) fail("F quantile needs 0 < p < 1");
  const bool use_lower = p <= 0.5;
  const double target = use_lower ? p : 1.0 - p;
  // tail(x) - target, increasing in x for the lower tail, decreasing for the upper
  auto tail = [&](double x) {
    double lo, up;
    if (std::isinf(n)) {
      igamma(0.5 * m, 0.5 * m * x, lo, up);
    } else {
      const double den = n + m * x;
      ibeta(0.5 * m, 0.5 * n, m * x / den, n / den, lo, up);
    }
    return use_lower ? lo : up;
  };
  double lo_x = 0.0, hi_x = 1.0;
  // grow hi until the lower-tail probability passes p
  while ((use_lower ? tail(hi_x) < target : tail(hi_x) > target)) {
    lo_x = hi_x;
    hi_x *= 2.0;
    if (hi_x > 1e300) fail("F quantile not bracketed");
  }
  double x = 0.5 * (lo_x + hi_x);
  for (int it = 0; it < 200; ++it) {
    const double t = tail(x);
    const double g = use_lower ? t - target : target - t;  // increasing in x
    if (g == 0.0) return x;
    if (g < 0.0) lo_x = x; else hi_x = x;
    const double dens = f_density(x, m, n);
    double next = dens > 0.0 ? x - g / dens : 0.5 * (lo_x + hi_x);
    if (!(next > lo_x && next < hi_x)) next = 0.5 * (lo_x + hi_x);
    if (std::fabs(next - x) <= 2.0 * kEps * std::fabs(x)) return next;
    x = next;
    if (hi_x - lo_x <= 2.0 * kEps * hi_x) return x;
  }
  return x;
}

// ---- order statistics ---------------------------------------------------

double quantile7(std::vector<double> v, double prob) {
  std::sort(v.begin(), v.end());
  const size_t n = v.size();
  const double h = (double)(n - 1) * prob;
  const size_t lo = (size_t)std::floor(h);
  const double frac = h - (double)lo;
  const double hi = lo + 1 < n ? v[lo + 1] : v[lo];
  return (1.0 - frac) * v[lo] + frac * hi;
}

double median_of(std::vector<double> v) {
  std::sort(v.begin(), v.end());
  const size_t n = v.size();
  if (n == 0) return std::numeric_limits<double>::quiet_NaN();
  return n % 2 ? v[n / 2] : 0.5 * (v[n / 2 - 1] + v[n / 2]);
}

// R's mean(x, trim): the mean of the finite values after dropping
// floor(trim n) from each end
double trimmed_mean(const std::vector<double>& x, double trim) {
  std::vector<double> v;
  for (double a : x)
    if (std::isfinite(a)) v.push_back(a);
  if (v.empty()) fail("trimmed mean of no finite values");
  std::sort(v.begin(), v.end());
  const size_t n = v.size(), k = (size_t)std::floor(trim * (double)n);
  const size_t lo = 2 * k >= n ? 0 : k, hi = 2 * k >= n ? n : n - k;
  double s = 0.0;
  for (size_t i = lo; i < hi; ++i) s += v[i];
  return s / (double)(hi - lo);
}

// 1-based ranks, ties sharing their mean rank
std::vector<double> average_ranks(const std::vector<double>& x) {
  const size_t n = x.size();
  std::vector<size_t> o(n);
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) { return x[a] < x[b]; });
  std::vector<double> r(n);
  for (size_t s = 0; s < n;) {
    size_t e = s + 1;
    while (e < n && x[o[e]] == x[o[s]]) ++e;
    const double mean_rank = 0.5 * (double)(s + 1 + e);
    for (size_t i = s; i < e; ++i) r[o[i]] = mean_rank;
    s = e;
  }
  return r;
}

std::vector<size_t> ascending_order(const std::vector<double>& x) {
  std::vector<size_t> o(x.size());
  std::iota(o.begin(), o.end(), 0);
  std::stable_sort(o.begin(), o.end(), [&](size_t a, size_t b) { return x[a] < x[b]; });
  return o;
}

// ---- lowess (Cleveland 1979; R's stats::lowess window and delta rules) ---

// The locally weighted line through points nleft..nright (sorted x) at xs,
// with robustness weights rw when `robust`; false when no point has weight.
bool lowess_point(const std::vector<double>& x, const std::vector<double>& y, double xs,
                  size_t nleft, size_t nright, const std::vector<double>& rw, bool robust,
                  std::vector<double>& w, double& ys) {
  const size_t n = x.size();
  const double range = x[n - 1] - x[0];
  const double h = std::max(xs - x[nleft], x[nright] - xs);
  const double h9 = 0.999 * h, h1 = 0.001 * h;
  double a = 0.0;
  size_t j = nleft;
  for (; j < n; ++j) {
    w[j] = 0.0;
    const double r = std::fabs(x[j] - xs);
    if (r <= h9) {
      if (r <= h1) {
        w[j] = 1.0;
      } else {
        const double q = r / h;
        const double t = 1.0 - q * q * q;
        w[j] = t * t * t;
      }
      if (robust) w[j] *= rw[j];
      a += w[j];
    } else if (x[j] > xs) {
      break;
    }
  }
  if (a <= 0.0 || j == nleft) return false;
  const size_t nrt = j - 1;
  for (size_t i = nleft; i <= nrt; ++i) w[i] /= a;
  if (h > 0.0) {
    double xm = 0.0;
    for (size_t i = nleft; i <= nrt; ++i) xm += w[i] * x[i];
    double b = xs - xm, c = 0.0;
    for (size_t i = nleft; i <= nrt; ++i) c += w[i] * (x[i] - xm) * (x[i] - xm);
    if (std::sqrt(c) > 0.001 * range) {
      b /= c;
      for (size_t i = nleft; i <= nrt; ++i) w[i] *= b * (x[i] - xm) + 1.0;
    }
  }
  ys = 0.0;
  for (size_t i = nleft; i <= nrt; ++i) ys += w[i] * y[i];
  return true;
}

// Fitted values at sorted x; span f, `iterations` robustness passes after the
// first fit, points within delta of a fitted point linearly interpolated.
std::vector<double> lowess_sorted(const std::vector<double>& x, const std::vector<double>& y,
                                  double f, int iterations, double delta) {
  const size_t n = x.size();
  std::vector<double> ys(n, 0.0), rw(n, 1.0), res(n, 0.0), w(n, 0.0);
  if (n < 2) {
    ys = y;
    return ys;
  }
  size_t ns = (size_t)std::max<long>(2, std::min<long>((long)n, (long)(f * (double)n + 1e-7)));
  for (int iter = 1; iter <= iterations + 1; ++iter) {
    size_t nleft = 0, nright = ns - 1;
    long last = -1;
    size_t i = 0;
    for (;;) {
      if (nright < n - 1) {
        const double d1 = x[i] - x[nleft], d2 = x[nright + 1] - x[i];
        if (d1 > d2) {
          ++nleft;
          ++nright;
          continue;
        }
      }
      double fit;
      if (!lowess_point(x, y, x[i], nleft, nright, rw, iter > 1, w, fit)) fit = y[i];
      ys[i] = fit;
      if (last < (long)i - 1) {
        const double denom = x[i] - x[(size_t)last];
        for (size_t j = (size_t)(last + 1); j < i; ++j) {
          const double alpha = (x[j] - x[(size_t)last]) / denom;
          ys[j] = alpha * ys[i] + (1.0 - alpha) * ys[(size_t)last];
        }
      }
      last = (long)i;
      const double cut = x[(size_t)last] + delta;
      for (i = (size_t)last + 1; i < n; ++i) {
        if (x[i] > cut) break;
        if (x[i] == x[(size_t)last]) {
          ys[i] = ys[(size_t)last];
          last = (long)i;
        }
      }
      i = std::max((size_t)last + 1, i - 1);
      if (last >= (long)n - 1) break;
    }
    for (size_t k = 0; k < n; ++k) res[k] = y[k] - ys[k];
    double sc = 0.0;
    for (size_t k = 0; k < n; ++k) sc += std::fabs(res[k]);
    sc /= (double)n;
    if (iter > iterations) break;
    std::vector<double> ar(n);
    for (size_t k = 0; k < n; ++k) ar[k] = std::fabs(res[k]);
    std::vector<double> srt = ar;
    const size_t m1 = n / 2;
    std::nth_element(srt.begin(), srt.begin() + (long)m1, srt.end());
    double cmad;
    if (n % 2 == 0) {
      const double hi = srt[m1];
      const size_t m2 = n - m1 - 1;
      std::nth_element(srt.begin(), srt.begin() + (long)m2, srt.end());
      cmad = 3.0 * (hi + srt[m2]);
    } else {
      cmad = 6.0 * srt[m1];
    }
    if (cmad < 1e-7 * sc) break;
    const double c9 = 0.999 * cmad, c1 = 0.001 * cmad;
    for (size_t k = 0; k < n; ++k) {
      const double r = ar[k];
      if (r <= c1) {
        rw[k] = 1.0;
      } else if (r <= c9) {
        const double u = r / cmad;
        rw[k] = (1.0 - u * u) * (1.0 - u * u);
      } else {
        rw[k] = 0.0;
      }
    }
  }
  return ys;
}

// stats::lowess(x, y, f, iter) in the input order (delta = 0.01 x range)
std::vector<double> lowess(const std::vector<double>& x, const std::vector<double>& y, double f,
                           int iterations) {
  const size_t n = x.size();
  std::vector<size_t> o = ascending_order(x);
  std::vector<double> xs(n), ysrc(n);
  for (size_t k = 0; k < n; ++k) {
    xs[k] = x[o[k]];
    ysrc[k] = y[o[k]];
  }
  const double delta = 0.01 * (xs[n - 1] - xs[0]);
  std::vector<double> fit = lowess_sorted(xs, ysrc, f, iterations, delta);
  std::vector<double> out(n);
  for (size_t k = 0; k < n; ++k) out[o[k]] = fit[k];
  return out;
}

// ---- natural cubic spline trend ------------------------------------------

// Fitted values and residual sum of squares of the least-squares fit of e on
// a natural cubic spline in x with `df` basis functions (intercept included;
// boundary knots at the range, df - 2 interior knots at equally spaced
// quantiles), and the basis rank (Householder QR, columns kept while their
// diagonal exceeds 1e-7 of the first).
void spline_trend(const std::vector<double>& x, const std::vector<double>& e, int df,
                  std::vector<double>& fitted, double& rss, int& rank) {
  const size_t n = x.size();
  std::vector<double> srt(x);
  std::sort(srt.begin(), srt.end());
  const double lower = srt.front(), upper = srt.back();
  int cols = df;
  std::vector<double> knots;
  if (df - 2 == 0 || lower == upper) {
    cols = 2;
  } else {
    knots.push_back(lower);
    for (int j = 1; j <= df - 2; ++j) knots.push_back(quantile7(srt, (double)j / (df - 1)));
    knots.push_back(upper);
    std::sort(knots.begin(), knots.end());
  }
  // column-major basis
  std::vector<double> A(n * (size_t)cols);
  auto cube = [](double v, double k) {
    const double t = std::max(v - k, 0.0);
    return t * t * t;
  };
  for (size_t i = 0; i < n; ++i) {
    A[i] = 1.0;
    A[n + i] = x[i];
    if (cols > 2) {
      const size_t K = knots.size();
      const double last = knots[K - 1];
      auto d = [&](double k) {
        const double span = last - k;
        return span == 0.0 ? 0.0 : (cube(x[i], k) - cube(x[i], last)) / span;
      };
      const double dprev = d(knots[K - 2]);
      for (size_t j = 0; j + 2 < K; ++j) A[(2 + j) * n + i] = d(knots[j]) - dprev;
    }
  }
  // Householder QR of A, applied to b = e
  std::vector<double> b(e);
  std::vector<double> diag((size_t)cols, 0.0);
  const int kmax = std::min<int>(cols, (int)n);
  for (int k = 0; k < kmax; ++k) {
    double* a = &A[(size_t)k * n];
    double norm = 0.0;
    for (size_t i = (size_t)k; i < n; ++i) norm += a[i] * a[i];
    norm = std::sqrt(norm);
    if (norm == 0.0) {
      diag[(size_t)k] = 0.0;
      continue;
    }
    const double alpha = a[k] > 0 ? -norm : norm;
    diag[(size_t)k] = alpha;
    const double head = a[k];
    a[k] -= alpha;  // v = a - alpha e_k, stored in place
    const double vnorm2 = norm * norm - head * head + a[k] * a[k];
    auto reflect = [&](double* c) {
      double dot = 0.0;
      for (size_t i = (size_t)k; i < n; ++i) dot += a[i] * c[i];
      const double s = 2.0 * dot / vnorm2;
      for (size_t i = (size_t)k; i < n; ++i) c[i] -= s * a[i];
    };
    for (int j = k + 1; j < cols; ++j) reflect(&A[(size_t)j * n]);
    reflect(b.data());
  }
  // numerical rank at 1e-7 of the first diagonal (the residual variance's
  // divisor); an exactly singular basis has no least-squares coefficients
  rank = 0;
  const double pivot = std::fabs(diag[0]);
  for (int k = 0; k < kmax; ++k) {
    if (diag[(size_t)k] == 0.0) fail("the natural spline basis of the variance trend is singular");
    if (std::fabs(diag[(size_t)k]) > 1e-7 * pivot) ++rank;
  }
  // Q^T e is in b: the fit keeps the first `cols` entries, the residual the rest
  rss = 0.0;
  for (size_t i = (size_t)cols; i < n; ++i) rss += b[i] * b[i];
  // fitted = e - Q (0, b[cols..]): apply the reflections in reverse to the residual part
  std::vector<double> r(n, 0.0);
  for (size_t i = (size_t)cols; i < n; ++i) r[i] = b[i];
  for (int k = kmax - 1; k >= 0; --k) {
    const double* a = &A[(size_t)k * n];
    const double alpha = diag[(size_t)k];
    if (alpha == 0.0) continue;
    double vnorm2 = 0.0;
    for (size_t i = (size_t)k; i < n; ++i) vnorm2 += a[i] * a[i];
    double dot = 0.0;
    for (size_t i = (size_t)k; i < n; ++i) dot += a[i] * r[i];
    const double s = 2.0 * dot / vnorm2;
    for (size_t i = (size_t)k; i < n; ++i) r[i] -= s * a[i];
  }
  fitted.assign(n, 0.0);
  for (size_t i = 0; i < n; ++i) fitted[i] = e[i] - r[i];
}

// ---- moment fits (limma's fitFDist) ---------------------------------------

double mean_of(const std::vector<double>& v) {
  double s = 0.0;
  for (double a : v) s += a;
  return s / (double)v.size();
}

// variances clamped at zero, then lifted to at least 1e-5 x their median (x 1
// when more than half are zero)
std::vector<double> off_zero(std::vector<double> x) {
  for (double& v : x) v = std::max(v, 0.0);
  double m = median_of(x);
  if (m == 0.0) m = 1.0;
  const double floor = 1e-5 * m;
  for (double& v : x) v = std::max(v, floor);
  return x;
}

// infinite covariate values moved one unit past the finite range
std::vector<double> finite_covariate(const std::vector<double>& c) {
  double lo = kInf, hi = -kInf;
  bool all = true;
  for (double v : c) {
    if (std::isnan(v)) fail("the variance trend's covariate holds NaN");
    if (std::isfinite(v)) {
      lo = std::min(lo, v);
      hi = std::max(hi, v);
    } else {
      all = false;
    }
  }
  if (all) return c;
  std::vector<double> out(c.size());
  for (size_t i = 0; i < c.size(); ++i) {
    const double v = c[i];
    if (std::isfinite(v))
      out[i] = v;
    else if (lo > hi)  // nothing finite: the signs
      out[i] = v > 0 ? 1.0 : -1.0;
    else
      out[i] = v > 0 ? hi + 1.0 : lo - 1.0;
  }
  return out;
}

// scale and df2 of a scaled-F prior by moments, no covariate
void fit_f(const std::vector<double>& x, double df1, double& scale, double& df2) {
  const size_t n = x.size();
  if (n == 1) {
    scale = x[0];
    df2 = 0.0;
    return;
  }
  std::vector<double> xo = off_zero(x);
  std::vector<double> e(n);
  for (size_t i = 0; i < n; ++i) e[i] = std::log(xo[i]) + log_minus_digamma(0.5 * df1);
  const double em = mean_of(e);
  double ss = 0.0;
  for (double v : e) ss += (v - em) * (v - em);
  const double evar = ss / (double)(n - 1) - trigamma(0.5 * df1);
  if (evar > 0.0) {
    df2 = 2.0 * trigamma_inverse(evar);
    scale = std::exp(em - log_minus_digamma(0.5 * df2));
  } else {
    df2 = kInf;
    scale = mean_of(xo);
  }
}

// the same with a natural-spline trend in the covariate (per-gene scale)
void fit_f_trend(const std::vector<double>& x, double df1, const std::vector<double>& cov_in,
                 std::vector<double>& scale, double& df2) {
  const size_t n = x.size();
  if (n == 1) {
    scale.assign(1, x[0]);
    df2 = 0.0;
    return;
  }
  const std::vector<double> cov = finite_covariate(cov_in);
  std::vector<double> u(cov);
  std::sort(u.begin(), u.end());
  u.erase(std::unique(u.begin(), u.end()), u.end());
  int sdf = 1 + (n >= 3) + (n >= 6) + (n >= 30);
  sdf = std::min<int>(sdf, (int)u.size());
  if (sdf < 2) {
    double s;
    fit_f(x, df1, s, df2);
    scale.assign(n, s);
    return;
  }
  std::vector<double> xo = off_zero(x);
  std::vector<double> e(n);
  for (size_t i = 0; i < n; ++i) e[i] = std::log(xo[i]) + log_minus_digamma(0.5 * df1);
  std::vector<double> fitted;
  double rss;
  int rank;
  spline_trend(cov, e, sdf, fitted, rss, rank);
  const double evar = ((size_t)rank < n ? rss / (double)(n - (size_t)rank) : 0.0) -
                      trigamma(0.5 * df1);
  scale.resize(n);
  if (evar > 0.0) {
    df2 = 2.0 * trigamma_inverse(evar);
    const double shift = log_minus_digamma(0.5 * df2);
    for (size_t i = 0; i < n; ++i) scale[i] = std::exp(fitted[i] - shift);
  } else {
    df2 = kInf;
    for (size_t i = 0; i < n; ++i) scale[i] = std::exp(fitted[i]);
  }
}

// ---- the robust fit (limma's fitFDistRobustly) ----------------------------

// Gauss-Legendre nodes and weights on (0, 1), weights summing to one
void gauss_legendre(int n, std::vector<double>& nodes, std::vector<double>& weights) {
  nodes.assign((size_t)n, 0.0);
  weights.assign((size_t)n, 0.0);
  for (int i = 0; i < (n + 1) / 2; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), pp = 0.0;
    for (int it = 0; it < 100; ++it) {
      double p1 = 1.0, p2 = 0.0;
      for (int j = 0; j < n; ++j) {
        const double p3 = p2;
        p2 = p1;
        p1 = ((2.0 * j + 1.0) * z * p2 - j * p3) / (j + 1.0);
      }
      pp = n * (z * p1 - p2) / (z * z - 1.0);
      const double z1 = z;
      z = z1 - p1 / pp;
      if (std::fabs(z - z1) <= kEps) break;
    }
    nodes[(size_t)i] = 0.5 * (1.0 - z);
    nodes[(size_t)(n - 1 - i)] = 0.5 * (1.0 + z);
    weights[(size_t)i] = 1.0 / ((1.0 - z * z) * pp * pp);
    weights[(size_t)(n - 1 - i)] = weights[(size_t)i];
  }
}

double link(double d) { return std::isinf(d) ? 1.0 : d / (1.0 + d); }
double unlink(double p) { return p / (1.0 - p); }

struct Quad {
  std::vector<double> nodes, weights;
};

// Mean and variance of log F(df1, df2) Winsorised at the (lo, hi) tail
// probabilities, the middle by quadrature over u = F / (1 + F)
void winsorized_moments(double df1, double df2, double plo, double phi, const Quad& q,
                        double& mean, double& var) {
  const double f0 = f_quantile(plo, df1, df2), f1 = f_quantile(1.0 - phi, df1, df2);
  const double z0 = std::log(f0), z1 = std::log(f1);
  const double u0 = link(f0), u1 = link(f1), width = u1 - u0;
  const size_t n = q.nodes.size();
  std::vector<double> zn(n), dn(n);
  for (size_t i = 0; i < n; ++i) {
    const double u = u0 + width * q.nodes[i];
    const double f = unlink(u);
    zn[i] = std::log(f);
    dn[i] = f_density(f, df1, df2) / ((1.0 - u) * (1.0 - u));
  }
  double m = 0.0;
  for (size_t i = 0; i < n; ++i) m += q.weights[i] * dn[i] * zn[i];
  m = width * m + z0 * plo + z1 * phi;
  double v = 0.0;
  for (size_t i = 0; i < n; ++i) v += q.weights[i] * dn[i] * (zn[i] - m) * (zn[i] - m);
  v = width * v + (z0 - m) * (z0 - m) * plo + (z1 - m) * (z1 - m) * phi;
  mean = m;
  var = v;
}

// Brent's root of f on [a, b] (f(a), f(b) of opposite signs): inverse
// quadratic or secant steps, bisection when they stall.
template <class F>
double brent_root(F&& f, double a, double b, double xtol, double rtol, int max_iter) {
  double fa = f(a), fb = f(b);
  if (fa == 0.0) return a;
  if (fb == 0.0) return b;
  if ((fa > 0) == (fb > 0)) fail("root not bracketed");
  double c = a, fc = fa, d = b - a, e = d;
  for (int it = 0; it < max_iter; ++it) {
    if ((fb > 0) == (fc > 0)) {
      c = a;
      fc = fa;
      d = e = b - a;
    }
    if (std::fabs(fc) < std::fabs(fb)) {
      a = b;
      b = c;
      c = a;
      fa = fb;
      fb = fc;
      fc = fa;
    }
    const double tol = 0.5 * (xtol + rtol * std::fabs(b));
    const double m = 0.5 * (c - b);
    if (fb == 0.0 || std::fabs(m) < tol) return b;
    if (std::fabs(e) >= tol && std::fabs(fa) > std::fabs(fb)) {
      double p, qq;
      const double s = fb / fa;
      if (a == c) {
        p = 2.0 * m * s;
        qq = 1.0 - s;
      } else {
        const double q1 = fa / fc, r = fb / fc;
        p = s * (2.0 * m * q1 * (q1 - r) - (b - a) * (r - 1.0));
        qq = (q1 - 1.0) * (r - 1.0) * (s - 1.0);
      }
      if (p > 0) qq = -qq; else p = -p;
      if (2.0 * p < std::min(3.0 * m * qq - std::fabs(tol * qq), std::fabs(e * qq))) {
        e = d;
        d = p / qq;
      } else {
        d = m;
        e = m;
      }
    } else {
      d = m;
      e = m;
    }
    a = b;
    fa = fb;
    b += std::fabs(d) > tol ? d : (m > 0 ? tol : -tol);
    fb = f(b);
  }
  return b;
}

struct RobustFit {
  std::vector<double> scale;    // 1 or n
  double df2;
  std::vector<double> df2_shrunk;  // n
};

// per-gene prior df: df2 discounted for genes further into the tail than
// their rank implies (limma's ProbNotOutlier mixture with df2.outlier), then
// made monotone in the tail probability
std::vector<double> shrink_df2(const std::vector<double>& log_tail, const std::vector<double>& fstat,
                               double df1, double df2) {
  const size_t n = fstat.size();
  const std::vector<double> rk = average_ranks(fstat);
  std::vector<double> lpn(n);
  bool any = false;
  for (size_t i = 0; i < n; ++i) {
    lpn[i] = std::min(log_tail[i] - (std::log((double)n - rk[i] + 0.5) - std::log((double)n)), 0.0);
    any = any || lpn[i] < 0.0;
  }
  if (!any) return std::vector<double>(n, df2);
  double min_lt = kInf, fmax = -kInf;
  for (size_t i = 0; i < n; ++i) {
    min_lt = std::min(min_lt, log_tail[i]);
    fmax = std::max(fmax, fstat[i]);
  }
  std::vector<double> out(n);
  const double ln_half = -M_LN2;
  if (min_lt == -kInf) {
    for (size_t i = 0; i < n; ++i) out[i] = std::exp(lpn[i]) * df2;
  } else {
    double df2_out = ln_half / min_lt * df2;
    const double new_lt = std::log(f_sf(fmax, df1, df2_out));
    df2_out *= ln_half / new_lt;
    for (size_t i = 0; i < n; ++i) out[i] = std::exp(lpn[i]) * df2 - std::expm1(lpn[i]) * df2_out;
  }
  // the leading run flattened to its smallest running mean, then a running max
  const std::vector<size_t> o = ascending_order(log_tail);
  double run = 0.0, best = kInf;
  size_t imin = 0;
  for (size_t k = 0; k < n; ++k) {
    run += out[o[k]];
    const double avg = run / (double)(k + 1);
    if (avg < best) {
      best = avg;
      imin = k;
    }
  }
  for (size_t k = 0; k <= imin; ++k) out[o[k]] = best;
  double mx = -kInf;
  for (size_t k = 0; k < n; ++k) {
    mx = std::max(mx, out[o[k]]);
    out[o[k]] = mx;
  }
  return out;
}

RobustFit robust_fit(const std::vector<double>& x_in, double df1, const std::vector<double>* cov,
                     double plo, double phi) {
  const size_t n = x_in.size();
  if (n < 2) fail("the robust fit needs at least two genes");
  if (cov)
    for (double c : *cov)
      if (!std::isfinite(c)) fail("the robust fit needs a finite covariate");
  RobustFit R;
  if (n == 2) {
    if (cov) {
      fit_f_trend(x_in, df1, *cov, R.scale, R.df2);
    } else {
      double s;
      fit_f(x_in, df1, s, R.df2);
      R.scale.assign(1, s);
    }
    R.df2_shrunk.assign(n, R.df2);
    return R;
  }
  std::vector<double> x(x_in);
  const double m = median_of(x);
  if (m <= 0.0) fail("the robust fit needs a positive median variance");
  for (double& v : x) v = std::max(v, m * 1e-12);
  std::vector<double> nr_scale;
  double nr_df2;
  if (cov) {
    fit_f_trend(x, df1, *cov, nr_scale, nr_df2);
  } else {
    double s;
    fit_f(x, df1, s, nr_df2);
    nr_scale.assign(1, s);
  }
  if (plo < 1.0 / (double)n && phi < 1.0 / (double)n) {
    R.scale = nr_scale;
    R.df2 = nr_df2;
    R.df2_shrunk.assign(n, nr_df2);
    return R;
  }
  std::vector<double> z(n);
  for (size_t i = 0; i < n; ++i) z[i] = std::log(x[i]);
  std::vector<double> ztrend;
  if (cov)
    ztrend = lowess(*cov, z, 0.4, 3);
  else
    ztrend.assign(1, trimmed_mean(z, phi));
  auto trend = [&](size_t i) { return ztrend.size() > 1 ? ztrend[i] : ztrend[0]; };
  std::vector<double> zres(n);
  for (size_t i = 0; i < n; ++i) zres[i] = z[i] - trend(i);
  const double lo = quantile7(zres, plo), hi = quantile7(zres, 1.0 - phi);
  std::vector<double> zw(n);
  for (size_t i = 0; i < n; ++i) zw[i] = std::min(std::max(zres[i], lo), hi);
  const double zwmean = mean_of(zw);
  double zwvar = 0.0;
  for (double v : zw) zwvar += (v - zwmean) * (v - zwmean);
  zwvar /= (double)(n - 1);
  Quad q;
  gauss_legendre(128, q.nodes, q.weights);
  double mom_mean_inf, mom_var_inf;
  winsorized_moments(df1, kInf, plo, phi, q, mom_mean_inf, mom_var_inf);
  const double funval_inf = std::log(zwvar / mom_var_inf);
  auto corrected_at = [&](double mom_mean) {
    std::vector<double> c(n);
    for (size_t i = 0; i < n; ++i) c[i] = trend(i) + zwmean - mom_mean;
    return c;
  };
  if (funval_inf <= 0.0) {
    // no finite df2 explains the spread: infinite prior df except for outliers
    const std::vector<double> corr = corrected_at(mom_mean_inf);
    R.scale.clear();
    if (ztrend.size() > 1)
      for (double c : corr) R.scale.push_back(std::exp(c));
    else
      R.scale.assign(1, std::exp(corr[0]));
    std::vector<double> fstat(n), tail(n);
    for (size_t i = 0; i < n; ++i) {
      fstat[i] = std::exp(z[i] - corr[i]);
      tail[i] = chisq_sf(fstat[i] * df1, df1);
    }
    const std::vector<double> rk = average_ranks(fstat);
    R.df2 = kInf;
    R.df2_shrunk.assign(n, kInf);
    bool any = false;
    for (size_t i = 0; i < n; ++i) {
      const double emp = ((double)n - rk[i] + 0.5) / (double)n;
      const double prob = std::min(tail[i] / emp, 1.0);
      if (prob < 1.0) {
        R.df2_shrunk[i] = prob * (double)n * df1;
        any = true;
      }
    }
    if (any) {
      const std::vector<size_t> o = ascending_order(tail);
      double mx = -kInf;
      for (size_t k = 0; k < n; ++k) {
        mx = std::max(mx, R.df2_shrunk[o[k]]);
        R.df2_shrunk[o[k]] = mx;
      }
    }
    return R;
  }
  if (std::isinf(nr_df2)) {
    R.scale = nr_scale;
    R.df2 = nr_df2;
    R.df2_shrunk.assign(n, nr_df2);
    return R;
  }
  auto objective = [&](double par) {
    double mm, vv;
    try {
      winsorized_moments(df1, unlink(par), plo, phi, q, mm, vv);
    } catch (const Failure&) {
      return funval_inf;
    }
    return std::log(zwvar / vv);
  };
  const double rbx = link(nr_df2);
  const double funval_low = objective(rbx);
  const double df2 =
      funval_low >= 0.0 ? nr_df2 : unlink(brent_root(objective, rbx, 1.0, 1e-15, 8.0 * kEps, 200));
  double mom_mean, mom_var;
  winsorized_moments(df1, df2, plo, phi, q, mom_mean, mom_var);
  const std::vector<double> corr = corrected_at(mom_mean);
  if (cov) {
    R.scale.resize(n);
    for (size_t i = 0; i < n; ++i) R.scale[i] = std::exp(corr[i]);
  } else {
    R.scale.assign(1, std::exp(corr[0]));
  }
  std::vector<double> fstat(n), log_tail(n);
  for (size_t i = 0; i < n; ++i) {
    fstat[i] = std::exp(z[i] - corr[i]);
    log_tail[i] = std::log(f_sf(fstat[i], df1, df2));
  }
  R.df2 = df2;
  R.df2_shrunk = shrink_df2(log_tail, fstat, df1, df2);
  return R;
}

// ---- squeezeVar, equal df -------------------------------------------------

struct Squeezed {
  std::vector<double> var_post, var_prior, df_prior;
};

Squeezed squeeze(const std::vector<double>& var, double df, const std::vector<double>* cov,
                 bool robust) {
  const size_t n = var.size();
  for (double v : var)
    if (!std::isfinite(v) || v < 0.0) fail("variances must be finite and non-negative");
  if (cov)
    for (double c : *cov)
      if (std::isnan(c)) fail("the covariate holds NaN");
  Squeezed S;
  if (n < 3) {
    S.var_post = var;
    S.var_prior = var;
    S.df_prior.assign(1, 0.0);
    return S;
  }
  if (robust) {
    RobustFit R = robust_fit(var, df, cov, 0.05, 0.1);
    S.var_prior = R.scale.size() == 1 ? std::vector<double>(n, R.scale[0]) : R.scale;
    S.df_prior = R.df2_shrunk;
  } else if (cov) {
    double df2;
    fit_f_trend(var, df, *cov, S.var_prior, df2);
    if (S.var_prior.size() == 1) S.var_prior.assign(n, S.var_prior[0]);
    S.df_prior.assign(1, df2);
  } else {
    double s, df2;
    fit_f(var, df, s, df2);
    S.var_prior.assign(1, s);
    S.df_prior.assign(1, df2);
  }
  auto at = [](const std::vector<double>& v, size_t i) { return v.size() == 1 ? v[0] : v[i]; };
  S.var_post.resize(n);
  double largest = -kInf, smallest = kInf;
  for (double d : S.df_prior) {
    largest = std::max(largest, d);
    smallest = std::min(smallest, d);
  }
  for (size_t i = 0; i < n; ++i) {
    const double dp = at(S.df_prior, i), vp = at(S.var_prior, i);
    if (std::isfinite(largest) || (smallest <= 1e100 && std::isfinite(dp)))
      S.var_post[i] = (df * var[i] + dp * vp) / (df + dp);
    else
      S.var_post[i] = vp;
  }
  return S;
}

double median_finite(const std::vector<double>& v) {
  std::vector<double> f;
  for (double a : v)
    if (std::isfinite(a)) f.push_back(a);
  return median_of(f);
}

}  // namespace
}  // namespace eb
}  // namespace nh

extern "C" double nh_sc_residual_df(size_t n_cells, size_t n_coef, size_t n_subjects) {
  const size_t used = n_coef + (n_subjects > 0 ? n_subjects - 1 : 0);
  return (double)std::max<size_t>(n_cells > used ? n_cells - used : 0, 1);
}

extern "C" int nh_shrink_sc_dispersion(const double* dispersion, const int32_t* convergence,
                                       const double* covariate, size_t n_genes,
                                       double df_residual, int32_t robust, double* phi_raw,
                                       double* phi_post, double* phi_prior,
                                       double* dispersion_shrunk, double* df_prior,
                                       size_t* n_df_prior, char* err, size_t err_len) {
  using namespace nh::eb;
  auto put = [&](const std::string& m) {
    if (err && err_len) {
      std::strncpy(err, m.c_str(), err_len - 1);
      err[err_len - 1] = 0;
    }
  };
  if (n_genes == 0) {
    put("shrink_sc_dispersion was given an empty dispersion vector");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (!dispersion || !convergence || !phi_raw || !phi_post || !phi_prior || !dispersion_shrunk ||
      !df_prior || !n_df_prior) {
    put("null argument");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (!(std::isfinite(df_residual) && df_residual > 0.0)) {
    put("shrink_sc_dispersion needs a positive finite df_residual");
    return NH_ERR_INVALID_ARGUMENT;
  }
  try {
    const size_t n = n_genes;
    std::vector<double> raw(n);
    for (size_t g = 0; g < n; ++g) raw[g] = dispersion[g] > 0.0 ? 1.0 / dispersion[g] : kInf;
    std::vector<size_t> usable;
    for (size_t g = 0; g < n; ++g)
      if (convergence[g] == 1 && std::isfinite(raw[g])) usable.push_back(g);
    auto unshrunk = [&]() {
      const double prior = median_finite(raw);
      for (size_t g = 0; g < n; ++g) {
        phi_raw[g] = raw[g];
        phi_post[g] = raw[g];
        phi_prior[g] = prior;
        dispersion_shrunk[g] = dispersion[g];
      }
      df_prior[0] = 0.0;
      *n_df_prior = 1;
      return NH_OK;
    };
    if (usable.size() < 3) return unshrunk();
    std::vector<double> phi_u, cov_u;
    for (size_t g : usable) {
      phi_u.push_back(std::max(raw[g], 1e-8));
      if (covariate) cov_u.push_back(covariate[g]);
    }
    // the ladder: trended robust, robust, plain; each rung only on failure
    struct Rung {
      bool trended, robust;
    };
    std::vector<Rung> rungs;
    if (covariate) rungs.push_back({true, robust != 0});
    rungs.push_back({false, robust != 0});
    if (robust) rungs.push_back({false, false});
    bool ok = false;
    Squeezed S;
    for (const Rung& r : rungs) {
      try {
        S = squeeze(phi_u, df_residual, r.trended ? &cov_u : nullptr, r.robust);
        bool finite = true;
        for (double v : S.var_post) finite = finite && !std::isnan(v);
        if (!finite) continue;
        ok = true;
        break;
      } catch (const Failure&) {
      }
    }
    if (!ok) return unshrunk();
    const double med_prior = median_finite(S.var_prior);
    for (size_t g = 0; g < n; ++g) {
      phi_raw[g] = raw[g];
      phi_post[g] = med_prior;
      phi_prior[g] = med_prior;
    }
    for (size_t k = 0; k < usable.size(); ++k) {
      phi_post[usable[k]] = S.var_post[k];
      phi_prior[usable[k]] = S.var_prior.size() == 1 ? S.var_prior[0] : S.var_prior[k];
    }
    for (size_t g = 0; g < n; ++g)
      dispersion_shrunk[g] = phi_post[g] > 0.0 ? 1.0 / phi_post[g] : kInf;
    for (size_t k = 0; k < S.df_prior.size(); ++k) df_prior[k] = S.df_prior[k];
    *n_df_prior = S.df_prior.size();
    return NH_OK;
  } catch (const Failure& f) {
    put(f.what());
    return f.code;
  } catch (const std::exception& e) {
    put(e.what());
    return NH_ERR_INVALID_ARGUMENT;
  }
}

extern "C" int nh_squeeze_var(const double* var, size_t n, double df, const double* covariate,
                              int32_t robust, double* var_post, double* var_prior,
                              size_t* n_var_prior, double* df_prior, size_t* n_df_prior,
                              char* err, size_t err_len) {
  using namespace nh::eb;
  auto put = [&](const std::string& m) {
    if (err && err_len) {
      std::strncpy(err, m.c_str(), err_len - 1);
      err[err_len - 1] = 0;
    }
  };
  if (n == 0) {
    put("squeeze_var was given an empty variance vector");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (!var || !var_post || !var_prior || !n_var_prior || !df_prior || !n_df_prior) {
    put("null argument");
    return NH_ERR_INVALID_ARGUMENT;
  }
  if (!(std::isfinite(df) && df > 0.0)) {
    put("squeeze_var needs a positive finite df");
    return NH_ERR_INVALID_ARGUMENT;
  }
  try {
    std::vector<double> v(var, var + n), c;
    if (covariate) c.assign(covariate, covariate + n);
    const Squeezed S = squeeze(v, df, covariate ? &c : nullptr, robust != 0);
    for (size_t i = 0; i < n; ++i) var_post[i] = S.var_post[i];
    for (size_t i = 0; i < S.var_prior.size(); ++i) var_prior[i] = S.var_prior[i];
    *n_var_prior = S.var_prior.size();
    for (size_t i = 0; i < S.df_prior.size(); ++i) df_prior[i] = S.df_prior[i];
    *n_df_prior = S.df_prior.size();
    return NH_OK;
  } catch (const Failure& f) {
    put(f.what());
    return f.code;
  } catch (const std::exception& e) {
    put(e.what());
    return NH_ERR_INVALID_ARGUMENT;
  }
}

// Test hooks: the file's F / chi-square functions, for the checks against
// scipy (which 0 F density, 1 F upper tail, 2 F quantile, 3 chi-square
// upper tail, 4 lowess at span a with b robustness passes (x, y given),
// 5 log x - digamma x, 6 trigamma, 7 inverse trigamma).
extern "C" int nh_shrink_selftest(int32_t which, const double* x, const double* y, size_t n,
                                  double a, double b, double* out) {
  using namespace nh::eb;
  try {
    switch (which) {
      case 0: for (size_t i = 0; i < n; ++i) out[i] = f_density(x[i], a, b); break;
      case 1: for (size_t i = 0; i < n; ++i) out[i] = f_sf(x[i], a, b); break;
      case 2: for (size_t i = 0; i < n; ++i) out[i] = f_quantile(x[i], a, b); break;
      case 3: for (size_t i = 0; i < n; ++i) out[i] = chisq_sf(x[i], a); break;
      case 4: {
        std::vector<double> xv(x, x + n), yv(y, y + n);
        std::vector<double> f = lowess(xv, yv, a, (int)b);
        for (size_t i = 0; i < n; ++i) out[i] = f[i];
        break;
      }
      case 5: for (size_t i = 0; i < n; ++i) out[i] = log_minus_digamma(x[i]); break;
      case 6: for (size_t i = 0; i < n; ++i) out[i] = trigamma(x[i]); break;
      case 7: for (size_t i = 0; i < n; ++i) out[i] = trigamma_inverse(x[i]); break;
      default: return NH_ERR_INVALID_ARGUMENT;
    }
  } catch (const Failure& f) {
    return f.code;
  }
  return NH_OK;
}

// Zero-count tables, f64: a restatement of src/sc/zeros.rs.
//
// build_zero_cells  <- zeros.rs:219-278 (groups of >= 32 cells of one subject
//                      sharing a design row, bitwise; built only when they
//                      hold half the cells)
// build_group       <- zeros.rs:291-341 (distinct log offsets with their
//                      multiplicities, base panels of width 4 halved up to 6
//                      times until every form's last two coefficients are
//                      small)
// fit_panel         <- zeros.rs:355-383, direct_sums <- :391-409 (Neumaier)
// ZeroGroup::eval   <- zeros.rs:141-193 (closed forms past +-40 in ln v)
// chebyshev / clenshaw <- zeros.rs:457-497
// Pinned by tests/test_oracle_zeros.py: the table against its own direct sums
// over the whole range (1e-13), the direct sums against numpy and against
// finite differences in s (each sum is the derivative of the one before),
// the grouping case of zeros.rs:543-557, continuity at the table ends and the
// tabled penalised fit and marginal likelihood against their dense forms.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "oracle.hpp"

namespace oracle {

namespace {

constexpr int kSums = 6;                    // L, A, B, C, D, E
constexpr int kDegree = 16;                 // zeros.rs:43
constexpr int kPoints = kDegree + 1;
constexpr double kBaseWidth = 4.0;          // zeros.rs:53
constexpr double kNoiseUlps = 32.0;         // zeros.rs:57
constexpr int kMaxDepth = 6;                // zeros.rs:60
constexpr double kPanelTol = 1e-14;         // zeros.rs:66
constexpr double kAsymptoteDepth = 40.0;    // zeros.rs:70
constexpr std::size_t kMinGroupCells = 32;  // zeros.rs:74
constexpr double kMinTabledShare = 0.5;     // zeros.rs:78
constexpr double kPi = 3.14159265358979323846;

struct Neumaier {
  double sum = 0.0, carry = 0.0;
  void add(double x) {
    const double t = sum + x;
    if (std::fabs(sum) >= std::fabs(x))
      carry += (sum - t) + x;
    else
      carry += (x - t) + sum;
    sum = t;
  }
  double total() const { return sum + carry; }
};

void chebyshev_coefficients(const double* values, double* c) {
  const double n = (double)kDegree;
  for (int k = 0; k < kPoints; ++k) {
    double acc = 0.0;
    for (int j = 0; j < kPoints; ++j) {
      const double weight = (j == 0 || j == kDegree) ? 0.5 : 1.0;
      acc += weight * values[j] * std::cos(kPi * (double)(j * k) / n);
    }
    c[k] = 2.0 * acc / n;
  }
  c[0] *= 0.5;
  c[kDegree] *= 0.5;
}

double clenshaw(const double* c, double x) {
  const double two_x = 2.0 * x;
  double b1 = 0.0, b2 = 0.0;
  for (int k = kPoints - 1; k >= 1; --k) {
    const double b0 = c[k] + two_x * b1 - b2;
    b2 = b1;
    b1 = b0;
  }
  return c[0] + x * b1 - b2;
}

using Distinct = std::vector<std::pair<double, double>>;  // (ln O, multiplicity)

// One panel [left, left + width): coefficients of the six forms and whether
// every form's tail is under the tolerance.
bool fit_panel(const Distinct& distinct, double left, double width, double* panel) {
  double values[kSums][kPoints];
  for (int j = 0; j < kPoints; ++j) {
    const double x = std::cos(kPi * (double)j / (double)kDegree);
    double sums[kSums];
    zero_direct_sums(distinct, left + 0.5 * width * (x + 1.0), sums);
    values[0][j] = std::log(sums[0]);
    values[1][j] = std::log(sums[1]);
    values[2][j] = std::log(sums[2]);
    for (int f = 3; f < kSums; ++f) values[f][j] = sums[f] / sums[2];
  }
  bool ok = true;
  for (int f = 0; f < kSums; ++f) {
    double* c = panel + f * kPoints;
    chebyshev_coefficients(values[f], c);
    double size = 0.0;
    for (int j = 0; j < kPoints; ++j) size = std::max(size, std::fabs(values[f][j]));
    const double tail = std::fabs(c[kDegree - 1]) + std::fabs(c[kDegree]);
    ok = ok && tail <= kPanelTol + kNoiseUlps * 2.220446049250313e-16 * size;
  }
  return ok;
}

double compensated_sum(const std::vector<double>& v) {
  Neumaier acc;
  for (double x : v) acc.add(x);
  return acc.total();
}

}  // namespace

void zero_direct_sums(const std::vector<std::pair<double, double>>& distinct, double s,
                      double* out) {
  Neumaier acc[kSums];
  for (const auto& lm : distinct) {
    const double l = lm.first, m = lm.second;
    const double v = std::exp(s + l);
    const double w = 1.0 / (1.0 + v);
    const double u = v * w;
    const double uw = u * w;
    acc[0].add(m * std::log1p(v));
    acc[1].add(m * u);
    acc[2].add(m * uw);
    acc[3].add(m * uw * (w - u));
    acc[4].add(m * uw * (w * w - 4.0 * u * w + u * u));
    acc[5].add(m * uw * (w * w * w - 11.0 * w * w * u + 11.0 * w * u * u - u * u * u));
  }
  for (int f = 0; f < kSums; ++f) out[f] = acc[f].total();
}

ZeroGroup build_zero_group(std::size_t row, const std::vector<double>& log_o) {
  std::vector<double> sorted = log_o;
  // total order: NaN is not expected in a log offset
  std::sort(sorted.begin(), sorted.end());
  Distinct distinct;
  for (double l : sorted) {
    if (!distinct.empty() && distinct.back().first == l)
      distinct.back().second += 1.0;
    else
      distinct.push_back({l, 1.0});
  }
  ZeroGroup g;
  g.row = row;
  g.n = (double)log_o.size();
  const double lo_o = sorted.front(), hi_o = sorted.back();
  g.s_lo = -kAsymptoteDepth - hi_o;
  g.s_hi = kAsymptoteDepth - lo_o;
  g.lo = g.s_lo;
  g.n_base = (std::size_t)std::ceil((g.s_hi - g.s_lo) / kBaseWidth);
  const std::size_t panel_doubles = (std::size_t)kSums * kPoints;
  for (std::size_t b = 0; b < g.n_base; ++b) {
    const double left = g.s_lo + (double)b * kBaseWidth;
    int depth = 0;
    std::vector<double> fitted;
    for (;;) {
      const int parts = 1 << depth;
      const double width = kBaseWidth / (double)parts;
      fitted.assign((std::size_t)parts * panel_doubles, 0.0);
      bool all = true;
      for (int p = 0; p < parts; ++p)
        all = fit_panel(distinct, left + (double)p * width, width,
                        &fitted[(std::size_t)p * panel_doubles]) && all;
      if (depth == kMaxDepth || all) break;
      ++depth;
    }
    g.index.push_back({(std::uint32_t)(g.panels.size() / panel_doubles), (std::uint8_t)depth});
    g.panels.insert(g.panels.end(), fitted.begin(), fitted.end());
  }
  std::vector<double> o(log_o.size()), inv_o(log_o.size());
  for (std::size_t i = 0; i < log_o.size(); ++i) {
    o[i] = std::exp(log_o[i]);
    inv_o[i] = std::exp(-log_o[i]);
  }
  g.sum_o = compensated_sum(o);
  g.sum_inv_o = compensated_sum(inv_o);
  g.sum_log_o = compensated_sum(log_o);
  return g;
}

void ZeroGroup::eval(int from, int m, double s, double* out) const {
  if (std::isnan(s)) {
    // every sum is NaN (the reference's saturating casts reach the same end)
    for (int j = 0; j < m; ++j) out[j] = s;
    return;
  }
  if (s < s_lo) {
    const double v = std::exp(s) * sum_o;
    for (int j = 0; j < m; ++j) out[j] = v;
    return;
  }
  if (s > s_hi) {
    const double r = std::exp(-s) * sum_inv_o;
    for (int j = 0; j < m; ++j) {
      switch (from + j) {
        case 0: out[j] = n * s + sum_log_o + r; break;
        case 1: out[j] = n - r; break;
        case 2:
        case 4: out[j] = r; break;
        default: out[j] = -r; break;
      }
    }
    return;
  }
  const double pos = (s - lo) / kBaseWidth;
  const std::size_t base = std::min((std::size_t)pos, n_base - 1);
  const std::uint32_t first = index[base].first;
  const int depth = index[base].second;
  const double parts = (double)(1 << depth);
  const double within = std::min(std::max((pos - (double)base) * parts, 0.0), parts - 1e-9);
  const std::size_t sub = (std::size_t)within;
  const double x = 2.0 * (within - (double)sub) - 1.0;
  const double* panel = &panels[((std::size_t)first + sub) * kSums * kPoints];
  double b_value = 0.0;
  for (int j = 0; j < m; ++j) {
    const int f = from + j;
    const double v = clenshaw(panel + f * kPoints, x);
    out[j] = f < 3 ? std::exp(v) : v;
    if (f == 2) b_value = out[j];
  }
  for (int j = 0; j < m; ++j)
    if (from + j >= 3) out[j] *= b_value;
}

bool build_zero_cells(const double* design, const double* log_offset, std::size_t n_cells,
                      const std::size_t* fid, std::size_t k, std::size_t n_coef, ZeroCells& z) {
  std::vector<std::vector<std::size_t>> members;
  z.subject_groups.assign(k + 1, 0);
  z.subject_loose.assign(k + 1, 0);
  z.loose.clear();
  z.groups.clear();
  std::size_t tabled = 0;
  for (std::size_t s = 0; s < k; ++s) {
    // rows keyed by their bit patterns; the map's order plays no part, as the
    // kept groups are sorted by first cell below
    std::map<std::vector<std::uint64_t>, std::vector<std::size_t>> by_row;
    for (std::size_t r = fid[s]; r < fid[s + 1]; ++r) {
      std::vector<std::uint64_t> key(n_coef);
      std::memcpy(key.data(), design + r * n_coef, n_coef * sizeof(double));
      by_row[key].push_back(r);
    }
    std::vector<std::size_t> loose;
    std::vector<std::vector<std::size_t>> kept;
    for (auto& kv : by_row) {
      if (kv.second.size() >= kMinGroupCells) {
        tabled += kv.second.size();
        kept.push_back(kv.second);
      } else {
        loose.insert(loose.end(), kv.second.begin(), kv.second.end());
      }
    }
    std::sort(kept.begin(), kept.end(),
              [](const std::vector<std::size_t>& a, const std::vector<std::size_t>& b) {
                return a[0] < b[0];
              });
    for (auto& c : kept) members.push_back(c);
    z.subject_groups[s + 1] = members.size();
    std::sort(loose.begin(), loose.end());
    z.loose.insert(z.loose.end(), loose.begin(), loose.end());
    z.subject_loose[s + 1] = z.loose.size();
  }
  if ((double)tabled < kMinTabledShare * (double)n_cells) return false;
  z.groups.reserve(members.size());
  for (const auto& cells : members) {
    std::vector<double> logs(cells.size());
    for (std::size_t i = 0; i < cells.size(); ++i) logs[i] = log_offset[cells[i]];
    z.groups.push_back(build_zero_group(cells[0], logs));
  }
  return true;
}

}  // namespace oracle

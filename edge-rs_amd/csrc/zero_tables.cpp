// Host build of the zero-count tables (src/sc/zeros.rs:219-409): grouping,
// then one table per group, groups on parallel host threads.
//
// Grouping sorts each subject's cells by their design row's bit pattern, so
// equal rows form runs; runs of >= 32 cells become groups (ordered by their
// first cell), the rest loose cells.  A group's table covers
// [-40 - max ln O, 40 - min ln O]; each base panel of width 4 is halved until
// the last two Chebyshev coefficients of all six forms are under
// 1e-14 + 32 ulp x the form's size, at most six times.  Node values are the
// six sums over the group's distinct offsets (with multiplicities), summed
// with Neumaier's compensation; coefficients come from a fixed 17 x 17 cosine
// matrix (the DCT-I of the Lobatto-node values).
#include "zero_tables.hpp"

#include <algorithm>
#include <array>
#include <cstring>
#include <numeric>

#include "../../include/nebula_hip.h"
#include "host_common.hpp"

namespace nh {

namespace {

constexpr int kMaxDepth = 6;
constexpr double kPanelTol = 1e-14;
constexpr double kNoiseUlps = 32.0;
constexpr double kEps = 2.220446049250313e-16;
constexpr double kPi = 3.14159265358979323846;

// Compensated running sum.
struct KahanBabuska {
  double s = 0.0, c = 0.0;
  void add(double x) {
    const double t = s + x;
    c += std::fabs(s) >= std::fabs(x) ? (s - t) + x : (x - t) + s;
    s = t;
  }
  double value() const { return s + c; }
};

struct Cheb {
  double node[kZPoints];              // cos(pi j / 16)
  double dct[kZPoints][kZPoints];     // coefficient k from values j
  Cheb() {
    for (int j = 0; j < kZPoints; ++j) node[j] = std::cos(kPi * j / kZDegree);
    for (int k = 0; k < kZPoints; ++k)
      for (int j = 0; j < kZPoints; ++j) {
        const double w = (j == 0 || j == kZDegree) ? 0.5 : 1.0;
        double v = 2.0 / kZDegree * w * std::cos(kPi * (double)(j * k) / kZDegree);
        if (k == 0 || k == kZDegree) v *= 0.5;
        dct[k][j] = v;
      }
  }
};
const Cheb& cheb() {
  static const Cheb c;
  return c;
}

struct Offsets {
  std::vector<double> value, mult;  // distinct ln O, ascending, and multiplicities
};

// L, A .. E at s over the distinct offsets.
void node_sums(const Offsets& o, double s, double* out) {
  KahanBabuska acc[kZSums];
  for (size_t i = 0; i < o.value.size(); ++i) {
    const double m = o.mult[i];
    const double v = std::exp(s + o.value[i]);
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
  for (int f = 0; f < kZSums; ++f) out[f] = acc[f].value();
}

// Coefficients of one panel into `panel` (6 x 17); whether every form's
// Chebyshev tail is under the tolerance.
bool fit_panel(const Offsets& o, double left, double width, double* panel) {
  const Cheb& ch = cheb();
  double val[kZSums][kZPoints];
  for (int j = 0; j < kZPoints; ++j) {
    double sums[kZSums];
    node_sums(o, left + 0.5 * width * (ch.node[j] + 1.0), sums);
    for (int f = 0; f < 3; ++f) val[f][j] = std::log(sums[f]);
    for (int f = 3; f < kZSums; ++f) val[f][j] = sums[f] / sums[2];
  }
  bool ok = true;
  for (int f = 0; f < kZSums; ++f) {
    double* c = panel + f * kZPoints;
    double size = 0.0;
    for (int j = 0; j < kZPoints; ++j) size = std::max(size, std::fabs(val[f][j]));
    for (int k = 0; k < kZPoints; ++k) {
      double acc = 0.0;
      for (int j = 0; j < kZPoints; ++j) acc += ch.dct[k][j] * val[f][j];
      c[k] = acc;
    }
    ok = ok && std::fabs(c[kZDegree - 1]) + std::fabs(c[kZDegree]) <=
                   kPanelTol + kNoiseUlps * kEps * size;
  }
  return ok;
}

// One group's record (index_off left for the caller), base-panel words with
// panel numbers local to the group, and its panels.
struct Built {
  ZGroup g{};
  std::vector<uint32_t> words;
  std::vector<double> coef;
};

Built build_group(uint32_t row, const std::vector<double>& log_o) {
  Built b;
  std::vector<double> sorted(log_o);
  std::sort(sorted.begin(), sorted.end());
  Offsets o;
  for (double l : sorted) {
    if (!o.value.empty() && o.value.back() == l) {
      o.mult.back() += 1.0;
    } else {
      o.value.push_back(l);
      o.mult.push_back(1.0);
    }
  }
  ZGroup& g = b.g;
  g.row = row;
  g.n = (double)log_o.size();
  g.s_lo = -kZAsymptote - sorted.back();
  g.s_hi = kZAsymptote - sorted.front();
  g.n_base = (uint32_t)std::ceil((g.s_hi - g.s_lo) / kZBaseWidth);
  std::vector<double> trial;
  for (uint32_t base = 0; base < g.n_base; ++base) {
    const double left = g.s_lo + (double)base * kZBaseWidth;
    int depth = 0;
    for (;;) {
      const int parts = 1 << depth;
      const double width = kZBaseWidth / parts;
      trial.assign((size_t)parts * kZPanel, 0.0);
      bool all = true;
      for (int p = 0; p < parts; ++p)
        all = fit_panel(o, left + p * width, width, &trial[(size_t)p * kZPanel]) && all;
      if (all || depth == kMaxDepth) break;
      ++depth;
    }
    b.words.push_back((uint32_t)(b.coef.size() / kZPanel) << 3 | (uint32_t)depth);
    b.coef.insert(b.coef.end(), trial.begin(), trial.end());
  }
  KahanBabuska so, si, sl;
  for (double l : log_o) {
    so.add(std::exp(l));
    si.add(std::exp(-l));
    sl.add(l);
  }
  g.sum_o = so.value();
  g.sum_inv_o = si.value();
  g.sum_log_o = sl.value();
  return b;
}

}  // namespace

bool build_zero_tables(const double* design, const double* log_offset, size_t n_cells, size_t nb,
                       const uint64_t* fid, size_t k, int threads, ZeroTables& out) {
  // groups, subject by subject: cells sorted by row bits (then index), runs
  std::vector<std::vector<uint32_t>> members;
  std::vector<uint32_t> subject_groups(k + 1, 0), subject_loose(k + 1, 0), loose;
  size_t tabled = 0;
  std::vector<uint32_t> order;
  auto row_less = [&](uint32_t a, uint32_t b) {
    const int c = std::memcmp(design + (size_t)a * nb, design + (size_t)b * nb, nb * sizeof(double));
    return c != 0 ? c < 0 : a < b;
  };
  auto row_equal = [&](uint32_t a, uint32_t b) {
    return std::memcmp(design + (size_t)a * nb, design + (size_t)b * nb, nb * sizeof(double)) == 0;
  };
  for (size_t s = 0; s < k; ++s) {
    order.resize(fid[s + 1] - fid[s]);
    std::iota(order.begin(), order.end(), (uint32_t)fid[s]);
    std::sort(order.begin(), order.end(), row_less);
    std::vector<std::vector<uint32_t>> kept;
    std::vector<uint32_t> sub_loose;
    for (size_t i = 0; i < order.size();) {
      size_t j = i + 1;
      while (j < order.size() && row_equal(order[i], order[j])) ++j;
      if (j - i >= kZMinGroupCells) {
        kept.emplace_back(order.begin() + (std::ptrdiff_t)i, order.begin() + (std::ptrdiff_t)j);
        tabled += j - i;
      } else {
        sub_loose.insert(sub_loose.end(), order.begin() + (std::ptrdiff_t)i,
                         order.begin() + (std::ptrdiff_t)j);
      }
      i = j;
    }
    std::sort(kept.begin(), kept.end(),
              [](const std::vector<uint32_t>& a, const std::vector<uint32_t>& b) { return a[0] < b[0]; });
    for (auto& m : kept) members.push_back(std::move(m));
    subject_groups[s + 1] = (uint32_t)members.size();
    std::sort(sub_loose.begin(), sub_loose.end());
    loose.insert(loose.end(), sub_loose.begin(), sub_loose.end());
    subject_loose[s + 1] = (uint32_t)loose.size();
  }
  if ((double)tabled < kZMinShare * (double)n_cells) return false;

  std::vector<Built> built(members.size());
  parallel_for(members.size(), threads, [&](size_t gi) {
    std::vector<double> lo(members[gi].size());
    for (size_t i = 0; i < lo.size(); ++i) lo[i] = log_offset[members[gi][i]];
    built[gi] = build_group(members[gi][0], lo);
  });
  ZeroTables z;
  z.subject_groups = std::move(subject_groups);
  z.subject_loose = std::move(subject_loose);
  z.loose = std::move(loose);
  z.tabled_cells = tabled;
  for (Built& b : built) {
    b.g.index_off = (uint32_t)z.index.size();
    const uint32_t panel0 = (uint32_t)(z.coef.size() / kZPanel);
    for (uint32_t w : b.words) z.index.push_back(((w >> 3) + panel0) << 3 | (w & 7u));
    z.coef.insert(z.coef.end(), b.coef.begin(), b.coef.end());
    z.groups.push_back(b.g);
  }
  out = std::move(z);
  return true;
}

}  // namespace nh

extern "C" int nh_zero_tables_probe(const double* design, const double* log_offset,
                                    size_t n_cells, size_t n_coef, const uint64_t* subject_start,
                                    size_t n_subjects, const double* s, size_t n_s, int32_t from,
                                    int32_t m, double* out, size_t out_len, int32_t* applies,
                                    uint64_t* n_groups, uint64_t* group_row,
                                    uint64_t* subject_groups, char* err, size_t err_len) {
  auto fail = [&](int code, const char* msg) {
    if (err && err_len) {
      std::strncpy(err, msg, err_len - 1);
      err[err_len - 1] = 0;
    }
    return code;
  };
  if (!design || !log_offset || !subject_start || !applies || !n_groups || !group_row ||
      !subject_groups || (n_s && (!s || !out)) || n_coef == 0 || n_subjects == 0)
    return fail(NH_ERR_INVALID_ARGUMENT, "null or empty argument");
  if (subject_start[0] != 0 || subject_start[n_subjects] != n_cells)
    return fail(NH_ERR_INVALID_ARGUMENT, "subject starts must run from 0 to n_cells");
  for (size_t q = 0; q < n_subjects; ++q)
    if (subject_start[q] > subject_start[q + 1])
      return fail(NH_ERR_INVALID_ARGUMENT, "subject starts must not decrease");
  const bool shape_ok = (from == 0 && (m == 2 || m == 3 || m == 4 || m == 6)) ||
                        (from == 2 && m == 4);
  if (n_s && !shape_ok) return fail(NH_ERR_INVALID_ARGUMENT, "unsupported (from, m)");
  nh::ZeroTables z;
  *applies = nh::build_zero_tables(design, log_offset, n_cells, n_coef, subject_start,
                                   n_subjects, 0, z) ? 1 : 0;
  *n_groups = z.groups.size();
  if (!*applies) return NH_OK;
  if (n_s && z.groups.size() * n_s * (size_t)m > out_len)
    return fail(NH_ERR_INVALID_ARGUMENT, "out_len too small");
  for (size_t q = 0; q <= n_subjects; ++q) subject_groups[q] = z.subject_groups[q];
  for (size_t g = 0; g < z.groups.size(); ++g) {
    group_row[g] = z.groups[g].row;
    for (size_t i = 0; i < n_s; ++i) {
      double* o = out + (g * n_s + i) * (size_t)m;
      const nh::ZGroup& G = z.groups[g];
      if (from == 2) nh::zt_eval<2, 4>(G, z.index.data(), z.coef.data(), s[i], o);
      else if (m == 2) nh::zt_eval<0, 2>(G, z.index.data(), z.coef.data(), s[i], o);
      else if (m == 3) nh::zt_eval<0, 3>(G, z.index.data(), z.coef.data(), s[i], o);
      else if (m == 4) nh::zt_eval<0, 4>(G, z.index.data(), z.coef.data(), s[i], o);
      else nh::zt_eval<0, 6>(G, z.index.data(), z.coef.data(), s[i], o);
    }
  }
  return NH_OK;
}

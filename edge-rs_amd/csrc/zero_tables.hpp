// Zero-count tables for designs whose cells fall into large groups sharing a
// design row within a subject (src/sc/zeros.rs).
//
// At a count of zero every term a sweep sums over a cell is a function of
// v = tau O, with O = exp(offset) the cell's and tau = exp(x'beta + log w) /
// gamma shared by the group, so the sum over a group is a function of
// s = ln tau alone: L = sum ln(1 + v) and its s-derivatives A..E.  Each is a
// piecewise Chebyshev series of degree 16 on panels of width 4 >> depth
// (L, A, B on the log scale, C, D, E as ratios to B), with closed forms past
// +-40 in ln v.  A sweep then visits, per subject, one lookup per group, the
// few loose cells one by one, and the positive counts.
//
// Device layout (one run, every gene): groups[] subject by subject, each with
// its base-panel index words (first panel << 3 | depth) at index_off and its
// panels' coefficients, 6 forms x 17 each, in coef[].
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

namespace nh {

constexpr int kZSums = 6;
constexpr int kZDegree = 16;
constexpr int kZPoints = kZDegree + 1;
constexpr int kZPanel = kZSums * kZPoints;   // doubles per panel
constexpr double kZBaseWidth = 4.0;
constexpr double kZAsymptote = 40.0;
constexpr uint32_t kZMinGroupCells = 32;
constexpr double kZMinShare = 0.5;

struct ZGroup {
  uint32_t row;        // first member cell: its design row is the group's
  uint32_t index_off;  // first base-panel index word
  uint32_t n_base;     // base panels
  uint32_t pad;
  double n, sum_o, sum_inv_o, sum_log_o, s_lo, s_hi;
};
static_assert(sizeof(ZGroup) == 64, "ZGroup is one 64-byte record");

// The sums FROM .. FROM + M - 1 (0 L, 1 A, 2 B, 3 C, 4 D, 5 E) of one group at
// s.  C, D and E are stored as ratios to B, so asking for them needs B in the
// same call (FROM <= 2).
template <int FROM, int M>
__host__ __device__ inline void zt_eval(const ZGroup& g, const uint32_t* index, const double* coef,
                                        double s, double* out) {
  static_assert(FROM + M <= kZSums && (FROM + M <= 3 || FROM <= 2), "C..E need B");
  if (!(s == s)) {
#pragma unroll
    for (int j = 0; j < M; ++j) out[j] = s;
    return;
  }
  if (s < g.s_lo) {
    const double v = exp(s) * g.sum_o;
#pragma unroll
    for (int j = 0; j < M; ++j) out[j] = v;
    return;
  }
  if (s > g.s_hi) {
    const double r = exp(-s) * g.sum_inv_o;
#pragma unroll
    for (int j = 0; j < M; ++j) {
      const int f = FROM + j;
      out[j] = f == 0 ? g.n * s + g.sum_log_o + r : (f == 1 ? g.n - r : (f == 2 || f == 4 ? r : -r));
    }
    return;
  }
  const double pos = (s - g.s_lo) / kZBaseWidth;
  uint32_t base = (uint32_t)pos;  // pos in [0, n_base] here
  if (base > g.n_base - 1) base = g.n_base - 1;
  const uint32_t word = index[g.index_off + base];
  const uint32_t depth = word & 7u;
  const double parts = (double)(1u << depth);
  double within = (pos - (double)base) * parts;
  within = within < 0.0 ? 0.0 : (within > parts - 1e-9 ? parts - 1e-9 : within);
  const uint32_t sub = (uint32_t)within;
  const double x = 2.0 * (within - (double)sub) - 1.0;
  const double* panel = coef + (size_t)((word >> 3) + sub) * kZPanel;
  const double two_x = 2.0 * x;
  double b_value = 0.0;
#pragma unroll
  for (int j = 0; j < M; ++j) {
    const double* c = panel + (FROM + j) * kZPoints;
    double b1 = 0.0, b2 = 0.0;
#pragma unroll
    for (int q = kZDegree; q >= 1; --q) {
      const double b0 = c[q] + two_x * b1 - b2;
      b2 = b1;
      b1 = b0;
    }
    const double v = c[0] + x * b1 - b2;
    out[j] = FROM + j < 3 ? exp(v) : v;
    if (FROM + j == 2) b_value = out[j];
  }
#pragma unroll
  for (int j = 0; j < M; ++j)
    if (FROM + j >= 3) out[j] *= b_value;
}

// A run's tables on the host, ready to upload.
struct ZeroTables {
  std::vector<ZGroup> groups;
  std::vector<uint32_t> subject_groups;  // k + 1
  std::vector<uint32_t> loose;           // cells in no group, ascending per subject
  std::vector<uint32_t> subject_loose;   // k + 1
  std::vector<uint32_t> index;
  std::vector<double> coef;
  size_t tabled_cells = 0;
};

// Groups the cells (rows compared bitwise) and tabulates every group of at
// least 32 cells; false, with `out` untouched, when such groups hold under
// half the cells (a sweep over every cell is then the better plan).
// `design` is the n_cells x nb centred design, `fid` the k + 1 subject starts.
bool build_zero_tables(const double* design, const double* log_offset, size_t n_cells, size_t nb,
                       const uint64_t* fid, size_t k, int threads, ZeroTables& out);

}  // namespace nh

// One penalised NB fit (edge-rs opt_pml, src/sc/pml.rs:1614-1800) per HIP
// workgroup, for gfx950.
//
// Data layout.  Every cell has one packed row in HBM: the design columns other
// than the intercept, then the log offset, zero-padded to R = 1, 2, 4 or 8
// values (16 bytes per cell in f32 at p = 4, one dwordx4 load).  The intercept
// column is 1 for every cell (nebula's centred design), so its sums are copies
// of sums the fit keeps anyway: score[int] = sum(resid), first[int] = weight,
// second[int][j] = first[j], second[int][int] = weight.  A fit therefore keeps
// 2 + Q per-subject and 1 + Q + Q(Q+1)/2 per-gene accumulators (Q = p - 1)
// instead of 3 + 2p + p(p+1)/2.
//
// Work split.  The workgroup's W waves cut the gene's cells into W contiguous
// blocks (64-cell aligned).  A wave walks its block subject by subject in tiles
// of 64 x U consecutive cells (U = 4 in f32, 2 in f64): each lane loads U
// packed rows, all independent, so a tile keeps U loads per lane in flight.
// The gene's positive counts are merged in with no dependent load: the wave
// holds a window of the next 64 x U entries of the gene's sorted CSR row in
// registers (a tile of 64 x U cells holds at most that many), scatters the
// entries that fall in the tile into a per-wave LDS row, advances its cursor
// by the ballot count and loads the next window before it computes the tile.
// Every cell is visited once with its own count; that one pass is edge-rs's
// dense sweep at y = 0 plus its sparse correction (pml.rs:1360-1363, exact
// because every per-cell term is linear in y).
//
// Precision.  Per cell the exp, log and reciprocal run in CT: float for the
// bulk Newton sweeps, double for the finishing sweeps.  Every accumulation is
// FP64 (on CDNA4 a DFMA issues at the rate of a non-packed FP32 FMA), so the
// long sums over 1e5..1e6 cells carry no summation error and the Schur
// complement vb - sum vwb vwb'/vw is formed from exact sums.  What FP32 leaves
// is the per-cell rounding of exp/log, which the FP64 finish removes: f64
// Newton steps from the f32 optimum and the log-likelihood and log-determinant
// at the stepped point (edge-rs newton_finish, src/sc/pml.rs:807-1090).  If
// the finish does not settle, the whole Newton loop reruns in f64 from the f32
// point with opt_pml's exact control flow (opt_pml_from, pml.rs:637).
//
// Reductions.  Lane partials are reduced with a fixed butterfly, wave partials
// are added in wave order and per-subject sums over subjects are taken by one
// wave per quantity with a fixed lane assignment, so results do not depend on
// timing.  Everything between sweeps (subject sums, Schur complement, the
// pivoted LDL' solve, damping) works on LDS, or on the global work area when
// the subject count makes the LDS one too large.
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "device_common.hpp"
#include "zero_tables.hpp"

namespace nh {

struct DeviceData {
  const float* row32;    // n_cells * rw: non-intercept design columns, log offset, padding
  const double* row64;   // the same in f64, n_cells * rw64
  const double* x64;     // n_cells * nb centred design (stage one)
  const double* off64;   // log offset (stage one)
  const uint32_t* fid;   // k + 1 subject boundaries
  const uint64_t* gptr;  // n_genes + 1 CSR row starts
  const uint32_t* gcell; // positive cells, increasing within a gene
  const float* gy32;     // positive counts
  const double* gy64;
  const double* gtot;    // n_genes * k subject totals
  // zero-count tables (tabled != 0): groups, per-subject group and loose-cell
  // ranges, loose cells, base-panel index words, panel coefficients
  const ZGroup* zgroup;
  const uint32_t* zsg;
  const uint32_t* zloose;
  const uint32_t* zsl;
  const uint32_t* zindex;
  const double* zcoef;
  uint32_t n_cells, nb, k, icpt, rw, rw64, tabled;
};

struct FitKnobs {
  double eps;        // nebula's absolute stopping tolerance
  double noise_rel;  // f32 resolution floor relative to |loglik|
  int max_iter;      // Newton budget (50)
  int max_backtrack; // 10
  // Stage-two evaluations: stop on the Newton decrement (the step's predicted
  // log-likelihood gain) where the measured rule would sweep the stepped
  // point -- in the FP32 loop below the resolution floor, in the FP64 finish
  // at its first step within eps (log-determinant moved to first order); and
  // keep an FP32 step whose measured loss is within FP32 rounding.  0 keeps
  // opt_pml's measured rules (stage three, the request API).
  int decrement_stop;
};

// Counters every fit adds to (device-wide, relaxed atomics).
struct FitCounters {
  unsigned long long sweeps_f32, sweeps_f64;
  unsigned long long cells_f32, nnz_f32;  // summed over FP32 sweeps
  unsigned long long cells_f64, nnz_f64;  // summed over FP64 sweeps
};

// SW_FINISH: the first sweep of the FP64 finish.  Per subject the resid and
// weight sums, the eta-derivatives of the weight sum (sum of d pw / d eta,
// and that times each stored column), from which the finish moves the
// log-determinant to the stepped point, and the gamma-derivative of the resid
// sum; per gene the log-likelihood, the score and the gamma-derivative of the
// score (Fitter::sensitivity).  The curvature blocks over the stored columns (pw x per subject,
// pw x x' per gene) are not summed: the finish's one Newton step takes them
// from the FP32 loop's last sweep, a point at most one small step away, where
// they enter only the step's second-order terms.
enum SweepMode { SW_FULL = 0, SW_FINISH = 1, SW_LAPLACE = 2 };

// Cells per lane in one tile of a sweep (tile = 64 x U cells).
#ifndef NH_TILE_F32
#define NH_TILE_F32 2
#endif
#ifndef NH_TILE_F64
#define NH_TILE_F64 1
#endif
#ifndef NH_TILE_F64_FINISH
#define NH_TILE_F64_FINISH 2
#endif
constexpr int kTileF32 = NH_TILE_F32;
constexpr int kTileF64 = NH_TILE_F64;
// the finish sweep keeps fewer sums, so two cells per lane fit in registers
// and their exp/log chains interleave
constexpr int kTileF64Finish = NH_TILE_F64_FINISH;
template <class CT, int MODE>
constexpr int tile_cells() {
  return sizeof(CT) == 4 ? kTileF32 : (MODE == SW_FINISH ? kTileF64Finish : kTileF64);
}
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// doubles in the FP64 sweeps' exp and log tables (fast_exp, fast_log)
constexpr int kF64TabDoubles = 64 + 2 * 129;
// floats in one wave's count staging row: a tile's counts in either type
constexpr int kStageFloats =
    kWave * cmax(4, cmax(kTileF32, 2 * cmax(kTileF64, kTileF64Finish)));

template <int NB>
struct Dims {
  static constexpr int Q = NB - 1;                // stored design columns
  // values per packed row: f32 rows padded to 1, 2, 4 or 8 (float4 loads),
  // f64 rows to the next even count (double2 loads), so p = 5, 6 read 48
  // bytes per cell, not 64
  static constexpr int R = Q + 1 <= 1 ? 1 : (Q + 1 <= 2 ? 2 : (Q + 1 <= 4 ? 4 : 8));
  static constexpr int R64 = Q + 1 <= 1 ? 1 : ((Q + 2) & ~1);
  template <class CT>
  static constexpr int RW = sizeof(CT) == 4 ? R : R64;
  static constexpr int QT = Q * (Q + 1) / 2;
  static constexpr int NS = 2 + Q;                // per subject: resid, weight, first[Q]
  static constexpr int NF = 4 + Q;                // SW_FINISH: resid, weight, dweight, dresid/dgamma,
                                                  //   dweight x[Q]
  static constexpr int NSL = NF > 4 ? NF : 4;     // slot width (Laplace keeps 4)
  static constexpr int NG = 1 + Q + QT;           // per gene: llp, score[Q], second[QT]
  static constexpr int NGF = 1 + 2 * Q;           // SW_FINISH: llp, score[Q], dscore/dgamma[Q]
  static constexpr int SUB = 3 + NB;              // dw, vw, w, vwb[NB]
  static constexpr int NTRI = NB * (NB + 1) / 2;
};

// Occupancy target of the fit kernels (waves per SIMD): the f64 accumulators of
// a sweep grow as p^2 / 2, so wide designs trade occupancy for no spills in
// the sweep loop (128 / 168 / 256 VGPRs).
template <int NB>
constexpr int kWavesPerEu = NB <= 4 ? 4 : (NB <= 6 ? 3 : 2);

__host__ __device__ inline uint32_t packed_row_width(uint32_t nb) {
  const uint32_t r = nb;  // nb - 1 columns + offset
  return r <= 1 ? 1u : (r <= 2 ? 2u : (r <= 4 ? 4u : 8u));
}
// the f64 rows' width (Dims::R64)
__host__ __device__ inline uint32_t packed_row_width64(uint32_t nb) {
  return nb <= 1 ? 1u : ((nb + 1u) & ~1u);
}

// Per-fit work arrays: k-sized and partial-sum buffers, in LDS when they fit.
template <int NB>
struct FitWork {
  using D = Dims<NB>;
  double* psub;     // (k + W) * NSL  wave partials per subject, slot s + w
  double* pgene;    // W * NG         wave partials per gene
  double* tsub;     // k * NSL        per-subject totals
  double* tgene;    // NG             per-gene totals
  double* llsub;    // k              prior and count terms of the log-likelihood
  double* subs;     // 2 * k * SUB    per side: dw, vw, w, vwb
  double* lw;       // k  current log w
  double* new_lw;   // k  trial
  double* step_lw;  // k
  double* damp_lw;  // k
  double* vw_step;  // k  vw of the last Newton step (log-det, pml.rs:1767)
  __host__ __device__ static size_t doubles(int k, int waves) {
    return (size_t)(k + waves) * D::NSL + (size_t)waves * D::NG + (size_t)k * D::NSL + D::NG +
           (size_t)k + 2 * (size_t)k * D::SUB + 5 * (size_t)k;
  }
  __device__ void carve(double* base, int k, int waves) {
    psub = base;
    base += (size_t)(k + waves) * D::NSL;
    pgene = base;
    base += (size_t)waves * D::NG;
    tsub = base;
    base += (size_t)k * D::NSL;
    tgene = base;
    base += D::NG;
    llsub = base;
    base += k;
    subs = base;
    base += 2 * (size_t)k * D::SUB;
    lw = base;
    base += k;
    new_lw = base;
    base += k;
    step_lw = base;
    base += k;
    damp_lw = base;
    base += k;
    vw_step = base;
  }
};

template <int NB, int WAVES>
struct FitLds {
  double beta[NB], new_beta[NB], step_beta[NB], damp_beta[NB];
  double db[2][NB];
  double vb[2][NB * NB];
  double vb2[NB * NB];
  double rhs[NB];
  double rhs2[2][NB];  // Fitter::sensitivity's right-hand sides
  double lm[NB * NB];  // LDL' work
  double lb[NB];
  int perm[NB];
  double ll[2];
  double scal[8];
  double f64tab[kF64TabDoubles];  // the FP64 sweep's exp and log tables (Fitter::setup)
  uint64_t ip0[WAVES];
  uint32_t s0[WAVES];
  float ysc[WAVES][kStageFloats];  // count staging: one tile's counts per wave
};

// What one fit hands back.
struct FitResult {
  double ll, ll_prev, log_det, second;
  int iterations, backtracks, finish_steps, exact;  // exact: f64 Newton loop ran
  int side;         // side holding the curvature blocks of the fit's last step
  int predictable;  // the SW_FINISH sums of that step are in the totals
                    // (Fitter::sensitivity may run)
};

// ---------------------------------------------------------------------------
// Eigen-style pivoted LDL' solve of a small symmetric system in LDS, in place
// (src/sc/pml.rs:1940-2021).  One thread.
template <int NB>
__device__ void ldlt_solve_lds(double* a, double* b, int* perm) {
  const int n = NB;
  for (int st = 0; st < n; ++st) {
    int piv = st;
    double best = fabs(a[st * n + st]);
    for (int i = st + 1; i < n; ++i) {
      double v = fabs(a[i * n + i]);
      if (v > best) {
        best = v;
        piv = i;
      }
    }
    perm[st] = piv;
    if (piv != st) {
      for (int j = 0; j < st; ++j) {
        double t = a[st * n + j];
        a[st * n + j] = a[piv * n + j];
        a[piv * n + j] = t;
      }
      for (int i = piv + 1; i < n; ++i) {
        double t = a[i * n + st];
        a[i * n + st] = a[i * n + piv];
        a[i * n + piv] = t;
      }
      double t = a[st * n + st];
      a[st * n + st] = a[piv * n + piv];
      a[piv * n + piv] = t;
      for (int i = st + 1; i < piv; ++i) {
        double u = a[i * n + st];
        a[i * n + st] = a[piv * n + i];
        a[piv * n + i] = u;
      }
    }
    if (st > 0) {
      double acc = 0.0;
      for (int j = 0; j < st; ++j) acc += a[st * n + j] * (a[j * n + j] * a[st * n + j]);
      a[st * n + st] -= acc;
      for (int i = st + 1; i < n; ++i) {
        double ac = 0.0;
        for (int j = 0; j < st; ++j) ac += a[i * n + j] * (a[j * n + j] * a[st * n + j]);
        a[i * n + st] -= ac;
      }
    }
    double d = a[st * n + st];
    if (d != 0.0)
      for (int i = st + 1; i < n; ++i) a[i * n + st] /= d;
  }
  for (int st = 0; st < n; ++st) {
    double t = b[st];
    b[st] = b[perm[st]];
    b[perm[st]] = t;
  }
  for (int i = 0; i < n; ++i) {
    double acc = b[i];
    for (int j = 0; j < i; ++j) acc -= a[i * n + j] * b[j];
    b[i] = acc;
  }
  for (int i = 0; i < n; ++i) {
    double d = a[i * n + i];
    b[i] = fabs(d) > 2.2250738585072014e-308 ? b[i] / d : 0.0;
  }
  for (int i = n - 1; i >= 0; --i) {
    double acc = b[i];
    for (int j = i + 1; j < n; ++j) acc -= a[j * n + i] * b[j];
    b[i] = acc;
  }
  for (int st = n - 1; st >= 0; --st) {
    double t = b[st];
    b[st] = b[perm[st]];
    b[perm[st]] = t;
  }
}

template <class CT>
__device__ __forceinline__ CT dexp(CT x);
template <>
__device__ __forceinline__ float dexp<float>(float x) { return __expf(x); }
template <>
__device__ __forceinline__ double dexp<double>(double x) { return exp(x); }
template <class CT>
__device__ __forceinline__ CT dlog(CT x);
template <>
__device__ __forceinline__ float dlog<float>(float x) { return __logf(x); }
template <>
__device__ __forceinline__ double dlog<double>(double x) { return log(x); }
template <class CT>
__device__ __forceinline__ CT drcp(CT x);
template <>
__device__ __forceinline__ float drcp<float>(float x) { return __frcp_rn(x); }
template <>
__device__ __forceinline__ double drcp<double>(double x) { return 1.0 / x; }

typedef float gfloat4 __attribute__((ext_vector_type(4)));
typedef float gfloat2 __attribute__((ext_vector_type(2)));
typedef double gdouble2 __attribute__((ext_vector_type(2)));

// One packed row of R values (16-byte vector loads where the row allows).
template <class CT, int R, class P>
__device__ __forceinline__ void load_row(P base, uint32_t cell, CT* out) {
  const auto p = base + (size_t)cell * R;
  if constexpr (sizeof(CT) == 4 && R >= 4) {
#pragma unroll
    for (int h = 0; h < R / 4; ++h) {
      const gfloat4 v = reinterpret_cast<__attribute__((address_space(1))) const gfloat4*>(p)[h];
      out[4 * h] = v.x;
      out[4 * h + 1] = v.y;
      out[4 * h + 2] = v.z;
      out[4 * h + 3] = v.w;
    }
  } else if constexpr (sizeof(CT) == 4 && R == 2) {
    const gfloat2 v = *reinterpret_cast<__attribute__((address_space(1))) const gfloat2*>(p);
    out[0] = v.x;
    out[1] = v.y;
  } else if constexpr (sizeof(CT) == 8 && R >= 2) {
#pragma unroll
    for (int h = 0; h < R / 2; ++h) {
      const gdouble2 v = reinterpret_cast<__attribute__((address_space(1))) const gdouble2*>(p)[h];
      out[2 * h] = v.x;
      out[2 * h + 1] = v.y;
    }
  } else {
    out[0] = p[0];
  }
}

// The same from a uniform base and a 32-bit byte offset (cell * R * size),
// so the load takes the base in scalar registers and one offset register.
typedef __attribute__((address_space(1))) const char gconst_char;
template <class CT, int R>
__device__ __forceinline__ void load_row_at(gconst_char* base, uint32_t cell, CT* out) {
  const uint32_t off = cell * (uint32_t)(R * sizeof(CT));
  load_row<CT, R>(reinterpret_cast<__attribute__((address_space(1))) const CT*>(base + off), 0u,
                  out);
}

// index of (a, b), a <= b, in the packed upper triangle of a Q x Q matrix
__host__ __device__ constexpr int tri_index(int q, int a, int b) {
  return a * q - a * (a - 1) / 2 + (b - a);
}

typedef __attribute__((address_space(3))) float lds_float;
typedef __attribute__((address_space(3))) double lds_double;
template <class CT>
struct Lds;
template <>
struct Lds<float> {
  using ptr = lds_float*;
};
template <>
struct Lds<double> {
  using ptr = lds_double*;
};

// Global-memory views of the inputs: loads through them are global_load
// instructions (counted on vmcnt only and returned in order), so a wait for
// one tile's rows leaves the next CSR window's loads in flight.  Generic
// pointers would compile to flat loads, which every wait must drain.
typedef __attribute__((address_space(1))) const float gconst_float;
typedef __attribute__((address_space(1))) const double gconst_double;
typedef __attribute__((address_space(1))) const uint32_t gconst_u32;
template <class CT>
struct Glb;
template <>
struct Glb<float> {
  using ptr = gconst_float*;
};
template <>
struct Glb<double> {
  using ptr = gconst_double*;
};

// Raw hardware transcendentals for the FP32 sweeps: v_exp_f32 / v_log_f32
// (base 2) and v_rcp_f32, about 1 ulp; the arguments here are normal numbers
// (t = e + gamma >= gamma > 0).  The table argument is unused in FP32.
__device__ __forceinline__ float fast_exp(float x, const lds_double*) {
  return __builtin_amdgcn_exp2f(x * 1.44269504f);
}
__device__ __forceinline__ float fast_log(float x, const lds_double*) {
  return __builtin_amdgcn_logf(x) * 0.693147181f;
}
__device__ __forceinline__ float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// Relative size, against |loglik|, of a predicted gain and a measured loss
// that an FP32 Newton step may show before it counts as an overshoot (the
// FP32 sums carry rounding of order 1e-7 of the summed magnitudes).
constexpr double kF32Trust = 1e-6;

// FP64 exp, log and reciprocal of the FP64 sweeps, with short dependency
// chains and no special-case branches (the sweep's arguments are finite;
// overflow gives inf, underflow 0, NaN stays NaN).  Tables in LDS
// (kF64TabDoubles doubles, filled by fill_f64_tables):
//   [0, 64)            2^(j/64)
//   [64 + 2j, +2)      1 / c_j and log c_j, c_j = 1 + j/128, j = 0..128
//                      (log c_128 = 0: its binade is moved up by one)
constexpr int kTabExp = 0, kTabLog = 64;
constexpr double kLn2Hi = 6.93147180369123816490e-01;  // 32 significant bits
constexpr double kLn2Lo = 1.90821492927058770002e-10;

// a * b + c with c a constant held in scalar registers.  Left to itself the
// compiler keeps such constants in vector registers and, to use the
// two-operand v_fmac_f64 that overwrites its addend, copies one before every
// multiply-add of a polynomial; v_fma_f64 reads the constant in place.
__device__ __forceinline__ double fma_sk(double a, double b, double c) {
  double d;
  asm("v_fma_f64 %0, %1, %2, %3" : "=v"(d) : "v"(a), "v"(b), "s"(c));
  return d;
}

// exp(x) = 2^(i/64) e^r, i = rint(64 x / ln 2), |r| <= ln 2 / 128; e^r - 1 by
// its Taylor series to r^5 (truncation < 4e-17 relative); within 2 ulp.
__device__ __forceinline__ double fast_exp(double x, const lds_double* tab) {
  const double kd = __builtin_rint(x * 92.332482616893657);  // 64 / ln 2
  double r = __builtin_fma(-kd, kLn2Hi * 0.015625, x);
  r = __builtin_fma(-kd, kLn2Lo * 0.015625, r);
  const int i = (int)kd;  // saturates for |x| beyond the double range of exp
  const double s = tab[kTabExp + (i & 63)];
  double q = fma_sk(r, 1.0 / 120, 1.0 / 24);
  q = fma_sk(r, q, 1.0 / 6);
  q = __builtin_fma(r, q, 0.5);
  const double em1 = __builtin_fma(r * r, q, r);
  return __builtin_ldexp(__builtin_fma(s, em1, s), i >> 6);
}

// log(t), t > 0: t = 2^e m, m in [1, 2); c_j the nearest multiple of 1/128,
// r = m / c_j - 1 (one fma), |r| <= 1/256; log(1 + r) by its series to r^6
// (truncation < 3e-18); absolute error about 1 ulp of the result's larger
// part.  inf and NaN pass through.
__device__ __forceinline__ double fast_log(double t, const lds_double* tab) {
  const uint64_t b = __builtin_bit_cast(uint64_t, t);
  const uint32_t hi = (uint32_t)(b >> 32);
  const uint32_t j = (((hi >> 12) & 255u) + 1u) >> 1;
  const int e = (int)(hi >> 20) - 1023 + (int)(j >> 7);
  const double m = __builtin_bit_cast(
      double, ((uint64_t)((hi & 0x000FFFFFu) | 0x3FF00000u) << 32) | (uint32_t)b);
  const double invc = tab[kTabLog + 2 * j], logc = tab[kTabLog + 2 * j + 1];
  const double r = __builtin_fma(m, invc, -1.0);
  double q = fma_sk(r, -1.0 / 6, 1.0 / 5);
  q = fma_sk(r, q, -1.0 / 4);
  q = fma_sk(r, q, 1.0 / 3);
  q = __builtin_fma(r, q, -0.5);
  const double lp = __builtin_fma(r * r, q, r);
  const double ed = (double)e;
  const double res = __builtin_fma(ed, kLn2Hi, logc) + __builtin_fma(ed, kLn2Lo, lp);
  return t < __builtin_inf() ? res : t;
}

// 1 / t: v_rcp_f64 and two Newton-Raphson steps (about 1 ulp).
__device__ __forceinline__ double fast_rcp(double t) {
  double r = __builtin_amdgcn_rcp(t);
  double e = __builtin_fma(-t, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-t, r, 1.0);
  return __builtin_fma(r, e, r);
}

// Fills the tables above (one workgroup's copy; threads 0..128).
__device__ __forceinline__ void fill_f64_tables(double* tab, int tid) {
  if (tid < 64) tab[kTabExp + tid] = exp2((double)tid * 0.015625);
  if (tid < 129) {
    // log c_j as -log of the rounded 1 / c_j, whose rounding then cancels
    // against the one r = m / c_j - 1 carries
    const double invc = 1.0 / (1.0 + (double)tid * 0.0078125);
    tab[kTabLog + 2 * tid] = invc;
    tab[kTabLog + 2 * tid + 1] = tid == 128 ? 0.0 : -log(invc);
  }
}

// One wave's part of a sweep: cells [cb, ce), subjects from s, CSR cursor ip
// (first entry at or past cb).  Writes the wave's per-subject sums to psub
// (slot s + wave) and its per-gene sums to pgene (row wave).  Everything it
// reads comes in as restrict-qualified arguments so loads can be hoisted and
// kept in flight; the count staging row is addressed as LDS.
template <class CT, int MODE, int NB>
__device__ __forceinline__ void sweep_wave(const CT* __restrict__ rows,
                                        const uint32_t* __restrict__ gcell,
                                        const CT* __restrict__ gy,
                                        const uint32_t* __restrict__ fid,
                                        const double* __restrict__ bsrc,
                                        const double* __restrict__ lwsrc,
                                        typename Lds<CT>::ptr ys, const lds_double* tab,
                                        double* __restrict__ psub,
                                        double* __restrict__ pgene, uint64_t ip, uint64_t g1,
                                        int s, int k, uint32_t cb, uint32_t ce, int icpt,
                                        double gamma, int lane, int wave) {
  using D = Dims<NB>;
  constexpr int Q = D::Q, R = D::template RW<CT>, NS = D::NS, NSL = D::NSL, NG = D::NG;
  constexpr int U = tile_cells<CT, MODE>();
  // every argument but lane is the wave's own: scalar registers
  gconst_char* const grows = (gconst_char*)uni(rows);
  fid = uni(fid);
  bsrc = uni(bsrc);
  lwsrc = uni(lwsrc);
  psub = uni(psub);
  pgene = uni(pgene);
  ip = uni(ip);
  g1 = uni(g1);
  s = uni(s);
  k = uni(k);
  cb = uni(cb);
  ce = uni(ce);
  icpt = uni(icpt);
  gamma = uni(gamma);
  wave = uni(wave);
  constexpr int TILE = kWave * U;
  constexpr int NSV = MODE == SW_LAPLACE ? 4 : (MODE == SW_FINISH ? D::NF : NS);
  constexpr int NGV = MODE == SW_LAPLACE ? 0 : (MODE == SW_FINISH ? D::NGF : NG);
  constexpr int QA = Q > 0 ? Q : 1;
  CT bq[QA];
#pragma unroll
  for (int q = 0; q < QA; ++q) bq[q] = q < Q ? (CT)bsrc[q < icpt ? q : q + 1] : (CT)0;
  const CT bint = (CT)bsrc[icpt];
  const CT g = (CT)gamma;
  double ag[NGV > 0 ? NGV : 1];
#pragma unroll
  for (int v = 0; v < (NGV > 0 ? NGV : 1); ++v) ag[v] = 0.0;
  if (cb < ce) {
    // window of the next TILE entries of the gene's CSR row.  Every load is
    // issued unconditionally at an in-bounds index (the last entry when past
    // the row; buffers hold at least one entry) and masked afterwards, so the
    // compiler can count outstanding loads and wait only for the ones needed.
    // Indices are 32-bit, relative to wbase = min(ip, last entry): the
    // cursor rel and the row end n_rel (one gene's entries fit 32 bits).
    const uint64_t e_last = g1 > 0 ? g1 - 1 : 0;
    const uint64_t wbase = ip < e_last ? ip : e_last;
    gconst_char* const ggcell = (gconst_char*)((gconst_u32*)uni(gcell) + wbase);
    gconst_char* const ggy = (gconst_char*)((typename Glb<CT>::ptr)uni(gy) + wbase);
    // entry ec of the window: 32-bit byte offsets from the scalar bases
    auto entry = [&](uint32_t ec, uint32_t& c, CT& y) {
      c = *(gconst_u32*)(ggcell + (ec << 2));
      y = *(typename Glb<CT>::ptr)(ggy + ec * (uint32_t)sizeof(CT));
    };
    const uint32_t n_rel = (uint32_t)(g1 - wbase);
    const uint32_t c_last = n_rel > 0 ? n_rel - 1 : 0;
    uint32_t rel = (uint32_t)(ip - wbase);
    uint32_t wc[U];
    CT wy[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const uint32_t e = rel + (uint32_t)(lane + kWave * j);
      const uint32_t ec = e < n_rel ? e : c_last;
      uint32_t c;
      CT y;
      entry(ec, c, y);
      wc[j] = e < n_rel ? c : 0xffffffffu;
      wy[j] = y;
    }
    for (; s < k && fid[s] < ce; ++s) {
      const uint32_t lo = max(fid[s], cb), hi = min(fid[s + 1], ce);
      const CT lws = (CT)lwsrc[s];
      const CT c0 = bint + lws;
      double as[NSV];
#pragma unroll
      for (int v = 0; v < NSV; ++v) as[v] = 0.0;
      for (uint32_t base = lo; base < hi; base += TILE) {
        const uint32_t tend = min(base + (uint32_t)TILE, hi);
        CT xr[U][R];
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t cell = base + (uint32_t)(lane + kWave * u);
          load_row_at<CT, R>(grows, cell < tend ? cell : tend - 1, xr[u]);
        }
        // the window's entries inside [base, tend) form a prefix of it
        uint32_t m = 0;
#pragma unroll
        for (int j = 0; j < U; ++j) {
          const bool in = wc[j] < tend;
          if (in) ys[wc[j] - base] = wy[j];
          m += (uint32_t)popc64(wave_ballot(in));
        }
        rel += m;
#pragma unroll
        for (int j = 0; j < U; ++j) {
          const uint32_t e = rel + (uint32_t)(lane + kWave * j);
          const uint32_t ec = e < n_rel ? e : c_last;
          uint32_t c;
          CT y;
          entry(ec, c, y);
          wc[j] = e < n_rel ? c : 0xffffffffu;
          wy[j] = y;
        }
        __builtin_amdgcn_wave_barrier();
        CT yv[U];
#pragma unroll
        for (int u = 0; u < U; ++u) yv[u] = ys[lane + kWave * u];
        __builtin_amdgcn_wave_barrier();
#pragma unroll
        for (int u = 0; u < U; ++u) ys[lane + kWave * u] = (CT)0;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          const uint32_t cell = base + (uint32_t)(lane + kWave * u);
          if (cell >= tend) continue;
          // the cell's terms written out here: as an inlined helper taking
          // the sums by pointer, p = 8 needed 256 VGPRs and spilled
          const CT y = yv[u];
          CT eta = xr[u][Q];  // log offset
#pragma unroll
          for (int q = 0; q < Q; ++q) eta += xr[u][q] * bq[q];
          const CT ex = fast_exp(eta + c0, tab);
          eta += bint;
          const CT t = ex + g;
          const CT inv = fast_rcp(t);
          if constexpr (MODE == SW_LAPLACE) {
            // (gamma + y) times the four derivative factors (pml.rs:1861-1874)
            double e64 = (double)ex, t64 = e64 + gamma, gs64 = gamma + (double)y;
            double p2 = e64 / (t64 * t64), p3 = p2 / t64, p4 = p3 / t64, p5 = p4 / t64;
            double e2 = e64 * e64;
            as[0] += gs64 * p2;
            as[1] += gs64 * (p3 * (gamma - e64));
            as[2] += gs64 * (p4 * (gamma * gamma + e2 - (4.0 * gamma) * e64));
            as[3] += gs64 * (p5 * (gamma * gamma * gamma - (11.0 * gamma * gamma) * e64 +
                                   (11.0 * gamma) * e2 - e2 * e64));
          } else {
            const CT u_ = ex * inv;
            const CT gs = g + y;
            const CT lt = fast_log(t, tab);
            const double pw = (double)(gs * u_ * inv);
            const double dres = (double)y - (double)gs * (double)u_;
            const double term = (double)(y * eta) - (double)gs * (double)lt;
            as[0] += dres;
            as[1] += pw;
            ag[0] += term;
            double xd[QA];
#pragma unroll
            for (int q = 0; q < Q; ++q) xd[q] = (double)xr[u][q];
            if constexpr (MODE == SW_FINISH) {
              // d pw / d eta = pw (gamma - e) / t; d resid / d gamma = (y - e) e / t^2
              const double dpw = pw * (((double)g - (double)ex) * (double)inv);
              const double dgam = ((double)y - (double)ex) * ((double)u_ * (double)inv);
              as[2] += dpw;
              as[3] += dgam;
#pragma unroll
              for (int q = 0; q < Q; ++q) {
                as[4 + q] += dpw * xd[q];
                ag[1 + q] += xd[q] * dres;
                ag[1 + Q + q] += xd[q] * dgam;
              }
            } else {
#pragma unroll
              for (int a = 0; a < Q; ++a) {
                const double pa = pw * xd[a];
                as[2 + a] += pa;
                ag[1 + a] += xd[a] * dres;
#pragma unroll
                for (int b = a; b < Q; ++b) ag[1 + Q + tri_index(Q, a, b)] += pa * xd[b];
              }
            }
          }
        }
      }
      // flush subject s: wave sum, slot s + wave
#pragma unroll
      for (int v = 0; v < NSV; ++v) {
        const double r = wave_sum(as[v]);
        if (lane == 0) psub[(size_t)(s + wave) * NSL + v] = r;
      }
    }
  }
  if constexpr (NGV > 0) {
#pragma unroll
    for (int v = 0; v < NGV; ++v) {
      const double r = wave_sum(ag[v]);
      if (lane == 0) pgene[(size_t)wave * NG + v] = r;
    }
  }
}

// One wave's part of the pass that starts a cold fit's log w: per subject of
// its block [cb, ce), sum over the cells of exp(x'beta + log offset) in FP32
// (a start value only), written to slot s + wave.
template <int NB>
__device__ __forceinline__ void sweep_wave_mean(const float* __restrict__ rows,
                                                const uint32_t* __restrict__ fid,
                                                const double* __restrict__ bsrc,
                                                double* __restrict__ psub, int s, int k,
                                                uint32_t cb, uint32_t ce, int icpt, int lane,
                                                int wave) {
  using D = Dims<NB>;
  constexpr int Q = D::Q, R = D::R, NSL = D::NSL;
  constexpr int QA = Q > 0 ? Q : 1;
  constexpr int U = 4;
  const gconst_float* grows = (const gconst_float*)rows;
  float bq[QA];
#pragma unroll
  for (int q = 0; q < QA; ++q) bq[q] = q < Q ? (float)bsrc[q < icpt ? q : q + 1] : 0.0f;
  const float bint = (float)bsrc[icpt];
  if (cb >= ce) return;
  for (; s < k && fid[s] < ce; ++s) {
    const uint32_t lo = max(fid[s], cb), hi = min(fid[s + 1], ce);
    float acc = 0.0f;
    for (uint32_t base = lo; base < hi; base += kWave * U) {
      float xr[U][R];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const uint32_t c = base + (uint32_t)(lane + kWave * u);
        load_row<float, R>(grows, c < hi ? c : hi - 1, xr[u]);
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        float eta = xr[u][Q] + bint;
#pragma unroll
        for (int q = 0; q < Q; ++q) eta += xr[u][q] * bq[q];
        const float e = fast_exp(eta, nullptr);
        if (base + (uint32_t)(lane + kWave * u) < hi) acc += e;
      }
    }
    const double r = wave_sum((double)acc);
    if (lane == 0) psub[(size_t)(s + wave) * NSL] = r;
  }
}

// One wave's part of a sweep on a design with zero-count tables, FP64
// throughout (src/sc/pml.rs:1510-1560, :1869-1890): per subject of its block
// [cb, ce), the zero-count sums of the subject's groups whose first cell lies
// in the block (one table lookup each, lanes over groups), the loose cells in
// the block, then what the positive counts in the block add -- every term is
// linear in the count through gamma + y, so the split is exact.  Writes the
// same per-subject slots (s + wave) and per-gene rows as sweep_wave.
template <int MODE, int NB>
__device__ __forceinline__ void sweep_wave_tabled(
    const double* __restrict__ rows, const uint32_t* __restrict__ gcell,
    const double* __restrict__ gy, const uint32_t* __restrict__ fid,
    const ZGroup* __restrict__ zgroup, const uint32_t* __restrict__ zsg,
    const uint32_t* __restrict__ zloose, const uint32_t* __restrict__ zsl,
    const uint32_t* __restrict__ zindex, const double* __restrict__ zcoef,
    const double* __restrict__ bsrc, const double* __restrict__ lwsrc, const lds_double* tab,
    double* __restrict__ psub, double* __restrict__ pgene, uint64_t ip, uint64_t g1, int s, int k,
    uint32_t cb, uint32_t ce, int icpt, double gamma, int lane, int wave) {
  using D = Dims<NB>;
  constexpr int Q = D::Q, R = D::R64, NS = D::NS, NSL = D::NSL, NG = D::NG;
  constexpr int NSV = MODE == SW_LAPLACE ? 4 : (MODE == SW_FINISH ? D::NF : NS);
  constexpr int NGV = MODE == SW_LAPLACE ? 0 : (MODE == SW_FINISH ? D::NGF : NG);
  constexpr int QA = Q > 0 ? Q : 1;
  double bq[QA];
#pragma unroll
  for (int q = 0; q < QA; ++q) bq[q] = q < Q ? bsrc[q < icpt ? q : q + 1] : 0.0;
  const double bint = bsrc[icpt];
  const double lng = log(gamma);
  double ag[NGV > 0 ? NGV : 1];
#pragma unroll
  for (int v = 0; v < (NGV > 0 ? NGV : 1); ++v) ag[v] = 0.0;
  // adds one item: resid d, weight p, log-likelihood term tm, d p / d eta dp
  // (SW_FINISH), or the four Laplace sums lp (SW_LAPLACE), at stored row xd
  auto fold = [&](double* as, const double* xd, double d, double p, double tm, double dp,
                  const double* lp) {
    if constexpr (MODE == SW_LAPLACE) {
#pragma unroll
      for (int v = 0; v < 4; ++v) as[v] += lp[v];
    } else {
      as[0] += d;
      as[1] += p;
      ag[0] += tm;
      if constexpr (MODE == SW_FINISH) {
        // the gamma-derivatives (slot 3, per-gene 1 + Q + q) stay zero: the
        // tabled fits make no start predictions
        as[2] += dp;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          as[4 + q] += dp * xd[q];
          ag[1 + q] += xd[q] * d;
        }
      } else {
#pragma unroll
        for (int a = 0; a < Q; ++a) {
          const double pa = p * xd[a];
          as[2 + a] += pa;
          ag[1 + a] += xd[a] * d;
#pragma unroll
          for (int b = a; b < Q; ++b) ag[1 + Q + tri_index(Q, a, b)] += pa * xd[b];
        }
      }
    }
  };
  // one cell: its zero-count terms (positive false) or what count y adds
  auto cell = [&](double* as, uint32_t c, double y, bool positive, double lws) {
    const double* rp = rows + (size_t)c * R;
    double xd[QA];
#pragma unroll
    for (int q = 0; q < QA; ++q) xd[q] = q < Q ? rp[q] : 0.0;
    double eta = rp[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) eta += xd[q] * bq[q];
    eta += bint;
    const double ex = fast_exp(eta + lws, tab);
    const double t = ex + gamma;
    const double inv = fast_rcp(t);
    const double w = positive ? y : gamma;
    if constexpr (MODE == SW_LAPLACE) {
      const double p2 = ex * inv * inv, p3 = p2 * inv, p4 = p3 * inv, p5 = p4 * inv;
      const double e2 = ex * ex;
      double lp[4];
      lp[0] = w * p2;
      lp[1] = w * (p3 * (gamma - ex));
      lp[2] = w * (p4 * (gamma * gamma + e2 - (4.0 * gamma) * ex));
      lp[3] = w * (p5 * (gamma * gamma * gamma - (11.0 * gamma * gamma) * ex +
                         (11.0 * gamma) * e2 - e2 * ex));
      fold(as, xd, 0.0, 0.0, 0.0, 0.0, lp);
    } else {
      const double u = ex * inv;
      const double lt = fast_log(t, tab);
      const double p = w * u * inv;
      const double d = positive ? y * (1.0 - u) : -gamma * u;
      const double tm = positive ? y * (eta - lt) : -gamma * lt;
      const double dp = MODE == SW_FINISH ? p * ((gamma - ex) * inv) : 0.0;
      fold(as, xd, d, p, tm, dp, nullptr);
    }
  };
  if (cb < ce) {
    for (; s < k && fid[s] < ce; ++s) {
      const uint32_t lo = max(fid[s], cb), hi = min(fid[s + 1], ce);
      const double lws = lwsrc[s];
      double as[NSV];
#pragma unroll
      for (int v = 0; v < NSV; ++v) as[v] = 0.0;
      // groups whose first cell is in [lo, hi): v = exp(ln tau + ln O) with
      // ln tau = log w + x'beta - ln gamma
      for (uint32_t gi = zsg[s] + (uint32_t)lane; gi < zsg[s + 1]; gi += kWave) {
        const ZGroup& G = zgroup[gi];
        if (G.row < lo || G.row >= hi) continue;
        const double* rp = rows + (size_t)G.row * R;
        double xd[QA];
#pragma unroll
        for (int q = 0; q < QA; ++q) xd[q] = q < Q ? rp[q] : 0.0;
        double st = lws - lng + bint;
#pragma unroll
        for (int q = 0; q < Q; ++q) st += xd[q] * bq[q];
        if constexpr (MODE == SW_LAPLACE) {
          double f[4];
          zt_eval<2, 4>(G, zindex, zcoef, st, f);  // gamma times each factor: B..E
          fold(as, xd, 0.0, 0.0, 0.0, 0.0, f);
        } else if constexpr (MODE == SW_FINISH) {
          double f[4];
          zt_eval<0, 4>(G, zindex, zcoef, st, f);
          fold(as, xd, -gamma * f[1], f[2], -gamma * (G.n * lng + f[0]), f[3], nullptr);
        } else {
          double f[3];
          zt_eval<0, 3>(G, zindex, zcoef, st, f);
          fold(as, xd, -gamma * f[1], f[2], -gamma * (G.n * lng + f[0]), 0.0, nullptr);
        }
      }
      // loose cells in [lo, hi): the first at or past lo, then in order
      uint32_t a = zsl[s], b = zsl[s + 1];
      const uint32_t l_end = b;
      while (a < b) {
        const uint32_t m = (a + b) / 2;
        if (zloose[m] < lo) a = m + 1; else b = m;
      }
      for (uint32_t i = a + (uint32_t)lane; i < l_end; i += kWave) {
        const uint32_t c = zloose[i];
        if (c >= hi) break;
        cell(as, c, 0.0, false, lws);
      }
      // positive counts with cells in [lo, hi): a prefix of the window at ip
      for (;;) {
        const uint64_t e = ip + (uint64_t)lane;
        const uint32_t c = e < g1 ? gcell[e] : 0xffffffffu;
        const bool in = c < hi;
        if (in) cell(as, c, gy[e], true, lws);
        const int m = popc64(wave_ballot(in));
        ip += (uint64_t)m;
        if (m < kWave) break;
      }
#pragma unroll
      for (int v = 0; v < NSV; ++v) {
        const double r = wave_sum(as[v]);
        if (lane == 0) psub[(size_t)(s + wave) * NSL + v] = r;
      }
    }
  }
  if constexpr (NGV > 0) {
#pragma unroll
    for (int v = 0; v < NGV; ++v) {
      const double r = wave_sum(ag[v]);
      if (lane == 0) pgene[(size_t)wave * NG + v] = r;
    }
  }
}

// ---------------------------------------------------------------------------
// The fit of one request, all threads of the workgroup cooperating.
// Every fit kernel holds its fit object and its FitLds in __shared__ memory:
// each step of the fit declares so (members and LDS arrays then read with
// ds_ instructions, not flat ones through private memory), binds `sh` to the
// kernel's FitLds and takes its thread indices from the hardware.
#define NH_FIT_LDS                                 \
  NH_ASSUME_LDS(this);                             \
  [[maybe_unused]] FitLds<NB, WAVES>& sh = *shp;   \
  NH_ASSUME_LDS(&sh);                              \
  [[maybe_unused]] const int tid = (int)threadIdx.x; \
  [[maybe_unused]] const int wave = tid / kWave;   \
  [[maybe_unused]] const int lane = tid % kWave;

template <int NB, int WAVES>
struct Fitter {
  using D = Dims<NB>;
  static constexpr int NT = WAVES * kWave;
  static constexpr int Q = D::Q, R = D::R, QT = D::QT, NS = D::NS, NSL = D::NSL, NG = D::NG;
  static constexpr int SUB = D::SUB, NTRI = D::NTRI;

  // The fit object is __shared__ in every fit kernel (see NH_FIT_LDS).
  DeviceData dd;
  FitLds<NB, WAVES>* shp;
  FitWork<NB> wk;
  FitCounters* counters;
  uint32_t gene;
  uint64_t g0, g1;   // CSR range
  double gamma, alpha, lambda;
  int k, icpt;
  uint32_t blk;      // cells per wave block

  // Set-up for gene g by one thread; the workgroup synchronises before the
  // fit starts.
  __device__ void init(const DeviceData& d, FitLds<NB, WAVES>* s, const FitWork<NB>& w,
                       FitCounters* ctr, uint32_t g, double gam, double alp, double lam) {
    dd = d;
    shp = s;
    wk = w;
    counters = ctr;
    gene = g;
    gamma = gam;
    alpha = alp;
    lambda = lam;
    k = (int)d.k;
    icpt = (int)d.icpt;
    g0 = d.gptr[g];
    g1 = d.gptr[g + 1];
    const uint32_t per = (d.n_cells + WAVES - 1) / WAVES;
    blk = (per + kWave - 1) / kWave * kWave;
  }

  __device__ double* side(int sd) const { return wk.subs + (size_t)sd * k * SUB; }

  // the FP64 exp / log tables (filled by setup); the per-subject steps use
  // the sweeps' exp and log, whose short code keeps these steps' register
  // use, and so their call-site saves, small
  __device__ const lds_double* tab() const { return (const lds_double*)&shp->f64tab[0]; }

  // design column of stored column q
  __device__ static constexpr int qcol(int q, int ic) { return q < ic ? q : q + 1; }

  __device__ uint32_t block_begin(int w) const { return min(dd.n_cells, (uint32_t)w * blk); }
  __device__ uint32_t block_end(int w) const { return min(dd.n_cells, (uint32_t)(w + 1) * blk); }

  // Per wave: the first positive entry at or past the block start (64-ary
  // search over the gene's sorted cells) and the subject holding it; clears
  // the wave's count staging row.
  __device__ __noinline__ void setup() {
    NH_FIT_LDS
    uint32_t cb = block_begin(wave);
    uint64_t lo = g0, hi = g1;
    while (hi - lo > (uint64_t)kWave) {
      uint64_t step = (hi - lo + kWave - 1) / kWave;
      uint64_t pos = lo + (uint64_t)lane * step;
      bool pred = pos < hi && dd.gcell[pos] < cb;
      int cnt = popc64(wave_ballot(pred));
      if (cnt == 0) {
        hi = lo;
      } else {
        uint64_t nlo = lo + (uint64_t)(cnt - 1) * step + 1;
        uint64_t nhi = lo + (uint64_t)cnt * step;
        hi = nhi < hi ? nhi : hi;
        lo = nlo;
      }
    }
    uint64_t pos = lo + (uint64_t)lane;
    bool pred = pos < hi && dd.gcell[pos] < cb;
    uint64_t ip = lo + (uint64_t)popc64(wave_ballot(pred));
    int a = 0, b = k;  // fid[a] <= cb < fid[b] (when cb < n_cells)
    while (b - a > 1) {
      int m = (a + b) / 2;
      if (dd.fid[m] <= cb) a = m; else b = m;
    }
#pragma unroll
    for (int u = 0; u < kStageFloats / kWave; ++u) sh.ysc[wave][lane + kWave * u] = 0.0f;
    fill_f64_tables(sh.f64tab, tid);
    if (lane == 0) {
      sh.ip0[wave] = ip;
      sh.s0[wave] = (uint32_t)a;
    }
    __syncthreads();
  }

  // One pass over the gene's cells at (beta in LDS `bsrc`, log w `lwsrc`),
  // leaving per-subject totals in wk.tsub and per-gene totals in wk.tgene.
  // The dense and the tabled passes are separate out-of-line functions, so
  // that neither's registers are allocated around a call to the other.
  template <class CT, int MODE>
  __device__ __forceinline__ void sweep(const double* bsrc, const double* lwsrc) {
    NH_FIT_LDS
    if (dd.tabled)
      sweep_tabled<CT, MODE>(bsrc, lwsrc);
    else
      sweep_dense<CT, MODE>(bsrc, lwsrc);
  }

  template <class CT, int MODE>
  __device__ __noinline__ void sweep_dense(const double* bsrc, const double* lwsrc) {
    NH_FIT_LDS
    const CT* rows = sizeof(CT) == 4 ? (const CT*)dd.row32 : (const CT*)dd.row64;
    const CT* gy = sizeof(CT) == 4 ? (const CT*)dd.gy32 : (const CT*)dd.gy64;
    sweep_wave<CT, MODE, NB>(rows, dd.gcell, gy, dd.fid, bsrc, lwsrc,
                             (typename Lds<CT>::ptr)(&sh.ysc[wave][0]),
                             (const lds_double*)&sh.f64tab[0], wk.psub, wk.pgene, sh.ip0[wave],
                             g1, (int)sh.s0[wave], k, block_begin(wave), block_end(wave), icpt,
                             gamma, lane, wave);
    sweep_totals<CT, MODE>();
  }

  // The pass on a design with zero-count tables: always FP64 (the table
  // lookups and the positive-count increments), counted as a sweep of CT.
  template <class CT, int MODE>
  __device__ __noinline__ void sweep_tabled(const double* bsrc, const double* lwsrc) {
    NH_FIT_LDS
    sweep_wave_tabled<MODE, NB>(dd.row64, dd.gcell, dd.gy64, dd.fid, dd.zgroup, dd.zsg, dd.zloose,
                                dd.zsl, dd.zindex, dd.zcoef, bsrc, lwsrc,
                                (const lds_double*)&sh.f64tab[0], wk.psub, wk.pgene,
                                sh.ip0[wave], g1, (int)sh.s0[wave], k, block_begin(wave),
                                block_end(wave), icpt, gamma, lane, wave);
    sweep_totals<CT, MODE>();
  }

  // Cold start of log w (dense designs): at the start's beta, each subject's
  // w solves the Poisson limit of its stationarity condition,
  // Y_s + alpha = (M_s + lambda) w_s with M_s = sum of exp(eta) over its
  // cells, the exact optimum when every cell's exp(eta) is small against
  // gamma; the Newton loop starts there instead of at log w = 0.
  __device__ __noinline__ void mean_start() {
    NH_FIT_LDS
    sweep_wave_mean<NB>(dd.row32, dd.fid, sh.beta, wk.psub, (int)sh.s0[wave], k,
                        block_begin(wave), block_end(wave), icpt, lane, wave);
    __syncthreads();
    const double* tot_y = dd.gtot + (size_t)gene * k;
    for (int s = tid; s < k; s += NT) {
      const uint32_t f0 = dd.fid[s], f1 = dd.fid[s + 1];
      const int w0 = (int)(f0 / blk), w1 = (int)((f1 - 1) / blk);
      double m = 0.0;
      for (int w = w0; w <= w1; ++w) m += wk.psub[(size_t)(s + w) * NSL];
      const double v = log((tot_y[s] + alpha) / (m + lambda));
      wk.lw[s] = isfinite(v) ? v : 0.0;
    }
    if (tid == 0) atomicAdd(&counters->cells_f32, (unsigned long long)dd.n_cells);
    __syncthreads();
  }

  // Per-subject totals from the waves' partials, waves added in order;
  // per-gene totals likewise; the sweep counters.
  template <class CT, int MODE>
  __device__ __forceinline__ void sweep_totals() {
    NH_FIT_LDS
    constexpr int NSV = MODE == SW_LAPLACE ? 4 : (MODE == SW_FINISH ? D::NF : NS);
    constexpr int NGV = MODE == SW_LAPLACE ? 0 : (MODE == SW_FINISH ? D::NGF : NG);
    __syncthreads();
    for (int idx = tid; idx < k * NSV; idx += NT) {
      const int ss = idx / NSV, v = idx % NSV;
      const uint32_t f0 = dd.fid[ss], f1 = dd.fid[ss + 1];
      const int w0 = (int)(f0 / blk), w1 = (int)((f1 - 1) / blk);
      double t = 0.0;
      for (int w = w0; w <= w1; ++w) t += wk.psub[(size_t)(ss + w) * NSL + v];
      wk.tsub[(size_t)ss * NSL + v] = t;
    }
    if constexpr (NGV > 0) {
      for (int v = tid; v < NGV; v += NT) {
        double t = 0.0;
        for (int w = 0; w < WAVES; ++w) t += wk.pgene[(size_t)w * NG + v];
        wk.tgene[v] = t;
      }
    }
    if (tid == 0) {
      if (sizeof(CT) == 4) {
        atomicAdd(&counters->sweeps_f32, 1ull);
        atomicAdd(&counters->cells_f32, (unsigned long long)dd.n_cells);
        atomicAdd(&counters->nnz_f32, (unsigned long long)(g1 - g0));
      } else {
        atomicAdd(&counters->sweeps_f64, 1ull);
        atomicAdd(&counters->cells_f64, (unsigned long long)dd.n_cells);
        atomicAdd(&counters->nnz_f64, (unsigned long long)(g1 - g0));
      }
    }
    __syncthreads();
  }

  // Sum over subjects of f(s), one wave per quantity: lane l takes subjects
  // l, l + 64, ...; the wave's fixed butterfly adds the lanes.
  template <class F>
  __device__ double subject_sum(F f) const {
    NH_FIT_LDS
    double acc = 0.0;
    for (int s = lane; s < k; s += kWave) acc += f(s);
    return wave_sum(acc);
  }

  // Gradient, curvature blocks and penalised log-likelihood of side `sd` from
  // the sweep totals (pml.rs:1564-1589), log w in `lwsrc`.  hvb null: every
  // block from this sweep's sums.  Otherwise the sweep was an SW_FINISH sweep,
  // and the blocks over the stored columns (each subject's vwb[j], at
  // hvwb[s * hstride + j], and the fixed-effect block hvb between stored
  // columns and between them and the intercept) are taken from there; the
  // gradient, the weights (vw, the intercept's blocks) and the log-likelihood
  // are this sweep's.
  __device__ __noinline__ void finalize(int sd, const double* lwsrc,
                                        const double* hvwb = nullptr, int hstride = 0,
                                        const double* hvb = nullptr) {
    NH_FIT_LDS
    const double* tot_y = dd.gtot + (size_t)gene * k;
    double* sub = side(sd);
    const double* hsub = hvwb;
    for (int s = tid; s < k; s += NT) {
      const double* t = wk.tsub + (size_t)s * NSL;
      const double lws = lwsrc[s];
      const double w = fast_exp(lws, tab());
      sub[s * SUB + 0] = t[0] + (alpha - lambda * w);
      sub[s * SUB + 1] = gamma * t[1] + lambda * w;
      sub[s * SUB + 2] = w;
      sub[s * SUB + 3 + icpt] = gamma * t[1];
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int j = qcol(q, icpt);
        sub[s * SUB + 3 + j] = hsub ? hsub[s * hstride + j] : gamma * t[2 + q];
      }
      wk.llsub[s] = lws * tot_y[s] + (alpha * lws - lambda * w);
    }
    __syncthreads();
    // quantities summed over subjects: 0 resid, 1 weight, 2..1+Q first, 2+Q ll
    for (int e = wave; e < Q + 3; e += WAVES) {
      double v = 0.0;
      if (e < 2)
        v = subject_sum([&](int s) { return wk.tsub[(size_t)s * NSL + e]; });
      else if (e < 2 + Q) {
        if (!hvb) v = subject_sum([&](int s) { return wk.tsub[(size_t)s * NSL + e]; });
      } else
        v = subject_sum([&](int s) { return wk.llsub[s]; });
      if (lane == 0) {
        if (e == 0) {
          sh.db[sd][icpt] = v;
        } else if (e == 1) {
          sh.vb[sd][icpt * NB + icpt] = gamma * v;
        } else if (e < 2 + Q) {
          const int j = qcol(e - 2, icpt);
          const double h = hvb ? hvb[icpt * NB + j] : gamma * v;
          sh.vb[sd][icpt * NB + j] = h;
          sh.vb[sd][j * NB + icpt] = h;
        } else {
          sh.ll[sd] = wk.tgene[0] + v;
        }
      }
    }
    for (int idx = tid; idx < Q + QT; idx += NT) {
      if (idx < Q) {
        sh.db[sd][qcol(idx, icpt)] = wk.tgene[1 + idx];
      } else {
        const int tq = idx - Q;
        int a = 0, r = tq;
        while (r >= Q - a) {
          r -= Q - a;
          ++a;
        }
        const int b = a + r;
        const int ja = qcol(a, icpt), jb = qcol(b, icpt);
        const double v = hvb ? hvb[ja * NB + jb] : gamma * wk.tgene[1 + Q + tq];
        sh.vb[sd][ja * NB + jb] = v;
        sh.vb[sd][jb * NB + ja] = v;
      }
    }
    __syncthreads();
  }

  // Schur complement of side sd into sh.vb2 (and the right-hand side into
  // sh.rhs when `with_rhs`), one wave per entry.
  __device__ __noinline__ void schur(int sd, bool with_rhs) {
    NH_FIT_LDS
    const double* sub = side(sd);
    const int n_e = NTRI + (with_rhs ? NB : 0);
    for (int e = wave; e < n_e; e += WAVES) {
      if (e < NTRI) {
        int a = 0, r = e;
        while (r >= NB - a) {
          r -= NB - a;
          ++a;
        }
        const int b = a + r;
        const double acc = subject_sum([&](int s) {
          return sub[s * SUB + 3 + a] * sub[s * SUB + 3 + b] / sub[s * SUB + 1];
        });
        if (lane == 0) {
          const double v = sh.vb[sd][a * NB + b] - acc;
          sh.vb2[a * NB + b] = v;
          sh.vb2[b * NB + a] = v;
        }
      } else {
        const int j = e - NTRI;
        const double acc = subject_sum([&](int s) {
          return sub[s * SUB + 3 + j] * (sub[s * SUB + 0] / sub[s * SUB + 1]);
        });
        if (lane == 0) sh.rhs[j] = sh.db[sd][j] - acc;
      }
    }
    __syncthreads();
  }

  // Newton step from side `sd` (pml.rs:1664-1710): Schur complement, LDL'
  // solve, per-subject steps; writes sh.new_beta and wk.new_lw.
  __device__ __noinline__ void newton_step(int sd) {
    NH_FIT_LDS
    const double* sub = side(sd);
    schur(sd, true);
    if (tid == 0) {
      for (int i = 0; i < NB * NB; ++i) sh.lm[i] = sh.vb2[i];
      for (int i = 0; i < NB; ++i) sh.lb[i] = sh.rhs[i];
      ldlt_solve_lds<NB>(sh.lm, sh.lb, sh.perm);
      for (int i = 0; i < NB; ++i) {
        sh.step_beta[i] = sh.lb[i];
        sh.new_beta[i] = sh.beta[i] + sh.lb[i];
        sh.damp_beta[i] = 1.0;
      }
    }
    __syncthreads();
    for (int s = tid; s < k; s += NT) {
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) acc += sub[s * SUB + 3 + j] * sh.step_beta[j];
      const double dwvw = sub[s * SUB + 0] / sub[s * SUB + 1];
      const double stp = dwvw - acc / sub[s * SUB + 1];
      wk.step_lw[s] = stp;
      wk.new_lw[s] = wk.lw[s] + stp;
      wk.damp_lw[s] = 1.0;
      wk.vw_step[s] = sub[s * SUB + 1];
    }
    __syncthreads();
  }

  // Newton decrement of the step newton_step(sd) just wrote: 0.5 g'D with g
  // the gradient at side sd and D the full step, the log-likelihood gain the
  // quadratic model predicts (the step solves I D = g, I positive definite).
  // NaN when a coordinate moves by more than 0.1, where the model is not
  // trusted.
  __device__ __noinline__ double decrement(int sd) {
    NH_FIT_LDS
    const double* sub = side(sd);
    if (wave == 0) {
      double acc = 0.0, big = 0.0;
      for (int s = lane; s < k; s += kWave) {
        acc += sub[s * SUB + 0] * wk.step_lw[s];
        big = fmax(big, fabs(wk.step_lw[s]));
      }
      acc = wave_sum(acc);
      big = wave_max(big);
      if (lane == 0) {
        double g = acc;
        for (int j = 0; j < NB; ++j) {
          g += sh.db[sd][j] * sh.step_beta[j];
          big = fmax(big, fabs(sh.step_beta[j]));
        }
        sh.scal[4] = big <= 0.1 ? 0.5 * g : NAN;
      }
    }
    __syncthreads();
    return sh.scal[4];
  }

  // Per-coordinate damping of the trial point (pml.rs:1729-1755).
  __device__ __noinline__ void damp(double min_step) {
    NH_FIT_LDS
    const double cut = 40.0;
    if (tid < NB) {
      const int j = tid;
      const double d = sh.step_beta[j];
      if (d < cut && d > -cut) {
        sh.damp_beta[j] /= 2.0;
        sh.new_beta[j] = sh.beta[j] + d * sh.damp_beta[j];
      } else if (d > 0.0) {
        sh.new_beta[j] = sh.beta[j] + min_step;
      } else {
        sh.new_beta[j] = sh.beta[j] - min_step;
      }
    }
    for (int s = tid; s < k; s += NT) {
      const double d = wk.step_lw[s];
      if (d < cut && d > -cut) {
        wk.damp_lw[s] /= 2.0;
        wk.new_lw[s] = wk.lw[s] + d * wk.damp_lw[s];
      } else if (d > 0.0) {
        wk.new_lw[s] = wk.lw[s] + min_step;
      } else {
        wk.new_lw[s] = wk.lw[s] - min_step;
      }
    }
    __syncthreads();
  }

  __device__ __noinline__ void accept() {
    NH_FIT_LDS
    if (tid < NB) sh.beta[tid] = sh.new_beta[tid];
    for (int s = tid; s < k; s += NT) wk.lw[s] = wk.new_lw[s];
    __syncthreads();
  }

  // max |gradient| over db and dw of side sd (pml.rs:1720-1724)
  __device__ __noinline__ double max_grad(int sd) {
    NH_FIT_LDS
    if (tid == 0) {
      double m = 0.0;
      for (int j = 0; j < NB; ++j) m = fmax(m, fabs(sh.db[sd][j]));
      double mw = 0.0;
      for (int s = 0; s < k; ++s) mw = fmax(mw, fabs(side(sd)[s * SUB + 0]));
      sh.scal[0] = fmax(m, mw) > 0.01 ? 1.0 : 0.0;
    }
    __syncthreads();
    return sh.scal[0];
  }

  // The Newton loop of optimise (pml.rs:1650-1766) with sweeps in CT.  In
  // f64 (noise_rel == 0) this is opt_pml's control flow exactly.  In f32 the
  // comparisons allow the resolution floor noise_rel * |ll|.  Starts by
  // sweeping the current point into side 0.  Returns the side holding the
  // final point; out gets ll, ll_prev, iterations, backtracks.
  template <class CT>
  __device__ int newton_loop(const FitKnobs& kn, double noise_rel, FitResult& out) {
    NH_FIT_LDS
    int cur = 0;
    sweep<CT, SW_FULL>(sh.beta, wk.lw);
    finalize(cur, wk.lw);
    double ll = sh.ll[cur], ll_prev = 0.0, likdif = 0.0;
    int step = 0, backtracks = 0;
    while (step == 0 ||
           (likdif > (noise_rel > 0.0 ? fmax(kn.eps, noise_rel * fabs(ll)) : kn.eps) &&
            step < kn.max_iter)) {
      ++step;
      newton_step(cur);
      const int tr = 1 - cur;
      const bool fast = noise_rel > 0.0 && kn.decrement_stop;
      double dec = NAN;
      if (fast) {
        // f32: a step predicted to gain less than the resolution floor is
        // taken without the sweep that would measure it; the f64 finish
        // sweeps the stepped point next.
        dec = decrement(cur);
        if (dec >= 0.0 && dec <= fmax(kn.eps, noise_rel * fabs(ll))) {
          accept();
          ll_prev = ll;
          break;
        }
      }
      ll_prev = ll;
      sweep<CT, SW_FULL>(sh.new_beta, wk.new_lw);
      finalize(tr, wk.new_lw);
      ll = sh.ll[tr];
      likdif = ll - ll_prev;
      const double noise = noise_rel * fabs(ll_prev);
      if (fast && likdif < -noise && isfinite(ll) && dec >= 0.0 &&
          dec <= kF32Trust * fabs(ll_prev) && likdif >= -kF32Trust * fabs(ll_prev)) {
        // A step predicted to gain at most kF32Trust |ll| that measures as a
        // loss no larger than that: the difference is below what FP32 sums
        // resolve on this gene, not an overshoot.  Halving the step would
        // only repeat the comparison; keep the step and let the f64 finish
        // decide from there.
        accept();
        cur = tr;
        break;
      }
      backtracks = 0;
      double min_step = 40.0;
      while (likdif < -noise || isinf(ll) || (noise_rel > 0.0 && isnan(ll))) {
        ++backtracks;
        min_step /= 2.0;
        if (backtracks > kn.max_backtrack) {
          likdif = 0.0;
          ll = ll_prev;
          if (max_grad(cur) > 0.0) ++backtracks;
          break;
        }
        damp(min_step);
        sweep<CT, SW_FULL>(sh.new_beta, wk.new_lw);
        finalize(tr, wk.new_lw);
        ll = sh.ll[tr];
        likdif = ll - ll_prev;
      }
      accept();
      cur = tr;
    }
    out.ll = ll;
    out.ll_prev = ll_prev;
    out.iterations = step;
    out.backtracks = backtracks;
    return cur;
  }

  // sum_s log |vw_s| at the point newton_step(sd) stepped to, from side sd's
  // vw and an SW_FINISH sweep's eta-derivatives of the weight sums.
  __device__ __noinline__ double stepped_log_det(int sd) {
    NH_FIT_LDS
    const double* sub = side(sd);
    if (wave == 0) {
      const double acc = subject_sum([&](int s) {
        const double* t = wk.tsub + (size_t)s * NSL;
        double dw = t[2] * (wk.step_lw[s] + sh.step_beta[icpt]);
#pragma unroll
        for (int q = 0; q < Q; ++q) dw += t[4 + q] * sh.step_beta[qcol(q, icpt)];
        const double w1 = fast_exp(wk.lw[s] + wk.step_lw[s], tab());
        const double vw = sub[s * SUB + 1] + gamma * dw + lambda * (w1 - sub[s * SUB + 2]);
        return fast_log(fabs(vw), tab());
      });
      if (lane == 0) sh.scal[5] = acc;
    }
    __syncthreads();
    return sh.scal[5];
  }

  // FP64 finish from the current point (newton_finish, pml.rs:807-1090):
  // Newton steps without damping, log-likelihood and log-determinant at the
  // stepped point.  `cur`: the side the first sweep goes to; with the
  // decrement stop the stored-column curvature blocks of the first step come
  // from hvwb (stride hstride) and hvb (see finalize).  Returns false when a
  // step worsens the objective beyond rounding or leaves it not finite, or
  // the budget runs out.
  __device__ bool finish(const FitKnobs& kn, FitResult& out, int cur, const double* hvwb,
                         int hstride, const double* hvb) {
    NH_FIT_LDS
    if (kn.decrement_stop) {
      sweep<double, SW_FINISH>(sh.beta, wk.lw);
      finalize(cur, wk.lw, hvwb, hstride, hvb);
    } else {
      sweep<double, SW_FULL>(sh.beta, wk.lw);
      finalize(cur, wk.lw);
    }
    for (int it = 1; it <= kn.max_iter; ++it) {
      const double ll_prev = sh.ll[cur];
      newton_step(cur);
      const int tr = 1 - cur;
      if (kn.decrement_stop && it == 1) {
        // A first step predicted to gain at most eps ends the fit as the
        // measured rule below would after sweeping the stepped point: the
        // log-likelihood there is ll + decrement (the quadratic model, exact
        // to third order in the step) and each subject's vw moves to first
        // order along d eta = step_lw + x . step_beta.
        const double dec = decrement(cur);
        if (dec >= 0.0 && dec <= kn.eps) {
          const double ld = stepped_log_det(cur);
          accept();
          out.ll = ll_prev + dec;
          out.ll_prev = ll_prev;
          out.log_det = ld;
          out.finish_steps = it;
          out.second = 0.0;
          out.side = cur;
          out.predictable = 1;
          return true;  // the information at the start point stays in vb2
        }
      }
      sweep<double, SW_FULL>(sh.new_beta, wk.new_lw);
      finalize(tr, wk.new_lw);
      const double ll = sh.ll[tr];
      const double likdif = ll - ll_prev;
      if (isnan(likdif) || likdif < -1e-12 * fabs(ll)) return false;
      accept();
      cur = tr;
      if (likdif <= kn.eps) {
        // log-det at the stepped point
        if (tid == 0) {
          double ld = 0.0;
          for (int s = 0; s < k; ++s) ld += log(fabs(side(cur)[s * SUB + 1]));
          sh.scal[1] = ld;
        }
        __syncthreads();
        out.ll = ll;
        out.ll_prev = ll_prev;
        out.log_det = sh.scal[1];
        out.finish_steps = it;
        out.second = 0.0;
        // information at the final point for the reply
        schur(cur, false);
        return true;
      }
    }
    return false;
  }

  // First-order change of the optimum (beta, log w) with the variance
  // components (implicit function theorem: the gradient g stays zero, so
  // d x* / d theta = I^-1 d g / d theta with I the information), at the point
  // whose curvature blocks side `sd` holds and whose SW_FINISH sums are in the
  // totals.  Column 0 per unit of `subject` (through alpha and lambda, whose
  // derivatives are closed forms), column 1 per unit of gamma (the sweep's
  // d resid / d gamma sums).  jb: 2 x NB (LDS), jl: 2 x k.  The same Schur
  // complement and pivoted LDL' solve as newton_step.
  __device__ __noinline__ void sensitivity(int sd, double subject, double* jb, double* jl) {
    NH_FIT_LDS
    const double* sub = side(sd);
    const double exps = exp(subject), em1 = exps - 1.0;
    const double dalpha = -exps / (em1 * em1);
    const double dlambda = -lambda * (0.5 + exps / em1);
    for (int s = tid; s < k; s += NT) {
      wk.step_lw[s] = dalpha - dlambda * sub[s * SUB + 2];
      wk.damp_lw[s] = wk.tsub[(size_t)s * NSL + 3];
    }
    schur(sd, false);  // ends with a barrier
    for (int e = wave; e < 2 * NB; e += WAVES) {
      const int col = e / NB, j = e % NB;
      const double* r = col == 0 ? wk.step_lw : wk.damp_lw;
      const double acc =
          subject_sum([&](int s) { return sub[s * SUB + 3 + j] * (r[s] / sub[s * SUB + 1]); });
      double rb = 0.0;
      if (col == 1) {
        if (j == icpt)
          rb = subject_sum([&](int s) { return wk.damp_lw[s]; });
        else
          rb = wk.tgene[1 + Q + (j < icpt ? j : j - 1)];
      }
      if (lane == 0) sh.rhs2[col][j] = rb - acc;
    }
    __syncthreads();
    if (tid == 0) {
      for (int col = 0; col < 2; ++col) {
        for (int i = 0; i < NB * NB; ++i) sh.lm[i] = sh.vb2[i];
        for (int i = 0; i < NB; ++i) sh.lb[i] = sh.rhs2[col][i];
        ldlt_solve_lds<NB>(sh.lm, sh.lb, sh.perm);
        for (int i = 0; i < NB; ++i) jb[col * NB + i] = sh.lb[i];
      }
    }
    __syncthreads();
    for (int idx = tid; idx < 2 * k; idx += NT) {
      const int col = idx / k, s = idx % k;
      const double* r = col == 0 ? wk.step_lw : wk.damp_lw;
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < NB; ++j) acc += sub[s * SUB + 3 + j] * jb[col * NB + j];
      jl[idx] = (r[s] - acc) / sub[s * SUB + 1];
    }
    __syncthreads();
  }

  // log-det from the last Newton step's vw (pml.rs:1767-1771)
  __device__ __noinline__ double log_det_step() {
    NH_FIT_LDS
    if (tid == 0) {
      double ld = 0.0;
      for (int s = 0; s < k; ++s) ld += log(fabs(wk.vw_step[s]));
      sh.scal[2] = ld;
    }
    __syncthreads();
    return sh.scal[2];
  }

  // Laplace correction at the current point (pml.rs:1828-1919), f64.
  __device__ __noinline__ double laplace(unsigned ord) {
    NH_FIT_LDS
    sweep<double, SW_LAPLACE>(sh.beta, wk.lw);
    if (tid == 0) {
      const bool high = ord > 2;
      double third_acc = 0.0, fourth_acc = 0.0, fifth_acc = 0.0;
      for (int s = 0; s < k; ++s) {
        const double* t = wk.tsub + (size_t)s * NSL;
        const double prior = lambda * exp(wk.lw[s]);
        const double vw = gamma * t[0] + prior;
        const double third = gamma * t[1] + prior;
        third_acc += third * third / (vw * vw * vw);
        if (high) {
          const double fourth = gamma * t[2] + prior;
          const double fifth = gamma * t[3] + prior;
          const double vs = vw * vw;
          fourth_acc += fourth / vs;
          fifth_acc += fifth * third / (vs * vs);
        }
      }
      double second = 5.0 * third_acc / 24.0;
      if (high) {
        second -= fourth_acc / 8.0;
        second += 7.0 * fifth_acc / 48.0;
      }
      sh.scal[3] = second;
    }
    __syncthreads();
    return sh.scal[3];
  }

  // The whole fit.  mode 0: f32 Newton then f64 finish (stage two); mode 1:
  // opt_pml in f64 from the given point (stage three).  beta in sh.beta and
  // log w in wk.lw on entry.  same_gene: this workgroup's previous fit was of
  // the same gene, so setup()'s per-wave CSR cursors, the zeroed count rows
  // (every sweep leaves them zero) and the FP64 tables are still in LDS.
  // cold: a stage-two fit with no earlier fit of the gene to start from; on a
  // dense design its log w starts at mean_start()'s values.  near_vwb /
  // near_vb (stage two, decrement stop): the start was predicted to first
  // order from a fit whose stored-column curvature blocks these are (vwb
  // per subject, stride NB); the FP32 loop is skipped and the FP64 finish
  // takes its first step with those blocks -- a start that close to the
  // optimum passes the finish's decrement test, and one that does not
  // continues with the finish's measured FP64 steps.
  __device__ FitResult run(int mode, unsigned ord, const FitKnobs& kn, bool same_gene = false,
                           bool cold = false, const double* near_vwb = nullptr,
                           const double* near_vb = nullptr) {
    NH_FIT_LDS
    FitResult r{};
    r.finish_steps = 0;
    r.exact = 0;
    if (!same_gene) setup();
    if (cold && mode == 0 && !dd.tabled) mean_start();
    bool done = false;
    if (mode == 0 && near_vb && kn.decrement_stop) {
      done = finish(kn, r, 0, near_vwb, NB, near_vb);
    } else if (mode == 0) {
      FitResult f32r{};
      const int c32 = newton_loop<float>(kn, kn.noise_rel, f32r);
      r.iterations = f32r.iterations;
      r.backtracks = f32r.backtracks;
      done = finish(kn, r, 1 - c32, side(c32) + 3, SUB, sh.vb[c32]);
    }
    if (!done) {
      FitResult e{};
      const int cur = newton_loop<double>(kn, 0.0, e);
      r.ll = e.ll;
      r.ll_prev = e.ll_prev;
      r.iterations += e.iterations;
      r.backtracks = e.backtracks;
      r.log_det = log_det_step();
      r.exact = 1;
      (void)cur;
    }
    r.second = ord > 1 ? laplace(ord) : 0.0;
    // the Laplace sweep overwrote the totals; tabled sweeps leave the
    // gamma-derivatives out
    if (ord > 1 || dd.tabled) r.predictable = 0;
    return r;
  }
};

}  // namespace nh

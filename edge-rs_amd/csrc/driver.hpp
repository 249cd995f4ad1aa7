// Per-run host state shared by the driver's stages.
#pragma once

#include "zero_tables.hpp"

#include <cstdint>
#include <vector>

#include "../../include/nebula_hip.h"
#include "host_common.hpp"

namespace nh {

struct Resident;

struct GeneInfo {
  double mean_count = 0.0;     // nebula's mct
  std::vector<double> totals;  // per-subject count totals
  std::vector<double> xty;     // sum over positive cells of x * y (centred design;
                               // filled on the device at the start of stage one)
  CountTail tail;              // lgamma tail histogram and n1 / n2
};

struct GenePlan {
  unsigned ord = 1;
  double sigma = 0, gamma = 0;  // stage one's variance components
  int convergence = 1;
  std::vector<double> beta_start;  // stage three's start (the reference's)
  std::vector<double> beta_one;    // stage two's cold fits: stage one's fixed effects
  int refit = 1;  // 0 none, 1 both, 2 subject only
  bool searched = false, found = false, failed = false;
  double v[2] = {0, 0};  // stage-two minimiser
  int evaluations = 0;
  double objective = 0;
  bool warm_rejected = false;  // learned start fell back to the cold search
};

struct Prepared {
  size_t n_genes = 0, n_cells = 0, n_coef = 0, k = 0, intercept = 0;
  nebula_params_t params;
  std::vector<size_t> fid;
  std::vector<double> log_offset, centred, sds;
  double log_mean_offset = 0, cv2 = 0, cells_per_subject = 0;
  int method = NH_METHOD_HL;
  bool tables = false;  // zero-count tables apply (and are in zt)
  ZeroTables zt;
  double t_tables = 0.0;  // seconds spent building them
  std::vector<size_t> cell_columns, kept;
  // kept genes' positive counts, CSR over kept genes
  std::vector<uint64_t> gene_ptr;
  RawVec<uint32_t> cells;
  RawVec<double> counts;
  std::vector<GeneInfo> genes;
  std::vector<GenePlan> plans;
  std::vector<float> features;  // kept x 12 warm-start features (after stage one)
};

// check_convergence (src/sc/pml.rs:1111-1139) on a device fit's reply:
// initial is the outer search's code (NH_CONV_SUCCESS when it succeeded),
// at_bound whether a variance component sits on its bound; information is
// the nb x nb information matrix of the fit.
struct DevReply;
constexpr int kFitMaxIter = 50, kFitMaxBacktrack = 10;
int fit_convergence(const DevReply& r, const double* information, size_t nb, bool at_bound,
                    int initial);

// Learned-start settings of one stage-two run.
struct LearnedStart {
  const float* params = nullptr;  // device copy of the MLP parameter blob
  double step = 0.01;             // BOBYQA initial step, relative to the predicted point
  double xtol_rel = 3e-2;         // rhoend / rhobeg of the warm search (rhoend 3e-4 x)
};

double digamma(double x);
double trigamma(double x);
void compute_features(Prepared& P);
void run_stage_two_batched(Prepared& P, Resident& R, const nebula_hip_options_t& o,
                           nebula_run_stats_t& st, const LearnedStart* learned);

}  // namespace nh

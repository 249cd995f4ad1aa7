/*
 * nebula_hip: NEBULA-HL stage two on AMD Instinct MI355X (gfx950), behind a
 * C ABI shaped like edge-rs's `nebula_sparse` / `nebula_sparse_gpu`.
 *
 * Every array is a plain pointer plus a length; nothing here names a torch or
 * HIP type, so a Rust `-sys` crate, ctypes or cgo binds it one to one.
 *
 * Reference interfaces replaced (edge-rs 0.4.0):
 *   nebula_params_t    <- NebulaParams           src/sc/nebula.rs:207-258
 *   nebula_fit_t       <- NebulaFit              src/sc/nebula.rs:301-344
 *   nebula_sparse_hip  <- nebula_sparse_gpu      src/gpu/stage_two.rs:233-265
 *                         (same inputs as nebula_sparse, src/sc/nebula.rs:434-445)
 *   nh_resident_*      <- ResidentBatch upload / submit / collect
 *                                                src/gpu/nebula_gpu.rs:244,466,605
 *   nh_request_t       <- PmlRequest             src/gpu/nebula_gpu.rs:66
 *   nh_reply_t         <- PmlReply               src/gpu/nebula_gpu.rs:86
 *   NH_ERR_*           <- EdgeErrors variants on this path, src/errors.rs:13-223
 *
 * Threading: calls on distinct handles are independent.  One call of
 * nebula_sparse_hip owns its device buffers and stream for its duration.
 */
#ifndef NEBULA_HIP_H
#define NEBULA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes: one per EdgeErrors variant the path can raise. */
enum {
  NH_OK = 0,
  NH_ERR_EMPTY_COUNTS = 1,            /* EdgeErrors::EmptyCounts */
  NH_ERR_INVALID_ARGUMENT = 2,        /* EdgeErrors::InvalidArgument */
  NH_ERR_MUST_BE_POSITIVE = 3,        /* EdgeErrors::MustBePositive */
  NH_ERR_LENGTH_MISMATCH = 4,         /* EdgeErrors::LengthMismatch */
  NH_ERR_SHAPE_MISMATCH = 5,          /* EdgeErrors::ShapeMismatch */
  NH_ERR_MALFORMED_SPARSE = 6,        /* EdgeErrors::MalformedSparse */
  NH_ERR_SUBJECTS_NOT_CONTIGUOUS = 7, /* EdgeErrors::SubjectsNotContiguous */
  NH_ERR_TOO_FEW_SUBJECTS = 8,        /* EdgeErrors::TooFewSubjects */
  NH_ERR_MISSING_INTERCEPT = 9,       /* EdgeErrors::MissingIntercept */
  NH_ERR_NO_GENES_AFTER_FILTERING = 10, /* EdgeErrors::NoGenesAfterFiltering */
  NH_ERR_INVALID_BOUNDS = 11,         /* EdgeErrors::InvalidBounds */
  NH_ERR_INVALID_DISPERSION = 12,     /* EdgeErrors::InvalidDispersion */
  NH_ERR_GPU = 13,                    /* EdgeErrors::Gpu */
  NH_ERR_COEF_OUT_OF_RANGE = 14       /* EdgeErrors::CoefOutOfRange (glm_sc_test) */
};

/* NEBULA method, NebulaMethod (src/sc/nebula.rs:195-205). */
enum { NH_METHOD_LN = 0, NH_METHOD_HL = 1 };

/* Convergence codes (src/sc/pml.rs:89-111, src/sc/nebula.rs:187). */
enum {
  NH_CONV_SUCCESS = 1,
  NH_CONV_CRITICAL_POINT = -10,
  NH_CONV_MAX_ITER = -20,
  NH_CONV_SINGULAR = -25,
  NH_CONV_NOT_FINITE = -30,
  NH_CONV_STALLED = -40,
  NH_CONV_OUTER_FAILED = -50,
  NH_CONV_AT_BOUND = -60
};

/* NebulaParams (src/sc/nebula.rs:207-258); defaults from
 * nebula_params_default() are the R package's (:260-275). */
typedef struct {
  double min_subject; /* min.0, lower bound on sigma^2 (1e-4) */
  double min_cell;    /* min.1, lower bound on the NB size (1e-4) */
  double max_subject; /* max.0 (10) */
  double max_cell;    /* max.1 (1000) */
  int32_t method;     /* NH_METHOD_LN (default) or NH_METHOD_HL */
  int32_t reml;       /* 0 (default) */
  double cutoff_cell; /* 20 */
  double kappa;       /* 800 */
  double cpc;         /* 0.005 */
  uint64_t mincp;     /* 5 */
  uint64_t min_subjects; /* 0 */
  double eps;         /* 1e-6 */
} nebula_params_t;

/* NebulaFit (src/sc/nebula.rs:301-344).  Arrays are owned by the library and
 * released with nebula_fit_free.  Row-major over kept genes. */
typedef struct {
  size_t n_genes;                  /* genes that survived the filter */
  size_t n_coef;
  double* coefficients;            /* n_genes * n_coef */
  double* covariance;              /* n_genes * n_coef(n_coef+1)/2, packed
                                      upper triangle: (i,j), i<=j, at j(j+1)/2+i */
  double* se;                      /* n_genes * n_coef */
  double* subject_overdispersion;  /* n_genes: sigma^2 */
  double* cell_overdispersion;     /* n_genes: phi = 1 / NB size */
  int32_t* convergence;            /* n_genes */
  uint8_t* sigma_at_bound;         /* n_genes, 0/1 */
  uint64_t* gene_index;            /* n_genes, zero-based input rows */
} nebula_fit_t;

/* Which stage-two engine nebula_sparse_hip_ex runs. */
enum {
  NH_ENGINE_PLAIN = 0,   /* host BOBYQA, one fit launch per search round */
  NH_ENGINE_BATCHED = 1, /* persistent work-queue kernel, search on device */
  NH_ENGINE_LEARNED = 2  /* batched, each search started from the MLP's
                            prediction with the cold search as fallback */
};

/* Options beyond the reference's arguments (no counterpart in edge-rs). */
typedef struct {
  int32_t engine;        /* NH_ENGINE_* (default NH_ENGINE_BATCHED) */
  int32_t learned;       /* 1: MLP warm start for the search (needs weights) */
  const float* mlp_weights; /* packed MLP parameters, see nh_mlp_param_count */
  int32_t wg_size;       /* threads per fit workgroup, 0 = autotuned */
  int32_t device;        /* HIP device ordinal */
  int32_t host_threads;  /* 0 = hardware concurrency, at most 16 */
  double learned_step;   /* warm search's initial BOBYQA step relative to the
                            predicted point (0 = default 0.01) */
  double learned_xtol;   /* warm search's rhoend / rhobeg (0 = default 3e-2) */
  int32_t zero_tables;   /* where zero-count tables apply (src/sc/zeros.rs:219-262):
                            1 (default) the sweeps read them when their groups
                            average at least 4,096 cells (below that a table
                            lookup's dependent loads cost more on the device
                            than sweeping the group's cells); 2 always; 0 never */
} nebula_hip_options_t;

/* Per-phase timings and counters of one call (EDGE_RS_GPU_TIMING's phases,
 * src/gpu/stage_two.rs:90-194). */
typedef struct {
  double t_setup, t_stage_one, t_stage_two, t_stage_three, t_total; /* seconds */
  double t_upload;
  uint64_t genes_searched;   /* genes entering stage two */
  uint64_t evaluations;      /* objective evaluations, all genes */
  uint64_t rounds;           /* launches of the stage-two fit kernel */
  uint64_t sweeps_f32;       /* passes over a gene's cells, FP32 */
  uint64_t sweeps_f64;       /* passes over a gene's cells, FP64 */
  uint64_t cell_sweeps;      /* sum over FP32 sweeps of the gene's cell count */
  uint64_t nnz_sweeps;       /* sum over FP32 sweeps of its positive counts */
  uint64_t warm_fallbacks;   /* learned starts rejected by the log-lik rule */
  double fit_kernel_ms;      /* device time of the stage-two kernels (events) */
  uint64_t fit_kernel_launches;
  uint64_t cell_sweeps_f64;  /* the same over FP64 sweeps */
  uint64_t nnz_sweeps_f64;
  double t_plan;             /* seconds of t_stage_two spent building the run's plan
                                (search descriptors, count tails, features, buffers,
                                the captured launch); 0 when a session replays one */
  uint64_t zero_groups;      /* zero-count table groups the sweeps read (0: none) */
  double t_tables;           /* seconds of t_setup spent building the tables */
} nebula_run_stats_t;

void nebula_params_default(nebula_params_t* params);
void nebula_hip_options_default(nebula_hip_options_t* options);

/* nebula_sparse_gpu's call shape: CSR counts over (n_genes, n_cells) with u32
 * indptr (n_genes + 1) and indices, f64 values; subject id per cell (each
 * subject one contiguous run); design n_cells * n_coef row-major with exactly
 * one constant column; offset per cell on the linear scale (NULL for ones);
 * params NULL for defaults.  Returns NH_OK or an NH_ERR_* code with a UTF-8
 * message in err.  n_coef above 8 and reml are rejected, as on the reference
 * GPU path (src/gpu/stage_two.rs:242-246, src/gpu/pml_kernel.rs:1208-1212). */
int nebula_sparse_hip(size_t n_genes, size_t n_cells, const uint32_t* indptr,
                      const uint32_t* indices, const double* data,
                      const uint64_t* subject_id, const double* design,
                      size_t n_coef, const double* offset,
                      const nebula_params_t* params, nebula_fit_t* out,
                      char* err, size_t err_len);

/* Same, with engine options and run statistics (either may be NULL). */
int nebula_sparse_hip_ex(size_t n_genes, size_t n_cells, const uint32_t* indptr,
                         const uint32_t* indices, const double* data,
                         const uint64_t* subject_id, const double* design,
                         size_t n_coef, const double* offset,
                         const nebula_params_t* params,
                         const nebula_hip_options_t* options,
                         nebula_fit_t* out, nebula_run_stats_t* stats,
                         char* err, size_t err_len);

void nebula_fit_free(nebula_fit_t* fit);

/* ---------------------------------------------------------------------------
 * Session API: nebula_sparse_hip split at edge-rs's device seam
 * (nebula_sparse_with's `fit` closure, src/sc/nebula.rs:494-511, and
 * search_all, src/gpu/stage_two.rs:362).  create = validation, set-up,
 * upload and stage one; stage_two can then run (and be timed) repeatedly with
 * the inputs resident on the device; finish = stage three into a NebulaFit.
 * ------------------------------------------------------------------------- */
typedef struct nh_session nh_session_t;

int nh_session_create(size_t n_genes, size_t n_cells, const uint32_t* indptr,
                      const uint32_t* indices, const double* data,
                      const uint64_t* subject_id, const double* design,
                      size_t n_coef, const double* offset,
                      const nebula_params_t* params,
                      const nebula_hip_options_t* options, nh_session_t** out,
                      char* err, size_t err_len);
/* Genes that survived the filter (the rows of every per-gene output). */
size_t nh_session_genes(const nh_session_t* s);
/* Stage one's variance components (kept x 2: sigma, gamma) and Laplace order. */
int nh_session_stage_one(const nh_session_t* s, double* sigma_gamma, int32_t* ord);
/* The 12 warm-start features of every kept gene (kept x 12 floats):
 * log rate, positive fraction, log variance, log moment dispersion, 10/50/90 %
 * quantiles and sd of the per-subject log rates, log nnz, log cells per
 * subject, stage one's log sigma^2 and log phi. */
int nh_session_features(const nh_session_t* s, float* out);
/* Per kept gene, 1 where the last NH_ENGINE_LEARNED run rejected the warm
 * result (the cold start point did better) and ran the cold search. */
int nh_session_warm_rejected(const nh_session_t* s, uint8_t* out);
/* Installs MLP parameters (nh_mlp_param_count floats) for NH_ENGINE_LEARNED. */
int nh_session_set_mlp(nh_session_t* s, const float* params, char* err, size_t err_len);
/* Runs stage two with engine NH_ENGINE_*.  minimisers: kept x 2 (subject,
 * cell), objectives: kept (the profile objective at the minimiser),
 * evaluations: kept; any may be NULL. */
int nh_session_stage_two(nh_session_t* s, int32_t engine, double* minimisers,
                         double* objectives, int32_t* evaluations,
                         nebula_run_stats_t* stats, char* err, size_t err_len);
/* Autotuner: runs stage two with engine twice per workgroup size (256, 512,
 * 1024 threads per gene, i.e. 4, 8, 16 waves), keeps the fastest for later
 * calls of this session; ms (3 values) gets each size's faster stage-two
 * time (NaN for a size that launches the same kernel as a smaller one). */
int nh_session_autotune(nh_session_t* s, int32_t engine, int32_t* wg_size, double* ms,
                        char* err, size_t err_len);
int nh_session_finish(nh_session_t* s, nebula_fit_t* out, double* t_stage_three,
                      char* err, size_t err_len);
/* Frees the session's kept stage-two plans (one per launch shape of the
 * batched and learned engines: device buffers and captured launches); the
 * next run of a shape builds its plan again.  No counterpart in edge-rs,
 * whose GPU path stages every run. */
int nh_session_release_plans(nh_session_t* s);
void nh_session_free(nh_session_t* s);

/* ---------------------------------------------------------------------------
 * Request-level batch API: one penalised fit (opt_pml) per request, all on the
 * device.  Mirrors ResidentBatch (src/gpu/nebula_gpu.rs:195-648).
 * ------------------------------------------------------------------------- */

typedef struct nh_resident nh_resident_t;

/* One penalised fit at fixed variance components, PmlRequest
 * (src/gpu/nebula_gpu.rs:66-84). */
typedef struct {
  uint32_t gene;         /* resident gene index */
  uint32_t ord;          /* Laplace order 1..3 */
  double subject;        /* nebula's sigma[0] = log(1 + sigma^2) */
  double cell;           /* NB size gamma */
  const double* beta_init;  /* n_coef */
  const double* log_w_init; /* n_subjects, or NULL for zeros */
} nh_request_t;

/* PmlReply (src/gpu/nebula_gpu.rs:86-130) plus PmlResult's fields
 * (src/sc/pml.rs:288-330).  Arrays are caller-allocated views. */
typedef struct {
  double log_likelihood;
  double log_likelihood_prev;
  double log_det;
  double second_order;
  int32_t convergence;
  int32_t iterations;
  int32_t backtracks;
  int32_t finish_steps;   /* FP64 Newton steps after the FP32 solve */
  double* beta;           /* n_coef */
  double* log_w;          /* n_subjects */
  double* information;    /* n_coef * n_coef */
} nh_reply_t;

/* Precision policy of a batch solve. */
enum {
  NH_SOLVE_F32_FINISH = 0, /* FP32 Newton, then FP64 finish (stage two) */
  NH_SOLVE_F64 = 1         /* opt_pml in FP64 on the device (stage three) */
};

/* Uploads the shared design (n_cells * n_coef, already centred, f64, with an
 * intercept column that is 1 in every cell, else NH_ERR_INVALID_ARGUMENT), log
 * offsets, subject boundaries (n_subjects + 1) and per-gene positive counts
 * (CSR over genes: gene_ptr n_genes+1, cells, counts) once. */
int nh_resident_upload(size_t n_cells, size_t n_coef, size_t n_subjects,
                       const double* design, const double* log_offset,
                       const uint64_t* subject_start, size_t n_genes,
                       const uint64_t* gene_ptr, const uint32_t* cells,
                       const double* counts, int32_t device,
                       nh_resident_t** out, char* err, size_t err_len);

/* ResidentBatch::submit (src/gpu/nebula_gpu.rs:466-530): checks and uploads
 * n_req requests and launches their fits on the handle's stream, returning
 * without waiting (the request arrays may be freed at once).  eps is
 * nebula's absolute stopping tolerance; mode is NH_SOLVE_*.  One launch may
 * be in flight per handle: a second submit before its collect is
 * NH_ERR_INVALID_ARGUMENT. */
int nh_resident_submit(nh_resident_t* r, const nh_request_t* requests, size_t n_req,
                       double eps, int32_t mode, char* err, size_t err_len);

/* ResidentBatch::collect (src/gpu/nebula_gpu.rs:605-648): waits for the
 * launch in flight and writes one reply per request, in request order;
 * n_req must be the submitted count.  A reply's convergence is
 * check_convergence (src/sc/pml.rs:1111-1139) on the fit alone: 1, -10
 * (critical point), -20 (iteration budget), -25 (singular information),
 * -30 (not finite) or -40 (stalled). */
int nh_resident_collect(nh_resident_t* r, nh_reply_t* replies, size_t n_req, char* err,
                        size_t err_len);

/* nh_resident_submit then nh_resident_collect. */
int nh_resident_solve(nh_resident_t* r, const nh_request_t* requests,
                      size_t n_req, double eps, int32_t mode,
                      nh_reply_t* replies, char* err, size_t err_len);

/* nh_resident_solve without allocation, for callers that cannot allocate
 * device memory (serving loops, code under stream capture): same arguments,
 * same kernels, same replies.  No counterpart in edge-rs.  Every device
 * buffer beyond the handle's resident data (requests, start points, replies,
 * result arrays, the fit's global-memory work area, its counters) lives in
 * one file-scope device array of 32 MiB per GPU, held until the process ends;
 * a batch needing more is NH_ERR_INVALID_ARGUMENT (about 250k requests at 4
 * coefficients and 20 subjects, ~300 at 8 coefficients and 180 subjects).
 * Host staging uses fixed stack buffers; no hipMalloc, hipFree or other
 * allocating routine is called.  Calls that overlap on one GPU share the
 * array and race; the array keeps the last call's contents. */
int nh_resident_solve_noalloc(nh_resident_t* r, const nh_request_t* requests,
                              size_t n_req, double eps, int32_t mode,
                              nh_reply_t* replies, char* err, size_t err_len);

void nh_resident_free(nh_resident_t* r);

/* BOBYQA as the device and the host search run it (NLopt 2.7.1 semantics).
 * Minimises a callback; trace receives every evaluated point (n values) and
 * its value, row by row (up to max_trace rows).  Returns the NLopt status. */
typedef double (*nh_objective_fn)(const double* x, void* ctx);
int nh_bobyqa(int n, const double* x0, const double* lower, const double* upper,
              double xtol_rel, int maxeval, nh_objective_fn f, void* ctx,
              double* x_out, double* f_out, int* evals_out);

/* The same stepper run on the device, objective = a built-in test function
 * (0 rosen, 1 corner, 2 ripple, 3 valley, 4 chain3) evaluated on the device.
 * trace: max_trace * (n + 1) doubles.  Returns the NLopt status, or
 * -100 - NH_ERR_* on an error (then err holds the message). */
int nh_bobyqa_device_trace(int problem, int n, const double* x0,
                           const double* lower, const double* upper,
                           double* trace, int max_trace, int* evals_out,
                           char* err, size_t err_len);

/* Stage one's search on the host against a caller's objective and gradient
 * f(x, g) (minimise_marginal, src/sc/nebula.rs:1075-1116, with the
 * quasi-Newton descent it runs where zero-count tables apply: L-BFGS-B,
 * src/numeric/lbfgsb.rs:876, ftol 1e-6 over the start's objective, pgtol
 * 1e-7, 300 iterations, 40 evaluations per line search, memory 10; then up
 * to three restarts off a bound the start was inside).  The driver runs the
 * same search with device sums for f and g.  Writes the minimiser, whether
 * the start's value was not finite, and the evaluation count.  Returns NH_OK,
 * NH_ERR_INVALID_ARGUMENT or NH_ERR_INVALID_BOUNDS. */
typedef double (*nh_value_grad_fn)(const double* x, double* g, void* ctx);
int nh_stage_one_search(int n, const double* start, const double* lower,
                        const double* upper, nh_value_grad_fn f, void* ctx,
                        double* x_out, int* failed, int* evals_out);

/* One gene's stage-two search control, the state machine both engines run
 * (StageTwoSearch, src/sc/nebula.rs:1568-1951), on the host against a caller's
 * objective f(subject, cell, order): kind 2 searches both variance components
 * with BOBYQA from start[0..2); kind 1 the subject variance alone from
 * start[0], cell held at start[1] (NEBULA-LN's one-component refit: bounded
 * Nelder-Mead, restart off a bound, check, three quadratic polish stencils).
 * ord 1 or 3 is the Laplace order of the first pass; a pass at order 3 that
 * meets a non-finite value is repeated at order 1.  Writes the minimiser
 * (x_out[0..2), the held cell for kind 1), whether one was found, whether
 * BOBYQA ended RoundoffLimited (nebula's -50), the final order and the
 * evaluation count.  Returns NH_OK, NH_ERR_INVALID_ARGUMENT or
 * NH_ERR_INVALID_BOUNDS (lower[i] >= upper[i]). */
typedef double (*nh_search_fn)(double subject, double cell, int order, void* ctx);
int nh_stage_two_search(int kind, const double* start, const double* lower,
                        const double* upper, int ord, nh_search_fn f, void* ctx,
                        double* x_out, int* found, int* failed, int* order_out,
                        int* evals_out);

/* The same control run on lane 0 of a device workgroup through the
 * persistent kernel's entry points, against a built-in objective of (s, c):
 * problem 0 (s-0.3)^2 + 0.01 (c-4)^2, 1 (s+c-1)^2, 2 s, 3 (s-0.3)^2 + 2,
 * 4 a shallow well at the lower bound and a deeper one at s = 0.05.  Calls
 * inj_from..inj_to (one-based) evaluated at order >= inj_order return +inf
 * (inj_kind 0), NaN (1) or the profile objective of a fit whose Laplace
 * correction is below -1 (2, +inf by VarianceObjective::assemble,
 * src/sc/nebula.rs:1452-1456).  trace: max_trace rows of (subject, cell,
 * order, value).  No counterpart in edge-rs (its state machine is tested on
 * the host only).  Returns NH_OK or an NH_ERR_* code. */
int nh_stage_two_search_device(int problem, int kind, int ord, const double* start,
                               const double* lower, const double* upper, int inj_from,
                               int inj_to, int inj_order, int inj_kind, int32_t device,
                               double* trace, int max_trace, double* x_out, int* found,
                               int* failed, int* order_out, int* evals_out, char* err,
                               size_t err_len);

/* A device-to-device copy of `mib` MiB, `reps` times (each launch reads and
 * writes every byte once), checked by reading the last word back: the
 * yardstick the HBM byte counters (FETCH_SIZE, WRITE_SIZE) are scaled by.
 * No counterpart in edge-rs.  Returns NH_OK or an NH_ERR_* code. */
int nh_selftest_copy(size_t mib, int32_t reps, int32_t device, char* err, size_t err_len);

/* The zero-count tables the device sweeps read (src/sc/zeros.rs:219-278's
 * ZeroCells::build and :141-193's Group::eval), built and evaluated on the
 * host: for a centred design (n_cells x n_coef), log offsets and subject
 * starts (n_subjects + 1), *applies is 1 when the tables apply and 0 when the
 * cells in groups of >= 32 identical rows are under half the run.  When they
 * apply: *n_groups groups; group_row (capacity n_cells) their first cells;
 * subject_groups (n_subjects + 1); and, for each group g and point i,
 * out[(g * n_s + i) * m + j] = sum from + j at s[i] (from, m one of (0, 2),
 * (0, 3), (0, 4), (2, 4), (0, 6)); out_len bounds out.  A test hook: no
 * counterpart in edge-rs's API.  Returns NH_OK or an NH_ERR_* code. */
int nh_zero_tables_probe(const double* design, const double* log_offset, size_t n_cells,
                         size_t n_coef, const uint64_t* subject_start, size_t n_subjects,
                         const double* s, size_t n_s, int32_t from, int32_t m, double* out,
                         size_t out_len, int32_t* applies, uint64_t* n_groups,
                         uint64_t* group_row, uint64_t* subject_groups, char* err,
                         size_t err_len);

/* The FP64 sweep's exp (which 0), log (1) and reciprocal (2), and the device
 * objective's log Gamma for x > 0 (3), applied on the device to x[0..n), for
 * numerics tests of the device fit (no counterpart in edge-rs, whose fits call
 * the Rust standard library's and statrs's f64 functions).
 * Returns NH_OK or an NH_ERR_* code. */
int nh_selftest_f64_math(const double* x, size_t n, int32_t which, int32_t device,
                         double* out, char* err, size_t err_len);

/* ---------------------------------------------------------------------------
 * Wald tests on a fit: glm_sc_test (src/sc/test.rs:98-201).
 * ------------------------------------------------------------------------- */

/* Tests coefficient `coef` (contrast NULL) or the contrast c'beta (contrast:
 * n_coef weights, not all zero; coef ignored) per gene, on a fit's
 * coefficients (n_genes x n_coef, row-major, coefficients_len values) and
 * packed upper-triangular covariance ((i, j), i <= j, at j(j+1)/2 + i;
 * covariance_len values), as nebula_fit_t holds them.  Writes n_genes
 * values to each of log_fc (c'beta), se (sqrt(c'Vc), full covariance; 0
 * when c'Vc <= 0), z (log_fc / se, 0 when se is 0), p_value (chi^2 with 1
 * df upper tail of z^2, 1 when se is 0) and fdr (Benjamini-Hochberg).
 * Errors: NH_ERR_MUST_BE_POSITIVE (a zero dimension), NH_ERR_LENGTH_MISMATCH,
 * NH_ERR_COEF_OUT_OF_RANGE, NH_ERR_INVALID_ARGUMENT (an all-zero contrast). */
int nh_glm_sc_test(const double* coefficients, size_t coefficients_len,
                   const double* covariance, size_t covariance_len, size_t n_genes,
                   size_t n_coef, size_t coef, const double* contrast, double* log_fc,
                   double* se, double* z, double* p_value, double* fdr, char* err,
                   size_t err_len);

/* ---------------------------------------------------------------------------
 * Empirical-Bayes shrinkage of the cell-level overdispersions, host code on a
 * fit's output (src/sc/shrink.rs:90-226).
 * ------------------------------------------------------------------------- */

/* sc_residual_df (src/sc/shrink.rs:90): n_cells - n_coef - (n_subjects - 1),
 * at least 1. */
double nh_sc_residual_df(size_t n_cells, size_t n_coef, size_t n_subjects);

/* shrink_sc_dispersion (src/sc/shrink.rs:118): phi = 1 / dispersion (infinite
 * at 0) of the genes with convergence == 1 and finite phi, floored at 1e-8,
 * moderated by limma's squeezeVar (equal-df family) on df_residual degrees of
 * freedom -- trended on `covariate` (n_genes values, NULL for none) when given,
 * robust when `robust` is non-zero (ShrinkParams::robust, default on), each
 * rung of the ladder trended -> untrended -> untrended non-robust reached only
 * when the one above fails, unshrunk (df_prior {0}) when all fail or fewer
 * than 3 genes are usable.  Writes n_genes values each to phi_raw, phi_post,
 * phi_prior and dispersion_shrunk (1 / phi_post), genes outside the fit taking
 * the median prior; df_prior (capacity n_genes) gets *n_df_prior values: one
 * per usable gene from the robust fit, else 1.  Errors (ShrinkResult's):
 * NH_ERR_INVALID_ARGUMENT for n_genes == 0 or df_residual not positive and
 * finite. */
int nh_shrink_sc_dispersion(const double* dispersion, const int32_t* convergence,
                            const double* covariate, size_t n_genes, double df_residual,
                            int32_t robust, double* phi_raw, double* phi_post,
                            double* phi_prior, double* dispersion_shrunk, double* df_prior,
                            size_t* n_df_prior, char* err, size_t err_len);

/* limma's squeezeVar as squeeze_var runs it for one shared df
 * (src/limma/squeeze_var.rs:286-391, the legacy family it picks then):
 * fitFDist (:426), its spline trend on `covariate` (:501), or the Winsorised
 * fit when `robust` is non-zero (:612, tail proportions 0.05 and 0.1, lowess
 * span 0.4 when trended).  var: n finite non-negative values; covariate NULL
 * or n values.  Writes var_post (n), var_prior (*n_var_prior values: 1 or n)
 * and df_prior (*n_df_prior values: 1 or n); fewer than 3 genes come back
 * unmoderated with df_prior {0}.  Returns NH_OK or NH_ERR_INVALID_ARGUMENT
 * (with the failing step's message). */
int nh_squeeze_var(const double* var, size_t n, double df, const double* covariate,
                   int32_t robust, double* var_post, double* var_prior, size_t* n_var_prior,
                   double* df_prior, size_t* n_df_prior, char* err, size_t err_len);

/* The distribution functions the shrinkage is built on, applied to x[0..n)
 * (a test hook, no counterpart in edge-rs's API): which 0 F(a, b) density,
 * 1 F(a, b) upper tail, 2 F(a, b) quantile at lower-tail probability x, 3
 * chi-square(a) upper tail, 4 stats::lowess(x, y, span a, b robustness
 * passes), 5 log x - digamma x, 6 trigamma, 7 inverse trigamma; b may be
 * infinite for 0-2.  Returns NH_OK or an NH_ERR_* code. */
int nh_shrink_selftest(int32_t which, const double* x, const double* y, size_t n, double a,
                       double b, double* out);

/* ---------------------------------------------------------------------------
 * Warm-start MLP (no counterpart in edge-rs): 12 features -> 64 -> 64 -> 64
 * -> 2 (log sigma^2, log phi), GELU (tanh form).  Parameter blob layout:
 * feature mean[12], feature sd[12], W1[64x12], b1[64], W2[64x64], b2[64],
 * W3[64x64], b3[64], W4[2x64], b4[2], target mean[2], target sd[2]; W
 * row-major (out x in).  The network predicts the correction to stage one's
 * estimate (features 10 and 11): targets are standardised (log sigma^2, log
 * phi) minus those features; predictions add them back.
 * ------------------------------------------------------------------------- */
size_t nh_mlp_param_count(void);
/* Trains on the device (FP32 kernels, Adam, cosine learning-rate decay,
 * minibatch `batch`, Huber loss with delta 1 on the standardised correction):
 * features n x 12, targets n x 2 = (log sigma^2, log phi).  Standardisation is
 * taken from the data; weights start Glorot-uniform from `seed`.  Writes the
 * blob to params and the final mean loss over all rows to final_loss. */
int nh_mlp_train(const float* features, const float* targets, size_t n, int32_t epochs,
                 int32_t batch, float lr, uint64_t seed, int32_t device, float* params,
                 float* final_loss, char* err, size_t err_len);
/* Training loss (mean Huber, delta 1, over the standardised corrections) and
 * its gradient with respect to the network part of the blob (W1..b4 in blob
 * order), computed by the training kernels on the device. */
int nh_mlp_loss_grad(const float* params, const float* features, const float* targets,
                     size_t n, int32_t device, float* loss, float* grads, char* err,
                     size_t err_len);
/* Inference with BF16 weights and activations on the matrix cores (FP32
 * accumulation): out n x 2 = (log sigma^2, log phi). */
int nh_mlp_predict(const float* params, const float* features, size_t n, int32_t device,
                   float* out, char* err, size_t err_len);

#ifdef __cplusplus
}
#endif

#endif /* NEBULA_HIP_H */

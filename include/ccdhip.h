/*
 * ccdhip.h -- C ABI of the MI355X-native CCD engine (libccdhip.so).
 *
 * Plain pointers and sizes only; every function returns 0 on success or a
 * negative CCDH_E* status, with a message available from ccdh_last_error().
 * No C++ exception crosses this boundary.  Handles are independent: calls on
 * different fit handles may come from different host threads; calls on one
 * handle must be serialised by the caller (the reference's ModelSpecifics is
 * single-threaded per instance in the same way).
 *
 * Reference interfaces each group replaces (Cyclops, src/cyclops/...):
 *   data handle ........ ModelData<RealType> + CompressedDataMatrix (ModelData.h:178-779,
 *                        CompressedDataMatrix.h:243-452)
 *   fit handle ......... one ModelSpecifics<Model,RealType> instance
 *                        (engine/AbstractModelSpecifics.h:36-229; factory
 *                        engine/AbstractModelSpecifics.cpp:128-154, GPU slot :32-35)
 *   per-coordinate ops . the AbstractModelSpecifics virtuals, one C function each
 *   ccdh_fit_plain ..... CyclicCoordinateDescent::update -> findMode
 *                        (CyclicCoordinateDescent.cpp:586-615, 909-1216) driving the
 *                        per-coordinate ops from the host (the "plain" GPU engine)
 *   ccdh_batch_fit ..... AbstractCrossValidationDriver::doCrossValidationStep's per-fold
 *                        ccd.update() calls (drivers/AbstractCrossValidationDriver.cpp:131-248),
 *                        run as device-resident sweeps for many fits at once
 *   ccdh_server_* ...... no reference counterpart; a request loop over the batched engine
 */
#ifndef CCDHIP_H_
#define CCDHIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define CCDH_OK 0
#define CCDH_E_INVALID (-1)   /* bad argument / unsupported model-format combination */
#define CCDH_E_DEVICE (-2)    /* HIP runtime error */
#define CCDH_E_NOMEM (-3)     /* device allocation failed */
#define CCDH_E_STATE (-4)     /* call out of order (e.g. gradient before statistics) */
#define CCDH_E_FIT (-5)       /* the optimizer raised an error (MM: "Non-increasing!") */

/* column formats: FormatType, CompressedDataMatrix.h:36-38 */
#define CCDH_DENSE 0
#define CCDH_SPARSE 1
#define CCDH_INDICATOR 2
#define CCDH_INTERCEPT 3

/* models: ModelType, Types.h:146-159 (the values match) */
#define CCDH_POISSON 2
#define CCDH_LOGISTIC 3
#define CCDH_CONDITIONAL_LOGISTIC 4
#define CCDH_COX 8

/* precision: PrecisionType, Types.h:168-173, plus the mixed path */
#define CCDH_FP64 0
#define CCDH_MIXED 1 /* FP32 xBeta / exp streams, FP64 reductions, scans and LL */

/* priors: priors::PriorType, Types.h:101-106 */
#define CCDH_PRIOR_NONE 0
#define CCDH_PRIOR_LAPLACE 1
#define CCDH_PRIOR_NORMAL 2

/* convergence criteria: ConvergenceType, Types.h:110-116 (ZHANG_OLES not offered) */
#define CCDH_CONV_GRADIENT 0
#define CCDH_CONV_LANGE 1
#define CCDH_CONV_MITTAL 2
#define CCDH_CONV_ONE_STEP 3

/* return flags: UpdateReturnFlags, Types.h:124-130 */
#define CCDH_SUCCESS 0
#define CCDH_FAIL 1
#define CCDH_MAX_ITERATIONS 2
#define CCDH_ILLCONDITIONED 3

typedef struct ccdh_data* ccdh_data_t;
typedef struct ccdh_fit* ccdh_fit_t;
typedef struct ccdh_server* ccdh_server_t;

/* Input data in column-major compressed form (host memory; copied to the GPU).
 * DENSE/INTERCEPT columns span all K rows (rows[] = 0..K-1); INDICATOR and
 * INTERCEPT columns carry no values.  Grouped (conditional logistic) rows must
 * be sorted by stratum; Cox rows by (stratum, -time, y). */
typedef struct {
  int model;               /* CCDH_POISSON ... CCDH_COX */
  int precision;           /* CCDH_FP64 or CCDH_MIXED */
  int K;                   /* rows */
  int J;                   /* columns */
  const int64_t* col_ptr;  /* [J+1] */
  const int32_t* rows;     /* [nnz] strictly increasing within a column */
  const double* vals;      /* [nnz] or NULL when no SPARSE/DENSE column */
  const int8_t* fmt;       /* [J] */
  const double* y;         /* [K] outcome */
  const double* offs;      /* [K] event time (Cox); may be NULL otherwise */
  const int32_t* strata;   /* [K] stratum id (CLR, Cox); NULL = one stratum */
  int has_offset;          /* column 0 is a DENSE offset with fixed coefficient 1 */
} ccdh_data_desc;

/* One fit's settings (CCDArguments / ModeFindingArguments, CcdInterface.h:70-146). */
typedef struct {
  const int32_t* prior_type;  /* [J] per-coordinate prior (NoPrior for offset/intercept) */
  const double* prior_var;    /* [J] prior variance (Laplace lambda = sqrt(2/var)) */
  const uint8_t* fixed;       /* [J] fixBeta */
  const double* beta0;        /* [J] starting coefficients (offset column 1) */
  const double* weights;      /* [K] or NULL: NULL means useCrossValidation=false */
  const double* heldout;      /* [K] or NULL: predictive log likelihood of these weights */
  double tolerance;           /* default 1e-6 */
  int max_iterations;         /* default 1000 */
  int convergence_type;       /* default CCDH_CONV_GRADIENT */
  double initial_bound;       /* default 2.0 */
  int max_bound_count;        /* default 5 */
} ccdh_fit_spec;

typedef struct {
  double log_likelihood;
  double log_prior;
  double pred_log_likelihood; /* 0 unless heldout weights were given */
  int iterations;
  int return_flag;
  int bound_count;
  int _pad;
  double seconds;             /* wall time of the fit on the host clock */
} ccdh_fit_result;

/* ---- library ---- */
const char* ccdh_last_error(void);
int ccdh_device_info(char* name, size_t len, int* cu_count, double* hbm_gb);

/* ---- data (ModelData) ---- */
int ccdh_data_create(const ccdh_data_desc* desc, ccdh_data_t* out);
int ccdh_data_destroy(ccdh_data_t data);
/* device bytes held by the data handle (X, outcome, index structures) */
int64_t ccdh_data_device_bytes(ccdh_data_t data);
/* model, rows, columns and offset flag of a data handle */
int ccdh_data_shape(ccdh_data_t data, int* model, int* K, int* J, int* has_offset);
/* CCDH_FP64 or CCDH_MIXED, as created */
int ccdh_data_precision(ccdh_data_t data);

/* Long-format input, as convertToCyclopsData takes it (R/NewDataConversion.R:95-235):
 * one outcome row per row id, covariate entries (covariate id, row id, value).  The
 * library sorts the rows (LR, Poisson: row id; CLR: stratum, row id; Cox: stratum,
 * time descending, y, row id), numbers the strata in that order (loadY,
 * ModelData.cpp:69-146), maps row ids to rows and covariate ids (ascending) to
 * columns, leaves out entries whose row id is not an outcome row and zero values,
 * refuses a repeated (row, covariate) entry, and makes a column SPARSE when a value
 * other than 0 and 1 occurs, INDICATOR otherwise (loadMultipleX, ModelData.cpp:148-243).
 * LR and Poisson get an intercept column first when add_intercept is set; Poisson
 * gets log(time) as the offset column before it (finalizeSqlCyclopsData with
 * useOffsetCovariate = -1; RcppModelData.cpp:482-556).  The sorts and maps run on
 * the device. */
typedef struct {
  int model;                    /* CCDH_POISSON ... CCDH_COX */
  int precision;                /* CCDH_FP64 or CCDH_MIXED */
  int64_t n_rows;               /* outcome rows */
  const int64_t* row_id;        /* [n_rows] distinct */
  const int64_t* stratum_id;    /* [n_rows]; required for CLR, optional for Cox, else ignored */
  const double* y;              /* [n_rows] */
  const double* time;           /* [n_rows]: Cox event time, Poisson exposure (> 0); else NULL */
  int64_t n_entries;            /* covariate entries */
  const int64_t* covariate_id;  /* [n_entries] */
  const int64_t* entry_row_id;  /* [n_entries] */
  const double* value;          /* [n_entries] or NULL: every entry 1 */
  int add_intercept;            /* LR / Poisson */
} ccdh_long_desc;

/* convertToCyclopsData + loadY / loadMultipleX / finalizeSqlCyclopsData (above). */
int ccdh_data_create_long(const ccdh_long_desc* desc, ccdh_data_t* out);
/* Copies a data set's compressed columns back: col_ptr [J+1], rows and vals [nnz]
 * (vals may be NULL), fmt [J], y, offs and strata [K]; any output may be NULL. */
int ccdh_data_get_csc(ccdh_data_t data, int64_t* col_ptr, int32_t* rows, double* vals,
                      int8_t* fmt, double* y, double* offs, int32_t* strata);
/* Introspection of a data set's derived per-column tables (no reference
 * counterpart: the reference walks its sorted columns directly; these are this
 * library's indexes over them, built on the host by ccdh_data_create and on the
 * device by ccdh_data_create_long).  Conditional logistic: seg_ptr [J+1], seg_off
 * and seg_key [n_seg] (runs of equal stratum inside each column: offset in the
 * column, stratum), nz_key [nnz] (stratum of each nonzero), strat_start
 * [n_strata+1]; every model: shard_row [65] (row shards), col_shard [J*65] (each
 * column's offset of each shard's first nonzero); not Cox: blk [J] (column
 * blocks: length at a block start, 0 inside a block, 1 alone).  Any pointer may
 * be NULL; call with only n_seg and n_strata first to size the arrays. */
int ccdh_data_get_tables(ccdh_data_t data, int64_t* n_seg, int* n_strata, int64_t* seg_ptr,
                         int32_t* seg_off, int32_t* seg_key, int32_t* nz_key,
                         int32_t* strat_start, int32_t* shard_row, int32_t* col_shard,
                         int8_t* blk);

/* ---- per-coordinate engine (AbstractModelSpecifics virtuals) ---- */
/* ModelSpecifics ctor + initialize(N,K,J,...) (ModelSpecifics.hpp:1899-1952) */
int ccdh_fit_create(ccdh_data_t data, ccdh_fit_t* out);
/* clone() (AbstractModelSpecifics.h:103): new state over the same resident data */
int ccdh_fit_clone(ccdh_fit_t fit, ccdh_fit_t* out);
int ccdh_fit_destroy(ccdh_fit_t fit);
/* setWeights(double*, double*, bool) (AbstractModelSpecifics.h:65) */
int ccdh_set_weights(ccdh_fit_t fit, const double* weights, int use_cv);
/* computeFixedTermsInLogLikelihood(bool) (:83) */
int ccdh_fixed_terms_loglik(ccdh_fit_t fit, int use_cv);
/* computeFixedTermsInGradientAndHessian(bool) (:85), i.e. XjY */
int ccdh_fixed_terms_grad_hess(ccdh_fit_t fit, int use_cv);
/* zeroXBeta() / axpyXBeta(double, int) (:120,122) */
int ccdh_zero_xbeta(ccdh_fit_t fit);
int ccdh_axpy_xbeta(ccdh_fit_t fit, double beta, int j);
/* computeXBeta(double*, bool) (:79; ModelSpecifics.hpp:313-364): xBeta from the
 * transpose of X (CompressedDataMatrix.cpp:151-208), each row summed in column
 * order.  As there, a matrix with INDICATOR columns and no DENSE column gives an
 * INDICATOR transpose, in which every entry counts 1.  Built on the first call. */
int ccdh_compute_xbeta(ccdh_fit_t fit, const double* beta);
/* computeMMGradientAndHessian(std::vector<GradientHessian>&, const std::vector<bool>&,
 * bool) (:70; ModelSpecifics.hpp:776-889): every column's MM gradient and Hessian
 * from the current statistics into gh_out[2j], gh_out[2j+1]; (0, 0) where fixed[j].
 * The increments weigh by nW whatever use_weights says (ModelSpecifics.h:500-517,
 * 571-584); Cox divides by each pid's own denominator (a tie group of events, or one
 * row), as the reference's Survival increments read denomPid.  Poisson is refused
 * with CCDH_E_INVALID, as the reference throws (ModelSpecifics.h:1546-1552). */
int ccdh_mm_grad_hess(ccdh_fit_t fit, const uint8_t* fixed, int use_weights, double* gh_out);
/* computeRemainingStatistics(bool) (:81) */
int ccdh_remaining_stats(ccdh_fit_t fit, int use_weights);
/* computeNumeratorForGradient(int, bool) (:72) */
int ccdh_numerator(ccdh_fit_t fit, int j, int use_weights);
/* computeGradientAndHessian(int, double*, double*, bool) (:67-68) */
int ccdh_grad_hess(ccdh_fit_t fit, int j, int use_weights, double* gradient, double* hessian);
/* updateXBeta(double, int, bool) (:77) */
int ccdh_update_xbeta(ccdh_fit_t fit, double delta, int j, int use_weights);
/* getLogLikelihood(bool) (:87) */
int ccdh_loglik(ccdh_fit_t fit, int use_cv, double* out);
/* getGradientObjective(bool) (:93) */
int ccdh_gradient_objective(ccdh_fit_t fit, int use_cv, double* out);
/* getPredictiveLogLikelihood(double*) (:89) */
int ccdh_pred_loglik(ccdh_fit_t fit, const double* weights, double* out);
/* getXBeta() (:114): copies K values */
int ccdh_get_xbeta(ccdh_fit_t fit, double* out);

/* ---- whole fits ---- */
/* Plain engine: the host CCD (findMode) calling the per-coordinate ops. */
int ccdh_fit_plain(ccdh_data_t data, const ccdh_fit_spec* spec, double* beta_out,
                   ccdh_fit_result* result);
/* The same driver with the MM algorithm (createControl(algorithm = "mm"), R/ModelFit.R:513-524;
 * CyclicCoordinateDescent.cpp:958-1005, 1026-1048, 1354-1391): each iteration steps every
 * free coordinate from one set of statistics (no bounding box), then recomputes xBeta
 * from the transpose and the statistics.  A decrease of the log posterior stops the fit
 * with CCDH_E_FIT ("Non-increasing!"); beta_out then holds the coefficients reached and
 * result->iterations the iteration.  Logistic, conditional logistic and Cox models (not
 * Poisson), FP64 data (on the mixed path the log posterior's FP32 noise trips that check
 * long before the fit converges). */
int ccdh_fit_mm(ccdh_data_t data, const ccdh_fit_spec* spec, double* beta_out,
                ccdh_fit_result* result);

/* Device time of the MM step's two main kernels, from HIP events on the fit's
 * stream: the per-column chunk pass (k_mm_chunks) and xBeta from the transpose
 * (k_xbeta_rows), with the sizes their bytes follow from.  No counterpart in the
 * reference (its CPU timers, printTiming). */
typedef struct {
  int64_t chunk_launches;
  double chunk_ms;     /* summed over launches */
  int64_t xbeta_launches;
  double xbeta_ms;
  int64_t nnz;         /* nonzeros of the data set's compressed columns */
  int64_t entries;     /* entries of its transpose */
  int64_t rows;
  int values;          /* 1: SPARSE / DENSE values are read (8 B per nonzero) */
} ccdh_mm_stats;
/* For one fit handle (all its MM steps so far) and for the last ccdh_fit_mm call on
 * this host thread. */
int ccdh_fit_mm_stats(ccdh_fit_t fit, ccdh_mm_stats* out);
int ccdh_mm_last_stats(ccdh_mm_stats* out);

/* Batched engine options; zero fields pick the defaults. */
typedef struct {
  int threads;           /* workgroup size: 256 (<= 256) or 512; 0: 256, or with autotune
                            256 x 8 or 512 x 4 (workgroup size x pass width), as timed */
  int sweeps_per_sync;   /* sweeps per launch: with max_gang 1 (0: 4); with gangs a cap on top of
                            the time quantum (0: none) */
  int use_graph;         /* replay the sweep launch from a hipGraph (only with max_gang 1) */
  int max_gang;          /* workgroups one fit may use (1: one each; 0: 64) */
  int64_t gang_min_nnz;  /* nonzeros per workgroup below which a fit's gang stops growing (0: 32768) */
  int pass_items;        /* long-column pass width, nonzeros per lane: 8 (0) or 4 */
  int autotune;          /* time pass width x gangs on this workload first; the choice is cached */
  double quantum_ms;     /* with gangs: a launch ends at the first coordinate boundary after
                            this long, and gang sizes are re-dealt between launches (0: 200) */
  int bundle;            /* fits per slot moving in lock step over interleaved state: 1, 2, 4
                            or 8 (0: 1; Cox: 1).  Consecutive fits with the same fixed
                            coordinates and controls share a slot. */
  int solo;              /* gangs: runs of short columns are swept by the lead workgroup alone,
                            without exchanges (0); -1: every column by the whole gang */
  int64_t spin_limit;    /* polls a gang rank waits for its peers before it gives up (0: 2^24).
                            A fit whose gang gives up restarts cold on one workgroup
                            (ccdh_batch_stats.gang_restarts); small values force that path. */
} ccdh_batch_options;

/* Batched engine: n_fits fits over one data handle, device-resident sweeps.
 * betas_out is [n_fits * J]; results [n_fits]. */
int ccdh_batch_fit(ccdh_data_t data, int n_fits, const ccdh_fit_spec* specs, double* betas_out,
                   ccdh_fit_result* results, const ccdh_batch_options* options);

/* Counters of the last ccdh_batch_fit on this host thread, for the roofline:
 * kernel time of the sweep launches (HIP events on the batch stream) and the
 * work they did, summed over fits. */
typedef struct {
  int64_t launches;        /* sweep launches (graph replays or direct) */
  int64_t sweeps;          /* completed CCD cycles, summed over fits */
  int64_t sweep_rounds;    /* max over fits of cycles per launch, summed over launches */
  int64_t gather_nnz;      /* column nonzeros read by gradient/Hessian gathers */
  int64_t gather_runs;     /* stratum runs (CLR) / touched pids (Cox) of those gathers */
  int64_t update_nnz;      /* column nonzeros updated (coordinates with delta != 0) */
  int64_t update_runs;
  int64_t scan_pids;       /* pids swept by the Cox risk-set scans */
  int64_t objective_rows;  /* rows read by the per-sweep convergence objective */
  int64_t ticks_single;    /* device wall-clock ticks (100 MHz), summed over fits: columns */
  int64_t ticks_long;      /*   done in one pass / in several passes / convergence objective
                              (Cox, lead rank: numerators / risk-set scan / objective) */
  int64_t ticks_objective;
  int64_t long_columns;    /* multi-pass column visits (Cox: ticks of the exchange and update) */
  double kernel_ms;        /* summed HIP-event time of the sweep launches */
  double wall_ms;          /* host wall time of the whole batch */
  /* the first 128 launches: HIP-event time, fits in flight, workgroups per fit */
  float launch_ms[128];
  int32_t launch_fits[128];
  int32_t launch_gang[128];
  /* what the batch ran with; with autotune, the candidates' rates (coordinates per ms) in the
     order timed: 256 x 8 on gangs, 256 x 8 on one workgroup, 512 x 4 on gangs (a caller's own
     workgroup size T: T x 8 on gangs, T x 8 on one workgroup, then T x 4 on gangs when T is
     256); when the last one timed wins, its launch is the batch's first (launch_ms[0]) */
  int32_t tuned_pass_items, tuned_max_gang;
  double tune_ms;
  double tune_rates[4];
  int32_t bundles;         /* slots the fits were bundled into */
  int32_t tuned_threads;   /* workgroup size of the sweep launches */
  int32_t gang_restarts;   /* slots restarted cold on one workgroup after a gang wait gave up */
  int32_t shared_setups;   /* one-fit slots whose specs passed another slot's weights array:
                              they read its device copy and took over its X'Y and group
                              weight sums */
} ccdh_batch_stats;
int ccdh_batch_last_stats(ccdh_batch_stats* out);

/* ---- cross-validation batch definition ---- */
/* CrossValidationSelector::getWeights after the permutations of `task`
 * (drivers/CrossValidationSelector.cpp:21-116,138-143): ids are the exchangeable
 * unit of each row (stratum for by_pid, ignored otherwise).  Seed -1 is the
 * selector's deterministic mode (no shuffle; drivers/AbstractSelector.cpp:32-48);
 * the clock seed -99 is refused (CCDH_E_INVALID), since each task is a separate call. */
int ccdh_cv_fold_weights(const int32_t* ids, int K, int by_pid, int n_fold, long seed, int task,
                         double* weights_out);
/* GridSearchCrossValidationDriver::computeGridPoint (GridSearchCrossValidationDriver.cpp:38-48) */
double ccdh_cv_grid_point(double lower, double upper, int size, int step);

/* ---- bootstrap batch definition ----
 * Seed -1 leaves the selector's std::mt19937 default-constructed
 * (drivers/AbstractSelector.cpp:32-48); -99 is refused as above. */
/* BootstrapSelector::permute + getWeights as BootstrapDriver::drive calls them
 * (drivers/BootstrapSelector.cpp:17-95, drivers/BootstrapDriver.cpp:47-71): the
 * weights of replicate `replicate` (0-based), each row weighing the number of
 * times its pid (ids[k]) was drawn.  exclude (nullable, n_exclude pids): only
 * pids with exclude[i] == 0 are drawn (the selector's wtsExclude). */
int ccdh_bootstrap_weights(const int32_t* ids, int K, long seed, int replicate,
                           const double* exclude, int n_exclude, double* weights_out);

/* ---- device-side ingestion ---- */
/* Cox risk-set groups under `weights` (NULL: all 1), built on the device as the
 * batched engine and the server build them for every fit
 * (setPidForAccumulation, engine/ModelSpecifics.hpp:1809-1868): pid_out[K] (N
 * for zero-weight rows), reset_out[N+1] (1 where a stratum starts), pid_first_out
 * [N+1] (first row of each group, K at N), shard_pid_out[65] (first group of each
 * row shard), *n_out = N.  Output pointers other than pid_out may be NULL. */
int ccdh_cox_groups(ccdh_data_t data, const double* weights, int32_t* pid_out, uint8_t* reset_out,
                    int32_t* pid_first_out, int32_t* shard_pid_out, int* n_out);

/* The same groups as ccdh_cox_groups, computed without allocating device memory.
 * No counterpart in the reference: for callers that cannot allocate (code that
 * must not call hipMalloc, such as a serving loop with a fixed memory budget).
 * Everything ccdh_cox_groups allocates per call is staged in one file-scope
 * device array of about 33 MiB that each GPU holds until the process ends, so
 * data sets of at most 1,048,576 rows are accepted (CCDH_E_INVALID beyond).
 * Calls in one process take turns; calls that overlap on one GPU from several
 * processes race on the array. */
int ccdh_cox_groups_noalloc(ccdh_data_t data, const double* weights, int32_t* pid_out,
                            uint8_t* reset_out, int32_t* pid_first_out, int32_t* shard_pid_out,
                            int* n_out);

/* ---- auto-search cross-validation ---- */
/* UniModalSearch::tried for each (xs[i], ys[i], sds[i]) in order, then ::step
 * (utils/HParSearch.h:30-65, utils/HParSearch.cpp:22-120) as
 * AutoSearchCrossValidationDriver::doCrossValidationLoop uses it
 * (drivers/AutoSearchCrossValidationDriver.cpp:63-165, UniModalSearch(10, 0.01,
 * log 1.5)): *cont = 0 when the search has converged, *next the next prior
 * variance, *expected the fitted value there.  sds may be null (all 0). */
int ccdh_search_step(const double* xs, const double* ys, const double* sds, int n, double stdstep,
                     double stop_by_y, double stop_by_x, double first_cut, int* cont, double* next,
                     double* expected);
/* QuadrLogFit (utils/HParSearch.cpp:22-63): c0 + c1 log x + c2 log^2 x, n >= 3. */
int ccdh_search_quadratic(const double* xs, const double* ys, const double* sds, int n,
                          double* c_out);

/* ---- serving: continuous batching of fit requests ----
 * No reference counterpart (Cyclops fits one model per call; its parallel
 * drivers run whole fits on std::threads, AbstractCrossValidationDriver.cpp:65-74).
 * A server owns one device memory arena (its HBM budget) and two host threads:
 * admission (FCFS; a request is admitted when its data + fit state fit in the
 * arena and a slot is free, then uploaded and set up on its own stream) and
 * scheduling (one sweep launch over every fit in flight, each fit running for a
 * time quantum and resuming at the next launch; finished fits leave, admitted
 * fits join at every launch boundary).  A fit's arithmetic does not depend on
 * the other fits in flight or on where its sweeps were paused. */
typedef struct {
  int precision;            /* CCDH_FP64 or CCDH_MIXED, for every request */
  int max_slots;            /* fits in flight at once (0: 1024) */
  int threads;              /* workgroup size (0: 256) */
  int max_gang;             /* workgroups one fit may use when the device has room (0: 64; 1: one) */
  double quantum_ms;        /* device time per fit per launch (0: 20 ms) */
  int64_t hbm_budget_bytes; /* arena size (0: half of the free device memory) */
  int64_t spin_limit;       /* polls a gang rank waits for its peers before it gives up (0: 2^24);
                               a request whose gang gives up restarts cold on one workgroup */
} ccdh_server_options;

typedef struct {
  double latency;           /* submit -> result ready, seconds (host clock) */
  double queued;            /* submit -> admitted (memory and slot granted) */
  double prepare;           /* admission: upload and set-up */
  int64_t footprint;        /* device bytes held while in flight */
  int64_t launches;         /* sweep launches the fit took part in */
} ccdh_request_info;

typedef struct {
  int64_t submitted, completed, failed;
  int64_t in_flight, queued;
  int64_t launches;         /* sweep launches so far */
  double kernel_ms;         /* summed HIP-event time of those launches */
  int64_t arena_bytes, arena_used, arena_peak;
  int64_t gang_restarts;    /* requests restarted cold on one workgroup after a gang wait gave up */
} ccdh_server_stats;

int ccdh_server_create(const ccdh_server_options* options, ccdh_server_t* out);
/* Queue a fit over host data; the server uploads it when the request is
 * admitted.  desc's arrays must stay valid until the request completes; the
 * spec's arrays are copied.  desc->precision must equal the server's. */
int ccdh_server_submit(ccdh_server_t server, const ccdh_data_desc* desc, const ccdh_fit_spec* spec,
                       int64_t* request_id);
/* Queue a fit over resident data (the handle must outlive the request). */
int ccdh_server_submit_resident(ccdh_server_t server, ccdh_data_t data, const ccdh_fit_spec* spec,
                                int64_t* request_id);
/* 1 when the request has completed (or failed), 0 while queued or in flight. */
int ccdh_server_poll(ccdh_server_t server, int64_t request_id);
/* Wait for a request (timeout_ms < 0: no limit; CCDH_E_STATE on timeout), copy
 * its coefficients (beta_out [J], may be NULL), result and timing, and release
 * it.  Returns the request's own status (CCDH_OK or the error it failed with). */
int ccdh_server_wait(ccdh_server_t server, int64_t request_id, double timeout_ms, double* beta_out,
                     ccdh_fit_result* result, ccdh_request_info* info);
int ccdh_server_get_stats(ccdh_server_t server, ccdh_server_stats* out);
/* Finish every queued and in-flight request, stop the threads, free the arena.
 * Results not yet collected are discarded. */
int ccdh_server_destroy(ccdh_server_t server);

#ifdef __cplusplus
}
#endif

#endif /* CCDHIP_H_ */

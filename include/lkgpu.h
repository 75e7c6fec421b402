/*
 * lkgpu.h -- C ABI of the MI355X (gfx950) Kriging likelihood path.
 *
 * Two layers, both plain C (pointers + sizes, no torch or HIP types):
 *
 *  1. The hot path itself: designs uploaded once to HBM, then batched
 *     log-likelihood (+ adjoint gradient) evaluations.  These replace the
 *     reference's KrigingImpl::populate_Model (src/lib/KrigingImpl.hpp:168-173,
 *     src/lib/KrigingImpl.cpp:73-125) plus the gradient block of
 *     Kriging::_logLikelihood (src/lib/Kriging.cpp:291-338) and
 *     KrigingImpl::compute_ll_grad_theta_vecs (src/lib/KrigingImpl.hpp:374-379).
 *
 *  2. The Kriging model object, mirroring the reference's own C ABI
 *     (bindings/Julia/jlibkriging/csrc/libkriging_c.h): same function names,
 *     argument meaning and error convention (0 on success, -1 on error, the
 *     message in lk_get_last_error()).
 *
 * All matrices are FP64 column-major with rows = observations (X is n x d).
 * Nothing here is thread-safe for concurrent calls on the same object, as
 * the reference's process-global configuration is not (LinearAlgebra /
 * Optim statics).
 */
#ifndef LKGPU_H
#define LKGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread
 * (replaces libkriging_c.h:10 lk_get_last_error). */
const char* lk_get_last_error(void);

/* Hash of the sources this library was built from (no counterpart in the
 * reference): the Python loader refuses a library whose hash differs from
 * the sources beside it. */
const char* lkgpu_build_id(void);

/* ------------------------------------------------------------------------ */
/* Layer 1: batched likelihood on one GPU                                   */
/* ------------------------------------------------------------------------ */

typedef struct lkgpu_context lkgpu_context;

/* Kernel ids (Covariance::resolve, src/lib/Covariance.cpp:219-230). */
#define LKGPU_KERNEL_GAUSS 0
#define LKGPU_KERNEL_EXP 1
#define LKGPU_KERNEL_MATERN3_2 2
#define LKGPU_KERNEL_MATERN5_2 3

/* Noise models (Kriging::NoiseModel, src/lib/include/libKriging/Kriging.hpp:45-49). */
#define LKGPU_NOISE_NONE 0
#define LKGPU_NOISE_NUGGET 1

/* Per-item status of a batched evaluation. */
#define LKGPU_OK 0
#define LKGPU_CHOL_FAILED_AFTER_RETRIES 1 /* "Exceed max numerical nugget" (LinearAlgebra.cpp:79-82) */
#define LKGPU_NUGGET_NOT_POSITIVE 2       /* LinearAlgebra.cpp:83-85 */

/* Create a context on HIP device `device` (one process per GPU). */
int lkgpu_context_create(int device, lkgpu_context** out);
void lkgpu_context_destroy(lkgpu_context* ctx);

/* The LinearAlgebra statics of the Cholesky retry ladder:
 * set_num_nugget (LinearAlgebra.cpp:35-37), check_chol_rcond (:55-57),
 * max_inc_choldiag (:63) and min_rcond (:100). */
int lkgpu_set_chol_config(lkgpu_context* ctx, double num_nugget, int rcond_check, int max_inc, double min_rcond);

/* Upload one design: X (n x d), y (n), trend matrix F (n x p, may be NULL
 * when p = 0), kernel and noise-model ids.  Returns its id in *id_out. */
int lkgpu_problem_create(lkgpu_context* ctx, const double* X, int n, int d, const double* y, const double* F, int p,
                         int kernel, int noise, int* id_out);
int lkgpu_problem_destroy(lkgpu_context* ctx, int id);

/* Estimation flags of one model (Kriging::_logLikelihood, Kriging.cpp:221-245). */
typedef struct {
  int est_sigma2; /* sigma2 concentrated out (default 1) */
  int est_nugget; /* nugget model: nugget estimated (default 1) */
  double sigma2;  /* value when !est_sigma2 */
  double nugget;  /* value when !est_nugget */
  double alpha;   /* member alpha, used when gamma has only d entries */
} lkgpu_model_flags;

/* Evaluate nb log-likelihoods in one batched pipeline.  Item i evaluates
 * design problem_ids[i] at gamma + i * gamma_n, where gamma_n is d
 * (theta) or d + 1 ([theta, alpha], nugget model), with flags[i] (or
 * flags[0] when flags_n == 1).  Outputs: ll_out[nb], grad_out[nb * gamma_n]
 * (when return_grad), status_out[nb] (LKGPU_* codes), inc_out[nb] (jitter
 * steps used).  status_out and inc_out may be NULL.
 * Replaces Kriging::logLikelihoodFun (Kriging.cpp:343-349) called nb times. */
int lkgpu_ll_grad_batch(lkgpu_context* ctx, int nb, const int* problem_ids, const double* gamma, int gamma_n,
                        const lkgpu_model_flags* flags, int flags_n, int return_grad, double* ll_out,
                        double* grad_out, int* status_out, int* inc_out);

/* lkgpu_ll_grad_batch without allocating.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code around HIP graph
 * capture, tight serving loops, harnesses that forbid allocation).  Same
 * arguments, results and retry ladder as lkgpu_ll_grad_batch; the designs
 * are the ones lkgpu_problem_create uploaded beforehand.  The slot tables and
 * every item's workspace (2 npad^2 + npad (2 p + 13) doubles plus the
 * reduction partials, npad = n rounded up to 64) live in a static device
 * array of 64 MiB per GPU, held until the process ends; a batch that needs
 * more, or has more than 64 items, fails with -1 and a message naming the
 * sizes (e.g. one item up to n = 1984, 48 items of n = 256).  The call makes
 * no device or pinned-host allocation and instantiates no graph: launches
 * go directly to the context's stream.  Every call on one GPU uses that
 * GPU's array, which keeps the last call's contents: calls that overlap on
 * one GPU (two threads, two contexts) race.  Nothing in the array needs
 * clearing: every value is written before it is read.  lkgpu_read_item
 * reads the items of the last call, whichever entry point made it. */
int lkgpu_ll_grad_batch_noalloc(lkgpu_context* ctx, int nb, const int* problem_ids, const double* gamma,
                                int gamma_n, const lkgpu_model_flags* flags, int flags_n, int return_grad,
                                double* ll_out, double* grad_out, int* status_out, int* inc_out);

/* Copy a factor of batch item `item` of the last lkgpu_ll_grad_batch call
 * to host memory (column-major, leading dimension n): the KModel fields of
 * the reference (src/lib/include/libKriging/KrigingImpl.hpp:26-38).
 * L and L^-1 come back lower triangular with zeros above the diagonal. */
#define LKGPU_READ_L 0     /* n x n   KModel.L */
#define LKGPU_READ_LINV 1  /* n x n   L^-1 */
#define LKGPU_READ_FSTAR 2 /* n x p   KModel.Fstar */
#define LKGPU_READ_YSTAR 3 /* n       KModel.ystar */
#define LKGPU_READ_ESTAR 4 /* n       KModel.Estar */
#define LKGPU_READ_X 5     /* n       L^-T E* = R^-1 (y - F beta) */
#define LKGPU_READ_RSTAR 6 /* p x p   KModel.Rstar (upper) */
/* after a leave-one-out evaluation (lk_kriging_leave_one_out_fun / _vec):
 * errorsLOO and sigma2LOO of Kriging::_leaveOneOut (Kriging.cpp:409-412) */
#define LKGPU_READ_LOO_ERRORS 7 /* n */
#define LKGPU_READ_LOO_SIGMA2 8 /* n */
int lkgpu_read_item(lkgpu_context* ctx, int item, int what, double* out);

/* Live timing of one pipeline stage: a HIP event pair is recorded on the
 * pipeline stream around the stage (inside the replayed graph), and every
 * batch call adds its elapsed time.  lkgpu_take_stage_time returns the mean
 * per batch call since the last take (and resets).  No reference
 * counterpart: the reference's Bench::tic/toc (src/lib/Bench.cpp:16-35)
 * times host steps. */
#define LKGPU_STAGE_NONE -1
#define LKGPU_STAGE_BUILD 0         /* covariance build (one launch) */
#define LKGPU_STAGE_CHOLESKY 1      /* all Cholesky launches */
#define LKGPU_STAGE_UPDATE0 2       /* first trailing update (one launch) */
#define LKGPU_STAGE_TRTRI 3         /* all L^-1 launches */
#define LKGPU_STAGE_SOLVES 4        /* GLS solves, LL, condition check */
#define LKGPU_STAGE_ADJOINT 5       /* fused R^-1 + adjoint gradient (one launch) */
int lkgpu_set_timed_stage(lkgpu_context* ctx, int stage);
int lkgpu_take_stage_time(lkgpu_context* ctx, double* avg_ms, int* count);

/* Wait for all work on the context's device (hipDeviceSynchronize). */
int lkgpu_synchronize(lkgpu_context* ctx);

/* Use (1) or bypass (0) hipGraph replay of the evaluation pipeline. */
int lkgpu_set_graph(lkgpu_context* ctx, int enable);

/* Most streams (1, 2 or 4; default 4, or LKGPU_STREAMS) an evaluation splits
 * its batch over (no reference counterpart): a batch of B slots runs as S
 * equal parts on S streams, S the largest allowed value with B a multiple of
 * 8 S, so that latency-bound launches of one part overlap the matrix-core
 * launches of the others.  Results do not depend on S (each slot's launches
 * are the same).  Runs with a timed stage use one stream. */
int lkgpu_set_streams(lkgpu_context* ctx, int streams);

/* Look-ahead Cholesky for a one-item evaluation of at least 128 64-row
 * blocks (n > 8128), on (1, the default; LKGPU_LOOKAHEAD=0 turns it off) or
 * off (0); no reference counterpart.  After panel p's update of panel
 * p + 1's columns, panel p + 1's diagonal and panel launches run beside
 * panel p's update of all later columns, which runs on a stream restricted
 * to all CUs but 32 (LKGPU_LA_RESERVE) so that the chain always finds free
 * CUs; every tile still takes the panels' updates in panel order, so
 * results are the same either way.  Such an evaluation is issued without
 * hipGraph replay (a stream's CU mask does not carry into a graph). */
int lkgpu_set_lookahead(lkgpu_context* ctx, int enable);

/* Size-bucket launch table (no reference counterpart).  json_text:
 *   {"buckets": [{"max_nblk": B, "panel_blocks": P, "streams": S}, ...]}
 * with B the bucket's largest matrix in 64-row blocks (the last bucket is
 * unbounded), P the Cholesky panel width in block columns and S the parts a
 * batch of that bucket alone is split into (1, 2, 4).  NULL restores the
 * defaults (P = 4, S = 4).  A context loads tuning_gfx950.json from the
 * library's directory at creation (or the file LKGPU_TUNING names; "none"
 * skips it), the table libkriging_amd/autotune.py writes.  A batch whose
 * slots span several buckets runs one part per bucket, each on its own
 * stream with launches sized to its own largest matrix; results do not
 * depend on the table (each slot's arithmetic is the same).
 * lkgpu_last_parts returns the parts of the last evaluation. */
int lkgpu_set_tuning(lkgpu_context* ctx, const char* json_text);
int lkgpu_last_parts(lkgpu_context* ctx);


/* ------------------------------------------------------------------------ */
/* Layer 2: the Kriging object (libkriging_c.h names and conventions)       */
/* Supported here: kernels gauss/exp/matern3_2/matern5_2, noise "none" or   */
/* "nugget", regmodel none/constant/linear, objective "LL", optim "none" or */
/* "BFGS[#]".  Other values fail with -1 / NULL and a message.              */
/* ------------------------------------------------------------------------ */

/* libkriging_c.h:19 */
void* lk_kriging_new(const char* kernel, const char* noise_model);
/* libkriging_c.h:20-43.  theta holds theta_n = rows x d values (row-major);
 * rows > 1 gives several BFGS starting points (the reference C ABI passes one
 * row). noise / noise_n must be NULL / 0 (no heterogeneous noise). */
void* lk_kriging_new_fit(const double* y, int n, const double* noise, int noise_n, const double* X, int nX, int d,
                         const char* kernel, const char* noise_model, const char* regmodel, int normalize,
                         const char* optim, const char* objective, const double* sigma2, int is_sigma2_estim,
                         const double* theta, int theta_n, int is_theta_estim, const double* beta, int beta_n,
                         int is_beta_estim, const double* nugget, int is_nugget_estim);
/* libkriging_c.h:44 */
void lk_kriging_delete(void* ptr);
/* libkriging_c.h:47-58 */
int lk_kriging_fit(void* ptr, const double* y, int n, const double* noise, int noise_n, const double* X, int nX, int d,
                   const char* regmodel, int normalize, const char* optim, const char* objective);
/* fit with starting / fixed parameters, as the Python binding's
 * fit(..., parameters=dict) (Kriging_binding.cpp:39-48); no counterpart in
 * libkriging_c.h, which only offers them through lk_kriging_new_fit */
int lk_kriging_fit_parameters(void* ptr, const double* y, int n, const double* X, int nX, int d, const char* regmodel,
                              int normalize, const char* optim, const char* objective, const double* sigma2,
                              int is_sigma2_estim, const double* theta, int theta_n, int is_theta_estim,
                              const double* beta, int beta_n, int is_beta_estim, const double* nugget,
                              int is_nugget_estim);
/* libkriging_c.h:60-71 (Kriging::predict, Kriging.cpp:2241-2285): mean,
 * stdev, the m x m covariance (at most 32768 points per call) and the
 * m x d derivatives of mean and stdev, column-major */
int lk_kriging_predict(void* ptr, const double* X_n, int m, int d, int return_stdev, int return_cov, int return_deriv,
                       double* mean_out, double* stdev_out, double* cov_out, double* mean_deriv_out,
                       double* stdev_deriv_out);
/* libkriging_c.h:109-116 (Kriging::logLikelihoodFun, Kriging.cpp:343-349);
 * return_hess is rejected */
int lk_kriging_log_likelihood_fun(void* ptr, const double* theta, int theta_n, int return_grad, int return_hess,
                                  double* ll_out, double* grad_out, double* hess_out);
/* nb parameter vectors (rows of theta, each theta_n long) in one batched
 * pipeline; no reference counterpart (the reference evaluates one at a time) */
int lk_kriging_log_likelihood_fun_batch(void* ptr, int nb, const double* theta, int theta_n, int return_grad,
                                        double* ll_out, double* grad_out);
/* libkriging_c.h:131 */
double lk_kriging_log_likelihood(void* ptr);
/* libkriging_c.h:118-123 (Kriging::leaveOneOutFun, Kriging.cpp:470-476);
 * theta has d entries; noise model none only */
int lk_kriging_leave_one_out_fun(void* ptr, const double* theta, int theta_n, int return_grad, double* loo_out,
                                 double* grad_out);
/* libkriging_c.h:125-130 (Kriging::logMargPostFun, Kriging.cpp:652-660):
 * theta has d entries (d + 1 with alpha for the nugget model) */
int lk_kriging_log_marg_post_fun(void* ptr, const double* theta, int theta_n, int return_grad, double* lmp_out,
                                 double* grad_out);
/* libkriging_c.h:134 (Kriging::logMargPost at the fitted parameters) */
double lk_kriging_log_marg_post(void* ptr);
/* libkriging_c.h:83-91 (Kriging::update, Kriging.cpp:2433-2633): n new
 * points (X_u n x d column-major); noise_u must be empty (no heterogeneous
 * noise on this path) */
int lk_kriging_update(void* ptr, const double* y_u, int n, const double* noise_u, int noise_n, const double* X_u,
                      int nX, int d, int refit);
/* libkriging_c.h:105-106 (Kriging::save / Kriging::load, Kriging.cpp:
 * 2691-2750): the reference's JSON version-2 model file; load returns a new
 * handle (NULL on error) whose factors are recomputed on the device */
int lk_kriging_save(void* ptr, const char* filename);
void* lk_kriging_load(const char* filename);
/* libkriging_c.h:43 (Kriging(const Kriging&, ExplicitCopySpecifier),
 * Kriging.cpp:164): a new independent handle with the same state (NULL on
 * error) */
void* lk_kriging_copy(void* ptr);
/* libkriging_c.h:107 (Kriging::summary, Kriging.cpp:2660-2670) */
const char* lk_kriging_summary(void* ptr);
/* libkriging_c.h:138 (KrigingImpl::covMat, KrigingImpl.cpp:60-71):
 * cov_out (n1 x n2, column-major) = sigma2 k(x1_i - x2_j) of the normalised
 * inputs, computed on the device */
int lk_kriging_cov_mat(void* ptr, const double* X1, int n1, int d1, const double* X2, int n2, int d2, double* cov_out);
/* libkriging_c.h:133 (Kriging::leaveOneOut at the fitted theta) */
double lk_kriging_leave_one_out(void* ptr);
/* libkriging_c.h:136 (Kriging::leaveOneOutVec, Kriging.cpp:478-485): n
 * leave-one-out predictions and standard errors */
int lk_kriging_leave_one_out_vec(void* ptr, const double* theta, int theta_n, double* yhat_out, double* stderr_out);

/* getters, libkriging_c.h:141-166 (normalised-space members) */
const char* lk_kriging_kernel(void* ptr);
const char* lk_kriging_optim(void* ptr);
const char* lk_kriging_objective(void* ptr);
int lk_kriging_is_normalize(void* ptr);
const char* lk_kriging_regmodel(void* ptr);
const char* lk_kriging_noise_model(void* ptr);
int lk_kriging_get_X(void* ptr, double* out, int* n, int* d);
int lk_kriging_get_centerX(void* ptr, double* out, int* d);
int lk_kriging_get_scaleX(void* ptr, double* out, int* d);
int lk_kriging_get_y(void* ptr, double* out, int* n);
double lk_kriging_get_centerY(void* ptr);
double lk_kriging_get_scaleY(void* ptr);
int lk_kriging_get_beta(void* ptr, double* out, int* n);
int lk_kriging_get_theta(void* ptr, double* out, int* n);
double lk_kriging_get_sigma2(void* ptr);
double lk_kriging_get_nugget(void* ptr);
/* libkriging_c.h:152-165: F (n x p trend matrix), T (n x n lower Cholesky
 * factor of R), M = T^-1 F (n x p) and z (n) of the committed model -- T, M,
 * z from one device evaluation at the committed parameters -- the
 * estimation flags, and the heterogeneous noise vector (always empty here:
 * the noise models of this path are none and nugget) */
int lk_kriging_get_F(void* ptr, double* out, int* n, int* d);
int lk_kriging_get_T(void* ptr, double* out, int* n, int* d);
int lk_kriging_get_M(void* ptr, double* out, int* n, int* d);
int lk_kriging_get_z(void* ptr, double* out, int* n);
int lk_kriging_is_beta_estim(void* ptr);
int lk_kriging_is_theta_estim(void* ptr);
int lk_kriging_is_sigma2_estim(void* ptr);
int lk_kriging_is_nugget_estim(void* ptr);
int lk_kriging_get_noise(void* ptr, double* out, int* n);

/* evaluations and batched launches of the last fit; no reference counterpart */
int lk_kriging_fit_statistics(void* ptr, int* evaluations, int* batches);
/* 1: the fit evaluates all pending starts in one batched launch (default);
 * 0: one launch sequence per evaluation (for comparison); no reference
 * counterpart */
int lk_kriging_set_batched_fit(void* ptr, int batched);

/* Process-global settings.  lk_optim_set/get cover the Optim statics
 * (src/lib/Optim.cpp:39-141: reparametrize, theta_lower_factor,
 * theta_upper_factor, variogram_bounds_heuristic, log_level, max_restart,
 * max_iteration, gradient_tolerance, objective_rel_tolerance) plus
 * lbfgs_memory; lk_linalg_set covers the LinearAlgebra statics of the
 * default device (num_nugget, chol_rcond_check, max_inc_choldiag, min_rcond;
 * LinearAlgebra.cpp:33-100); lk_random_reset_seed is Random::reset_seed
 * (Random.cpp:17-20). */
int lk_optim_set(const char* name, double value);
int lk_optim_get(const char* name, double* value);
int lk_linalg_set(const char* name, double value);
/* the current value of a LinearAlgebra setting (as lk_linalg_set names
 * them): get_num_nugget, chol_rcond_checked (LinearAlgebra.cpp:37-62) */
int lk_linalg_get(const char* name, double* value);

/* The public LinearAlgebra routines the R binding exports
 * (bindings/R/rlibkriging/src/linalg.cpp:39-71), on the default device.
 * Matrices are n x n column-major host buffers.
 * lk_linalg_chol_safe: LinearAlgebra::safe_chol_lower (LinearAlgebra.cpp:
 *   43-98): L (lower, zero above) of X + the retry ladder's cumulative
 *   jitter (checked against rcond when chol_rcond_check is on); inc_out (may
 *   be NULL) receives the jitter steps; -1 with the reference's messages
 *   when the ladder is exhausted or num_nugget is not positive.
 * lk_linalg_rcond_chol: LinearAlgebra::rcond_chol (LinearAlgebra.cpp:
 *   106-115), rcond(L)^2 by the 1-norm estimator of LAPACK's dtrcon, for
 *   the lower triangle of L (0 for a zero diagonal entry).
 * lk_linalg_rcond_approx_chol: LinearAlgebra::rcond_approx_chol
 *   (LinearAlgebra.cpp:120-139), (min / max of L's diagonal)^2.
 * lk_linalg_chol_block: LinearAlgebra::chol_block (LinearAlgebra.cpp:
 *   257-299): the factor of C from Loo (no x no), the factor of C's leading
 *   block (diagonal >= 0); L_ou = Loo^-1 C_ou, L_uu = safe_chol(C_uu -
 *   L_ou' L_ou), and safe_chol(C) when that fails. */
int lk_linalg_chol_safe(const double* X, int n, double* L_out, int* inc_out);
int lk_linalg_rcond_chol(const double* L, int n, double* rcond_out);
int lk_linalg_rcond_approx_chol(const double* L, int n, double* rcond_out);
int lk_linalg_chol_block(const double* C, int n, const double* Loo, int no, double* L_out);
int lk_random_reset_seed(unsigned seed);

/* The bound-constrained optimiser the fit uses, on a user objective
 * (replaces lbfgsb::Optimizer::minimize, src/lib/Kriging.cpp:1980-1987).
 * fn returns f(x) and writes g(x).  x is overwritten with the solution;
 * task_out receives the termination string. */
typedef double (*lk_objective_fn)(const double* x, double* g, void* user);
int lk_lbfgsb_minimize(int n, double* x, const double* lower, const double* upper, lk_objective_fn fn, void* user,
                       int max_iter, double pgtol, double factr, int m, double* f_out, int* iters_out,
                       int* evals_out, char* task_out, int task_len);

/* ======================================================================
 * Layer 3: NestedKriging, mirroring bindings/Julia/jlibkriging/csrc/
 * libkriging_c.h:394-456 (implementation libkriging_c.cpp:1663-1870).
 * Plain Kriging submodels only: n_warping must be 0 and the LLVecchia
 * objective is rejected (-1 with a message).  All groups' submodel fits run
 * in shared batched pipelines; the NK aggregation runs on the device.
 * Partitions: "kmeans" / "random" seeded by `seed` through std::mt19937
 * (reproducible per seed, not Armadillo's generator).
 * ====================================================================== */
void* lk_nested_kriging_new_fit(const double* y, int n, const double* X, int nX, int d, const char* kernel,
                                int nb_groups, const char* aggregation, const char* partition, int seed,
                                const char** warping, int n_warping, const char* regmodel, const char* optim,
                                const char* objective, const double* sigma2, int is_sigma2_estim, const double* theta,
                                int theta_n, int is_theta_estim, const double* beta, int beta_n, int is_beta_estim);
void lk_nested_kriging_delete(void* ptr);
int lk_nested_kriging_fit(void* ptr, const double* y, int n, const double* X, int nX, int d, int nb_groups,
                          const char** warping, int n_warping, const char* regmodel, const char* optim,
                          const char* objective, const double* sigma2, int is_sigma2_estim, const double* theta,
                          int theta_n, int is_theta_estim, const double* beta, int beta_n, int is_beta_estim);
int lk_nested_kriging_predict(void* ptr, const double* X_n, int m, int d, int return_stdev, double* mean_out,
                              double* stdev_out);
const char* lk_nested_kriging_summary(void* ptr);
const char* lk_nested_kriging_kernel(void* ptr);
const char* lk_nested_kriging_aggregation(void* ptr);
int lk_nested_kriging_nb_groups(void* ptr);
int lk_nested_kriging_get_theta(void* ptr, double* out, int* n);
double lk_nested_kriging_get_sigma2(void* ptr);
double lk_nested_kriging_get_beta0(void* ptr);

/* Extensions with no counterpart in the reference's C ABI:
 * lk_nested_kriging_new is the NestedKriging(covType) constructor
 * (NestedKriging.hpp:62) plus the aggregation / partition / seed options;
 * lk_nested_kriging_fit_groups fits on caller-given groups (group_of[i] in
 * 0..nb_groups-1, every group non-empty), e.g. uneven group sizes;
 * lk_nested_kriging_groups writes the group index of every point;
 * lk_nested_kriging_submodel returns a borrowed lk_kriging_* handle of
 * submodel i (valid until the next fit or delete; never delete it);
 * lk_nested_kriging_set_predict_chunk is NestedKriging::set_predict_chunk
 * (0 = sized from a device-memory budget); lk_nested_kriging_predict_fallbacks
 * counts the points of the last NK predict whose aggregation system was not
 * positive definite (solved by LU, or by the pseudo-inverse when singular). */
void* lk_nested_kriging_new(const char* kernel, const char* aggregation, const char* partition, int seed);
int lk_nested_kriging_fit_groups(void* ptr, const double* y, int n, const double* X, int nX, int d,
                                 const int* group_of, int nb_groups, const char* regmodel, const char* optim,
                                 const char* objective, const double* sigma2, int is_sigma2_estim,
                                 const double* theta, int theta_n, int is_theta_estim, const double* beta,
                                 int beta_n, int is_beta_estim);
int lk_nested_kriging_groups(void* ptr, int* group_of_out);
void* lk_nested_kriging_submodel(void* ptr, int i);
int lk_nested_kriging_set_predict_chunk(void* ptr, int chunk);
int lk_nested_kriging_predict_fallbacks(void* ptr);
/* LL+grad evaluations and batched pipelines of the groups' optimisation in
 * the last fit (all groups share each pipeline); no reference counterpart. */
int lk_nested_kriging_fit_statistics(void* ptr, int* evaluations, int* batches);
/* The groups' own fitted hyper-parameters before unification
 * (NestedKriging.cpp:308-322 combines them): theta_out nb_groups x d
 * (group-major: group g at theta_out[g d ..]), sigma2_out and beta0_out
 * nb_groups each (beta0 0 without a constant trend); any may be NULL.  No
 * reference counterpart: the input of the unification, for checking it. */
int lk_nested_kriging_group_fits(void* ptr, double* theta_out, double* sigma2_out, double* beta0_out);
/* HIP-event timing of the NK pair-contraction launches (k_nk_pair) of the
 * following predicts (no reference counterpart: the kernel's roofline
 * figure); lk_nested_kriging_pair_time returns the milliseconds, launches
 * and points accumulated since the previous call and resets them. */
int lk_nested_kriging_time_pairs(void* ptr, int enable);
int lk_nested_kriging_pair_time(void* ptr, double* ms_out, int* launches_out, int* points_out);

#ifdef __cplusplus
}
#endif
#endif

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

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread
 * (replaces libkriging_c.h:10 lk_get_last_error). */
const char* lk_get_last_error(void);

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
int lkgpu_read_item(lkgpu_context* ctx, int item, int what, double* out);

/* Use (1) or bypass (0) hipGraph replay of the evaluation pipeline. */
int lkgpu_set_graph(lkgpu_context* ctx, int enable);

#ifdef __cplusplus
}
#endif
#endif

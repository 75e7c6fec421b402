/*
 * lk_oracle.h -- CPU restatement of libKriging's likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (libkriging_amd/) may
 * link, load or call this code.  It is used by tests/ as the checker, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg
 * (kind "port") as the timed CPU path.
 *
 * All matrices are FP64 column-major (rows = observations), as in the
 * reference (AGENTS.md:66-71).  Every function cites the reference file:line
 * it restates.  Parity pins are listed in the header comment of
 * oracle/lk_oracle.c.
 */
#ifndef LK_ORACLE_H
#define LK_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

enum { LKO_GAUSS = 0, LKO_EXP = 1, LKO_MATERN32 = 2, LKO_MATERN52 = 3 };
enum { LKO_NOISE_NONE = 0, LKO_NOISE_NUGGET = 1 };

/* LinearAlgebra statics (src/lib/LinearAlgebra.cpp:33,53,63,100) */
typedef struct {
  double num_nugget; /* 1e-10 */
  int rcond_check;   /* 1 */
  int max_inc;       /* 10 */
  double min_rcond;  /* 1e-18 */
} lko_chol_cfg;

/* One Kriging problem: the state KrigingImpl keeps after fit_setup_impl
 * (KrigingImpl.cpp:764-841) plus the estimation flags of Kriging::fit. */
typedef struct {
  int n, d, p;
  const double* X; /* n x d */
  const double* y; /* n */
  const double* F; /* n x p */
  int kernel;
  int noise;       /* LKO_NOISE_* */
  int est_sigma2;
  int est_nugget;
  int est_beta;
  double sigma2;   /* used when !est_sigma2 */
  double nugget;   /* used when !est_nugget */
  double alpha;    /* member alpha, used when gamma has only d entries */
} lko_problem;

/* Outputs of one populate_Model call (KrigingImpl.hpp:26-38 KModel). All
 * pointers may be NULL when not wanted, except where stated. */
typedef struct {
  double* R;     /* n x n, un-jittered covariance (KModel.R) */
  double* L;     /* n x n, lower factor of the jittered R */
  double* Rinv;  /* n x n */
  double* Fstar; /* n x p */
  double* ystar; /* n */
  double* Rstar; /* p x p upper */
  double* Estar; /* n */
  double* betahat; /* p */
  double SSEstar;
  int inc_cond;  /* jitter steps used by the retry ladder */
  double rcond;  /* rcond(L)^2 of the accepted factor */
} lko_model;

double lko_cov(int kernel, const double* dx, const double* theta, int d);
void lko_dlncov_dtheta(int kernel, const double* dx, const double* theta, int d, double* out);

void lko_build_R(int kernel, const double* X, int n, int d, const double* theta, double factor,
                 const double* diag, double* R);
/* NestedKriging::corrMat: out (n1 x n2) = k(x1_i - x2_j) */
void lko_corr_mat(int kernel, const double* X1, int n1, const double* X2, int n2, int d, const double* theta,
                  double* out);

int lko_potrf_lower(double* A, int n);
double lko_trcon_lower(const double* L, int n);
double lko_rcond_chol(const double* L, int n);
int lko_safe_chol_lower(const double* R, int n, const lko_chol_cfg* cfg, double* L, int* inc_out,
                        double* rcond_out);
void lko_inv_sympd(const double* L, int n, double* Rinv);
void lko_solve_lower(const double* L, int n, double* B, int nrhs);
void lko_solve_upper_t(const double* L, int n, double* B, int nrhs);

int lko_populate_model(const lko_problem* P, const double* theta, double alpha,
                       const lko_chol_cfg* cfg, lko_model* m);

int lko_log_likelihood(const lko_problem* P, const double* gamma, int gamma_n, int want_grad,
                       const lko_chol_cfg* cfg, double* ll_out, double* grad_out, int* inc_out);

void lko_theta_bounds(const double* X, const double* y, int n, int d, double lower_factor,
                      double upper_factor, int variogram, double* lower, double* upper);

int lko_predict(const lko_problem* P, const double* theta, double alpha, double sigma2,
                const lko_chol_cfg* cfg, const double* Xn, int m, double* mean, double* stdev);

const char* lko_last_error(void);

/* Route the factorisation, triangular solves and condition estimate through
 * LAPACK's dpotrf / dtrtrs / dtrcon (Fortran calling convention without
 * hidden string lengths); NULLs restore the restated loops. */
void lko_set_lapack(void* dpotrf, void* dtrtrs, void* dtrcon);

#ifdef __cplusplus
}
#endif
#endif

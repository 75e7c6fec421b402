/*
 * lk_oracle.c -- CPU restatement of libKriging's likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py.  The product never links it.
 *
 * What it restates (reference = libKriging v1.1.0):
 *   kernels + d ln k / d theta   src/lib/Covariance.cpp:24-179
 *   R build                      src/lib/LinearAlgebra.cpp:142-204 (cholCov)
 *   jitter-retry Cholesky        src/lib/LinearAlgebra.cpp:43-98
 *   rcond of the factor          src/lib/LinearAlgebra.cpp:106-115 (arma::rcond -> LAPACK dtrcon '1')
 *   R^-1                         src/lib/LinearAlgebra.cpp:708-710 (inv_sympd)
 *   GLS model                    src/lib/KrigingImpl.cpp:73-125 (populate_Model)
 *   LL value + gradient          src/lib/Kriging.cpp:214-338 (_logLikelihood)
 *   gradient pair loop           src/lib/KrigingImpl.cpp:855-885
 *   theta bounds                 src/lib/Optim.cpp:179-209
 *   predict (mean, stdev)        src/lib/KrigingImpl.cpp:145-243 + Kriging.cpp:2241-2285
 *
 * How each part is pinned (tests/test_oracle.py):
 *   - LL and dLL/dtheta (gauss, constant trend, no nugget): the reference's own
 *     golden vectors tests/golden/data1-scal-* and data2-grad-{1..10}-* at 1e-12.
 *   - kernels: closed forms at special points (dx = 0 -> 1, gauss/exp
 *     factorisation over dimensions, Matern product form), and central finite
 *     differences of ln k for the theta derivatives.
 *   - Cholesky, solves, inverse: LAPACK dpotrf/dtrtrs through scipy on random
 *     SPD matrices; L L^T = R; R R^-1 = I.
 *   - rcond: LAPACK dtrcon('1','L','N') through scipy (the routine
 *     arma::rcond calls for a triangular matrix) on random factors.
 *   - jitter ladder: a singular R (duplicated design rows) whose accepted
 *     factor must equal chol(R + sum_k 1e-10*10^k I) for the reported k.
 *   - nugget alpha gradient, other kernels, linear trend, fixed sigma2:
 *     central finite differences of this same LL (the reference's own
 *     NuggetKrigingLogLikTest.cpp:11-191 check) plus an independent numpy
 *     formulation of the LL.
 *   - theta bounds: a numpy restatement on small n.
 *   - predict: interpolation at design points (mean = y, stdev = 0) and an
 *     independent numpy formulation.
 */
#define _GNU_SOURCE
#include "lk_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];
const char* lko_last_error(void) { return g_err; }

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------ */
/* Kernels: Covariance.cpp:24-36 (gauss), 65-75 (exp), 104-116 (matern3_2),  */
/* 149-161 (matern5_2).  Same operation order as the reference.              */
/* ------------------------------------------------------------------------ */
double lko_cov(int kernel, const double* dx, const double* theta, int d) {
  const double s3 = sqrt(3.0), s5 = sqrt(5.0);
  double s = 0.0;
  switch (kernel) {
    case LKO_GAUSS:
      for (int k = 0; k < d; k++) {
        double v = dx[k] / theta[k];
        s += v * v;
      }
      return exp(-0.5 * s);
    case LKO_EXP:
      for (int k = 0; k < d; k++) s += fabs(dx[k] / theta[k]);
      return exp(-s);
    case LKO_MATERN32:
      for (int k = 0; k < d; k++) {
        double u = s3 * fabs(dx[k] / theta[k]);
        s += u - log1p(u);
      }
      return exp(-s);
    case LKO_MATERN52:
      for (int k = 0; k < d; k++) {
        double u = s5 * fabs(dx[k] / theta[k]);
        s += u - log1p(u + (u * u) / 3.0);
      }
      return exp(-s);
  }
  return NAN;
}

/* d ln k / d theta_k: Covariance.cpp:38-50, 77-88, 118-130, 163-179 */
void lko_dlncov_dtheta(int kernel, const double* dx, const double* theta, int d, double* out) {
  const double s3 = sqrt(3.0), s5 = sqrt(5.0);
  for (int k = 0; k < d; k++) {
    double t = theta[k];
    switch (kernel) {
      case LKO_GAUSS:
        out[k] = (dx[k] * dx[k]) / (t * t * t);
        break;
      case LKO_EXP:
        out[k] = fabs(dx[k]) / (t * t);
        break;
      case LKO_MATERN32: {
        double u = s3 * fabs(dx[k] / t);
        out[k] = (u * u) / (1.0 + u) / t;
        break;
      }
      case LKO_MATERN52: {
        double u = s5 * fabs(dx[k] / t);
        double a = 1.0 + u;
        double b = (u * u) / 3.0;
        out[k] = (a * b) / (a + b) / t;
        break;
      }
      default:
        out[k] = NAN;
    }
  }
}

/* R(i,j) = factor * k(x_j - x_i) for j < i (the dX column i*n+j holds
 * x_j - x_i, LinearAlgebra.cpp:457-468), symmetric; the whole matrix is
 * scaled by `factor` before the diagonal is set to 1 or `diag`
 * (LinearAlgebra.cpp:151-157,190-194). */
void lko_build_R(int kernel, const double* X, int n, int d, const double* theta, double factor,
                 const double* diag, double* R) {
  double dx[64];
  double* dxp = d <= 64 ? dx : (double*)malloc(sizeof(double) * d);
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < i; j++) {
      for (int k = 0; k < d; k++) dxp[k] = X[j + (size_t)k * n] - X[i + (size_t)k * n];
      double c = lko_cov(kernel, dxp, theta, d) * factor;
      R[i + (size_t)j * n] = c;
      R[j + (size_t)i * n] = c;
    }
  }
  for (int i = 0; i < n; i++) R[i + (size_t)i * n] = diag ? diag[i] : 1.0;
  if (dxp != dx) free(dxp);
}

/* Cross-correlation matrix out[i + j*n1] = k(x1_i - x2_j; theta), as
 * NestedKriging::corrMat (src/lib/NestedKriging.cpp:105-118): the kernel on
 * the difference vector, no coincident-point special case.  X1 is n1 x d and
 * X2 n2 x d, both column-major. */
void lko_corr_mat(int kernel, const double* X1, int n1, const double* X2, int n2, int d, const double* theta,
                  double* out) {
  double dx[64];
  double* dxp = d <= 64 ? dx : (double*)malloc(sizeof(double) * d);
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n1; i++) {
      for (int k = 0; k < d; k++) dxp[k] = X1[i + (size_t)k * n1] - X2[j + (size_t)k * n2];
      out[i + (size_t)j * n1] = lko_cov(kernel, dxp, theta, d);
    }
  if (dxp != dx) free(dxp);
}

/* Optional LAPACK routines: the ones the reference reaches through
 * Armadillo (arma::chol -> dpotrf, arma::solve(trimatl / trimatu) ->
 * dtrtrs, arma::rcond -> dtrcon).  When lko_set_lapack has set them, the
 * factorisation, the triangular solves and the condition estimate call them
 * (the pair loops stay serial C, as in the reference); NULL restores the
 * restated loops.  The CPU baseline times the oracle this way. */
typedef void (*lko_dpotrf_t)(char*, int*, double*, int*, int*);
typedef void (*lko_dtrtrs_t)(char*, char*, char*, int*, int*, double*, int*, double*, int*, int*);
typedef void (*lko_dtrcon_t)(char*, char*, char*, int*, double*, int*, double*, double*, int*, int*);
static lko_dpotrf_t g_dpotrf = NULL;
static lko_dtrtrs_t g_dtrtrs = NULL;
static lko_dtrcon_t g_dtrcon = NULL;

void lko_set_lapack(void* dpotrf, void* dtrtrs, void* dtrcon) {
  g_dpotrf = (lko_dpotrf_t)dpotrf;
  g_dtrtrs = (lko_dtrtrs_t)dtrtrs;
  g_dtrcon = (lko_dtrcon_t)dtrcon;
}

/* Lower Cholesky in place (LAPACK dpotrf semantics: returns j+1 when the
 * j-th pivot is <= 0 or NaN, 0 on success; the strict upper triangle is set to
 * zero as arma::chol(L, X, "lower") returns a triangular matrix). */
int lko_potrf_lower(double* A, int n) {
  if (g_dpotrf) {
    char uplo = 'L';
    int info = 0, nn = n;
    g_dpotrf(&uplo, &nn, A, &nn, &info);
    if (info != 0) return info;
    for (int j = 1; j < n; j++)
      for (int i = 0; i < j; i++) A[i + (size_t)j * n] = 0.0;
    return 0;
  }
  for (int j = 0; j < n; j++) {
    double* cj = A + (size_t)j * n;
    for (int k = 0; k < j; k++) {
      const double* ck = A + (size_t)k * n;
      double ljk = ck[j];
      for (int i = j; i < n; i++) cj[i] -= ck[i] * ljk;
    }
    double ajj = cj[j];
    if (!(ajj > 0.0)) return j + 1;
    ajj = sqrt(ajj);
    cj[j] = ajj;
    double r = 1.0 / ajj;
    for (int i = j + 1; i < n; i++) cj[i] *= r;
  }
  for (int j = 1; j < n; j++)
    for (int i = 0; i < j; i++) A[i + (size_t)j * n] = 0.0;
  return 0;
}

/* B <- L^-1 B (forward substitution, L lower n x n, B n x nrhs). */
void lko_solve_lower(const double* L, int n, double* B, int nrhs) {
  if (g_dtrtrs) {
    char uplo = 'L', trans = 'N', diag = 'N';
    int nn = n, nr = nrhs, info = 0;
    if (nr > 0) g_dtrtrs(&uplo, &trans, &diag, &nn, &nr, (double*)L, &nn, B, &nn, &info);
    return;
  }
  for (int c = 0; c < nrhs; c++) {
    double* b = B + (size_t)c * n;
    for (int j = 0; j < n; j++) {
      const double* lj = L + (size_t)j * n;
      double v = b[j] / lj[j];
      b[j] = v;
      if (v != 0.0)
        for (int i = j + 1; i < n; i++) b[i] -= lj[i] * v;
    }
  }
}

/* B <- L^-T B (back substitution with the upper factor L^T). */
void lko_solve_upper_t(const double* L, int n, double* B, int nrhs) {
  if (g_dtrtrs) {
    char uplo = 'L', trans = 'T', diag = 'N';
    int nn = n, nr = nrhs, info = 0;
    if (nr > 0) g_dtrtrs(&uplo, &trans, &diag, &nn, &nr, (double*)L, &nn, B, &nn, &info);
    return;
  }
  for (int c = 0; c < nrhs; c++) {
    double* b = B + (size_t)c * n;
    for (int i = n - 1; i >= 0; i--) {
      const double* li = L + (size_t)i * n;
      double s = b[i];
      for (int k = i + 1; k < n; k++) s -= li[k] * b[k];
      b[i] = s / li[i];
    }
  }
}

/* x <- L^-T x for a single vector (used by the estimator). */
static void solve_lt_vec(const double* L, int n, double* x) { lko_solve_upper_t(L, n, x, 1); }
static void solve_l_vec(const double* L, int n, double* x) { lko_solve_lower(L, n, x, 1); }

static double sign1(double v) { return v >= 0.0 ? 1.0 : -1.0; }

/* 1-norm estimate of L^-1 by the Hager / Higham iteration (N. J. Higham,
 * ACM TOMS 14 (1988) 381-396, Algorithm 4.1, with the alternating-sign
 * extra estimate), the estimator LAPACK dtrcon uses for rcond in the 1-norm. */
static double inv_norm1_estimate(const double* L, int n) {
  double* x = (double*)malloc(sizeof(double) * n);
  double* xi = (double*)malloc(sizeof(double) * n);
  double est;
  for (int i = 0; i < n; i++) x[i] = 1.0 / n;
  solve_l_vec(L, n, x);
  if (n == 1) {
    est = fabs(x[0]);
    free(x);
    free(xi);
    return est;
  }
  est = 0.0;
  for (int i = 0; i < n; i++) est += fabs(x[i]);
  for (int i = 0; i < n; i++) xi[i] = sign1(x[i]);
  memcpy(x, xi, sizeof(double) * n);
  solve_lt_vec(L, n, x);
  int j = 0;
  for (int i = 1; i < n; i++)
    if (fabs(x[i]) > fabs(x[j])) j = i;
  int iter = 2;
  for (;;) {
    memset(x, 0, sizeof(double) * n);
    x[j] = 1.0;
    solve_l_vec(L, n, x);
    double estold = est;
    est = 0.0;
    for (int i = 0; i < n; i++) est += fabs(x[i]);
    int changed = 0;
    for (int i = 0; i < n; i++)
      if (sign1(x[i]) != xi[i]) {
        changed = 1;
        break;
      }
    if (!changed || est <= estold) break;
    for (int i = 0; i < n; i++) xi[i] = sign1(x[i]);
    memcpy(x, xi, sizeof(double) * n);
    solve_lt_vec(L, n, x);
    int jlast = j;
    j = 0;
    for (int i = 1; i < n; i++)
      if (fabs(x[i]) > fabs(x[j])) j = i;
    if (x[jlast] != fabs(x[j]) && iter < 5) {
      iter++;
      continue;
    }
    break;
  }
  double altsgn = 1.0;
  for (int i = 0; i < n; i++) {
    x[i] = altsgn * (1.0 + (double)i / (double)(n - 1));
    altsgn = -altsgn;
  }
  solve_l_vec(L, n, x);
  double temp = 0.0;
  for (int i = 0; i < n; i++) temp += fabs(x[i]);
  temp = 2.0 * (temp / (double)(3 * n));
  if (temp > est) est = temp;
  free(x);
  free(xi);
  return est;
}

/* rcond of a lower triangular matrix in the 1-norm (dtrcon '1','L','N'). */
double lko_trcon_lower(const double* L, int n) {
  if (g_dtrcon && n > 0) {
    char norm = '1', uplo = 'L', diag = 'N';
    int nn = n, info = 0;
    double rc = 0.0;
    double* work = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    int* iwork = (int*)malloc(sizeof(int) * (size_t)n);
    g_dtrcon(&norm, &uplo, &diag, &nn, (double*)L, &nn, &rc, work, iwork, &info);
    free(work);
    free(iwork);
    return rc;
  }
  double anorm = 0.0;
  for (int j = 0; j < n; j++) {
    double s = 0.0;
    for (int i = j; i < n; i++) s += fabs(L[i + (size_t)j * n]);
    if (s > anorm || isnan(s)) anorm = s;
  }
  if (n == 0) return 1.0;
  if (anorm == 0.0) return 0.0;
  double ainv = inv_norm1_estimate(L, n);
  if (ainv == 0.0) return 0.0;
  return (1.0 / anorm) / ainv;
}

/* LinearAlgebra::rcond_chol (LinearAlgebra.cpp:106-115): rcond(L)^2. */
double lko_rcond_chol(const double* L, int n) {
  double r = lko_trcon_lower(L, n);
  return r * r;
}

/* LinearAlgebra::safe_chol_lower_retry (LinearAlgebra.cpp:68-98): accept the
 * factor iff potrf succeeds and (check off or rcond(L)^2 >= min_rcond), else
 * add num_nugget*10^k to the diagonal (cumulative over retries) and retry,
 * throwing after max_inc+1 attempts.  Returns 0 on success, -1 on failure. */
int lko_safe_chol_lower(const double* R, int n, const lko_chol_cfg* cfg, double* L, int* inc_out,
                        double* rcond_out) {
  size_t nn = (size_t)n * n;
  double* X = (double*)malloc(sizeof(double) * nn);
  memcpy(X, R, sizeof(double) * nn);
  int inc = 0;
  for (;;) {
    memcpy(L, X, sizeof(double) * nn);
    int ok = lko_potrf_lower(L, n) == 0;
    int wrong = cfg->rcond_check;
    double rc = NAN;
    if (ok) {
      rc = lko_rcond_chol(L, n);
      wrong = wrong && (rc < cfg->min_rcond);
    }
    if (!ok || wrong) {
      if (inc > cfg->max_inc) {
        snprintf(g_err, sizeof g_err, "Exceed max numerical nugget (%d x 1e%g) added to force chol matrix", inc,
                 log10(cfg->num_nugget));
        free(X);
        return -1;
      } else if (cfg->num_nugget <= 0.0) {
        snprintf(g_err, sizeof g_err, "Cannot add numerical nugget which is not strictly positive: %g",
                 cfg->num_nugget);
        free(X);
        return -1;
      }
      double add = cfg->num_nugget * pow(10.0, inc);
      for (int i = 0; i < n; i++) X[i + (size_t)i * n] += add;
      inc++;
      continue;
    }
    if (inc_out) *inc_out = inc;
    if (rcond_out) *rcond_out = rc;
    free(X);
    return 0;
  }
}

/* inv_sympd (LinearAlgebra.cpp:708-710): R^-1 = L^-T (L^-1 I). */
void lko_inv_sympd(const double* L, int n, double* Rinv) {
  size_t nn = (size_t)n * n;
  memset(Rinv, 0, sizeof(double) * nn);
  for (int i = 0; i < n; i++) Rinv[i + (size_t)i * n] = 1.0;
  lko_solve_lower(L, n, Rinv, n);
  lko_solve_upper_t(L, n, Rinv, n);
}

/* p x p upper Cholesky (chol_upper, LinearAlgebra.cpp:712-714): U^T U = A. */
static int chol_upper_small(const double* A, int p, double* U) {
  /* U = L^T with L the lower factor */
  double* Lw = (double*)malloc(sizeof(double) * p * p);
  memcpy(Lw, A, sizeof(double) * p * p);
  int info = lko_potrf_lower(Lw, p);
  for (int j = 0; j < p; j++)
    for (int i = 0; i < p; i++) U[i + j * p] = (i <= j) ? Lw[j + i * p] : 0.0;
  free(Lw);
  return info;
}

/* KrigingImpl::populate_Model (KrigingImpl.cpp:73-125) with the noise-model
 * dispatch of Kriging::populate_Model (Kriging.cpp:166-190): factor = alpha
 * for the nugget model (1 otherwise), unit diagonal. */
int lko_populate_model(const lko_problem* P, const double* theta, double alpha, const lko_chol_cfg* cfg,
                       lko_model* m) {
  const int n = P->n, p = P->p;
  size_t nn = (size_t)n * n;
  double factor = (P->noise == LKO_NOISE_NUGGET) ? alpha : 1.0;
  lko_build_R(P->kernel, P->X, n, P->d, theta, factor, NULL, m->R);
  if (lko_safe_chol_lower(m->R, n, cfg, m->L, &m->inc_cond, &m->rcond) != 0) return -1;
  if (m->Rinv) lko_inv_sympd(m->L, n, m->Rinv);
  memcpy(m->Fstar, P->F, sizeof(double) * n * p);
  lko_solve_lower(m->L, n, m->Fstar, p);
  memcpy(m->ystar, P->y, sizeof(double) * n);
  lko_solve_lower(m->L, n, m->ystar, 1);
  /* R* = chol_upper(F*^T F*) */
  double* G = (double*)calloc((size_t)p * p, sizeof(double));
  for (int a = 0; a < p; a++)
    for (int b = 0; b < p; b++) {
      double s = 0.0;
      for (int i = 0; i < n; i++) s += m->Fstar[i + (size_t)a * n] * m->Fstar[i + (size_t)b * n];
      G[a + b * p] = s;
    }
  if (p > 0 && chol_upper_small(G, p, m->Rstar) != 0) {
    free(G);
    snprintf(g_err, sizeof g_err, "chol(F*'F*) failed");
    return -1;
  }
  /* beta = R*^-1 (R*^-T (F*^T y*)) */
  double* bt = (double*)calloc((size_t)(p > 0 ? p : 1), sizeof(double));
  for (int a = 0; a < p; a++) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += m->Fstar[i + (size_t)a * n] * m->ystar[i];
    bt[a] = s;
  }
  for (int a = 0; a < p; a++) { /* forward with R*^T (lower) */
    double s = bt[a];
    for (int k = 0; k < a; k++) s -= m->Rstar[k + a * p] * bt[k];
    bt[a] = s / m->Rstar[a + a * p];
  }
  for (int a = p - 1; a >= 0; a--) { /* back with R* (upper) */
    double s = bt[a];
    for (int k = a + 1; k < p; k++) s -= m->Rstar[a + k * p] * bt[k];
    bt[a] = s / m->Rstar[a + a * p];
  }
  /* E* = L^-1 (y - F beta), SSE = E*'E* */
  for (int i = 0; i < n; i++) {
    double s = 0.0;
    for (int a = 0; a < p; a++) s += P->F[i + (size_t)a * n] * bt[a];
    m->Estar[i] = P->y[i] - s;
  }
  lko_solve_lower(m->L, n, m->Estar, 1);
  double sse = 0.0;
  for (int i = 0; i < n; i++) sse += m->Estar[i] * m->Estar[i];
  m->SSEstar = sse;
  for (int a = 0; a < p; a++) m->betahat[a] = P->est_beta ? bt[a] : 0.0;
  free(G);
  free(bt);
  (void)nn;
  return 0;
}

/* Kriging::_logLikelihood (Kriging.cpp:214-338) for the None and Nugget
 * noise models.  gamma = theta (d) or [theta, alpha] (d+1, nugget).  The
 * gradient has gamma_n entries. */
int lko_log_likelihood(const lko_problem* P, const double* gamma, int gamma_n, int want_grad,
                       const lko_chol_cfg* cfg, double* ll_out, double* grad_out, int* inc_out) {
  const int n = P->n, d = P->d, p = P->p;
  size_t nn = (size_t)n * n;
  const double* theta = gamma;
  double extra = (gamma_n > d) ? gamma[d] : P->alpha;
  if (P->noise == LKO_NOISE_NUGGET && !P->est_sigma2 && !P->est_nugget)
    extra = P->sigma2 / (P->sigma2 + P->nugget);

  lko_model m;
  memset(&m, 0, sizeof m);
  m.R = (double*)malloc(sizeof(double) * nn);
  m.L = (double*)malloc(sizeof(double) * nn);
  m.Rinv = want_grad ? (double*)malloc(sizeof(double) * nn) : NULL;
  m.Fstar = (double*)malloc(sizeof(double) * (size_t)n * (p > 0 ? p : 1));
  m.ystar = (double*)malloc(sizeof(double) * n);
  m.Rstar = (double*)malloc(sizeof(double) * (p > 0 ? p * p : 1));
  m.Estar = (double*)malloc(sizeof(double) * n);
  m.betahat = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  int rc = lko_populate_model(P, theta, extra, cfg, &m);
  if (rc != 0) goto done;
  if (inc_out) *inc_out = m.inc_cond;

  double logdet = 0.0;
  for (int i = 0; i < n; i++) logdet += log(m.L[i + (size_t)i * n]);
  double sigma2_grad, ll;
  if (P->noise == LKO_NOISE_NUGGET) {
    double a = extra, s2 = P->sigma2, nug = P->nugget;
    if (P->est_sigma2) {
      if (P->est_nugget) {
        double var = m.SSEstar / n;
        s2 = a * var;
        nug = (1.0 - a) * var;
      } else {
        s2 = P->nugget * a / (1.0 - a);
      }
    } else {
      if (P->est_nugget) nug = P->sigma2 * (1.0 - a) / a;
    }
    double tv = s2 + nug;
    ll = -0.5 * (n * log(2 * M_PI * tv) + 2 * logdet + m.SSEstar / tv);
    sigma2_grad = tv;
  } else {
    if (P->est_sigma2) {
      sigma2_grad = m.SSEstar / n;
      ll = -0.5 * (n * log(2 * M_PI * sigma2_grad) + 2 * logdet + n);
    } else {
      sigma2_grad = P->sigma2;
      double q = 0.0;
      for (int i = 0; i < n; i++) q += m.Estar[i] * m.Estar[i];
      ll = -0.5 * (n * log(2 * M_PI * sigma2_grad) + 2 * logdet + q / sigma2_grad);
    }
  }
  *ll_out = ll;

  if (want_grad) {
    /* x = L^-T E* (Kriging.cpp:294) */
    double* x = (double*)malloc(sizeof(double) * n);
    memcpy(x, m.Estar, sizeof(double) * n);
    lko_solve_upper_t(m.L, n, x, 1);
    double* t1 = (double*)calloc(d, sizeof(double));
    double* t2 = (double*)calloc(d, sizeof(double));
    double* g = (double*)malloc(sizeof(double) * d);
    double* dx = (double*)malloc(sizeof(double) * d);
    /* KrigingImpl.cpp:871-884: lower triangle of R and R^-1 only */
    for (int i = 0; i < n; i++) {
      for (int j = 0; j < i; j++) {
        for (int k = 0; k < d; k++) dx[k] = P->X[j + (size_t)k * n] - P->X[i + (size_t)k * n];
        lko_dlncov_dtheta(P->kernel, dx, theta, d, g);
        double Rij = m.R[i + (size_t)j * n];
        double xi = x[i], xj = x[j];
        double Rinv_ij = m.Rinv[i + (size_t)j * n];
        for (int k = 0; k < d; k++) {
          double gk = Rij * g[k];
          t1[k] += 2.0 * xi * gk * xj;
          t2[k] -= 2.0 * Rinv_ij * gk;
        }
      }
    }
    for (int k = 0; k < d; k++) grad_out[k] = (t1[k] / sigma2_grad + t2[k]) / 2.0;
    if (gamma_n > d && P->noise == LKO_NOISE_NUGGET) {
      double a = extra;
      if (P->est_sigma2 && P->est_nugget) {
        /* Kriging.cpp:310-316: dR/dalpha = R/alpha with zero diagonal */
        double q = 0.0, tr = 0.0;
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            if (i == j) continue;
            double D = m.R[i + (size_t)j * n] / a;
            q += x[i] * D * x[j];
            tr += m.Rinv[i + (size_t)j * n] * D;
          }
        double term1 = -q / sigma2_grad;
        grad_out[d] = -0.5 * (term1 + tr);
      } else if (P->est_sigma2 && !P->est_nugget) {
        /* Kriging.cpp:317-323: unit diagonal, chain factor nugget/(1-alpha)^2 */
        double tv = sigma2_grad, q = 0.0, tr = 0.0;
        for (int j = 0; j < n; j++)
          for (int i = 0; i < n; i++) {
            double D = (i == j) ? 1.0 : m.R[i + (size_t)j * n] / a;
            q += x[i] * D * x[j];
            tr += (m.Rinv[i + (size_t)j * n] / tv) * D;
          }
        double term1 = -q / (tv * tv);
        grad_out[d] = -0.5 * (term1 + tr) * P->nugget / (1.0 - a) / (1.0 - a);
      } else {
        grad_out[d] = 0.0;
      }
    }
    free(x);
    free(t1);
    free(t2);
    free(g);
    free(dx);
  }
done:
  free(m.R);
  free(m.L);
  free(m.Rinv);
  free(m.Fstar);
  free(m.ystar);
  free(m.Rstar);
  free(m.Estar);
  free(m.betahat);
  return rc;
}

/* Optim::theta_bounds (Optim.cpp:179-209) with the variogram heuristic. */
void lko_theta_bounds(const double* X, const double* y, int n, int d, double lower_factor, double upper_factor,
                      int variogram, double* lower, double* upper) {
  for (int k = 0; k < d; k++) {
    double mn = X[(size_t)k * n], mx = mn;
    for (int i = 1; i < n; i++) {
      double v = X[i + (size_t)k * n];
      if (v < mn) mn = v;
      if (v > mx) mx = v;
    }
    double maxdx = mx - mn;
    lower[k] = lower_factor * maxdx;
    upper[k] = upper_factor * maxdx;
  }
  if (!variogram || n <= 0) return;
  /* weights w_ij = (y_i - y_j)^2 / |x_i - x_j|^2 over ordered pairs i != j
   * (NaN from coincident points replaced by 0), normalised to sum 1;
   * steepest_dX_mean_k = sum_ij w_ij |dx_ijk|. */
  double wsum = 0.0;
  double* acc = (double*)calloc(d, sizeof(double));
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      if (i == j) continue;
      double s = 0.0;
      for (int k = 0; k < d; k++) {
        double v = X[i + (size_t)k * n] - X[j + (size_t)k * n];
        s += v * v;
      }
      double dy = y[i] - y[j];
      double w = (dy * dy) / s;
      if (isnan(w)) w = 0.0;
      wsum += w;
    }
  if (wsum > 0.0) {
    for (int i = 0; i < n; i++)
      for (int j = 0; j < n; j++) {
        if (i == j) continue;
        double s = 0.0;
        for (int k = 0; k < d; k++) {
          double v = X[i + (size_t)k * n] - X[j + (size_t)k * n];
          s += v * v;
        }
        double dy = y[i] - y[j];
        double w = (dy * dy) / s;
        if (isnan(w)) w = 0.0;
        w /= wsum;
        for (int k = 0; k < d; k++) acc[k] += fabs(X[i + (size_t)k * n] - X[j + (size_t)k * n]) * w;
      }
    for (int k = 0; k < d; k++) {
      double lo = lower_factor * acc[k];
      if (lo > lower[k]) lower[k] = lo;
      if (lower[k] > upper[k]) lower[k] = upper[k];
      if (upper[k] < lower[k]) upper[k] = lower[k];
    }
  }
  free(acc);
}

/* predict mean and stdev (KrigingImpl.cpp:145-243): R_on = factor*k(x_o - x_n),
 * forced to 1 at coincident points; Rstar_on = L^-1 R_on; mean = F_n beta +
 * Rstar_on' z; Ecirc = (F_n - Rstar_on' F*) R*^-1; stdev^2 = var_scale *
 * max(0, 1 - |Rstar_on|^2 + |Ecirc|^2).  Model from populate_Model at
 * (theta, alpha) with z = E* (estimated beta); sigma2 is the committed one.
 * Inputs are already normalised (normalize = false). */
int lko_predict(const lko_problem* P, const double* theta, double alpha, double sigma2, const lko_chol_cfg* cfg,
                const double* Xn, int m, double* mean, double* stdev) {
  const int n = P->n, d = P->d, p = P->p;
  size_t nn = (size_t)n * n;
  lko_model mo;
  memset(&mo, 0, sizeof mo);
  mo.R = (double*)malloc(sizeof(double) * nn);
  mo.L = (double*)malloc(sizeof(double) * nn);
  mo.Fstar = (double*)malloc(sizeof(double) * (size_t)n * (p > 0 ? p : 1));
  mo.ystar = (double*)malloc(sizeof(double) * n);
  mo.Rstar = (double*)malloc(sizeof(double) * (p > 0 ? p * p : 1));
  mo.Estar = (double*)malloc(sizeof(double) * n);
  mo.betahat = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
  int rc = lko_populate_model(P, theta, alpha, cfg, &mo);
  if (rc == 0) {
    double factor = (P->noise == LKO_NOISE_NUGGET) ? alpha : 1.0;
    double var_scale = (P->noise == LKO_NOISE_NUGGET) ? sigma2 / alpha : sigma2;
    double* r = (double*)malloc(sizeof(double) * n);
    double* dx = (double*)malloc(sizeof(double) * d);
    double* fn = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
    double* e = (double*)malloc(sizeof(double) * (p > 0 ? p : 1));
    const double eps = 2.220446049250313e-16;
    for (int q = 0; q < m; q++) {
      for (int i = 0; i < n; i++) {
        int zero = 1;
        for (int k = 0; k < d; k++) {
          dx[k] = P->X[i + (size_t)k * n] - Xn[q + (size_t)k * m];
          if (fabs(dx[k]) > eps) zero = 0;
        }
        r[i] = zero ? 1.0 : lko_cov(P->kernel, dx, theta, d) * factor;
      }
      lko_solve_lower(mo.L, n, r, 1);
      /* trend row: constant = 1, linear = [1, x] (Trend.cpp:34-61) */
      for (int a = 0; a < p; a++) fn[a] = (a == 0) ? 1.0 : Xn[q + (size_t)(a - 1) * m];
      double mu = 0.0;
      for (int a = 0; a < p; a++) mu += fn[a] * mo.betahat[a];
      double s = 0.0, rr = 0.0;
      for (int i = 0; i < n; i++) {
        s += r[i] * mo.Estar[i];
        rr += r[i] * r[i];
      }
      mean[q] = mu + s;
      for (int a = 0; a < p; a++) {
        double t = 0.0;
        for (int i = 0; i < n; i++) t += r[i] * mo.Fstar[i + (size_t)a * n];
        e[a] = fn[a] - t;
      }
      /* Ecirc = e R*^-1 (row vector): solve R*^T Ecirc^T = e^T */
      double ee = 0.0;
      for (int a = 0; a < p; a++) {
        double t = e[a];
        for (int k = 0; k < a; k++) t -= mo.Rstar[k + a * p] * e[k];
        e[a] = t / mo.Rstar[a + a * p];
        ee += e[a] * e[a];
      }
      double v = 1.0 - rr + ee;
      if (isnan(v) || v < 0) v = 0.0;
      stdev[q] = sqrt(v * var_scale);
    }
    free(r);
    free(dx);
    free(fn);
    free(e);
  }
  free(mo.R);
  free(mo.L);
  free(mo.Fstar);
  free(mo.ystar);
  free(mo.Rstar);
  free(mo.Estar);
  free(mo.betahat);
  return rc;
}

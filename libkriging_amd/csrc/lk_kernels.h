// Device-side data structures and kernel launchers of the Kriging likelihood
// pipeline (gfx950).  Internal to liblkgpu.so: the public boundary is the C
// ABI in include/lkgpu.h.
#pragma once
#include <hip/hip_runtime.h>

namespace lk {

constexpr int NB = 64;     // tile / panel width (rows and columns of one block)
constexpr int DMAX = 32;   // largest input dimension d supported by the kernels
constexpr int PMAX = 33;   // largest trend width p (linear trend with d = 32)

enum Kernel { K_GAUSS = 0, K_EXP = 1, K_MATERN32 = 2, K_MATERN52 = 3 };
enum Noise { NOISE_NONE = 0, NOISE_NUGGET = 1 };

// Status of one slot after an evaluation.
enum SlotStatus {
  ST_OK = 0,
  ST_CHOL_FAILED = 1,   // a non-positive or NaN pivot
  ST_RCOND_SUSPECT = 2, // exact rcond below threshold: the 1-norm estimator decides
  ST_RCOND_LOW = 3,     // estimator confirmed rcond below threshold
  ST_INACTIVE = 4,
};

// Static description of one batch slot: which design it evaluates and the
// device memory it works in.  Rebuilt only when the slot -> problem mapping
// changes.
struct Slot {
  const double* X;  // n x d, column-major (ld = n)
  const double* y;  // n
  const double* F;  // n x p, column-major (ld = n)
  double* A;        // npad x npad: R, then its lower Cholesky factor L
  double* W;        // npad x npad: L^-1 in the lower triangle; scratch above
  double* V;        // npad x (p + 4): [F* | y* | r | E* | x]
  double* red;      // reduction scratch (see RED_* offsets)
  const double* M;  // n x n (ld = n) input matrix of the matrix / factor modes, else null
  int n, npad, nblk, d, p;
  int kernel, noise;
  int pad_;
};

// Per-evaluation parameters of one slot (uploaded before each batch).
struct SlotParams {
  double theta[DMAX];
  double alpha;      // nugget model: R = alpha*C + (1 - alpha)*I
  double sigma2;     // fixed sigma2 (when !est_sigma2)
  double nugget;     // fixed nugget (when !est_nugget)
  double num_nugget; // jitter unit of the retry ladder
  double min_rcond;
  int inc;           // jitter steps already applied to the diagonal
  int active;
  int est_sigma2;
  int est_nugget;
  int rcond_check;
  int want_grad;
  int gamma_has_alpha; // gradient also over alpha
  int objective;       // OBJ_LL, OBJ_LOO or OBJ_LMP (batch-wide)
  int mode;            // MODE_KERNEL, MODE_MATRIX or MODE_FACTOR
  int pad_;
};

// What a slot's matrix A starts as.  MODE_KERNEL: R built from the design
// (the likelihood path).  MODE_MATRIX: the lower triangle of the slot's input
// matrix M, jitter ladder on its diagonal (LinearAlgebra::safe_chol_lower).
// MODE_FACTOR: M is already a lower Cholesky factor: no factorisation, the
// inverse, solves and the condition estimate (always the 1-norm estimator,
// as arma::rcond) from it.
enum SlotMode { MODE_KERNEL = 0, MODE_MATRIX = 1, MODE_FACTOR = 2 };

// Per-slot outputs (downloaded after each batch).
struct SlotOut {
  double ll;
  double logdet;     // sum log L_ii
  double sse;        // E*'E*
  double sigma2;     // variance used to normalise the gradient
  double norm1_L;
  double norm1_Linv;
  double rcond2;     // (1 / (|L|_1 |L^-1|_1))^2, or the estimator's value
  double grad[DMAX + 1];
  double beta[PMAX];
  double logdet_rstar; // sum log R*_aa (log-marginal-posterior objective)
  int status;
  int pad_[3];
};

// Columns of Slot::V (npad rows each): [0, p) F*, p y*, p+1 r, p+2 E*, p+3 x,
// p+4 .. p+6 estimator scratch; the leave-one-out objective adds
// [p+7, 2p+7) G' (G = R*^-T F*' L^-1, p x n, stored transposed), then
// sigma2LOO (first diag B), errorsLOO, w, sqrt(a), t = L^-1 w and B w.
__host__ __device__ inline int v_loo_g(int p) { return p + 7; }
__host__ __device__ inline int v_loo_s2(int p) { return 2 * p + 7; }
__host__ __device__ inline int v_loo_err(int p) { return 2 * p + 8; }
__host__ __device__ inline int v_loo_w(int p) { return 2 * p + 9; }
__host__ __device__ inline int v_loo_sa(int p) { return 2 * p + 10; }
__host__ __device__ inline int v_loo_t(int p) { return 2 * p + 11; }
__host__ __device__ inline int v_loo_bw(int p) { return 2 * p + 12; }
__host__ __device__ inline int v_cols(int p) { return 2 * p + 13; }

// objective of a batched evaluation
enum Objective { OBJ_LL = 0, OBJ_LOO = 1, OBJ_LMP = 2, OBJ_COUNT = 3 };

// Layout of Slot::red (doubles).
constexpr int RED_LOGDET = 0;                 // nblk partial log-dets
__host__ __device__ inline int red_sse_off(int nblk) { return nblk; }          // nblk partial SSE
__host__ __device__ inline int red_colL_off(int nblk) { return 2 * nblk; }     // nblk partial |L|_1 column maxima
__host__ __device__ inline int red_colW_off(int nblk) { return 3 * nblk; }     // nblk partial |L^-1|_1 column maxima
__host__ __device__ inline int red_grad_off(int nblk) { return 4 * nblk; }     // tiles x GRAD_STRIDE adjoint partials
constexpr int GRAD_STRIDE = 2 * DMAX + 4;
// R* = chol_upper(F*'F*), p x p column-major, after the adjoint partials
__host__ __device__ inline int red_rstar_off(int nblk) { return 4 * nblk + nblk * (nblk + 1) / 2 * GRAD_STRIDE; }
// per diagonal block k, after R*: 64 max|L_kk| max|L_kk^-1|, a bound on
// cond(L_kk) up to the factor 64 (written by the diagonal-block kernel, read
// by the panel kernel)
__host__ __device__ inline int red_cond_off(int nblk) { return red_rstar_off(nblk) + PMAX * PMAX; }
// per row block r >= 1 and column c of the blocks left of it: sum_i |L(i, c)|
// over the block's rows, written by the panel kernel (r * npad + c), read
// for ||L||_1 by k_solve_x
__host__ __device__ inline int red_lcol_off(int nblk) { return red_cond_off(nblk) + nblk; }
__host__ __device__ inline size_t red_size(int nblk) { return (size_t)red_lcol_off(nblk) + (size_t)nblk * nblk * NB; }
// Panels whose diagonal block's condition estimate exceeds this get one step
// of iterative refinement: multiplying by the explicit inverse of an
// ill-conditioned block is not backward stable, the refined product is.
constexpr double PANEL_REFINE_COND = 1e4;

struct Batch {
  const Slot* slots;        // device
  const SlotParams* params; // device
  SlotOut* out;             // device
  int nslots;
  int nblk_max;             // grid extent (max nblk over slots)
  int dmax;                 // max d over slots (selects the unrolled kernel variant)
  hipStream_t stream;
};

// Launchers (each returns the hipGetLastError of the launch).
hipError_t launch_build(const Batch& b);
hipError_t launch_potrf_diag(const Batch& b, int k);
hipError_t launch_potrf_panel(const Batch& b, int k);
// A(r, c) -= sum_{k in [kb, ke)} L(r, k) L(c, k)^T, tile columns c in [c0, c1]
// (c1 < 0: to the last column), rows r >= c
hipError_t launch_potrf_update(const Batch& b, int kb, int ke, int c0, int c1);
// the trailing update (every column from c0 on) on 128 x 64 tiles
hipError_t launch_potrf_update_trailing(const Batch& b, int kb, int ke, int c0);
// the same update of a one-slot batch by `workgroups` persistent workgroups
hipError_t launch_potrf_update_persistent(const Batch& b, int kb, int ke, int c0, int c1, int workgroups);
hipError_t launch_trtri_level(const Batch& b, int s, int step);
hipError_t launch_solve_fy(const Batch& b);
hipError_t launch_gls(const Batch& b);
hipError_t launch_solve_e(const Batch& b);
hipError_t launch_solve_x(const Batch& b);
hipError_t launch_finalize_ll(const Batch& b);
hipError_t launch_lauum_adjoint(const Batch& b);

hipError_t launch_finalize_grad(const Batch& b);
hipError_t launch_rcond_estimate(const Batch& b);

int trtri_levels(int nblk);

// Allocation-free evaluations: the static device scratch array (64 MiB per
// device) that holds the slot tables and workspaces, and the most items one
// call may carry.
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_ITEMS = 64;
// device address of the scratch array on `device` (the current device)
hipError_t noalloc_scratch(int device, char** ptr);

// Prediction at new points from a committed model (KrigingImpl::predict_impl,
// src/lib/KrigingImpl.cpp:145-243): mean and variance in normalised units.
struct PredictArgs {
  const double* X;     // n x d design (normalised), ld = n
  const double* Linv;  // npad x npad, lower triangle = L^-1
  const double* MZ;    // npad x (p + 1): [M = F* | z]
  const double* coef;  // beta (p) then R* (p x p, column-major upper)
  const double* Xn;    // m x d new points (normalised), ld = m
  double* Ron;         // npad x mpad workspace: R(x_obs, x_new)
  double* C;           // npad x mpad workspace: L^-1 R_on
  double* out;         // 2 m: mean (normalised), variance factor
  double* ecirc;       // m x p (ld = m): E circ rows, when cov or derivatives are wanted (else null)
  // covariance (KrigingImpl.cpp:245-253): m x m, ld = m, before var_scale
  double* cov;
  double nn_factor;    // R_nn_factor (alpha for the nugget model); diagonal 1
  // derivatives (KrigingImpl.cpp:255-291): dR_on / dx_new and L^-1 of it,
  // npad x (d mpad) with column k mpad + j for point j and coordinate k;
  // dmean, dvar: m x d (ld = m), normalised units
  double* DR;
  double* DW;
  double* dmean;
  double* dvar;
  double theta[DMAX];
  double factor;       // R_on_factor (alpha for the nugget model)
  int n, d, npad, p, m, mpad;
  int kernel, linear_trend;
  int want_var;        // derivatives of the variance too
};
hipError_t launch_predict(const PredictArgs& a, hipStream_t stream);
// covariance between two point sets: out(i, j) = scale k(x1_i - x2_j)
struct CovMatArgs {
  const double* X1;  // n1 x d, ld = n1
  const double* X2;  // n2 x d, ld = n2
  double* out;       // n1 x n2, ld = n1
  double theta[DMAX];
  double scale;
  int n1, n2, d, kernel;
  int unit_diag;  // X1 == X2: out(i, i) = 1 (the correlation matrix's diagonal)
  int pad_;
};
hipError_t launch_cov_mat(const CovMatArgs& a, hipStream_t stream);
// C = Linv R_on alone (PredictArgs' Linv, Ron, C, npad, mpad; both padded
// to whole 64-row blocks, Linv's diagonal blocks zero above the diagonal)
hipError_t launch_trmm_lower(const PredictArgs& a, hipStream_t stream);
// Schur complement of a block Cholesky update: S = Cuu - U' U, with U
// kdim x mpad (ld = kdim, both padded to whole 64-row blocks, zero padding),
// Cuu and S m x m (ld = m); S is written in full (both triangles)
struct SchurArgs {
  const double* U;
  const double* Cuu;
  double* S;
  int kdim, m, mpad, pad_;
};
hipError_t launch_schur(const SchurArgs& a, hipStream_t stream);
// L (n x n, ld = n) of a block Cholesky update: Loo's lower triangle (no x
// no, ld = ldo), U' below it (U no x (n - no), ld = ldu), Luu's lower
// triangle (ld = ldl) in the trailing block, zeros above the diagonal
struct BlockAssembleArgs {
  const double* Loo;
  const double* U;
  const double* Luu;
  double* L;
  int n, no, ldo, ldu, ldl, pad_;
};
hipError_t launch_block_assemble(const BlockAssembleArgs& a, hipStream_t stream);
hipError_t launch_predict_cov(const PredictArgs& a, hipStream_t stream);    // after launch_predict (same chunk)
hipError_t launch_predict_deriv(const PredictArgs& a, hipStream_t stream);  // after launch_predict (same chunk)
// log-marginal-posterior gradient: the fused R^-1 / adjoint pass with
// B = R^-1 - G'G in place of R^-1 (after launch_loo_g)
hipError_t launch_lmp_adjoint(const Batch& b);
// leave-one-out objective (lk_loo.hip), after the L^-1 path's solves
hipError_t launch_loo_g(const Batch& b);
hipError_t launch_loo_value(const Batch& b);
hipError_t launch_loo_wv(const Batch& b);
hipError_t launch_loo_bw(const Batch& b);
hipError_t launch_loo_bs(const Batch& b);
hipError_t launch_loo_adjoint(const Batch& b);
hipError_t launch_loo_finalize_grad(const Batch& b);
// R^-1 = L^-T L^-1 in full into A (after an L^-1-path evaluation)
hipError_t launch_rinv_store(const Batch& b);

// ---------------------------------------------------------------------------
// NestedKriging "NK" aggregation (NestedKriging::predict_nk,
// src/lib/NestedKriging.cpp:353-465).  The design is stored group by group,
// each group's rows padded to a multiple of 64 (pad rows: X = 0, weights 0),
// so that every 64-row tile belongs to exactly one group.
// ---------------------------------------------------------------------------
struct NkGroup {
  int off;      // first row of the group in the group-ordered arrays
  int n;        // points in the group
  int npad;     // n rounded up to 64
  int pad_;
  long long loff;  // offset of the group's L^-1 (npad x npad) in the pool
};

struct NkArgs {
  const double* Xg;        // Npad x d, group-ordered design, ld = Npad
  const double* Linv;      // pool of per-group L^-1 (lower, column-major)
  double* alpha;           // Npad: R_g^-1 (y_g - beta0), group-ordered
  const NkGroup* groups;   // p
  const int* tile_group;   // Npad / 64: group of each 64-row tile
  const int2* pairs;       // npairs (i < j) group pairs
  const double* Xc;        // qc x d points of this chunk, ld = qc
  double* C;               // Npad x qpad: k(x_a, x_t), then R_g^-1 k_g
  double* U;               // Npad x qpad: L_g^-1 k_g
  double* M;               // qc x p: submodel means
  double* Kd;              // qc x p: k_g' R_g^-1 k_g
  double* KM;              // qc x p x p: cov(M_i, M_j) / sigma2 (column-major per point)
  double* KF;              // qc x p x p: factor workspace
  double* out;             // 2 qc: mean, variance
  int* info;               // qc: 0 Cholesky, 1 LU fallback, 2 singular
  double theta[DMAX];
  double itheta[DMAX];     // 1 / theta (1 beyond d): wave-uniform kernel arguments
  double beta0, sigma2, jitter;
  int Npad, d, p, qc, qpad, kernel, npairs;
};
hipError_t launch_nk_prepare(const NkArgs& a, hipStream_t stream);  // C, U, M, Kd, R_g^-1 k_g
// points per workgroup of the pair kernel: the chunk's points are padded to
// a multiple of it (qpad)
constexpr int NK_QT = 128;
hipError_t launch_nk_pairs(const NkArgs& a, hipStream_t stream);    // off-diagonal KM
hipError_t launch_nk_solve(const NkArgs& a, hipStream_t stream);    // per-point aggregation
hipError_t launch_nk_alpha(const NkArgs& a, const double* r, hipStream_t stream);  // alpha = Linv' r

}  // namespace lk

// Kriging concentrated log-likelihood + adjoint gradient, FP64, gfx950.
//
// One "slot" is one LL(+grad) evaluation of one design at one parameter
// vector.  Every kernel below works on a batch of slots (grid.y or grid.x =
// slot) so that 64 multistart evaluations, or 256 uneven NestedKriging
// submodels, run in the same launches.  Matrices live in HBM as FP64
// column-major npad x npad (npad = n rounded up to 64, the padded part set to
// the identity so that every kernel works on whole 64 x 64 tiles).
//
// Pipeline of one evaluation (reference: Kriging::_logLikelihood,
// src/lib/Kriging.cpp:214-338, and KrigingImpl::populate_Model,
// src/lib/KrigingImpl.cpp:73-125):
//   build           R = factor * k(x_i - x_j; theta), unit (+jitter) diagonal
//   potrf diag/panel/update   blocked right-looking Cholesky, FP64 MFMA
//   trtri           L^-1 by recursive doubling, FP64 MFMA
//   solve_fy, gls, solve_e, solve_x   F* = L^-1 F, y* = L^-1 y, GLS beta,
//                   E* = L^-1 (y - F beta), x = L^-T E*, |L|_1, |L^-1|_1
//   finalize_ll     log-det, SSE, LL, rcond check
//   lauum_adjoint   R^-1 = L^-T L^-1 tile by tile on MFMA; each tile is
//                   contracted at once with dR/dtheta_k recomputed from X
//                   (all k, plus the nugget direction), never stored
//   finalize_grad   deterministic reduction of the per-tile partial sums
#include <hip/hip_runtime.h>

#include <math.h>

#include <type_traits>

#include "lk_device.hpp"
#include "lk_kernels.h"

namespace lk {

// ---------------------------------------------------------------------------
// (1) covariance build: lower 64 x 64 tiles (diagonal tiles in full)
// grid (ntiles(nblk_max), nslots), 256 threads
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(TPB) void k_build(const Slot* __restrict__ slots, const SlotParams* __restrict__ params,
                                               SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  if (!P.active) {
    if (ts.tile == 0 && threadIdx.x == 0) out[b].status = ST_INACTIVE;
    return;
  }
  if (ts.tile == 0 && threadIdx.x == 0) out[b].status = ST_OK;
  const int ntiles = S.nblk * (S.nblk + 1) / 2;
  if (ts.tile >= ntiles) return;
  int ti, tj;
  tri_decode(ts.tile, ti, tj);
  const int i0 = ti * NB, j0 = tj * NB;
  const int n = S.n, d = S.d, tid = threadIdx.x;
  if (P.mode != MODE_KERNEL) {
    // the input matrix's lower triangle (diagonal tiles: zero above), the
    // padded part the identity; the matrix mode adds the retry ladder's
    // cumulative jitter to the diagonal in the reference's order
    // (LinearAlgebra.cpp:88: X.diag() += num_nugget * 10^inc, per retry)
    const int ii = tid & 63, gi = i0 + ii;
    for (int r = 0; r < NB / 4; ++r) {
      const int jj = (tid >> 6) + 4 * r, gj = j0 + jj;
      double v;
      if (gi >= n || gj >= n) {
        v = (gi == gj) ? 1.0 : 0.0;
      } else if (gi < gj) {
        v = 0.0;
      } else {
        v = S.M[(size_t)gj * n + gi];
        // a given factor's columns are taken with a positive diagonal: the
        // 1-norm condition number does not change with column signs
        if (P.mode == MODE_FACTOR && S.M[(size_t)gj * n + gj] < 0.0) v = -v;
        if (gi == gj && P.mode == MODE_MATRIX) {
          double p10 = 1.0;
          for (int q = 0; q < P.inc; ++q) {
            v += P.num_nugget * p10;
            p10 *= 10.0;
          }
        }
      }
      S.A[(size_t)gj * S.npad + gi] = v;
    }
    return;
  }

  __shared__ double sXi[DT * NB];
  __shared__ double sXj[DT * NB];
  for (int q = tid; q < d * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    sXi[q] = (i0 + r < n) ? S.X[(size_t)k * n + i0 + r] : 0.0;
    sXj[q] = (j0 + r < n) ? S.X[(size_t)k * n + j0 + r] : 0.0;
  }
  double it[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) it[k] = (k < d) ? 1.0 / P.theta[k] : 1.0;
  const double factor = (S.noise == NOISE_NUGGET) ? P.alpha : 1.0;
  // diagonal: 1 plus the cumulative jitter of the retry ladder, added in the
  // same order as LinearAlgebra.cpp:88 (X.diag() += num_nugget * 10^inc)
  double dval = 1.0;
  {
    double p10 = 1.0;
    for (int q = 0; q < P.inc; ++q) {
      dval += P.num_nugget * p10;
      p10 *= 10.0;
    }
  }
  __syncthreads();

  const int ii = tid & 63;
  const int gi = i0 + ii;
  // specialised per covariance kernel (one kernel's code runs)
  auto fill = [&](auto ktag) __attribute__((always_inline)) {
    constexpr int KT = decltype(ktag)::value;
    double dx[DT];
    for (int r = 0; r < NB / 4; ++r) {
      const int jj = (tid >> 6) + 4 * r;
      const int gj = j0 + jj;
      double v;
      if (gi >= n || gj >= n) {
        v = (gi == gj) ? 1.0 : 0.0;
      } else if (gi == gj) {
        v = dval;
      } else {
#pragma unroll
        for (int k = 0; k < DT; ++k) dx[k] = (k < d) ? sXj[k * NB + jj] - sXi[k * NB + ii] : 0.0;
        v = kern_rho<DT>(KT, dx, it, DT <= 8 ? d : DT) * factor;  // padded coordinates skipped up to 8 dimensions (beyond, the branches cost more than they save: n = 8192, d = 10 build 0.216 vs 0.143 ms)
      }
      S.A[(size_t)gj * S.npad + gi] = v;
    }
  };
  switch (S.kernel) {
    case K_GAUSS: fill(std::integral_constant<int, K_GAUSS>{}); break;
    case K_EXP: fill(std::integral_constant<int, K_EXP>{}); break;
    case K_MATERN32: fill(std::integral_constant<int, K_MATERN32>{}); break;
    default: fill(std::integral_constant<int, K_MATERN52>{}); break;
  }
}

// ---------------------------------------------------------------------------
// (2a) Cholesky of the diagonal block k and its inverse, one 512-thread
// workgroup per slot: waves 0-3 run the factor sweep, waves 4-7 the inverse
// sweep.  Both go 4 columns / rows at a time (16 sequential steps instead of
// 64): a 4 x 4 diagonal sub-block is factored redundantly by every factor
// lane (lane = row) from LDS broadcasts, each lane solves its row's 4 L
// values, and the rank-4 update of the later columns runs on MFMA
// (V_MFMA_F64_16X16X4_F64, one 16 x 16 tile per instruction): its operands
// are one value per lane from LDS instead of a broadcast read per updated
// element, which kept the LDS return path the bound of the step.  Factor
// wave w holds columns 16w .. 16w+15 of the block as four 16 x 16 MFMA
// tiles; inverse wave w holds rows 16w .. 16w+15 of X = L_kk^-1 the same
// way.  Pivot square roots come from v_rsq_f64 refined by one third-order
// step.  Writes L_kk to A and L_kk^-1 to the same block of W (both with
// zero strict upper part), the partial log-det, and flags a non-positive /
// NaN pivot (dpotrf failure, LinearAlgebra.cpp:71).
// ---------------------------------------------------------------------------
__device__ inline void pivot_sqrt(double a, double& root, double& rroot) {
  // y = v_rsq_f64(a) (relative error below 2^-22), then one third-order
  // step: with e = 1 - a y^2, a^-1/2 = y (1 - e)^-1/2 = y (1 + e/2 + 3e^2/8
  // + O(e^3)); two dependent operations fewer than two Newton steps
  const double y = __builtin_amdgcn_rsq(a);
  const double t = a * y;
  const double e = fma(-t, y, 1.0);
  const double p = fma(e, 0.375, 0.5);
  const double iv = fma(y * e, p, y);
  double s = a * iv;
  const double r = fma(-s, s, a);
  root = fma(r, 0.5 * iv, s);
  rroot = iv;
}

// columns (factor) / rows (inverse) per step of the diagonal-block sweeps:
// the K of one MFMA
constexpr int DCW = 4;

struct DiagLds {
  // LDS layouts keep a wave's 64 lanes on consecutive doubles for every
  // per-lane access (stride 1 writes / reads, or one broadcast address)
  double P[2][DCW * NB];  // current values of the step's columns, [column][row]
  double Q[4][DCW * NB];  // their L values, [column][row], one copy per factor wave
  double L[NB * (NB + 2)];  // L_kk column-major, stride NB + 2 (16-byte aligned column pairs)
  double X[NB * (NB + 1)];  // L_kk^-1 rows staged for coalesced stores, stride NB + 1
  double R[2][DCW * NB];    // the step's rows of L_kk^-1 in the inverse sweep, [row][column]
  double Z[DCW * NB];       // the inverse owner's accumulated step rows, [row][column]
  double diag[NB], rinv[NB];
};

// MFMA tile layouts (lane l, register e; the instruction returns D(a, b) with
// a = (l >> 4) + 4e, b = l & 15 for D = P Q, P (16 x 4) given as lane l ->
// P(l & 15, l >> 4) and Q (4 x 16) as lane l -> Q(l >> 4, l & 15)):
//   factor tile rb of wave W: element (16 rb + (l & 15), 16 W + (l >> 4) + 4e)
//     (D transposed: the update's P is the wave's columns, Q the tile's rows)
//   inverse tile cb of wave W: element (16 W + (l >> 4) + 4e, 16 cb + (l & 15))

// Before the barrier of step p: the owner of the step's columns publishes
// their current values (tile register e = (j0 & 15) / 4 holds column
// j0 + (l >> 4) of the tile's 16 rows).  Rows above 16 W are not written:
// no lane reads them (their L values are forced to zero).
template <int W>
__device__ inline void factor_publish(int p, const d4 (&T)[4], DiagLds& sh, int l) {
  const int j0 = DCW * p;
  if (W == j0 / 16) {
    double* P = sh.P[p & 1];
    const int e = (j0 & 15) / 4;
#pragma unroll
    for (int rb = W; rb < 4; ++rb) P[(l >> 4) * NB + 16 * rb + (l & 15)] = T[rb][e];
  }
}

// After the barrier: the DCW x DCW diagonal sub-block (factored by every lane
// from LDS broadcasts), this row's L values in the step's columns, then the
// rank-DCW update of the wave's columns right of the step on MFMA; the owner
// also writes the step's columns of L (zero above the diagonal) to sh.L, and
// wave 0 their diagonal and reciprocal, which the inverse sweep reads two
// steps later.
template <int W>
__device__ inline bool factor_step(int p, d4 (&T)[4], DiagLds& sh, int l) {
  bool ok = true;
  const int j0 = DCW * p;
  const double* P = sh.P[p & 1];
  double* Q = sh.Q[W];
  // DCW x DCW diagonal sub-block, factored by every lane (left-looking)
  double L4[DCW][DCW], iv[DCW];
#pragma unroll
  for (int r = 0; r < DCW; ++r)
#pragma unroll
    for (int c = 0; c <= r; ++c) L4[r][c] = P[c * NB + j0 + r];
#pragma unroll
  for (int c = 0; c < DCW; ++c) {
    double d = L4[c][c];
#pragma unroll
    for (int m = 0; m < c; ++m) d = fma(-L4[c][m], L4[c][m], d);
    ok = ok && (d > 0.0);
    pivot_sqrt(d, L4[c][c], iv[c]);
#pragma unroll
    for (int r = c + 1; r < DCW; ++r) {
      double t = L4[r][c];
#pragma unroll
      for (int m = 0; m < c; ++m) t = fma(-L4[r][m], L4[c][m], t);
      L4[r][c] = t * iv[c];
    }
  }
  // this row's L values in columns j0 .. j0+DCW-1
  double x[DCW];
  {
    double v[DCW];
#pragma unroll
    for (int s = 0; s < DCW; ++s) v[s] = P[s * NB + l];
#pragma unroll
    for (int s = 0; s < DCW; ++s) {
      double t = v[s];
#pragma unroll
      for (int m = 0; m < s; ++m) t = fma(-x[m], L4[s][m], t);
      x[s] = t * iv[s];
    }
    // rows inside the block reproduce L4 below the diagonal (same
    // operations); the diagonal takes the corrected root, rows above the
    // step are zero (their published values are not meaningful).  Entries
    // above the diagonal inside the block are left as computed: they only
    // reach the strict upper part of the tiles' columns, which no
    // lower-part value reads and the stores zero.
#pragma unroll
    for (int s = 0; s < DCW; ++s) {
      x[s] = (l == j0 + s) ? L4[s][s] : x[s];
      x[s] = (l < j0) ? 0.0 : x[s];
    }
  }
  if (W == j0 / 16) {
#pragma unroll
    for (int s = 0; s < DCW; ++s) sh.L[(j0 + s) * (NB + 2) + l] = (l >= j0 + s) ? x[s] : 0.0;
  }
  // every wave computed the same x for all 64 rows: each keeps its own copy
  // and reads the MFMA operands back without a workgroup barrier (one
  // wave's LDS operations complete in order; the fences keep the compiler
  // from moving the reads above the writes)
#pragma unroll
  for (int s = 0; s < DCW; ++s) Q[s * NB + l] = x[s];
  if (W == 0 && l == 0) {
#pragma unroll
    for (int s = 0; s < DCW; ++s) {
      sh.diag[j0 + s] = L4[s][s];
      sh.rinv[j0 + s] = iv[s];
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  // tile rb -= X_rb X_W^T over the wave's columns right of the step (the
  // P operand is zero for the columns up to j0 + DCW - 1, which stay)
  if (W >= j0 / 16 && 16 * W + 15 >= j0 + DCW) {
    const int c = 16 * W + (l & 15);
    const double pv = (c >= j0 + DCW) ? -Q[(l >> 4) * NB + c] : 0.0;
#pragma unroll
    for (int rb = W; rb < 4; ++rb) {
      const double qv = Q[(l >> 4) * NB + 16 * rb + (l & 15)];
      T[rb] = __builtin_amdgcn_mfma_f64_16x16x4f64(pv, qv, T[rb], 0, 0, 0);
    }
  }
  return ok;
}

// One step of the inverse sweep for wave W (owns rows 16W .. 16W+15 of
// X = L_kk^-1), split at its barrier.  Step p solves rows j0 = DCW p ..
// j0 + DCW - 1 (owner; lane = column) after reading the wave's MFMA operand
// L(16W + (l & 15), j0 + (l >> 4)) for the update that follows the barrier.
template <int W>
__device__ inline void inverse_solve(int p, const d4 (&U)[4], double& pv, DiagLds& sh, int l) {
  const int j0 = DCW * p;
  double* Rr = sh.R[p & 1];
  {
    const int i = 16 * W + (l & 15);
    pv = (i >= j0 + DCW) ? sh.L[(j0 + (l >> 4)) * (NB + 2) + i] : 0.0;
  }
  if (W == j0 / 16) {
    // the step's rows of the accumulated sums: lane group s holds row j0 + s
    // (register (j0 & 15) / 4 of each tile); regroup to lane = column
    const int e = (j0 & 15) / 4;
#pragma unroll
    for (int cb = 0; cb < 4; ++cb) sh.Z[(l >> 4) * NB + 16 * cb + (l & 15)] = (cb <= W) ? U[cb][e] : 0.0;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double x[DCW];
#pragma unroll
    for (int s = 0; s < DCW; ++s) {
      double t = ((l == j0 + s) ? 1.0 : 0.0) - sh.Z[s * NB + l];
#pragma unroll
      for (int m = 0; m < s; ++m) t = fma(-sh.L[(j0 + m) * (NB + 2) + j0 + s], x[m], t);
      x[s] = t * sh.rinv[j0 + s];
      Rr[s * NB + l] = x[s];
      sh.X[(j0 + s) * (NB + 1) + l] = x[s];
    }
  }
}

// After the barrier: tile cb += L(rows of the wave, step columns) X(step
// rows, columns of the tile) for the wave's rows below the step (zero P
// rows elsewhere); tiles right of the wave's rows stay zero (X is lower
// triangular)
template <int W>
__device__ inline void inverse_update(int p, d4 (&U)[4], double pv, const DiagLds& sh, int l) {
  const int j0 = DCW * p;
  if (16 * W + 15 < j0 + DCW) return;
  const double* Rr = sh.R[p & 1];
#pragma unroll
  for (int cb = 0; cb <= W; ++cb) {
    const double qv = Rr[(l >> 4) * NB + 16 * cb + (l & 15)];
    U[cb] = __builtin_amdgcn_mfma_f64_16x16x4f64(pv, qv, U[cb], 0, 0, 0);
  }
}

// The two sweeps pipelined over the workgroup's 8 waves: waves 0-3 run the
// factor sweep, waves 4-7 the inverse sweep two steps behind it (step p of
// the inverse reads the L columns and reciprocals that factor step p wrote,
// one barrier earlier).  Every wave passes the same NB / DCW + 2 barriers.
// With the factor given (factor mode), sh.L, sh.diag and sh.rinv are filled
// before the sweep and the factor waves only keep the barrier count.
template <int W8>
__device__ inline bool diag_sweeps(d4 (&T)[4], DiagLds& sh, int l, bool factor_given) {
  constexpr bool FW = W8 < 4;
  constexpr int W = FW ? W8 : W8 - 4;
  constexpr int NR = NB / DCW;
  bool ok = true;
  d4 U[4];
#pragma unroll
  for (int cb = 0; cb < 4; ++cb) U[cb] = d4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int r = 0; r < NR + 2; ++r) {
    double pv = 0.0;
    if (FW) {
      if (r < NR && !factor_given) factor_publish<W>(r, T, sh, l);
    } else if (r >= 2) {
      inverse_solve<W>(r - 2, U, pv, sh, l);
    }
    __syncthreads();
    if (FW) {
      if (r < NR && !factor_given) ok = factor_step<W>(r, T, sh, l) && ok;
    } else if (r >= 2) {
      inverse_update<W>(r - 2, U, pv, sh, l);
    }
  }
  return ok;
}

constexpr int DIAG_TPB = 2 * TPB;  // 8 waves: 4 factor, 4 inverse

__global__ __launch_bounds__(DIAG_TPB) void k_potrf_diag(const Slot* __restrict__ slots,
                                                         const SlotParams* __restrict__ params,
                                                         SlotOut* __restrict__ out, int k) {
  const int b = blockIdx.x;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || k >= S.nblk) return;
  const int tid = threadIdx.x, l = tid & 63, w = tid >> 6;
  const bool fw = w < 4;  // factor wave (else inverse wave w - 4)
  const int k0 = k * NB;
  const int ld = S.npad;
  double* Ab = S.A + (size_t)k0 * ld + k0;
  double* Wb = S.W + (size_t)k0 * ld + k0;
  __shared__ DiagLds sh;
  __shared__ int sbad;
  __shared__ double smax[2][4];
  if (tid == 0) sbad = 0;
  const bool given = params[b].mode == MODE_FACTOR;

  d4 T[4];
#pragma unroll
  for (int rb = 0; rb < 4; ++rb) T[rb] = d4{0.0, 0.0, 0.0, 0.0};
  if (fw) {
    if (given) {
      // the block is already L_kk: its columns, diagonal and reciprocals for
      // the inverse sweep; a non-positive or NaN diagonal entry is a failure
#pragma unroll
      for (int q = 0; q < 16; ++q) {
        const int c = 16 * w + q;
        const double v = Ab[(size_t)c * ld + l];
        sh.L[c * (NB + 2) + l] = (l >= c) ? v : 0.0;
        if (l == c) {
          sh.diag[c] = v;
          sh.rinv[c] = 1.0 / v;
        }
      }
    } else {
      // the wave's tiles on and below the diagonal (rows from 16 w)
#pragma unroll
      for (int rb = 0; rb < 4; ++rb)
        if (rb >= w)
#pragma unroll
          for (int e = 0; e < 4; ++e) T[rb][e] = Ab[(size_t)(16 * w + (l >> 4) + 4 * e) * ld + 16 * rb + (l & 15)];
    }
  }
  __syncthreads();
  bool ok;
  switch (w) {
    case 0: ok = diag_sweeps<0>(T, sh, l, given); break;
    case 1: ok = diag_sweeps<1>(T, sh, l, given); break;
    case 2: ok = diag_sweeps<2>(T, sh, l, given); break;
    case 3: ok = diag_sweeps<3>(T, sh, l, given); break;
    case 4: ok = diag_sweeps<4>(T, sh, l, given); break;
    case 5: ok = diag_sweeps<5>(T, sh, l, given); break;
    case 6: ok = diag_sweeps<6>(T, sh, l, given); break;
    default: ok = diag_sweeps<7>(T, sh, l, given); break;
  }
  if (given && fw && l < 16) ok = sh.diag[16 * w + l] > 0.0;
  // every factor lane saw every pivot; in factor mode a lane checks one
  // diagonal entry: any failure is shared through LDS before the uniform exit
  if (fw && !ok) sbad = 1;
  __syncthreads();
  if (sbad) {
    if (tid == 0) out[b].status = ST_CHOL_FAILED;
    return;
  }
  if (fw) {
    // L_kk -> A (lower, zero above), from its LDS copy (lane = row)
    double lmax = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = 16 * w + q;
      const double v = sh.L[c * (NB + 2) + l];
      Ab[(size_t)c * ld + l] = v;
      lmax = fmax(lmax, fabs(v));
    }
    lmax = wave_max(lmax);
    if (l == 0) smax[0][w] = lmax;
    if (w == 0) {
      const double ld_part = wave_sum(log(sh.diag[l]));
      if (l == 0) S.red[RED_LOGDET + k] = ld_part;
    }
  } else {
    // L_kk^-1 -> W: rows staged as sh.X[row][col]; columns written coalesced
    // (lane = row)
    const int wi = w - 4;
    double xmax = 0.0;
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int c = 16 * wi + q;
      const double v = sh.X[l * (NB + 1) + c];
      Wb[(size_t)c * ld + l] = v;
      xmax = fmax(xmax, fabs(v));
    }
    xmax = wave_max(xmax);
    if (l == 0) smax[1][wi] = xmax;
  }
  __syncthreads();
  // condition estimate of L_kk for the panel kernel's refinement decision
  if (tid == 0) {
    const double lm = fmax(fmax(smax[0][0], smax[0][1]), fmax(smax[0][2], smax[0][3]));
    const double xm = fmax(fmax(smax[1][0], smax[1][1]), fmax(smax[1][2], smax[1][3]));
    S.red[red_cond_off(S.nblk) + k] = (double)NB * lm * xm;
  }
}

// ---------------------------------------------------------------------------
// (2b) panel: L(r, k) = A(r, k) * L_kk^-T for row blocks r > k, as a product
// with the explicit inverse from the diagonal-block kernel.  When the
// diagonal block is ill-conditioned (its estimate above PANEL_REFINE_COND)
// the product is refined once, X += (A - X L_kk^T) L_kk^-T: the product
// alone is accurate to cond(L_kk) eps relative to |L_kk^-1|, the refined one
// to the level of a substitution (dtrsm, as LAPACK's dpotrf uses).  The
// unrefined X waits in W's strict upper triangle (row block k, column block
// r: unused until the inverse is formed) while the residual is formed.
// grid (nblk_max - k - 1, nslots), 256 threads, three waves per SIMD (a
// few registers spilled; 64 x n = 1024 Cholesky 1.019 -> 1.007 ms against
// two, n = 8192 7.35 -> 7.25 ms)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB, 3) void k_potrf_panel(const Slot* __restrict__ slots,
                                                     const SlotParams* __restrict__ params,
                                                     SlotOut* __restrict__ out, int k) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  const int r = k + 1 + ts.tile;
  if (!slot_live(params[b], out[b]) || r >= S.nblk || params[b].mode == MODE_FACTOR) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int i0 = r * NB, k0 = k * NB, ld = S.npad;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  // A(i0 + ii, k) times Dinv(c, k) at W[k * ld + k0 + c]
  tile_mainloop<false, false, false>(S.A, ld, i0, S.W, ld, k0, k0, k0 + NB, sbuf, acc, tid, lane, wr, wc);
  double* A21 = S.A + (size_t)k0 * ld + i0;
  if (S.red[red_cond_off(S.nblk) + k] > PANEL_REFINE_COND) {
    // X^T into W(k0 + c, i0 + r); the residual A21 - X L_kk^T into A21
    acc_store_t(acc, S.W + (size_t)i0 * ld + k0, ld, lane, wr, wc);
    d4 res[2][2];
    acc_load(res, A21, ld, lane, wr, wc);
    __syncthreads();
    tile_mainloop<true, false, true>(S.W, ld, i0, S.A, ld, k0, k0, k0 + NB, sbuf, res, tid, lane, wr, wc);
    acc_store(res, A21, ld, lane, wr, wc);
    __syncthreads();
    // X += residual Dinv^T
    tile_mainloop<false, false, false>(S.A, ld, i0, S.W, ld, k0, k0, k0 + NB, sbuf, acc, tid, lane, wr, wc);
  }
  acc_store(acc, A21, ld, lane, wr, wc);
  // this block's column sums of |L| for ||L||_1: rows bi in the lane, the 16
  // lanes of a column group by shuffles, the two row waves through LDS
  __shared__ double scol[NB];
  double cs[2][4];
#pragma unroll
  for (int bj = 0; bj < 2; ++bj)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double v = fabs(acc[0][bj][e]) + fabs(acc[1][bj][e]);
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
      cs[bj][e] = v;
    }
  if (wr == 1 && (lane & 15) == 0) {
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) scol[acc_col(lane, wc, bj, e)] = cs[bj][e];
  }
  __syncthreads();
  if (wr == 0 && (lane & 15) == 0) {
    double* part = S.red + red_lcol_off(S.nblk) + (size_t)r * ld + k0;
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int c = acc_col(lane, wc, bj, e);
        part[c] = cs[bj][e] + scol[c];
      }
  }
}

// ---------------------------------------------------------------------------
// (2c) deferred update: A(r, c) -= sum_{k in [kb, ke)} L(r, k) L(c, k)^T
// over tile columns c in [c0, c1] (c1 < 0: every column from c0 on) and
// rows r >= c.  Two-level blocking: inside a panel of P block columns the
// update reaches only the panel's own columns (K = 64); the rest of the
// matrix is updated once per panel with K = 64 P, so the trailing matrix is
// read and written 1 / P as often.
// grid (tiles of the widest slot, nslots), 256 threads, four waves per SIMD
// (128 registers, no spill: the 64 x n = 1024 batch's Cholesky 1.035 ->
// 1.021 ms against three)
// ---------------------------------------------------------------------------
// tile number `tile` of the update's tile set -> (r, c); false when out of range
__device__ inline bool update_tile_rc(int tile, int nb, int c0, int c1, int& r, int& c) {
  if (c1 < 0) {
    const int T = nb - c0;
    if (tile >= T * (T + 1) / 2) return false;
    int ti, tj;
    tri_decode(tile, ti, tj);
    r = c0 + ti;
    c = c0 + tj;
    return true;
  }
  int t = tile;
  c = c0;
  const int cl = min(c1, nb - 1);
  while (c <= cl && t >= nb - c) {
    t -= nb - c;
    ++c;
  }
  if (c > cl) return false;
  r = c + t;
  return true;
}

__device__ inline void update_tile(const Slot& S, int r, int c, int kb, int ke, double* sbuf) {
  const int i0 = r * NB, j0 = c * NB, ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  d4 acc[2][2];
  double* C = S.A + (size_t)j0 * ld + i0;
  acc_load(acc, C, ld, lane, wr, wc);
  tile_mainloop<false, false, true>(S.A, ld, i0, S.A, ld, j0, kb * NB, ke * NB, sbuf, acc, tid, lane, wr, wc);
  acc_store(acc, C, ld, lane, wr, wc);
}

__global__ __launch_bounds__(TPB, 4) void k_potrf_update(const Slot* __restrict__ slots,
                                                      const SlotParams* __restrict__ params,
                                                      SlotOut* __restrict__ out, int kb, int ke, int c0, int c1) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || params[b].mode == MODE_FACTOR) return;
  const int nb = S.nblk;
  if (c0 >= nb || ke > nb) return;
  int r, c;
  if (!update_tile_rc(ts.tile, nb, c0, c1, r, c)) return;
  __shared__ double sbuf[4 * KC * LDSS];
  update_tile(S, r, c, kb, ke, sbuf);
}

// The trailing update (every column from c0 on) on 128 x 64 tiles: column c's
// row blocks c, c+1, ... taken in pairs (the last one alone when their count
// is odd).  Tile t -> (c, first row block).  grid (tiles, nslots)
__device__ inline bool update_tile_w(int tile, int nb, int c0, int& r, int& c) {
  int t = tile;
  for (c = c0; c < nb; ++c) {
    const int pairs = (nb - c + 1) / 2;
    if (t < pairs) {
      r = c + 2 * t;
      return true;
    }
    t -= pairs;
  }
  return false;
}

__global__ __launch_bounds__(TPB, 2) void k_potrf_update_w(const Slot* __restrict__ slots,
                                                        const SlotParams* __restrict__ params,
                                                        SlotOut* __restrict__ out, int kb, int ke, int c0) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || params[b].mode == MODE_FACTOR) return;
  const int nb = S.nblk;
  if (c0 >= nb || ke > nb) return;
  int r, c;
  if (!update_tile_w(ts.tile, nb, c0, r, c)) return;
  const int i0 = r * NB, j0 = c * NB, ld = S.npad;
  const int rows_valid = (r + 1 < nb) ? WROWS : NB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  __shared__ double sbuf[2 * KC * (LDSW + LDSS)];
  d4 acc[4][2];
  double* C = S.A + (size_t)j0 * ld + i0;
  accw_load(acc, C, ld, lane, wr, wc, rows_valid);
  tile_mainloop_w<true>(S.A, ld, i0, S.A, ld, j0, kb * NB, ke * NB, sbuf, acc, tid, lane, wr, wc, rows_valid);
  accw_store(acc, C, ld, lane, wr, wc, rows_valid);
}

// The same update for one slot by a persistent grid: workgroup g takes tiles
// g, g + G, g + 2G, ... (G = gridDim.x).  Launched with about two
// workgroups per CU, it leaves room on every CU (LDS and registers) for the
// diagonal-block, panel and look-ahead launches of the next panel, which
// run beside it on another stream.  Each tile gets the same arithmetic as in
// k_potrf_update.
__global__ __launch_bounds__(TPB) void k_potrf_update_persistent(const Slot* __restrict__ slots,
                                                                 const SlotParams* __restrict__ params,
                                                                 SlotOut* __restrict__ out, int kb, int ke, int c0,
                                                                 int c1, int tiles) {
  const Slot S = slots[0];
  if (!slot_live(params[0], out[0]) || params[0].mode == MODE_FACTOR) return;
  const int nb = S.nblk;
  if (c0 >= nb || ke > nb) return;
  __shared__ double sbuf[4 * KC * LDSS];
  for (int t = blockIdx.x; t < tiles; t += gridDim.x) {
    int r, c;
    if (!update_tile_rc(t, nb, c0, c1, r, c)) return;  // tiles are in increasing order: the rest are out too
    update_tile(S, r, c, kb, ke, sbuf);
  }
}

// ---------------------------------------------------------------------------
// (3) L^-1 by recursive doubling.  At level s (block size s tiles), for each
// aligned pair (top = 2ps, bot = top + s):
//   step 0:  T = L(bot, top) * Linv(top, top)          -> W upper (transposed)
//   step 1:  Linv(bot, top) = -Linv(bot, bot) * T      -> W lower
// grid (npairs_max * s * s, nslots), 256 threads, four waves per SIMD (124
// registers, no spill: 64 x n = 1024 L^-1 0.664 -> 0.651 ms against three)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB, 4) void k_trtri(const Slot* __restrict__ slots, const SlotParams* __restrict__ params,
                                               SlotOut* __restrict__ out, int s, int step) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b])) return;
  const int per_pair = s * s;
  const int pr = ts.tile / per_pair;
  const int rem = ts.tile % per_pair;
  // tiles in order of falling K length (longest first, the short ones fill
  // the tail), and consecutive workgroups, which the dispatcher deals to the
  // 8 XCDs in turn, with equal K: step 0's K falls with tj, step 1's grows
  // with ti
  const int ti = (step == 0) ? rem % s : s - 1 - rem / s;
  const int tj = (step == 0) ? rem / s : rem % s;
  const int top = 2 * pr * s, bot = top + s;
  const int it = bot + ti, jt = top + tj;  // tile row / col of the output
  if (bot >= S.nblk || it >= S.nblk) return;
  const int ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int i0 = it * NB, j0 = jt * NB;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  if (step == 0) {
    // T(i, j) = sum_{k = jt*NB}^{(top + s)*NB - 1} L(i, k) Linv(k, j)
    tile_mainloop<false, true, false>(S.A, ld, i0, S.W, ld, j0, jt * NB, (top + s) * NB, sbuf, acc, tid, lane, wr,
                                      wc);
    // T(i, j) stored at W[(i0 + row) * ld + j0 + col] (strict upper triangle of W)
    acc_store_t(acc, S.W + (size_t)i0 * ld + j0, ld, lane, wr, wc);
  } else {
    // Linv(i, j) = -sum_{k = bot*NB}^{(it + 1)*NB - 1} Linv(i, k) T(k, j)
    tile_mainloop<false, false, true>(S.W, ld, i0, S.W, ld, j0, bot * NB, (it + 1) * NB, sbuf, acc, tid, lane, wr,
                                      wc);
    acc_store(acc, S.W + (size_t)j0 * ld + i0, ld, lane, wr, wc);
  }
}

// ---------------------------------------------------------------------------
// (4) triangular products with the explicit inverse.
// V columns: [0, p) F*, p y*, p+1 r, p+2 E*, p+3 x, p+4.. p+6 estimator scratch
// ---------------------------------------------------------------------------
// [F* | y*] = Linv [F | y]
__global__ __launch_bounds__(MV_TPB) void k_solve_fy(const Slot* __restrict__ slots,
                                                  const SlotParams* __restrict__ params,
                                                  SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || ts.tile >= S.nblk) return;
  const int rb = ts.tile;
  // [F | y] -> V columns 0 .. p in one pass over this row block of Linv
  lower_matvec_block(S, rb, S.F, S.n, S.n, S.p + 1, S.V, S.npad, nullptr, S.y);
}

// GLS: R* = chol_upper(F*'F*), beta = R*^-1 R*^-T F*'y*, r = y - F beta
// (KrigingImpl.cpp:105-121).  One workgroup per slot.
__global__ __launch_bounds__(TPB) void k_gls(const Slot* __restrict__ slots, const SlotParams* __restrict__ params,
                                             SlotOut* __restrict__ out) {
  const int b = blockIdx.x;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b])) return;
  const int tid = threadIdx.x, p = S.p, n = S.n, ld = S.npad;
  __shared__ double G[PMAX * (PMAX + 1)];  // p x (p + 1): [F*'F* | F*'y*]
  __shared__ double part[TPB];
  __shared__ double beta[PMAX];
  for (int a = 0; a < p; ++a)
    for (int c = a; c <= p; ++c) {  // column c == p is F*'y*
      double s = 0.0;
      for (int i = tid; i < n; i += TPB) s = fma(S.V[(size_t)a * ld + i], S.V[(size_t)c * ld + i], s);
      part[tid] = s;
      __syncthreads();
      for (int h = TPB / 2; h > 0; h >>= 1) {
        if (tid < h) part[tid] += part[tid + h];
        __syncthreads();
      }
      if (tid == 0) G[a + c * PMAX] = part[0];
      __syncthreads();
    }
  if (tid == 0) {
    // upper Cholesky U'U = G (stored in place in the upper triangle)
    bool ok = true;
    for (int j = 0; j < p; ++j) {
      double s = G[j + j * PMAX];
      for (int q = 0; q < j; ++q) s -= G[q + j * PMAX] * G[q + j * PMAX];
      ok = ok && (s > 0.0);
      const double ujj = sqrt(s);
      G[j + j * PMAX] = ujj;
      for (int c = j + 1; c < p; ++c) {
        double t = G[j + c * PMAX];
        for (int q = 0; q < j; ++q) t -= G[q + j * PMAX] * G[q + c * PMAX];
        G[j + c * PMAX] = t / ujj;
      }
    }
    // beta: U' z = h, U beta = z
    double z[PMAX];
    for (int a = 0; a < p; ++a) {
      double t = G[a + p * PMAX];
      for (int q = 0; q < a; ++q) t -= G[q + a * PMAX] * z[q];
      z[a] = t / G[a + a * PMAX];
    }
    for (int a = p - 1; a >= 0; --a) {
      double t = z[a];
      for (int q = a + 1; q < p; ++q) t -= G[a + q * PMAX] * z[q];
      z[a] = t / G[a + a * PMAX];
    }
    for (int a = 0; a < p; ++a) {
      beta[a] = z[a];
      out[b].beta[a] = z[a];
    }
    double* rs = S.red + red_rstar_off(S.nblk);
    double ldr = 0.0;
    for (int c = 0; c < p; ++c) {
      ldr += log(G[c + c * PMAX]);
      for (int a = 0; a < p; ++a) rs[a + c * p] = (a <= c) ? G[a + c * PMAX] : 0.0;
    }
    out[b].logdet_rstar = ldr;
    if (!ok) out[b].status = ST_CHOL_FAILED;
  }
  __syncthreads();
  // residual r = y - F beta into V column p + 1 (zero in padded rows)
  double* r = S.V + (size_t)(p + 1) * ld;
  for (int i = tid; i < ld; i += TPB) {
    double v = 0.0;
    if (i < n) {
      double s = 0.0;
      for (int a = 0; a < p; ++a) s = fma(S.F[(size_t)a * n + i], beta[a], s);
      v = S.y[i] - s;
    }
    r[i] = v;
  }
}

// E* = Linv r, SSE partials
__global__ __launch_bounds__(MV_TPB) void k_solve_e(const Slot* __restrict__ slots,
                                                 const SlotParams* __restrict__ params,
                                                 SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || ts.tile >= S.nblk) return;
  const double* r = S.V + (size_t)(S.p + 1) * S.npad;
  double* e = S.V + (size_t)(S.p + 2) * S.npad;
  lower_matvec_block(S, ts.tile, r, S.npad, S.n, 1, e, S.npad, S.red + red_sse_off(S.nblk));
}

// x = Linv^T E*, and the 1-norms (max column abs sums) of L and Linv
__global__ __launch_bounds__(MV_TPB) void k_solve_x(const Slot* __restrict__ slots,
                                                 const SlotParams* __restrict__ params,
                                                 SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || ts.tile >= S.nblk) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, ld = S.npad, n = S.n;
  const double* e = S.V + (size_t)(S.p + 2) * ld;
  double* x = S.V + (size_t)(S.p + 3) * ld;
  double mW = 0.0, mL = 0.0;
  // ||L||_1: L's columns are read only inside the diagonal block; below it
  // the panel kernel's per-block column sums stand in for them (a factor
  // given as input had no panel step: then all of L is read)
  const int lend = (params[b].mode == MODE_FACTOR) ? ld : (ts.tile + 1) * NB;
  const double* lpart = S.red + red_lcol_off(S.nblk);
  // wave w takes the block's columns c0 .. c0 + CW - 1 together: one pass
  // down the rows with one load of E* and CW coalesced column loads in
  // flight per row group (the reductions run once per column at the end)
  constexpr int CW = NB / MV_W;  // columns per wave
  const int c0 = ts.tile * NB + CW * w;
  double s[CW], aw[CW];
#pragma unroll
  for (int q = 0; q < CW; ++q) {
    s[q] = 0.0;
    aw[q] = 0.0;
  }
  for (int i = c0 + lane; i < ld; i += 64) {
    const double ev = gload(e + i);
    double wv[CW];
#pragma unroll
    for (int q = 0; q < CW; ++q) wv[q] = gload(S.W + (size_t)(c0 + q) * ld + i);
#pragma unroll
    for (int q = 0; q < CW; ++q)
      if (i >= c0 + q) {  // the lower triangle only
        s[q] = fma(wv[q], ev, s[q]);
        aw[q] += fabs(wv[q]);
      }
  }
#pragma unroll
  for (int q = 0; q < CW; ++q) {
    const int c = c0 + q;
    const double* __restrict__ Ac = S.A + (size_t)c * ld;
    double al = 0.0;
    for (int i = c + lane; i < lend; i += 64) al += fabs(gload(Ac + i));
    if (lend < ld)
      for (int r = ts.tile + 1 + lane; r < S.nblk; r += 64) al += lpart[(size_t)r * ld + c];
    const double sx = wave_sum(s[q]);
    const double awx = wave_sum(aw[q]);
    al = wave_sum(al);
    if (lane == 0) x[c] = sx;
    if (c < n) {
      mW = fmax(mW, awx);
      mL = fmax(mL, al);
    }
  }
  __shared__ double sm[2][MV_W];
  if (lane == 0) {
    sm[0][w] = mW;
    sm[1][w] = mL;
  }
  __syncthreads();
  if (tid == 0) {
    double a = sm[0][0], l = sm[1][0];
    for (int q = 1; q < MV_W; ++q) {
      a = fmax(a, sm[0][q]);
      l = fmax(l, sm[1][q]);
    }
    S.red[red_colW_off(S.nblk) + ts.tile] = a;
    S.red[red_colL_off(S.nblk) + ts.tile] = l;
  }
}

// ---------------------------------------------------------------------------
// (5) LL value and condition check (Kriging.cpp:247-289,
// LinearAlgebra.cpp:68-115).  One thread per slot.
// ---------------------------------------------------------------------------
__global__ void k_finalize_ll(const Slot* __restrict__ slots, const SlotParams* __restrict__ params,
                              SlotOut* __restrict__ out, int nslots) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nslots) return;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  SlotOut& O = out[b];
  if (!slot_live(P, O)) return;
  const int nb = S.nblk;
  double logdet = 0.0, sse = 0.0, nL = 0.0, nW = 0.0;
  for (int q = 0; q < nb; ++q) {
    logdet += S.red[RED_LOGDET + q];
    sse += S.red[red_sse_off(nb) + q];
    nL = fmax(nL, S.red[red_colL_off(nb) + q]);
    nW = fmax(nW, S.red[red_colW_off(nb) + q]);
  }
  const double n = (double)S.n;
  const double two_pi = 2.0 * 3.14159265358979323846;
  O.logdet = logdet;
  O.sse = sse;
  O.norm1_L = nL;
  O.norm1_Linv = nW;
  const double rc = (1.0 / nL) / nW;
  O.rcond2 = rc * rc;
  if (P.rcond_check && !(O.rcond2 >= P.min_rcond) && O.status == ST_OK)
    O.status = ST_RCOND_SUSPECT;
  // a given factor: rcond from the 1-norm estimator, as arma::rcond
  if (P.mode == MODE_FACTOR && O.status == ST_OK) O.status = ST_RCOND_SUSPECT;
  double ll, s2g;
  if (P.objective == OBJ_LMP) {
    // log marginal likelihood of Kriging::_logMargPost (Kriging.cpp:560-583),
    // S_2 = y'R^-1 y - y'P y = SSE; the prior term is added by the host
    const double nmp = n - S.p;
    if (S.noise == NOISE_NUGGET) {
      if (P.est_sigma2 && P.est_nugget) s2g = sse / nmp;
      else if (P.est_sigma2 || P.est_nugget) s2g = P.sigma2 / P.alpha;
      else s2g = P.sigma2 + P.nugget;
    } else {
      s2g = P.est_sigma2 ? sse / nmp : P.sigma2;
    }
    O.ll = -logdet - O.logdet_rstar - 0.5 * nmp * log(s2g * nmp);
    O.sigma2 = s2g;
    return;
  }
  if (S.noise == NOISE_NUGGET) {
    const double a = P.alpha;
    double s2 = P.sigma2, nug = P.nugget;
    if (P.est_sigma2) {
      if (P.est_nugget) {
        const double var = sse / n;
        s2 = a * var;
        nug = (1.0 - a) * var;
      } else {
        s2 = P.nugget * a / (1.0 - a);
      }
    } else if (P.est_nugget) {
      nug = P.sigma2 * (1.0 - a) / a;
    }
    const double tv = s2 + nug;
    ll = -0.5 * (n * log(two_pi * tv) + 2.0 * logdet + sse / tv);
    s2g = tv;
  } else if (P.est_sigma2) {
    s2g = sse / n;
    ll = -0.5 * (n * log(two_pi * s2g) + 2.0 * logdet + n);
  } else {
    s2g = P.sigma2;
    ll = -0.5 * (n * log(two_pi * s2g) + 2.0 * logdet + sse / s2g);
  }
  O.ll = ll;
  O.sigma2 = s2g;
}

// ---------------------------------------------------------------------------
// (6) fused R^-1 tile + adjoint contraction.  Tile (i >= j) of
// R^-1 = Linv^T Linv is accumulated on MFMA; then every strictly-lower
// element (gi > gj) contributes, with dR_ij/dtheta_k = R_ij g_k(x_j - x_i)
// recomputed from X (KrigingImpl.cpp:871-884, Kriging.cpp:303-316):
//   s1_k += (x_i x_j - sigma2 Rinv_ij) R_ij g_k   (the finaliser's
//           s1 / sigma2 - s2 in one sum; the s2 partials are zero)
//   a1   += x_i x_j rho_ij        a2   += Rinv_ij rho_ij   (nugget models
//           whose parameter vector has alpha only)
// and diagonal elements add x_i^2 and Rinv_ii.  Partials per tile go to
// red[grad_off + tile * GRAD_STRIDE]: [s1 (DT) | s2 (DT) | a1 a2 dg1 dg2].
// LMP: the log-marginal-posterior gradient (KrigingImpl.cpp:887-921) has the
// same contraction with B = R^-1 - R^-1 F (F'R^-1 F)^-1 F'R^-1 = R^-1 - G'G in
// place of R^-1 (G = R*^-T F*' L^-1, columns of V from the leave-one-out
// pass k_loo_g), subtracted from each tile before the contraction.
// grid (nblk_max (nblk_max + 1) / 2, nslots), 256 threads.  Waves per SIMD:
// 4 up to 8 dimensions (128 registers: 0.717 -> 0.680 ms per 64 x n = 1024
// launch against 3 waves; the diagonal elements' sums are taken before the
// element loop; the 8-dimension instance spills a few values), 2
// for 9..16 dimensions (1 before: 9.5 -> 8 ms at n = 8192, d = 10).  Beyond
// 8 dimensions the epilogue runs its four 16 x 16 blocks as a loop (the
// sixteen elements are not unrolled there): with the block picked by
// selects instead of an index into the accumulators, which had moved them
// to scratch memory, n = 8192, d = 10 went from 7.61 to 3.69 ms per launch
// (24 -> 50 TFLOP/s) and 64 x n = 1024, d = 10 from 1.55 to 1.06 ms.  9..12
// dimensions use their own instance, which skips the padded coordinates
// (166 registers, three waves: d = 10 at 3.53 / 0.96 ms); up to 6
// dimensions the 6-dimension instance (64 x n = 1024, d = 6: 0.667 -> 0.630
// ms against the 8-dimension one, which computes two padded coordinates)
// ---------------------------------------------------------------------------
template <int DT, bool LMP>
__global__ __launch_bounds__(TPB, (DT == 16 || DT == 12 ? 2 : (DT <= 8 ? 4 : 1))) void k_lauum_adjoint(const Slot* __restrict__ slots,
                                                       const SlotParams* __restrict__ params,
                                                       SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  if (!slot_live(P, out[b]) || !P.want_grad) return;
  const int ntiles = S.nblk * (S.nblk + 1) / 2;
  if (ts.tile >= ntiles) return;
  int it, jt;
  tri_decode(ts.tile, it, jt);
  const int ld = S.npad, n = S.n, d = S.d;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int i0 = it * NB, j0 = jt * NB;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  // Linv(k, i0 + ii) Linv(k, j0 + jj) over k >= it * NB
  // (the epilogue reads only elements on or below the diagonal)
  tile_mainloop<true, true, false, true>(S.W, ld, i0, S.W, ld, j0, it * NB, S.nblk * NB, sbuf, acc, tid, lane, wr,
                                         wc, !upper_quadrant(it == jt, wr, wc));

  // epilogue: X rows of both blocks and the x vector in LDS (reuse sbuf)
  // the X rows need 2 DT NB doubles: in the staging buffer when it is large
  // enough (it is idle now), else in an array of their own
  constexpr bool kReuse = 4 * KC * LDSS >= 2 * DT * NB;
  __shared__ double sXown[kReuse ? 1 : 2 * DT * NB];
  double* sXi = kReuse ? sbuf : sXown;  // d x NB
  double* sXj = sXi + DT * NB;
  __shared__ double sxi[NB], sxj[NB];
  const double* xv = S.V + (size_t)(S.p + 3) * ld;
  // all DT coordinate rows, zero from d on: the padded coordinate
  // differences are then zero without a select per element
  for (int q = tid; q < DT * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    sXi[q] = (k < d && i0 + r < n) ? S.X[(size_t)k * n + i0 + r] : 0.0;
    sXj[q] = (k < d && j0 + r < n) ? S.X[(size_t)k * n + j0 + r] : 0.0;
  }
  if (tid < NB) {
    sxi[tid] = xv[i0 + tid];
    sxj[tid] = xv[j0 + tid];
  }
  __shared__ double sG[LMP ? 2 * PMAX * NB : 1];
  if (LMP) {
    const double* Gt = S.V + (size_t)v_loo_g(S.p) * ld;
    for (int q = tid; q < S.p * NB; q += TPB) {
      const int a = q / NB, r = q % NB;
      sG[q] = Gt[(size_t)a * ld + i0 + r];
      sG[PMAX * NB + q] = Gt[(size_t)a * ld + j0 + r];
    }
  }
  // 1 / theta_k: in registers up to 8 dimensions; beyond, read from LDS
  // (broadcast) so that the epilogue's registers leave room for more waves
  constexpr bool kItLds = DT > 8;
  __shared__ double sIt[kItLds ? DT : 1];
  double itr[kItLds ? 1 : DT];
  if (kItLds) {
    if (tid < DT) sIt[tid] = 1.0 / ((tid < d) ? P.theta[tid] : 1.0);
  } else {
#pragma unroll
    for (int k = 0; k < (kItLds ? 1 : DT); ++k) itr[k] = 1.0 / ((k < d) ? P.theta[k] : 1.0);
  }
  __syncthreads();
  const double* it_ = kItLds ? sIt : itr;
  const double factor = (S.noise == NOISE_NUGGET) ? P.alpha : 1.0;
  const int kern = S.kernel;
  // one sum per theta_k: sum (x_i x_j - sigma2 Rinv_ij) dR_ij/dtheta_k, the
  // finaliser's s1 / sigma2 - s2 with sigma2 (k_finalize_ll, earlier in the
  // pipeline) folded in per element; the s2 partials are written as zero
  const double s2g = out[b].sigma2;
  double sg[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) sg[k] = 0.0;
  double a1 = 0.0, a2 = 0.0;
  // the diagonal elements' sums first (diagonal tiles only), reduced at once
  // so that their accumulators are not live across the element loop
  __shared__ double sred[4][2 * DT + 4];
  if (it == jt) {
    double dg1 = 0.0, dg2 = 0.0;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int r = acc_row(lane, wr, bi), c = acc_col(lane, wc, bj, e);
          const int gi = i0 + r, gj = j0 + c;
          if (gi >= n || gi != gj) continue;
          double rinv = acc[bi][bj][e];
          if (LMP)
            for (int a = 0; a < S.p; ++a) rinv = fma(-sG[a * NB + r], sG[PMAX * NB + a * NB + c], rinv);
          dg1 += sxi[r] * sxj[c];
          dg2 += rinv;
        }
    const double v2 = wave_sum(dg1), v3 = wave_sum(dg2);
    if (lane == 0) {
      sred[w][2 * DT + 2] = v2;
      sred[w][2 * DT + 3] = v3;
    }
  } else if (lane == 0) {
    sred[w][2 * DT + 2] = 0.0;
    sred[w][2 * DT + 3] = 0.0;
  }

  // the alpha direction's sums only for a nugget model whose gamma has alpha
  const bool need_alpha = S.noise == NOISE_NUGGET && P.gamma_has_alpha;
  // unit inverse length-scales for kern_rho on scaled differences (folded)
  [[maybe_unused]] double kOnes[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) kOnes[k] = 1.0;
  // one element (r, c) of the tile, R^-1 value rinv
  auto element = [&](auto ktag, int r, int c, double rinv) __attribute__((always_inline)) {
    constexpr int KT = decltype(ktag)::value;
    const int gi = i0 + r, gj = j0 + c;
    if (gi >= n || gj > gi) return;
    if (LMP)
      for (int a = 0; a < S.p; ++a) rinv = fma(-sG[a * NB + r], sG[PMAX * NB + a * NB + c], rinv);
    const double xx = sxi[r] * sxj[c];
    if (gi == gj) return;
    // in the 12-dimension instance the coordinates from d on are skipped
    // (uniform branches): fewer registers, three waves per SIMD
    constexpr bool kSkip = DT == 12;
    double dx[DT];
#pragma unroll
    for (int k = 0; k < DT; ++k) dx[k] = sXj[k * NB + c] - sXi[k * NB + r];
    // rho first, then each g_k recomputed from dx_k while accumulating:
    // no g[] held across the exp (fewer live registers, more waves)
    const double rho = kern_rho<DT>(KT, dx, it_, kSkip ? d : DT);
    const double Rij = rho * factor;
    const double wf = fma(-s2g, rinv, xx) * Rij;  // one product per element, not per k
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      if (kSkip && k >= d) break;
      sg[k] = fma(wf, kern_g(KT, dx[k], it_[k]), sg[k]);
    }
    a1 = fma(xx, rho, a1);
    a2 = fma(rinv, rho, a2);
  };
  if constexpr (DT <= 8) {
    // the element loop specialised per covariance kernel: the executed code
    // is one kernel's, contiguous, instead of all four's interleaved per
    // element (0.611 -> 0.563 ms per 64 x n = 1024, d = 6 launch)
    auto elements = [&](auto ktag) __attribute__((always_inline)) {
      constexpr int KT = decltype(ktag)::value;
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const int r = acc_row(lane, wr, bi), c = acc_col(lane, wc, bj, e);
            const int gi = i0 + r, gj = j0 + c;
            double rinv = acc[bi][bj][e];
            if (gi >= n || gj > gi) continue;
            if (LMP)
              for (int a = 0; a < S.p; ++a) rinv = fma(-sG[a * NB + r], sG[PMAX * NB + a * NB + c], rinv);
            const double xx = sxi[r] * sxj[c];
            if (gi == gj) continue;
            // scaled coordinate differences dx_k / theta_k (the products
            // kern_rho forms): rho from them, and g_k = d ln rho / d theta_k
            // as g(dx_k / theta_k; 1) / theta_k, the 1 / theta_k applied once
            // per tile (one product per coordinate fewer)
            double dx[DT];
#pragma unroll
            for (int k = 0; k < DT; ++k) dx[k] = (sXj[k * NB + c] - sXi[k * NB + r]) * it_[k];
            // rho first, then each g_k recomputed from dx_k while accumulating:
            // no g[] held across the exp (fewer live registers, more waves);
            // the padded coordinates are computed (zeros): skipping them by
            // branches here makes the unrolled elements spill
            const double rho = kern_rho<DT>(KT, dx, kOnes);
            const double Rij = rho * factor;
            const double wf = fma(-s2g, rinv, xx) * Rij;  // one product per element, not per k
#pragma unroll
            for (int k = 0; k < DT; ++k) sg[k] = fma(wf, kern_g(KT, dx[k], 1.0), sg[k]);
            if (need_alpha) {
              a1 = fma(xx, rho, a1);
              a2 = fma(rinv, rho, a2);
            }
          }
    };
    switch (kern) {
      case K_GAUSS: elements(std::integral_constant<int, K_GAUSS>{}); break;
      case K_EXP: elements(std::integral_constant<int, K_EXP>{}); break;
      case K_MATERN32: elements(std::integral_constant<int, K_MATERN32>{}); break;
      default: elements(std::integral_constant<int, K_MATERN52>{}); break;
    }
  } else {
    // beyond 8 dimensions the sixteen elements do not unroll: the four
    // 16 x 16 blocks in a loop, each picked from the accumulators by
    // selects (a loop index into acc would move it to scratch memory)
    // (specialised per covariance kernel, as above)
    auto blocks = [&](auto ktag) __attribute__((always_inline)) {
#pragma unroll 1
      for (int q = 0; q < 4; ++q) {
        const d4 blk = (q == 0) ? acc[0][0] : ((q == 1) ? acc[0][1] : ((q == 2) ? acc[1][0] : acc[1][1]));
#pragma unroll
        for (int e = 0; e < 4; ++e) element(ktag, acc_row(lane, wr, q >> 1), acc_col(lane, wc, q & 1, e), blk[e]);
      }
    };
    switch (kern) {
      case K_GAUSS: blocks(std::integral_constant<int, K_GAUSS>{}); break;
      case K_EXP: blocks(std::integral_constant<int, K_EXP>{}); break;
      case K_MATERN32: blocks(std::integral_constant<int, K_MATERN32>{}); break;
      default: blocks(std::integral_constant<int, K_MATERN52>{}); break;
    }
  }

  // up to 8 dimensions g_k carried 1 / theta_k as a common factor
  if constexpr (DT <= 8) {
#pragma unroll
    for (int k = 0; k < DT; ++k) sg[k] *= it_[k];
  }
  // workgroup reduction of the 2 * DT + 4 partials
  double* dst = S.red + red_grad_off(S.nblk) + (size_t)ts.tile * GRAD_STRIDE;
#pragma unroll
  for (int k = 0; k < DT; ++k) {
    const double v1 = wave_sum(sg[k]);
    if (lane == 0) {
      sred[w][k] = v1;
      sred[w][DT + k] = 0.0;
    }
  }
  {
    const double v0 = wave_sum(a1), v1 = wave_sum(a2);
    if (lane == 0) {
      sred[w][2 * DT + 0] = v0;
      sred[w][2 * DT + 1] = v1;
    }
  }
  __syncthreads();
  if (tid < 2 * DT + 4) {
    const double v = sred[0][tid] + sred[1][tid] + sred[2][tid] + sred[3][tid];
    int q;
    if (tid < DT) q = tid;                       // s1 (sigma2 folded in, see above)
    else if (tid < 2 * DT) q = DMAX + tid - DT;  // s2 (zero)
    else q = 2 * DMAX + (tid - 2 * DT);          // a1 a2 dg1 dg2
    dst[q] = v;
  }
}

// ---------------------------------------------------------------------------
// (7) gradient (Kriging.cpp:303-326): sums of the tile partials in tile order
// ---------------------------------------------------------------------------
constexpr int FG_TPB = 1024;  // 16 waves
// a slot of at least FG_SPLIT_MIN tiles (n >= 4096) is summed in two levels:
// FG_GROUPS workgroups each add a contiguous run of its tiles in tile order
// and leave the sum in the run's first tile row, then k_finalize_grad adds
// those FG_GROUPS rows in order (one workgroup alone reads the 4.5 MB of an
// n = 8192 slot's partials at a fraction of the chip's bandwidth)
constexpr int FG_SPLIT_MIN = 2048;
constexpr int FG_GROUPS = 128;
constexpr int FG1_TPB = 256;
__device__ inline int fg_group_start(int ntiles, int g) { return (int)(((long long)ntiles * g) / FG_GROUPS); }

__global__ __launch_bounds__(FG1_TPB) void k_grad_groups(const Slot* __restrict__ slots,
                                                         const SlotParams* __restrict__ params,
                                                         SlotOut* __restrict__ out) {
  const int g = blockIdx.x, b = blockIdx.y;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  if (!slot_live(P, out[b]) || !P.want_grad) return;
  const int ntiles = S.nblk * (S.nblk + 1) / 2;
  if (ntiles < FG_SPLIT_MIN) return;
  constexpr int NW = FG1_TPB / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double* src = S.red + red_grad_off(S.nblk);
  const int g0 = fg_group_start(ntiles, g), g1 = fg_group_start(ntiles, g + 1);
  const int t0 = g0 + (g1 - g0) * w / NW, t1 = g0 + (g1 - g0) * (w + 1) / NW;
  __shared__ double part[NW][GRAD_STRIDE];
  for (int q = lane; q < GRAD_STRIDE; q += 64) {
    double sm[4] = {0.0, 0.0, 0.0, 0.0};
    int t = t0;
    for (; t + 3 < t1; t += 4) {
#pragma unroll
      for (int u = 0; u < 4; ++u) sm[u] += src[(size_t)(t + u) * GRAD_STRIDE + q];
    }
    for (; t < t1; ++t) sm[0] += src[(size_t)t * GRAD_STRIDE + q];
    part[w][q] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
  }
  // every read of this run is done before its first row is overwritten;
  // the runs of the other workgroups are disjoint
  __syncthreads();
  for (int q = threadIdx.x; q < GRAD_STRIDE; q += FG1_TPB) {
    double s = part[0][q];
#pragma unroll
    for (int v = 1; v < NW; ++v) s += part[v][q];
    src[(size_t)g0 * GRAD_STRIDE + q] = s;
  }
}

__global__ __launch_bounds__(FG_TPB) void k_finalize_grad(const Slot* __restrict__ slots,
                                                          const SlotParams* __restrict__ params,
                                                          SlotOut* __restrict__ out, int nslots) {
  const int b = blockIdx.x;
  if (b >= nslots) return;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  SlotOut& O = out[b];
  if (!slot_live(P, O) || !P.want_grad) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ntiles = S.nblk * (S.nblk + 1) / 2;
  constexpr int NW = FG_TPB / 64;
  __shared__ double tot[GRAD_STRIDE];
  __shared__ double part[NW][GRAD_STRIDE];
  // wave w sums the tiles of its sixteenth, lane = component (a tile's
  // partials are contiguous: coalesced rows), eight running sums in tile
  // order (eight loads in flight); the sixteenths are added in order:
  // deterministic for a given tile count
  if (ntiles >= FG_SPLIT_MIN) {
    // the FG_GROUPS run sums of k_grad_groups, in run order
    const double* src = S.red + red_grad_off(S.nblk);
    constexpr int RW = FG_GROUPS / NW;  // runs per wave
    static_assert(FG_GROUPS % NW == 0 && FG_SPLIT_MIN >= 16 * FG_GROUPS, "whole runs per wave, >= 16 tiles a run");
    for (int q = lane; q < GRAD_STRIDE; q += 64) {
      double sm = 0.0;
#pragma unroll
      for (int u = 0; u < RW; ++u) sm += src[(size_t)fg_group_start(ntiles, w * RW + u) * GRAD_STRIDE + q];
      part[w][q] = sm;
    }
  } else {
    const double* src = S.red + red_grad_off(S.nblk);
    const int t0 = (int)(((long long)ntiles * w) / NW);
    const int t1 = (int)(((long long)ntiles * (w + 1)) / NW);
    for (int q = lane; q < GRAD_STRIDE; q += 64) {
      double sm[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) sm[u] = 0.0;
      int t = t0;
      for (; t + 7 < t1; t += 8) {
#pragma unroll
        for (int u = 0; u < 8; ++u) sm[u] += src[(size_t)(t + u) * GRAD_STRIDE + q];
      }
      for (; t < t1; ++t) sm[0] += src[(size_t)t * GRAD_STRIDE + q];
      part[w][q] = ((sm[0] + sm[1]) + (sm[2] + sm[3])) + ((sm[4] + sm[5]) + (sm[6] + sm[7]));
    }
  }
  __syncthreads();
  for (int q = threadIdx.x; q < GRAD_STRIDE; q += FG_TPB) {
    double s = part[0][q];
    for (int v = 1; v < NW; ++v) s += part[v][q];
    tot[q] = s;
  }
  __syncthreads();
  const int q = threadIdx.x;
  if (q != 0) return;
  const int d = S.d;
  const double s2g = O.sigma2;
  for (int k = 0; k < d; ++k) {
    // term1 = 2 s1, term2 = -2 s2, grad = (term1 / sigma2 + term2) / 2
    O.grad[k] = tot[k] / s2g - tot[DMAX + k];
  }
  if (P.gamma_has_alpha && S.noise == NOISE_NUGGET) {
    const double a1 = tot[2 * DMAX], a2 = tot[2 * DMAX + 1], dg1 = tot[2 * DMAX + 2], dg2 = tot[2 * DMAX + 3];
    const double a = P.alpha;
    double ga;
    if (P.objective == OBJ_LMP) {
      // -tr(Wb)/2 + y'Wb'Q_output/(2 sigma2), Wb = D B, D = R / alpha with
      // zero diagonal (Kriging.cpp:612-618); the prior term is the host's
      ga = a1 / s2g - a2;
    } else if (P.est_sigma2 && P.est_nugget) {
      // -0.5 * (-(x' D x) / var + <Rinv, D>), D = R / alpha with zero diagonal
      ga = -0.5 * (-(2.0 * a1) / s2g + 2.0 * a2);
    } else if (P.est_sigma2 && !P.est_nugget) {
      const double tv = s2g;
      const double q1 = 2.0 * a1 + dg1;
      const double tr = (2.0 * a2 + dg2) / tv;
      ga = -0.5 * (-q1 / (tv * tv) + tr) * P.nugget / (1.0 - a) / (1.0 - a);
    } else {
      ga = 0.0;
    }
    O.grad[d] = ga;
  }
}

// ---------------------------------------------------------------------------
// (8) 1-norm estimate of Linv (Higham, ACM TOMS 14 (1988) 381-396, Alg. 4.1
// with the alternating-sign extra estimate), run only for slots whose exact
// rcond fell below the threshold.  The estimate never exceeds the exact norm,
// so an exact rcond above the threshold implies the estimate-based one is
// too.  One workgroup per slot; uses V columns p+4 .. p+6 as scratch.
// ---------------------------------------------------------------------------
__device__ void est_lower_mv(const Slot& S, const double* __restrict__ in, double* __restrict__ o) {
  // o = Linv * in  (row i: sum_{k <= i} W[k*ld + i] in[k]) ; thread per row
  const int ld = S.npad, n = S.n;
  for (int i = threadIdx.x; i < n; i += TPB) {
    double s = 0.0;
    for (int k = 0; k <= i; ++k) s = fma(S.W[(size_t)k * ld + i], in[k], s);
    o[i] = s;
  }
}
__device__ void est_upper_mv(const Slot& S, const double* __restrict__ in, double* __restrict__ o) {
  // o = Linv^T * in  (column c: sum_{i >= c} W[c*ld + i] in[i]) ; wave per column
  const int ld = S.npad, n = S.n, lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  for (int c = w; c < n; c += TPB / 64) {
    double s = 0.0;
    for (int i = c + lane; i < n; i += 64) s = fma(S.W[(size_t)c * ld + i], in[i], s);
    s = wave_sum(s);
    if (lane == 0) o[c] = s;
  }
}

__global__ __launch_bounds__(TPB) void k_rcond_estimate(const Slot* __restrict__ slots,
                                                        const SlotParams* __restrict__ params,
                                                        SlotOut* __restrict__ out) {
  const int b = blockIdx.x;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  SlotOut& O = out[b];
  if (!P.active || O.status != ST_RCOND_SUSPECT) return;
  const int n = S.n, ld = S.npad, tid = threadIdx.x;
  double* x = S.V + (size_t)(S.p + 4) * ld;
  double* xi = S.V + (size_t)(S.p + 5) * ld;
  double* y = S.V + (size_t)(S.p + 6) * ld;
  __shared__ double sh[TPB];
  __shared__ int shi[TPB];
  __shared__ double sest, sestold;
  __shared__ int sj, sflag;

  auto block_sum_abs = [&](const double* v) -> double {
    double s = 0.0;
    for (int i = tid; i < n; i += TPB) s += fabs(v[i]);
    sh[tid] = s;
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
      if (tid < h) sh[tid] += sh[tid + h];
      __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
  };
  auto block_argmax_abs = [&](const double* v) -> int {
    double best = -1.0;
    int bi = 0;
    for (int i = tid; i < n; i += TPB)
      if (fabs(v[i]) > best) {
        best = fabs(v[i]);
        bi = i;
      }
    sh[tid] = best;
    shi[tid] = bi;
    __syncthreads();
    for (int h = TPB / 2; h > 0; h >>= 1) {
      if (tid < h) {
        const double o = sh[tid + h];
        const int oi = shi[tid + h];
        if (o > sh[tid] || (o == sh[tid] && oi < shi[tid])) {
          sh[tid] = o;
          shi[tid] = oi;
        }
      }
      __syncthreads();
    }
    const int r = shi[0];
    __syncthreads();
    return r;
  };

  double est;
  for (int i = tid; i < n; i += TPB) x[i] = 1.0 / n;
  __syncthreads();
  est_lower_mv(S, x, y);
  __syncthreads();
  if (n == 1) {
    est = fabs(y[0]);
  } else {
    est = block_sum_abs(y);
    for (int i = tid; i < n; i += TPB) xi[i] = (y[i] >= 0.0) ? 1.0 : -1.0;
    __syncthreads();
    est_upper_mv(S, xi, x);
    __syncthreads();
    int j = block_argmax_abs(x);
    int iter = 2;
    for (;;) {
      for (int i = tid; i < n; i += TPB) x[i] = (i == j) ? 1.0 : 0.0;
      __syncthreads();
      est_lower_mv(S, x, y);
      __syncthreads();
      const double estold = est;
      est = block_sum_abs(y);
      if (tid == 0) sflag = 0;
      __syncthreads();
      for (int i = tid; i < n; i += TPB)
        if (((y[i] >= 0.0) ? 1.0 : -1.0) != xi[i]) sflag = 1;
      __syncthreads();
      const int changed = sflag;
      __syncthreads();
      if (!changed || est <= estold) break;
      for (int i = tid; i < n; i += TPB) xi[i] = (y[i] >= 0.0) ? 1.0 : -1.0;
      __syncthreads();
      est_upper_mv(S, xi, x);
      __syncthreads();
      const int jlast = j;
      j = block_argmax_abs(x);
      if (x[jlast] != fabs(x[j]) && iter < 5) {
        ++iter;
        continue;
      }
      break;
    }
    for (int i = tid; i < n; i += TPB) x[i] = ((i & 1) ? -1.0 : 1.0) * (1.0 + (double)i / (double)(n - 1));
    __syncthreads();
    est_lower_mv(S, x, y);
    __syncthreads();
    double temp = block_sum_abs(y);
    temp = 2.0 * (temp / (double)(3 * n));
    if (temp > est) est = temp;
  }
  (void)sest;
  (void)sestold;
  (void)sj;
  if (tid == 0) {
    const double rc = (1.0 / O.norm1_L) / est;
    O.rcond2 = rc * rc;
    O.status = (P.mode != MODE_FACTOR && O.rcond2 < P.min_rcond) ? ST_RCOND_LOW : ST_OK;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static inline int tri(int t) { return t * (t + 1) / 2; }

// ---------------------------------------------------------------------------
// (9) prediction: R_on, C = L^-1 R_on, per-point reductions
// ---------------------------------------------------------------------------
// R_on(i, j) = factor * k(x_i - xn_j), 1 where the points coincide
// (|dx_k| <= eps for all k, KrigingImpl.cpp:202-205); zero padding.
// grid (mpad / 64, npad / 64), 256 threads
template <int DT>
__global__ __launch_bounds__(TPB) void k_cross_cov(PredictArgs a) {
  const int j0 = blockIdx.x * NB, i0 = blockIdx.y * NB, tid = threadIdx.x;
  const int n = a.n, m = a.m, d = a.d;
  __shared__ double sXo[DT * NB];
  __shared__ double sXn[DT * NB];
  for (int q = tid; q < d * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    sXo[q] = (i0 + r < n) ? a.X[(size_t)k * n + i0 + r] : 0.0;
    sXn[q] = (j0 + r < m) ? a.Xn[(size_t)k * m + j0 + r] : 0.0;
  }
  double it[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) it[k] = (k < d) ? 1.0 / a.theta[k] : 1.0;
  __syncthreads();
  const int ii = tid & 63, gi = i0 + ii;
  double dx[DT];
  for (int r = 0; r < NB / 4; ++r) {
    const int jj = (tid >> 6) + 4 * r, gj = j0 + jj;
    double v = 0.0;
    if (gi < n && gj < m) {
      bool same = true;
#pragma unroll
      for (int k = 0; k < DT; ++k) {
        dx[k] = (k < d) ? sXo[k * NB + ii] - sXn[k * NB + jj] : 0.0;
        same = same && fabs(dx[k]) <= 2.220446049250313e-16;
      }
      v = same ? 1.0 : kern_rho<DT>(a.kernel, dx, it) * a.factor;
    }
    a.Ron[(size_t)gj * a.npad + gi] = v;
  }
}

// Covariance between two point sets, LinearAlgebra::covMat_rect
// (LinearAlgebra.cpp:600-660) as used by KrigingImpl::covMat:
// out(i, j) = scale k(x1_i - x2_j; theta), n1 x n2 column-major.
// grid (ceil(n2 / 64), ceil(n1 / 64)), 256 threads
template <int DT>
__global__ __launch_bounds__(TPB) void k_cov_mat(CovMatArgs a) {
  const int j0 = blockIdx.x * NB, i0 = blockIdx.y * NB, tid = threadIdx.x;
  const int n1 = a.n1, n2 = a.n2, d = a.d;
  __shared__ double s1[DT * NB];
  __shared__ double s2[DT * NB];
  for (int q = tid; q < d * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    s1[q] = (i0 + r < n1) ? a.X1[(size_t)k * n1 + i0 + r] : 0.0;
    s2[q] = (j0 + r < n2) ? a.X2[(size_t)k * n2 + j0 + r] : 0.0;
  }
  double it[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) it[k] = (k < d) ? 1.0 / a.theta[k] : 1.0;
  __syncthreads();
  const int ii = tid & 63, gi = i0 + ii;
  double dx[DT];
  for (int r = 0; r < NB / 4; ++r) {
    const int jj = (tid >> 6) + 4 * r, gj = j0 + jj;
    if (gi >= n1 || gj >= n2) continue;
#pragma unroll
    for (int k = 0; k < DT; ++k) dx[k] = (k < d) ? s1[k * NB + ii] - s2[k * NB + jj] : 0.0;
    a.out[(size_t)gj * n1 + gi] = (a.unit_diag && gi == gj) ? 1.0 : kern_rho<DT>(a.kernel, dx, it) * a.scale;
  }
}

hipError_t launch_cov_mat(const CovMatArgs& a, hipStream_t stream) {
  if (a.n1 <= 0 || a.n2 <= 0) return hipSuccess;
  const dim3 grid((a.n2 + NB - 1) / NB, (a.n1 + NB - 1) / NB);
  if (a.d <= 4)
    hipLaunchKernelGGL(k_cov_mat<4>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 8)
    hipLaunchKernelGGL(k_cov_mat<8>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 16)
    hipLaunchKernelGGL(k_cov_mat<16>, grid, dim3(TPB), 0, stream, a);
  else
    hipLaunchKernelGGL(k_cov_mat<32>, grid, dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

// C(it, jt) = sum_{kt <= it} Linv(it, kt) R_on(kt, jt); grid (mpad / 64, npad / 64)
__global__ __launch_bounds__(TPB) void k_trmm_lower(PredictArgs a) {
  const int jt = blockIdx.x, it = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int i0 = it * NB, j0 = jt * NB, ld = a.npad;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  // Linv(i0 + ii, k) R_on(k, j0 + jj) over k < (it + 1) * NB
  tile_mainloop<false, true, false>(a.Linv, ld, i0, a.Ron, ld, j0, 0, (it + 1) * NB, sbuf, acc, tid, lane, wr, wc);
  acc_store(acc, a.C + (size_t)j0 * ld + i0, ld, lane, wr, wc);
}

hipError_t launch_trmm_lower(const PredictArgs& a, hipStream_t stream) {
  if (a.mpad <= 0 || a.npad <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_trmm_lower, dim3(a.mpad / NB, a.npad / NB), dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

// S = Cuu - U' U over the lower tiles of S (grid T (T + 1) / 2, T = mpad /
// 64); an off-diagonal tile also writes its transpose
__global__ __launch_bounds__(TPB) void k_schur(SchurArgs a) {
  int ti, tj;
  tri_decode(blockIdx.x, ti, tj);
  const int i0 = ti * NB, j0 = tj * NB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  tile_mainloop<true, true, true>(a.U, a.kdim, i0, a.U, a.kdim, j0, 0, a.kdim, sbuf, acc, tid, lane, wr, wc);
  const int m = a.m;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int gi = i0 + acc_row(lane, wr, bi), gj = j0 + acc_col(lane, wc, bj, e);
        if (gi >= m || gj >= m || (ti == tj && gi < gj)) continue;
        const double v = a.Cuu[(size_t)gj * m + gi] + acc[bi][bj][e];
        a.S[(size_t)gj * m + gi] = v;
        if (gi != gj) a.S[(size_t)gi * m + gj] = v;
      }
}

hipError_t launch_schur(const SchurArgs& a, hipStream_t stream) {
  if (a.m <= 0) return hipSuccess;
  const int T = a.mpad / NB;
  hipLaunchKernelGGL(k_schur, dim3(tri(T)), dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

// grid (ceil(n / 64), ceil(n / 64)), 256 threads; lane = row (coalesced
// stores, the transposed reads of U are strided)
__global__ __launch_bounds__(TPB) void k_block_assemble(BlockAssembleArgs a) {
  const int i = blockIdx.x * NB + (threadIdx.x & 63);
  if (i >= a.n) return;
  for (int r = 0; r < NB / 4; ++r) {
    const int j = blockIdx.y * NB + (threadIdx.x >> 6) + 4 * r;
    if (j >= a.n) break;
    double v = 0.0;
    if (i >= j) {
      if (j < a.no)
        v = (i < a.no) ? a.Loo[(size_t)j * a.ldo + i] : a.U[(size_t)(i - a.no) * a.ldu + j];
      else
        v = a.Luu[(size_t)(j - a.no) * a.ldl + (i - a.no)];
    }
    a.L[(size_t)j * a.n + i] = v;
  }
}

hipError_t launch_block_assemble(const BlockAssembleArgs& a, hipStream_t stream) {
  if (a.n <= 0) return hipSuccess;
  const int T = (a.n + NB - 1) / NB;
  hipLaunchKernelGGL(k_block_assemble, dim3(T, T), dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

// One wave per new point j: mean = f_j beta + C(:, j)' z, and
// 1 - |C(:, j)|^2 + |E circ_j|^2 with E circ_j = (f_j - C(:, j)' M) R*^-1
// (KrigingImpl.cpp:227-242), clipped at 0.  grid ceil(m / 4), 256 threads
__global__ __launch_bounds__(TPB) void k_predict_reduce(PredictArgs a) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  __shared__ double sfh[4][PMAX];
  if (j >= a.m) return;
  const int n = a.n, ld = a.npad, p = a.p;
  const double* c = a.C + (size_t)j * ld;
  const double* z = a.MZ + (size_t)p * ld;
  double mz = 0.0, ss = 0.0;
  for (int i = lane; i < n; i += 64) {
    const double v = c[i];
    mz = fma(v, z[i], mz);
    ss = fma(v, v, ss);
  }
  mz = wave_sum(mz);
  ss = wave_sum(ss);
  for (int q = 0; q < p; ++q) {
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s = fma(c[i], a.MZ[(size_t)q * ld + i], s);
    s = wave_sum(s);
    if (lane == 0) sfh[w][q] = s;
  }
  if (lane != 0) return;
  const double* beta = a.coef;
  const double* Rs = a.coef + p;
  double mean = mz, ee = 0.0;
  double e[PMAX];
  for (int q = 0; q < p; ++q) {
    const double f = (q == 0) ? 1.0 : a.Xn[(size_t)(q - 1) * a.m + j];
    mean = fma(f, beta[q], mean);
    e[q] = f - sfh[w][q];
  }
  for (int q = 0; q < p; ++q) {  // R*^T c = e (forward substitution)
    double t = e[q];
    for (int r = 0; r < q; ++r) t -= Rs[r + q * p] * e[r];
    e[q] = t / Rs[q + q * p];
    ee = fma(e[q], e[q], ee);
  }
  if (a.ecirc != nullptr)
    for (int q = 0; q < p; ++q) a.ecirc[(size_t)q * a.m + j] = e[q];
  double v = 1.0 - ss + ee;
  if (isnan(v) || v < 0.0) v = 0.0;
  a.out[j] = mean;
  a.out[a.m + j] = v;
}

// ---------------------------------------------------------------------------
// (10) predictive covariance: Sigma = R_nn - C' C + Ecirc Ecirc' (before
// var_scale), C = L^-1 R_on.  R_nn as LinearAlgebra::covMat_sym_X
// (LinearAlgebra.cpp:531-596): factor * k(xn_i - xn_j) off the diagonal, 1
// on it.  Sigma is symmetric: grid (T (T + 1) / 2), T = mpad / 64, over the
// lower tiles, each off-diagonal tile stored at both of its places
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(TPB) void k_predict_cov(PredictArgs a) {
  // lower tiles only; an off-diagonal tile also writes its transpose
  int ti, tj;
  tri_decode(blockIdx.x, ti, tj);
  const int i0 = ti * NB, j0 = tj * NB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int m = a.m, d = a.d;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  tile_mainloop<true, true, true>(a.C, a.npad, i0, a.C, a.npad, j0, 0, a.npad, sbuf, acc, tid, lane, wr, wc);
  double* sXi = sbuf;
  double* sXj = sbuf + DT * NB;
  for (int q = tid; q < DT * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    sXi[q] = (k < d && i0 + r < m) ? a.Xn[(size_t)k * m + i0 + r] : 0.0;
    sXj[q] = (k < d && j0 + r < m) ? a.Xn[(size_t)k * m + j0 + r] : 0.0;
  }
  __syncthreads();
  double it[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) it[k] = (k < d) ? 1.0 / a.theta[k] : 1.0;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = acc_row(lane, wr, bi), c = acc_col(lane, wc, bj, e);
        const int gi = i0 + r, gj = j0 + c;
        if (gi >= m || gj >= m) continue;
        double rnn = 1.0;
        if (gi != gj) {
          double dx[DT];
#pragma unroll
          for (int k = 0; k < DT; ++k) dx[k] = sXi[k * NB + r] - sXj[k * NB + c];
          rnn = kern_rho<DT>(a.kernel, dx, it) * a.nn_factor;
        }
        double ec = 0.0;
        for (int q = 0; q < a.p; ++q) ec = fma(a.ecirc[(size_t)q * m + gi], a.ecirc[(size_t)q * m + gj], ec);
        const double v = acc[bi][bj][e] + rnn + ec;
        a.cov[(size_t)gj * m + gi] = v;
        acc[bi][bj][e] = v;
      }
  if (ti == tj) return;
  // the transpose of an off-diagonal tile, staged through LDS so that its
  // columns are stored with 64 consecutive lanes (coalesced)
  static_assert(4 * KC * LDSS >= NB * (NB + 1), "a 64 x 65 tile fits the staging buffer");
  double* sT = sbuf;
  __syncthreads();  // every wave is done with sXi / sXj
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        sT[acc_row(lane, wr, bi) * (NB + 1) + acc_col(lane, wc, bj, e)] = acc[bi][bj][e];
  __syncthreads();
  const int gj = j0 + lane;
  for (int r = w * (NB / 4); r < (w + 1) * (NB / 4); ++r) {
    const int gi = i0 + r;
    if (gi < m && gj < m) a.cov[(size_t)gi * m + gj] = sT[r * (NB + 1) + lane];
  }
}

// d ln k / d dx_k (Covariance.cpp:52-63, 90-101, 132-146, 181-198), dx = x_new - x_obs
__device__ inline double dlncov_dx(int kern, double dx, double th) {
  const double sg = dx > 0.0 ? 1.0 : (dx < 0.0 ? -1.0 : 0.0);
  if (kern == K_GAUSS) return -dx / (th * th);
  if (kern == K_EXP) return -sg / th;
  if (kern == K_MATERN32) {
    const double u = SQRT3 * fabs(dx / th);
    return -SQRT3 * sg * u / (1.0 + u) / th;
  }
  const double u = SQRT5 * fabs(dx / th);
  const double aa = 1.0 + u, bb = u / 3.0;
  return -SQRT5 * sg * (aa * bb) / (aa + u * bb) / th;
}

// DR(i, k mpad + j) = R_on(i, j) d ln k / d dx_k (xn_j - x_i); zero padding.
// R_on as k_cross_cov (1 at coincident points, where the derivative is 0).
// grid (mpad / 64, npad / 64)
template <int DT>
__global__ __launch_bounds__(TPB) void k_predict_dron(PredictArgs a) {
  const int j0 = blockIdx.x * NB, i0 = blockIdx.y * NB, tid = threadIdx.x;
  const int n = a.n, m = a.m, d = a.d;
  __shared__ double sXo[DT * NB];
  __shared__ double sXn[DT * NB];
  for (int q = tid; q < d * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    sXo[q] = (i0 + r < n) ? a.X[(size_t)k * n + i0 + r] : 0.0;
    sXn[q] = (j0 + r < m) ? a.Xn[(size_t)k * m + j0 + r] : 0.0;
  }
  double it[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) it[k] = (k < d) ? 1.0 / a.theta[k] : 1.0;
  __syncthreads();
  const int ii = tid & 63, gi = i0 + ii;
  double dx[DT];
  for (int r = 0; r < NB / 4; ++r) {
    const int jj = (tid >> 6) + 4 * r, gj = j0 + jj;
    double v = 0.0;
    if (gi < n && gj < m) {
      bool same = true;
#pragma unroll
      for (int k = 0; k < DT; ++k) {
        dx[k] = (k < d) ? sXo[k * NB + ii] - sXn[k * NB + jj] : 0.0;
        same = same && fabs(dx[k]) <= 2.220446049250313e-16;
      }
      v = same ? 1.0 : kern_rho<DT>(a.kernel, dx, it) * a.factor;
    }
    for (int k = 0; k < d; ++k) {
      const double g = (gi < n && gj < m) ? dlncov_dx(a.kernel, sXn[k * NB + jj] - sXo[k * NB + ii], a.theta[k]) : 0.0;
      a.DR[((size_t)k * a.mpad + gj) * a.npad + gi] = v * g;
    }
  }
}

// one wave per point j: per coordinate k, with W = L^-1 DR,
//   dmean_k = (dF_j beta)_k + W(:, k j)' z
//   dvar_k  = -2 C(:, j)' W(:, k j) + 2 Ecirc_j DEcirc(:, k),
//   DEcirc(:, k) = solve_lower(R*', dF_j(k, :)' - M' W(:, k j))
// (KrigingImpl.cpp:276-288).  grid ceil(m / 4)
__global__ __launch_bounds__(TPB) void k_predict_deriv(PredictArgs a) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = blockIdx.x * 4 + w;
  if (j >= a.m) return;
  const int n = a.n, ld = a.npad, p = a.p, m = a.m;
  const double* z = a.MZ + (size_t)p * ld;
  const double* cj = a.C + (size_t)j * ld;
  const double* beta = a.coef;
  const double* Rs = a.coef + p;
  for (int k = 0; k < a.d; ++k) {
    const double* wk = a.DW + ((size_t)k * a.mpad + j) * ld;
    double sz = 0.0, sc = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double wv = wk[i];
      sz = fma(wv, z[i], sz);
      sc = fma(wv, cj[i], sc);
    }
    sz = wave_sum(sz);
    sc = wave_sum(sc);
    double dm = sz;
    if (a.linear_trend) dm += beta[k + 1];  // dF(k, q) = 1 for q = k + 1
    double dv = 0.0;
    if (a.want_var) {
      double e[PMAX];
      for (int q = 0; q < p; ++q) {
        double s = 0.0;
        for (int i = lane; i < n; i += 64) s = fma(wk[i], a.MZ[(size_t)q * ld + i], s);
        s = wave_sum(s);
        const double df = (a.linear_trend && q == k + 1) ? 1.0 : 0.0;
        e[q] = df - s;
      }
      double ed = 0.0;
      for (int q = 0; q < p; ++q) {  // R*' x = e (forward substitution)
        double t = e[q];
        for (int r = 0; r < q; ++r) t -= Rs[r + q * p] * e[r];
        e[q] = t / Rs[q + q * p];
        ed = fma(a.ecirc[(size_t)q * m + j], e[q], ed);
      }
      dv = -2.0 * sc + 2.0 * ed;
    }
    if (lane == 0) {
      a.dmean[(size_t)k * m + j] = dm;
      if (a.want_var) a.dvar[(size_t)k * m + j] = dv;
    }
  }
}

hipError_t launch_predict_cov(const PredictArgs& a, hipStream_t stream) {
  const int T = a.mpad / NB;
  const dim3 grid(T * (T + 1) / 2);
  if (a.d <= 4)
    hipLaunchKernelGGL(k_predict_cov<4>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 8)
    hipLaunchKernelGGL(k_predict_cov<8>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 16)
    hipLaunchKernelGGL(k_predict_cov<16>, grid, dim3(TPB), 0, stream, a);
  else
    hipLaunchKernelGGL(k_predict_cov<32>, grid, dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_predict_deriv(const PredictArgs& a, hipStream_t stream) {
  const dim3 grid(a.mpad / NB, a.npad / NB);
  if (a.d <= 4)
    hipLaunchKernelGGL(k_predict_dron<4>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 8)
    hipLaunchKernelGGL(k_predict_dron<8>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 16)
    hipLaunchKernelGGL(k_predict_dron<16>, grid, dim3(TPB), 0, stream, a);
  else
    hipLaunchKernelGGL(k_predict_dron<32>, grid, dim3(TPB), 0, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // W = L^-1 DR: the lower TRMM over d mpad columns
  PredictArgs t = a;
  t.Ron = a.DR;
  t.C = a.DW;
  t.mpad = a.mpad * a.d;
  hipLaunchKernelGGL(k_trmm_lower, dim3(t.mpad / NB, a.npad / NB), dim3(TPB), 0, stream, t);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_predict_deriv, dim3((a.m + 3) / 4), dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_predict(const PredictArgs& a, hipStream_t stream) {
  dim3 grid(a.mpad / NB, a.npad / NB);
  if (a.d <= 4)
    hipLaunchKernelGGL(k_cross_cov<4>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 8)
    hipLaunchKernelGGL(k_cross_cov<8>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 16)
    hipLaunchKernelGGL(k_cross_cov<16>, grid, dim3(TPB), 0, stream, a);
  else
    hipLaunchKernelGGL(k_cross_cov<32>, grid, dim3(TPB), 0, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_trmm_lower, grid, dim3(TPB), 0, stream, a);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_predict_reduce, dim3((a.m + 3) / 4), dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_build(const Batch& B) {
  dim3 grid(tri(B.nblk_max), B.nslots);
  if (B.dmax <= 4)
    hipLaunchKernelGGL(k_build<4>, grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else if (B.dmax <= 8)
    hipLaunchKernelGGL(k_build<8>, grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else if (B.dmax <= 16)
    hipLaunchKernelGGL(k_build<16>, grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else
    hipLaunchKernelGGL(k_build<32>, grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_potrf_diag(const Batch& B, int k) {
  hipLaunchKernelGGL(k_potrf_diag, dim3(B.nslots), dim3(DIAG_TPB), 0, B.stream, B.slots, B.params, B.out, k);
  return hipGetLastError();
}
hipError_t launch_potrf_panel(const Batch& B, int k) {
  const int T = B.nblk_max - k - 1;
  if (T <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_potrf_panel, dim3(T, B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out, k);
  return hipGetLastError();
}
hipError_t launch_potrf_update(const Batch& B, int kb, int ke, int c0, int c1) {
  const int nb = B.nblk_max;
  if (c0 >= nb) return hipSuccess;
  int tiles = 0;
  if (c1 < 0) {
    tiles = tri(nb - c0);
  } else {
    for (int c = c0; c <= c1 && c < nb; ++c) tiles += nb - c;
  }
  if (tiles <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_potrf_update, dim3(tiles, B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out, kb, ke,
                     c0, c1);
  return hipGetLastError();
}
hipError_t launch_potrf_update_trailing(const Batch& B, int kb, int ke, int c0) {
  const int nb = B.nblk_max;
  if (c0 >= nb) return hipSuccess;
  int tiles = 0;
  for (int c = c0; c < nb; ++c) tiles += (nb - c + 1) / 2;
  hipLaunchKernelGGL(k_potrf_update_w, dim3(tiles, B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out, kb, ke,
                     c0);
  return hipGetLastError();
}
hipError_t launch_potrf_update_persistent(const Batch& B, int kb, int ke, int c0, int c1, int workgroups) {
  if (B.nslots != 1) return hipErrorInvalidValue;
  const int nb = B.nblk_max;
  if (c0 >= nb) return hipSuccess;
  int tiles = 0;
  if (c1 < 0) {
    tiles = tri(nb - c0);
  } else {
    for (int c = c0; c <= c1 && c < nb; ++c) tiles += nb - c;
  }
  if (tiles <= 0) return hipSuccess;
  const int g = workgroups < 1 ? 1 : (workgroups < tiles ? workgroups : tiles);
  hipLaunchKernelGGL(k_potrf_update_persistent, dim3(g), dim3(TPB), 0, B.stream, B.slots, B.params, B.out, kb, ke,
                     c0, c1, tiles);
  return hipGetLastError();
}
int trtri_levels(int nblk) {
  int lv = 0;
  while ((1 << lv) < nblk) ++lv;
  return lv;
}
hipError_t launch_trtri_level(const Batch& B, int s, int step) {
  const int npairs = (B.nblk_max + 2 * s - 1) / (2 * s);
  if (npairs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_trtri, dim3(npairs * s * s, B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out, s,
                     step);
  return hipGetLastError();
}
hipError_t launch_solve_fy(const Batch& B) {
  hipLaunchKernelGGL(k_solve_fy, dim3(B.nblk_max, B.nslots), dim3(MV_TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_gls(const Batch& B) {
  hipLaunchKernelGGL(k_gls, dim3(B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_solve_e(const Batch& B) {
  hipLaunchKernelGGL(k_solve_e, dim3(B.nblk_max, B.nslots), dim3(MV_TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_solve_x(const Batch& B) {
  hipLaunchKernelGGL(k_solve_x, dim3(B.nblk_max, B.nslots), dim3(MV_TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_finalize_ll(const Batch& B) {
  hipLaunchKernelGGL(k_finalize_ll, dim3((B.nslots + 63) / 64), dim3(64), 0, B.stream, B.slots, B.params, B.out,
                     B.nslots);
  return hipGetLastError();
}
template <bool LMP>
static hipError_t launch_lauum_adjoint_t(const Batch& B) {
  dim3 grid(tri(B.nblk_max), B.nslots);
  const int dmax_batch = B.dmax;
  if (dmax_batch <= 4)
    hipLaunchKernelGGL((k_lauum_adjoint<4, LMP>), grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else if (dmax_batch <= 6)
    hipLaunchKernelGGL((k_lauum_adjoint<6, LMP>), grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else if (dmax_batch <= 8)
    hipLaunchKernelGGL((k_lauum_adjoint<8, LMP>), grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else if (dmax_batch <= 12)
    hipLaunchKernelGGL((k_lauum_adjoint<12, LMP>), grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else if (dmax_batch <= 16)
    hipLaunchKernelGGL((k_lauum_adjoint<16, LMP>), grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else
    hipLaunchKernelGGL((k_lauum_adjoint<32, LMP>), grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_lauum_adjoint(const Batch& B) { return launch_lauum_adjoint_t<false>(B); }
hipError_t launch_lmp_adjoint(const Batch& B) { return launch_lauum_adjoint_t<true>(B); }
hipError_t launch_finalize_grad(const Batch& B) {
  if (tri(B.nblk_max) >= FG_SPLIT_MIN) {
    hipLaunchKernelGGL(k_grad_groups, dim3(FG_GROUPS, B.nslots), dim3(FG1_TPB), 0, B.stream, B.slots, B.params,
                       B.out);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_finalize_grad, dim3(B.nslots), dim3(FG_TPB), 0, B.stream, B.slots, B.params, B.out, B.nslots);
  return hipGetLastError();
}
hipError_t launch_rcond_estimate(const Batch& B) {
  hipLaunchKernelGGL(k_rcond_estimate, dim3(B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}

}  // namespace lk

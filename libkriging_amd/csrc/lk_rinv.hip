// Gradient path without L^-1, FP64, gfx950.
//
// The log-likelihood gradient needs R^-1 (the trace terms of
// Kriging.cpp:291-326) and x = R^-1 (y - F beta).  Instead of forming L^-1
// (trtri), then R^-1 = L^-T L^-1 (lauum) and the solves with L^-1, this path
// computes Z = R^-1 directly from the Cholesky factor by a backward sweep over
// block columns, from L^T Z = L^-1 (and its transpose Z L = L^-T):
//   T(i, j)  = sum_{k > j} Z(i, k) L(k, j)             i > j  (K = n - 64 (j+1))
//   Z(i, j)  = -T(i, j) L_jj^-1
//   P_j      = sum_{i > j} L(i, j)^T T(i, j)
//   Z(j, j)  = L_jj^-T (I + P_j) L_jj^-1
// (the diagonal inverses L_jj^-1 come from the diagonal-block factorisation).
// Column j needs only columns > j: one launch for the off-diagonal tiles of
// column j (long-K MFMA products, plus the P_j partial of each tile) and one
// for the diagonal tile.  The trace part of the adjoint (sum of
// Z_ij dR_ij/dtheta_k over the lower triangle) is contracted in the epilogue
// of each tile as it is produced; Z is kept (lower tiles, full diagonal
// tiles) for the solves: [Z F | Z y] in one symmetric product, then GLS beta,
// x = Z y - Z F beta and SSE = (y - F beta)' x.  The x part of the adjoint
// (sum x_i x_j dR_ij/dtheta_k) is a separate pass that recomputes the kernel
// values.  Condition check: |L^-1|_1 <= sqrt(n) max_i sqrt(Z_ii); when that
// bound leaves rcond below the threshold the slot is marked and the exact
// L^-1 path decides (lk_runtime.cpp).
#include <hip/hip_runtime.h>

#include <math.h>

#include "lk_device.hpp"
#include "lk_kernels.h"

namespace lk {

namespace {

// fields of a tile's partial record (red[grad_off + t * GRAD_STRIDE])
constexpr int F_S1 = 0, F_S2 = DMAX, F_A1 = 2 * DMAX, F_A2 = 2 * DMAX + 1, F_DG1 = 2 * DMAX + 2,
              F_DG2 = 2 * DMAX + 3;

__device__ inline int tri_index(int i, int j) { return i * (i + 1) / 2 + j; }

// X rows of tile blocks i0 and j0 (DT coordinates, zero beyond d and n)
template <int DT>
__device__ inline void stage_x(const Slot& S, int i0, int j0, double* sXi, double* sXj, int tid) {
  const int n = S.n, d = S.d;
  for (int q = tid; q < DT * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    sXi[q] = (k < d && i0 + r < n) ? S.X[(size_t)k * n + i0 + r] : 0.0;
    sXj[q] = (k < d && j0 + r < n) ? S.X[(size_t)k * n + j0 + r] : 0.0;
  }
}

// trace part of the adjoint for the accumulator tile (values Z(i0 + row,
// j0 + col)): strictly lower elements with row < n add Z R g_k to s2 and
// Z rho to a2; diagonal elements (diag_tile) add Z_ii to dg2 and update the
// block's max Z_ii.  Writes the record's s2, a2, dg2 fields.
template <int DT>
__device__ inline void trace_epilogue(const Slot& S, const SlotParams& P, const d4 (&acc)[2][2], int i0, int j0,
                                      bool diag_tile, double* rec, double* zmax, double* sbuf, int tid, int lane,
                                      int w, int wr, int wc) {
  double* sXi = sbuf;
  double* sXj = sbuf + DT * NB;
  __syncthreads();  // sbuf is free again
  stage_x<DT>(S, i0, j0, sXi, sXj, tid);
  __syncthreads();
  const int n = S.n, d = S.d;
  double it[DT], it2[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) {
    it[k] = (k < d) ? 1.0 / P.theta[k] : 1.0;
    it2[k] = it[k] * it[k];
  }
  const double factor = (S.noise == NOISE_NUGGET) ? P.alpha : 1.0;
  double s2[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) s2[k] = 0.0;
  double a2 = 0.0, dg2 = 0.0, zm = 0.0;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = acc_row(lane, wr, bi), c = acc_col(lane, wc, bj, e);
        const int gi = i0 + r, gj = j0 + c;
        const double z = acc[bi][bj][e];
        if (gi >= n || gj > gi) continue;
        if (gi == gj) {
          dg2 += z;
          zm = fmax(zm, z);
          continue;
        }
        double dx[DT], g[DT];
#pragma unroll
        for (int k = 0; k < DT; ++k) dx[k] = sXj[k * NB + c] - sXi[k * NB + r];
        const double rho = kern_rho_g<DT>(S.kernel, dx, it, it2, g);
        const double Rij = rho * factor;
#pragma unroll
        for (int k = 0; k < DT; ++k) s2[k] = fma(z, Rij * g[k], s2[k]);
        a2 = fma(z, rho, a2);
      }
  __shared__ double sred[4][DT + 3];
#pragma unroll
  for (int k = 0; k < DT; ++k) {
    const double v = wave_sum(s2[k]);
    if (lane == 0) sred[w][k] = v;
  }
  {
    const double v0 = wave_sum(a2), v1 = wave_sum(dg2), v2 = wave_max(zm);
    if (lane == 0) {
      sred[w][DT] = v0;
      sred[w][DT + 1] = v1;
      sred[w][DT + 2] = v2;
    }
  }
  __syncthreads();
  if (tid < DT) rec[F_S2 + tid] = (sred[0][tid] + sred[1][tid]) + (sred[2][tid] + sred[3][tid]);
  if (tid == DT) rec[F_A2] = (sred[0][DT] + sred[1][DT]) + (sred[2][DT] + sred[3][DT]);
  if (tid == DT + 1) rec[F_DG2] = diag_tile ? (sred[0][DT + 1] + sred[1][DT + 1]) + (sred[2][DT + 1] + sred[3][DT + 1]) : 0.0;
  if (tid == DT + 2 && diag_tile && zmax != nullptr)
    *zmax = fmax(fmax(sred[0][DT + 2], sred[1][DT + 2]), fmax(sred[2][DT + 2], sred[3][DT + 2]));
}

}  // namespace

// ---------------------------------------------------------------------------
// (Z1) off-diagonal tiles of column j: grid (nblk_max - 1 - j, nslots)
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(TPB) void k_zsweep_offdiag(const Slot* __restrict__ slots,
                                                        const SlotParams* __restrict__ params,
                                                        SlotOut* __restrict__ out, int j) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  if (!slot_live(P, out[b]) || !P.want_grad) return;
  const int nb = S.nblk;
  const int i = j + 1 + ts.tile;
  if (i >= nb) return;
  const int ld = S.npad, i0 = i * NB, j0 = j * NB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  // T(a, c) = sum_{k in [j0 + NB, i0)} Z(i0 + a, k) L(k, j0 + c)      (Z lower tiles of row i)
  //         + sum_{k in [i0, nb NB)}   Z(k, i0 + a) L(k, j0 + c)      (Z tiles of column i)
  tile_mainloop<false, true, false>(S.W, ld, i0, S.A, ld, j0, j0 + NB, i0, sbuf, acc, tid, lane, wr, wc);
  tile_mainloop<true, true, false>(S.W, ld, i0, S.A, ld, j0, i0, nb * NB, sbuf, acc, tid, lane, wr, wc);
  // T -> strict upper tile (row j, col i) of W: W[(i0 + c) ld + j0 + a]
  acc_store(acc, S.W + (size_t)i0 * ld + j0, ld, lane, wr, wc);
  __threadfence_block();
  __syncthreads();
  // Z(i, j) = -T L_jj^-1
  acc_zero(acc);
  tile_mainloop<false, true, true>(S.W + (size_t)i0 * ld, ld, j0, S.W + j0, ld, j0, 0, NB, sbuf, acc, tid, lane, wr,
                                   wc);
  acc_store(acc, S.W + (size_t)j0 * ld + i0, ld, lane, wr, wc);
  double* rec = S.red + red_grad_off(nb) + (size_t)tri_index(i, j) * GRAD_STRIDE;
  trace_epilogue<DT>(S, P, acc, i0, j0, false, rec, nullptr, sbuf, tid, lane, w, wr, wc);
  __syncthreads();
  // P_j partial of this tile: L(i, j)^T T -> strict upper tile (row j, col i) of A
  acc_zero(acc);
  tile_mainloop<true, true, false>(S.A + i0, ld, j0, S.W + j0, ld, i0, 0, NB, sbuf, acc, tid, lane, wr, wc);
  acc_store(acc, S.A + (size_t)i0 * ld + j0, ld, lane, wr, wc);
}

// ---------------------------------------------------------------------------
// (Z2) diagonal tile of column j: Z(j, j) = L_jj^-T (I + P_j) L_jj^-1, stored
// symmetric (from its lower triangle) over the L_jj^-1 block of W.
// grid (1, nslots)
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(TPB) void k_zsweep_diag(const Slot* __restrict__ slots,
                                                     const SlotParams* __restrict__ params,
                                                     SlotOut* __restrict__ out, int j) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  if (!slot_live(P, out[b]) || !P.want_grad) return;
  const int nb = S.nblk;
  if (j >= nb) return;
  const int ld = S.npad, j0 = j * NB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  __shared__ double sbuf[4 * KC * LDSS];
  double* sc1 = S.red + red_ztmp_off(nb);
  double* sc2 = sc1 + NB * NB;
  d4 acc[2][2];
  acc_zero(acc);
  // I + P_j, P_j = sum of the partials in A's strict upper tiles (row j, col i > j)
  for (int i = j + 1; i < nb; ++i) {
    const double* Pi = S.A + (size_t)(i * NB) * ld + j0;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int e = 0; e < 4; ++e)
          acc[bi][bj][e] += Pi[(size_t)acc_col(lane, wc, bj, e) * ld + acc_row(lane, wr, bi)];
  }
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        if (acc_row(lane, wr, bi) == acc_col(lane, wc, bj, e)) acc[bi][bj][e] += 1.0;
  acc_store(acc, sc1, NB, lane, wr, wc);
  __threadfence_block();
  __syncthreads();
  // N = (I + P_j) L_jj^-1
  acc_zero(acc);
  tile_mainloop<false, true, false>(sc1, NB, 0, S.W + j0, ld, j0, 0, NB, sbuf, acc, tid, lane, wr, wc);
  acc_store(acc, sc2, NB, lane, wr, wc);
  __threadfence_block();
  __syncthreads();
  // Z = L_jj^-T N
  acc_zero(acc);
  tile_mainloop<true, true, false>(S.W + j0, ld, j0, sc2, NB, 0, 0, NB, sbuf, acc, tid, lane, wr, wc);
  // symmetric store from the lower triangle (through sc1, free again)
  acc_store(acc, sc1, NB, lane, wr, wc);
  __threadfence_block();
  __syncthreads();
  for (int q = tid; q < NB * NB; q += TPB) {
    const int a = q % NB, c = q / NB;
    S.W[(size_t)(j0 + c) * ld + j0 + a] = (a >= c) ? sc1[c * NB + a] : sc1[a * NB + c];
  }
  double* rec = S.red + red_grad_off(nb) + (size_t)tri_index(j, j) * GRAD_STRIDE;
  trace_epilogue<DT>(S, P, acc, j0, j0, true, rec, S.red + red_colW_off(nb) + j, sbuf, tid, lane, w, wr, wc);
}

// ---------------------------------------------------------------------------
// (Z3) [Z F | Z y] for block row i: lower tiles of row i (k <= i, the
// diagonal tile in full) and the transposed tiles below it (k > i).
// grid (nblk_max, nslots); output to V columns 0 .. p
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_zsymv(const Slot* __restrict__ slots, const SlotParams* __restrict__ params,
                                               SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  if (!slot_live(P, out[b]) || !P.want_grad) return;
  const int i = ts.tile;
  if (i >= S.nblk) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int ld = S.npad, n = S.n, p = S.p, i0 = i * NB;
  const int ncols = p + 1;
  __shared__ double red[4][4][NB];
  __shared__ double colsum[4][NB];
  auto vin = [&](int c, int k) -> double {  // column c of [F | y], zero beyond n
    if (k >= n) return 0.0;
    return (c < p) ? S.F[(size_t)c * n + k] : S.y[k];
  };
  for (int c0 = 0; c0 < ncols; c0 += 4) {
    const int nc = min(4, ncols - c0);
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // row part: W[k ld + i0 + lane], k < i0 + NB
#pragma unroll 4
    for (int k = w; k < min(i0 + NB, n); k += 4) {
      const double zv = S.W[(size_t)k * ld + i0 + lane];
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < nc) acc[c] = fma(zv, vin(c0 + c, k), acc[c]);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) red[w][c][lane] = acc[c];
    // column part: for the 16 rows a = 16 w + q of this wave, sum over k >= i0 + NB
    for (int q = 0; q < 16; ++q) {
      const int a = 16 * w + q;
      double s[4] = {0.0, 0.0, 0.0, 0.0};
      for (int k = i0 + NB + lane; k < n; k += 64) {
        const double zv = S.W[(size_t)(i0 + a) * ld + k];
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (c < nc) s[c] = fma(zv, vin(c0 + c, k), s[c]);
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double v = wave_sum(s[c]);
        if (lane == 0 && c < nc) colsum[c][a] = v;
      }
    }
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < nc)
          S.V[(size_t)(c0 + c) * ld + i0 + lane] =
              ((red[0][c][lane] + red[1][c][lane]) + (red[2][c][lane] + red[3][c][lane])) + colsum[c][lane];
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// (Z4) GLS from [Z F | Z y] (KrigingImpl.cpp:105-121 with F*'F* = F'ZF and
// F*'y* = F'Zy): R* = chol_upper(F'ZF), beta, x = Zy - ZF beta (V column
// p + 3), SSE = (y - F beta)' x.  One workgroup per slot.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_gls_z(const Slot* __restrict__ slots, const SlotParams* __restrict__ params,
                                               SlotOut* __restrict__ out) {
  const int b = blockIdx.x;
  const Slot S = slots[b];
  const SlotParams& Pp = params[b];
  if (!slot_live(Pp, out[b]) || !Pp.want_grad) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, p = S.p, n = S.n, ld = S.npad;
  __shared__ double G[PMAX * (PMAX + 1)];
  __shared__ double sh4[4];
  __shared__ double beta[PMAX];
  __shared__ int sok;
  const double* ZF = S.V;
  const double* Zy = S.V + (size_t)p * ld;
  for (int a = 0; a < p; ++a)
    for (int c = a; c <= p; ++c) {  // column c == p is F'Zy
      const double* zc = (c < p) ? ZF + (size_t)c * ld : Zy;
      double s = 0.0;
      for (int i = tid; i < n; i += TPB) s = fma(S.F[(size_t)a * n + i], zc[i], s);
      const double t = block_sum4(s, sh4, lane, w);
      if (tid == 0) G[a + c * PMAX] = t;
    }
  __syncthreads();
  if (tid == 0) {
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
    for (int c = 0; c < p; ++c)
      for (int a = 0; a < p; ++a) rs[a + c * p] = (a <= c) ? G[a + c * PMAX] : 0.0;
    sok = ok ? 1 : 0;
  }
  __syncthreads();
  if (!sok) {
    if (tid == 0) out[b].status = ST_CHOL_FAILED;
    return;
  }
  double* x = S.V + (size_t)(p + 3) * ld;
  double sse = 0.0;
  for (int i = tid; i < ld; i += TPB) {
    double xv = 0.0;
    if (i < n) {
      double fb = 0.0, zfb = 0.0;
      for (int a = 0; a < p; ++a) {
        fb = fma(S.F[(size_t)a * n + i], beta[a], fb);
        zfb = fma(ZF[(size_t)a * ld + i], beta[a], zfb);
      }
      xv = Zy[i] - zfb;
      sse = fma(S.y[i] - fb, xv, sse);
    }
    x[i] = xv;
  }
  const double t = block_sum4(sse, sh4, lane, w);
  if (tid == 0) {
    S.red[red_sse_off(S.nblk)] = t;
    for (int q = 1; q < S.nblk; ++q) S.red[red_sse_off(S.nblk) + q] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// (Z5) |L|_1 column maxima per block column (for the condition check).
// grid (nblk_max, nslots)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_norm1_L(const Slot* __restrict__ slots,
                                                 const SlotParams* __restrict__ params, SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || !params[b].want_grad || ts.tile >= S.nblk) return;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, ld = S.npad, n = S.n;
  double mL = 0.0;
  for (int cc = w; cc < NB; cc += 4) {
    const int c = ts.tile * NB + cc;
    if (c >= n) break;
    double al = 0.0;
#pragma unroll 4
    for (int i = c + lane; i < n; i += 64) al += fabs(S.A[(size_t)c * ld + i]);
    al = wave_sum(al);
    mL = fmax(mL, al);
  }
  __shared__ double sm[4];
  if (lane == 0) sm[w] = mL;
  __syncthreads();
  if (tid == 0) S.red[red_colL_off(S.nblk) + ts.tile] = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
}

// ---------------------------------------------------------------------------
// (Z6) x part of the adjoint for lower tile (it, jt): s1_k += x_i x_j R_ij
// g_k, a1 += x_i x_j rho_ij over strictly lower elements, dg1 += x_i^2 on the
// diagonal.  grid (nblk_max (nblk_max + 1) / 2, nslots)
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(TPB) void k_adjoint_x(const Slot* __restrict__ slots,
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
  const int n = S.n, d = S.d, ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i0 = it * NB, j0 = jt * NB;
  __shared__ double sXi[DT * NB], sXj[DT * NB];
  __shared__ double sxi[NB], sxj[NB];
  stage_x<DT>(S, i0, j0, sXi, sXj, tid);
  const double* xv = S.V + (size_t)(S.p + 3) * ld;
  if (tid < NB) {
    sxi[tid] = xv[i0 + tid];
    sxj[tid] = xv[j0 + tid];
  }
  __syncthreads();
  double itv[DT], it2[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) {
    itv[k] = (k < d) ? 1.0 / P.theta[k] : 1.0;
    it2[k] = itv[k] * itv[k];
  }
  const double factor = (S.noise == NOISE_NUGGET) ? P.alpha : 1.0;
  double s1[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) s1[k] = 0.0;
  double a1 = 0.0, dg1 = 0.0;
  const int r = lane;  // row within the tile
  for (int c = w; c < NB; c += 4) {
    const int gi = i0 + r, gj = j0 + c;
    if (gi >= n || gj > gi) continue;
    const double xx = sxi[r] * sxj[c];
    if (gi == gj) {
      dg1 += xx;
      continue;
    }
    double dx[DT], g[DT];
#pragma unroll
    for (int k = 0; k < DT; ++k) dx[k] = sXj[k * NB + c] - sXi[k * NB + r];
    const double rho = kern_rho_g<DT>(S.kernel, dx, itv, it2, g);
    const double Rij = rho * factor;
#pragma unroll
    for (int k = 0; k < DT; ++k) s1[k] = fma(xx, Rij * g[k], s1[k]);
    a1 = fma(xx, rho, a1);
  }
  __shared__ double sred[4][DT + 2];
#pragma unroll
  for (int k = 0; k < DT; ++k) {
    const double v = wave_sum(s1[k]);
    if (lane == 0) sred[w][k] = v;
  }
  {
    const double v0 = wave_sum(a1), v1 = wave_sum(dg1);
    if (lane == 0) {
      sred[w][DT] = v0;
      sred[w][DT + 1] = v1;
    }
  }
  __syncthreads();
  double* rec = S.red + red_grad_off(S.nblk) + (size_t)ts.tile * GRAD_STRIDE;
  if (tid < DT) rec[F_S1 + tid] = (sred[0][tid] + sred[1][tid]) + (sred[2][tid] + sred[3][tid]);
  if (tid == DT) rec[F_A1] = (sred[0][DT] + sred[1][DT]) + (sred[2][DT] + sred[3][DT]);
  if (tid == DT + 1) rec[F_DG1] = (sred[0][DT + 1] + sred[1][DT + 1]) + (sred[2][DT + 1] + sred[3][DT + 1]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
#define LK_DT_DISPATCH(KER, grid, B, ...)                                                                   \
  do {                                                                                                      \
    if ((B).dmax <= 4)                                                                                      \
      hipLaunchKernelGGL(KER<4>, grid, dim3(TPB), 0, (B).stream, (B).slots, (B).params, (B).out, ##__VA_ARGS__); \
    else if ((B).dmax <= 8)                                                                                 \
      hipLaunchKernelGGL(KER<8>, grid, dim3(TPB), 0, (B).stream, (B).slots, (B).params, (B).out, ##__VA_ARGS__); \
    else if ((B).dmax <= 16)                                                                                \
      hipLaunchKernelGGL(KER<16>, grid, dim3(TPB), 0, (B).stream, (B).slots, (B).params, (B).out, ##__VA_ARGS__); \
    else                                                                                                    \
      hipLaunchKernelGGL(KER<32>, grid, dim3(TPB), 0, (B).stream, (B).slots, (B).params, (B).out, ##__VA_ARGS__); \
  } while (0)

hipError_t launch_zsweep_offdiag(const Batch& B, int j) {
  const int T = B.nblk_max - 1 - j;
  if (T <= 0) return hipSuccess;
  LK_DT_DISPATCH(k_zsweep_offdiag, dim3(T, B.nslots), B, j);
  return hipGetLastError();
}

hipError_t launch_zsweep_diag(const Batch& B, int j) {
  LK_DT_DISPATCH(k_zsweep_diag, dim3(1, B.nslots), B, j);
  return hipGetLastError();
}

hipError_t launch_zsymv(const Batch& B) {
  hipLaunchKernelGGL(k_zsymv, dim3(B.nblk_max, B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}

hipError_t launch_gls_z(const Batch& B) {
  hipLaunchKernelGGL(k_gls_z, dim3(B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}

hipError_t launch_norm1_L(const Batch& B) {
  hipLaunchKernelGGL(k_norm1_L, dim3(B.nblk_max, B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}

hipError_t launch_adjoint_x(const Batch& B) {
  const int nt = B.nblk_max * (B.nblk_max + 1) / 2;
  LK_DT_DISPATCH(k_adjoint_x, dim3(nt, B.nslots), B);
  return hipGetLastError();
}

}  // namespace lk

// Device helpers shared by the likelihood kernels and the NestedKriging
// kernels (gfx950): wave reductions, the FP64 MFMA 64 x 64 tile machinery
// and the covariance functions.
#pragma once
#include <hip/hip_runtime.h>

#include <math.h>

#include "lk_kernels.h"

namespace lk {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));  // loads as one 16-byte vector

#ifndef LK_KC
#define LK_KC 16
#endif
constexpr int KC = LK_KC;       // K-chunk staged in LDS per step
#ifndef LK_PREFETCH_DEPTH
#define LK_PREFETCH_DEPTH 3
#endif
// LDS row strides (doubles) of a KC x NB staging tile.  LDS serves 16 lanes
// of a 64-bit access per pass over 32 four-byte banks (the measured bank
// conflicts follow that model).  A transposed operand is staged by 16 lanes
// writing 8 k rows 2 apart and 2 columns: row steps of 4 S words must be
// distinct mod 32 for the 8 rows, i.e. S odd: S = NB + 1 (NB + 2 put rows k
// and k + 8 on one bank pair, 2-way; NB + 4, 4-way).  A plain operand is
// staged as 16-byte column pairs, conflict-free with an even stride that
// keeps them 16-byte aligned: S = NB + 2 (NB + 1 splits them, 2-way).  The
// MFMA operand reads are conflict-free with either.  LDSS is the pitch the
// buffers are sized with.
constexpr int LDSS = NB + 2;
template <bool TRANS>
struct LdsStride {
  static constexpr int v = TRANS ? NB + 1 : NB + 2;
};
constexpr int TPB = 256;        // threads of the tile kernels (4 waves)

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
// Loads through a global-address-space pointer.  Matrix pointers reach the
// kernels inside slot descriptors, so the compiler cannot prove that they
// point to global memory and emits FLAT loads, which count against the LDS
// counter too: every LDS wait of the MFMA loop would then also wait for the
// prefetch loads in flight.  GLOBAL loads count against vmcnt only.
template <class T>
__device__ inline T gload(const T* p) {
  return *(const __attribute__((address_space(1))) T*)p;
}

__device__ inline double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// lower-triangular tile index t -> (i, j), i >= j, row-major over i
__device__ inline void tri_decode(int t, int& i, int& j) {
  int r = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
  while ((r + 1) * (r + 2) / 2 <= t) ++r;
  while (r * (r + 1) / 2 > t) --r;
  i = r;
  j = t - r * (r + 1) / 2;
}

// (tile, slot) of this workgroup in a (tiles, slots) grid, XCD-aware.
// Workgroups go to the 8 XCDs round-robin by linear id, each XCD with its own
// L2: with slots a multiple of 8, slot s is placed on XCD s % 8 and its tiles
// in consecutive dispatch order there, so one slot's matrices are shared in
// one L2 instead of being spread over all eight.  Bijective on the grid.
struct TileSlot {
  int tile, slot;
};
__device__ inline TileSlot tile_slot() {
  const int T = gridDim.x, S = gridDim.y;
  if ((S & 7) != 0) return TileSlot{(int)blockIdx.x, (int)blockIdx.y};
  const int L = blockIdx.x + T * blockIdx.y;
  const int x = L & 7, idx = L >> 3;
  return TileSlot{idx % T, x + 8 * (idx / T)};
}

__device__ inline bool slot_live(const SlotParams& P, const SlotOut& O) {
  return P.active && O.status != ST_CHOL_FAILED && O.status != ST_INACTIVE;
}

// ---------------------------------------------------------------------------
// MFMA tile machinery: a 64 x 64 output tile per 256-thread workgroup, each
// wave owning a 32 x 32 quarter as 2 x 2 blocks of V_MFMA_F64_16X16X4_F64.
// Operands are staged in LDS as [k][row] (row contiguous).  The instruction
// computes D = P Q with P (16 x 4) given as lane l -> P(l & 15, l >> 4),
// Q (4 x 16) as lane l -> Q(l >> 4, l & 15), and returns D(l >> 4 + 4e, l & 15)
// in register e of lane l (checked on gfx950).  We feed P = B^T and Q = A^T,
// so D = (A B)^T and lane l, register e holds C(row = l & 15, col = l >> 4 + 4e):
// sixteen consecutive rows of one column per lane group, i.e. coalesced
// column-major loads and stores of the C tile.
// ---------------------------------------------------------------------------

// Operand chunks are KC x NB slices of op(M) staged as dst[kk * S + ii] =
// op(M)(i0 + ii, k0 + kk):
//   TRANS = false: op(M)(i, k) = M[k * ld + i]   (i contiguous in memory)
//   TRANS = true : op(M)(i, k) = M[i * ld + k]   (k contiguous in memory)
template <bool NEG, int KCH = KC, int SA = LDSS, int SB = LDSS>
__device__ inline void mma_chunk(const double* __restrict__ sA, const double* __restrict__ sB, d4 (&acc)[2][2],
                                 int lane, int wr, int wc) {
  const int r = lane & 15, kq = lane >> 4;
#pragma unroll
  for (int k0 = 0; k0 < KCH; k0 += 4) {
    const double* pa = sA + (k0 + kq) * SA + wr * 32 + r;
    const double* pb = sB + (k0 + kq) * SB + wc * 32 + r;
    double a0 = pa[0], a1 = pa[16];
    const double b0 = pb[0], b1 = pb[16];
    if (NEG) {
      a0 = -a0;
      a1 = -a1;
    }
    acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a0, acc[0][0], 0, 0, 0);
    acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a0, acc[0][1], 0, 0, 0);
    acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, a1, acc[1][0], 0, 0, 0);
    acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, a1, acc[1][1], 0, 0, 0);
  }
}

// Register-staged copy of one KC x NB chunk: the global loads of the next
// chunk are issued before the MFMAs of the current one and written to the
// other LDS buffer afterwards (software pipelining over K).
constexpr int CHUNK_V2 = KC * NB / (2 * TPB);  // double2 per thread per KC x NB chunk
static_assert(KC % 8 == 0 && CHUNK_V2 >= 1, "chunk layout");
struct ChunkRegs {
  d2 v[CHUNK_V2];
};

// TRANS = false: a thread copies pairs of rows (ii, ii + 1) of k rows
//   tid / 32 + 8 q; TRANS = true: pairs of k (kk, kk + 1) of rows
//   tid / (KC / 2) + (2 TPB / KC) q.  Both read 16-byte vectors along the
//   contiguous dimension of memory.
template <bool TRANS>
__device__ inline void chunk_load(ChunkRegs& r, const double* __restrict__ M, int ld, int i0, int k0, int tid) {
  if (!TRANS) {
    const int ii = 2 * (tid & 31);
#pragma unroll
    for (int q = 0; q < CHUNK_V2; ++q)
      r.v[q] = gload(reinterpret_cast<const d2*>(M + (size_t)(k0 + (tid >> 5) + 8 * q) * ld + i0 + ii));
  } else {
    constexpr int H = KC / 2;
    const int kk = 2 * (tid % H);
#pragma unroll
    for (int q = 0; q < CHUNK_V2; ++q)
      r.v[q] = gload(reinterpret_cast<const d2*>(M + (size_t)(i0 + tid / H + (TPB / H) * q) * ld + k0 + kk));
  }
}

template <bool TRANS>
__device__ inline void chunk_store(const ChunkRegs& r, double* __restrict__ dst, int tid) {
  constexpr int S = LdsStride<TRANS>::v;  // row stride of this operand's staging
  if (!TRANS) {
    const int ii = 2 * (tid & 31);
#pragma unroll
    for (int q = 0; q < CHUNK_V2; ++q) {
      const int kk = (tid >> 5) + 8 * q;
      dst[kk * S + ii] = r.v[q].x;
      dst[kk * S + ii + 1] = r.v[q].y;
    }
  } else {
    constexpr int H = KC / 2;
    const int kk = 2 * (tid % H);
#pragma unroll
    for (int q = 0; q < CHUNK_V2; ++q) {
      const int ii = tid / H + (TPB / H) * q;
      dst[kk * S + ii] = r.v[q].x;
      dst[(kk + 1) * S + ii] = r.v[q].y;
    }
  }
}

// acc += op(A)(ai0 + 0..63, k) op(B)(bj0 + 0..63, k)^T summed over
// k in [kbeg, kend) (a multiple of KC), with op as above.  sbuf holds two
// (A, B) chunk buffers: 4 * KC * LDSS doubles.  One barrier per chunk.
// With SKIP, a wave with mma_on false stages chunks but issues no MFMA (its
// quadrant of the tile is not needed; its acc is left as it was); without
// it mma_on is ignored and the loop carries no such branch.
template <bool TA, bool TB, bool NEG, bool SKIP = false>
__device__ inline void tile_mainloop(const double* __restrict__ A, int lda, int ai0, const double* __restrict__ B,
                                     int ldb, int bj0, int kbeg, int kend, double* __restrict__ sbuf,
                                     d4 (&acc)[2][2], int tid, int lane, int wr, int wc, bool mma_on = true);

// the wave of a diagonal tile whose quadrant lies strictly above the
// diagonal (rows 0..31, columns 32..63): symmetric and triangular products
// need only the lower part of such tiles
__device__ inline bool upper_quadrant(bool diagonal_tile, int wr, int wc) { return diagonal_tile && wr == 0 && wc == 1; }

__device__ inline void acc_zero(d4 (&acc)[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
}

// row / col inside the 64 x 64 tile of accumulator element (bi, bj, e)
__device__ inline int acc_row(int lane, int wr, int bi) { return wr * 32 + bi * 16 + (lane & 15); }
__device__ inline int acc_col(int lane, int wc, int bj, int e) { return wc * 32 + bj * 16 + (lane >> 4) + 4 * e; }

// C tile (column-major, leading dim ld) <-> accumulators
__device__ inline void acc_load(d4 (&acc)[2][2], const double* __restrict__ C, int ld, int lane, int wr, int wc) {
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        acc[bi][bj][e] = gload(C + (size_t)acc_col(lane, wc, bj, e) * ld + acc_row(lane, wr, bi));
}

__device__ inline void acc_store(const d4 (&acc)[2][2], double* __restrict__ C, int ld, int lane, int wr, int wc) {
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        C[(size_t)acc_col(lane, wc, bj, e) * ld + acc_row(lane, wr, bi)] = acc[bi][bj][e];
}

template <bool TA, bool TB, bool NEG, bool SKIP>
__device__ inline void tile_mainloop(const double* __restrict__ A, int lda, int ai0, const double* __restrict__ B,
                                     int ldb, int bj0, int kbeg, int kend, double* __restrict__ sbuf,
                                     d4 (&acc)[2][2], int tid, int lane, int wr, int wc, bool mma_on) {
  if (kbeg >= kend) return;
  double* sA[2] = {sbuf, sbuf + 2 * KC * LDSS};
  double* sB[2] = {sbuf + KC * LDSS, sbuf + 3 * KC * LDSS};
  // Chunk t is staged from register set t % D into LDS buffer t % 2 and
  // chunk t + D is then loaded into the same registers, so D chunks are in
  // flight and each load's latency is covered by the MFMAs of D chunks.
  // Unrolled over lcm(D, 2) chunks so that register sets stay static.
  constexpr int D = LK_PREFETCH_DEPTH;
  static_assert(D >= 1 && D <= 4, "prefetch depth");
  ChunkRegs ra[D], rb[D];
  const int nch = (kend - kbeg) / KC;
#pragma unroll
  for (int q = 0; q < D; ++q)
    if (q < nch) {
      chunk_load<TA>(ra[q], A, lda, ai0, kbeg + q * KC, tid);
      chunk_load<TB>(rb[q], B, ldb, bj0, kbeg + q * KC, tid);
    }
  constexpr int U = (D % 2 == 0) ? D : 2 * D;
  for (int t0 = 0; t0 < nch; t0 += U) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = t0 + u;
      if (t >= nch) break;
      const int lb = u & 1;
      chunk_store<TA>(ra[u % D], sA[lb], tid);
      chunk_store<TB>(rb[u % D], sB[lb], tid);
      __syncthreads();
      if (t + D < nch) {
        chunk_load<TA>(ra[u % D], A, lda, ai0, kbeg + (t + D) * KC, tid);
        chunk_load<TB>(rb[u % D], B, ldb, bj0, kbeg + (t + D) * KC, tid);
      }
      if (!SKIP || mma_on) mma_chunk<NEG, KC, LdsStride<TA>::v, LdsStride<TB>::v>(sA[lb], sB[lb], acc, lane, wr, wc);
    }
  }
  // the caller may reuse sbuf: wait until every wave has finished reading it
  __syncthreads();
}

// ---------------------------------------------------------------------------
// Wide tiles: 128 rows (two 64-row blocks of A) x 64 columns per 256-thread
// workgroup, both operands plain (i contiguous in memory); wave (wr, wc) owns
// rows 64 wr .. 64 wr + 63 and columns 32 wc .. 32 wc + 31 as 4 x 2 MFMA
// blocks.  Per K-chunk a wave issues 8 MFMAs for 6 LDS operand reads (4 for
// 4 in the 64 x 64 tile) and the workgroup stages 3 NB x KC operand slices
// for twice the flops: half the barriers and 3/4 of the staging per flop.
// ---------------------------------------------------------------------------
constexpr int WROWS = 2 * NB;
constexpr int LDSW = WROWS + 2;  // row stride of the 128-row staging (16-byte aligned pairs)

template <int S>
__device__ inline void chunk_store_plain(const ChunkRegs& r, double* __restrict__ dst, int tid) {
  const int ii = 2 * (tid & 31);
#pragma unroll
  for (int q = 0; q < CHUNK_V2; ++q) {
    const int kk = (tid >> 5) + 8 * q;
    dst[kk * S + ii] = r.v[q].x;
    dst[kk * S + ii + 1] = r.v[q].y;
  }
}

__device__ inline void acc_zero_w(d4 (&acc)[4][2]) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
}
__device__ inline int accw_row(int lane, int wr, int bi) { return wr * 64 + bi * 16 + (lane & 15); }
__device__ inline int accw_col(int lane, int wc, int bj, int e) { return wc * 32 + bj * 16 + (lane >> 4) + 4 * e; }

// rows >= rows_valid are neither loaded nor stored (a half tile)
__device__ inline void accw_load(d4 (&acc)[4][2], const double* __restrict__ C, int ld, int lane, int wr, int wc,
                                 int rows_valid) {
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = accw_row(lane, wr, bi);
        acc[bi][bj][e] = row < rows_valid ? gload(C + (size_t)accw_col(lane, wc, bj, e) * ld + row) : 0.0;
      }
}
__device__ inline void accw_store(const d4 (&acc)[4][2], double* __restrict__ C, int ld, int lane, int wr, int wc,
                                  int rows_valid) {
#pragma unroll
  for (int bi = 0; bi < 4; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int row = accw_row(lane, wr, bi);
        if (row < rows_valid) C[(size_t)accw_col(lane, wc, bj, e) * ld + row] = acc[bi][bj][e];
      }
}

// acc (+)= A(ai0 + 0..127, k) B(bj0 + 0..63, k)^T over k in [kbeg, kend)
// (a multiple of KC), both operands plain (A[k * lda + i]).  When
// rows_valid <= 64 the second row block is not read (its rows are staged as
// a copy of the first and their results are discarded by the caller).
// sbuf: 2 (KC LDSW + KC LDSS) doubles.
template <bool NEG>
__device__ inline void tile_mainloop_w(const double* __restrict__ A, int lda, int ai0, const double* __restrict__ B,
                                       int ldb, int bj0, int kbeg, int kend, double* __restrict__ sbuf,
                                       d4 (&acc)[4][2], int tid, int lane, int wr, int wc, int rows_valid) {
  if (kbeg >= kend) return;
  double* sA[2] = {sbuf, sbuf + KC * (LDSW + LDSS)};
  double* sB[2] = {sbuf + KC * LDSW, sbuf + KC * (2 * LDSW + LDSS)};
  const int ai1 = rows_valid > NB ? ai0 + NB : ai0;
  constexpr int D = 2;  // chunks in flight (3 x 8 registers each)
  ChunkRegs ra0[D], ra1[D], rb[D];
  const int nch = (kend - kbeg) / KC;
#pragma unroll
  for (int q = 0; q < D; ++q)
    if (q < nch) {
      chunk_load<false>(ra0[q], A, lda, ai0, kbeg + q * KC, tid);
      chunk_load<false>(ra1[q], A, lda, ai1, kbeg + q * KC, tid);
      chunk_load<false>(rb[q], B, ldb, bj0, kbeg + q * KC, tid);
    }
  const int r = lane & 15, kq = lane >> 4;
  for (int t0 = 0; t0 < nch; t0 += 2) {
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = t0 + u;
      if (t >= nch) break;
      chunk_store_plain<LDSW>(ra0[u], sA[u], tid);
      chunk_store_plain<LDSW>(ra1[u], sA[u] + NB, tid);
      chunk_store_plain<LDSS>(rb[u], sB[u], tid);
      __syncthreads();
      if (t + D < nch) {
        chunk_load<false>(ra0[u], A, lda, ai0, kbeg + (t + D) * KC, tid);
        chunk_load<false>(ra1[u], A, lda, ai1, kbeg + (t + D) * KC, tid);
        chunk_load<false>(rb[u], B, ldb, bj0, kbeg + (t + D) * KC, tid);
      }
#pragma unroll
      for (int k0 = 0; k0 < KC; k0 += 4) {
        const double* pa = sA[u] + (k0 + kq) * LDSW + wr * 64 + r;
        const double* pb = sB[u] + (k0 + kq) * LDSS + wc * 32 + r;
        double av[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) av[i] = NEG ? -pa[16 * i] : pa[16 * i];
        const double b0 = pb[0], b1 = pb[16];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          acc[i][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(b0, av[i], acc[i][0], 0, 0, 0);
          acc[i][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(b1, av[i], acc[i][1], 0, 0, 0);
        }
      }
    }
  }
  __syncthreads();
}

// store the transpose: element (row, col) of the tile goes to C[row * ld + col]
__device__ inline void acc_store_t(const d4 (&acc)[2][2], double* __restrict__ C, int ld, int lane, int wr,
                                   int wc) {
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e)
        C[(size_t)acc_row(lane, wr, bi) * ld + acc_col(lane, wc, bj, e)] = acc[bi][bj][e];
}

// ---------------------------------------------------------------------------
// kernel functions k(dx; theta) and d ln k / d theta_k, restated from
// src/lib/Covariance.cpp:24-179 (gauss, exp, matern3_2, matern5_2).  The
// Matern forms use the exact identity exp(-sum(u - log1p(q(u)))) =
// exp(-sum u) * prod q'(u) with q' = 1 + q, one exp per pair instead of one
// exp and d log1p.
// ---------------------------------------------------------------------------
constexpr double SQRT3 = 1.7320508075688772;
constexpr double SQRT5 = 2.2360679774997896;

// rho(dx) for the kernel; it[k] = 1 / theta_k.  Coordinates from dlim on
// are skipped (a uniform branch, no work); those below DT and beyond the
// design's d must then hold dx = 0 (they add exact zeros: the result is the
// same either way)
template <int DT>
__device__ inline double kern_rho(int kern, const double* dx, const double* it, int dlim = DT) {
  double s = 0.0;
  if (kern == K_GAUSS) {
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      if (k >= dlim) break;
      const double v = dx[k] * it[k];
      s = fma(v, v, s);
    }
    return exp(-0.5 * s);
  } else if (kern == K_EXP) {
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      if (k >= dlim) break;
      s += fabs(dx[k] * it[k]);
    }
    return exp(-s);
  } else if (kern == K_MATERN32) {
    double pr = 1.0;
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      if (k >= dlim) break;
      const double u = SQRT3 * fabs(dx[k] * it[k]);
      s += u;
      pr *= 1.0 + u;
    }
    return exp(-s) * pr;
  } else {
    double pr = 1.0;
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      if (k >= dlim) break;
      const double u = SQRT5 * fabs(dx[k] * it[k]);
      s += u;
      pr *= 1.0 + u + (u * u) * (1.0 / 3.0);
    }
    return exp(-s) * pr;
  }
}

// g_k = d ln rho / d theta_k for one coordinate (Covariance.cpp:38-179),
// it = 1 / theta_k: the same expressions as kern_rho_g, one k at a time
__device__ inline double kern_g(int kern, double dx, double it) {
  if (kern == K_GAUSS) return (dx * dx) * ((it * it) * it);
  if (kern == K_EXP) return fabs(dx) * (it * it);
  if (kern == K_MATERN32) {
    const double u = SQRT3 * fabs(dx * it);
    return (u * u) / (1.0 + u) * it;
  }
  const double u = SQRT5 * fabs(dx * it);
  const double aa = 1.0 + u;
  const double bb = (u * u) * (1.0 / 3.0);
  return (aa * bb) / (aa + bb) * it;
}

// rho(dx) and g_k = d ln rho / d theta_k (Covariance.cpp:38-179); it = 1 / theta,
// it2 = it^2, entries k >= d hold dx = 0 (their g is 0)
template <int DT>
__device__ inline double kern_rho_g(int kern, const double* dx, const double* it, const double* it2, double* g) {
  double s = 0.0, pr = 1.0;
  if (kern == K_GAUSS) {
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      const double v = dx[k] * it[k];
      s = fma(v, v, s);
      g[k] = (dx[k] * dx[k]) * (it2[k] * it[k]);
    }
    return exp(-0.5 * s);
  } else if (kern == K_EXP) {
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      s += fabs(dx[k] * it[k]);
      g[k] = fabs(dx[k]) * it2[k];
    }
    return exp(-s);
  } else if (kern == K_MATERN32) {
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      const double u = SQRT3 * fabs(dx[k] * it[k]);
      s += u;
      pr *= 1.0 + u;
      g[k] = (u * u) / (1.0 + u) * it[k];
    }
    return exp(-s) * pr;
  } else {
#pragma unroll
    for (int k = 0; k < DT; ++k) {
      const double u = SQRT5 * fabs(dx[k] * it[k]);
      const double aa = 1.0 + u;
      const double bb = (u * u) * (1.0 / 3.0);
      s += u;
      pr *= aa + bb;
      g[k] = (aa * bb) / (aa + bb) * it[k];
    }
    return exp(-s) * pr;
  }
}

// The matrix-vector kernels stream the L^-1 (and L) triangles once; they
// run 16 waves per workgroup and issue MV_U independent loads per lane
// before the first dependent FMA, so a workgroup keeps ~64 KB in flight
// instead of one 512-byte line per wave (one workgroup per 64-row block:
// for one large matrix the last row block's workgroup reads its whole
// row alone, at what its loads in flight allow).
constexpr int MV_TPB = 1024;
constexpr int MV_W = MV_TPB / 64;
constexpr int MV_U = 8;
static_assert(4 * MV_U <= 64, "one lane per (input column, batch column)");

// out[:, c] = Linv * in[:, c] for rows of block rb (grid (nblk_max, nslots))
// in columns come from `src` (n rows, ld = n, zero beyond n) ; out to V.
// Column c of the input is src[c * lds + k] for c < ncols_src and src2[k]
// for c == ncols_src (when src2 is given), so [F | y] go through in one pass.
// Wave w sums columns k in batches [w U + j W U, w U + j W U + U); the
// partial sums of the waves are added in a fixed order (deterministic).
__device__ inline void lower_matvec_block(const Slot& S, int rb, const double* __restrict__ src, int lds, int nsrc,
                                          int ncols, double* __restrict__ dst, int ldd, double* __restrict__ part_sq,
                                          const double* __restrict__ src2 = nullptr) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int i0 = rb * NB, ld = S.npad, n = S.n;
  const int kend = min(i0 + NB, n);  // columns k <= row, and k < n (src is zero beyond n)
  __shared__ double red[MV_W][4][NB];
  const double* __restrict__ Wr = S.W + i0 + lane;
  double ssq = 0.0;
  for (int c0 = 0; c0 < ncols; c0 += 4) {
    const int nc = min(4, ncols - c0);
    // columns past nc re-read column c0 (same cache lines) and are dropped
    const double* colp[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const int cs = c0 + (c < nc ? c : 0);
      colp[c] = (src2 != nullptr && cs == ncols - 1) ? src2 : src + (size_t)cs * lds;
    }
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    // lane c MV_U + u (< 4 MV_U) fetches input column c at k = kb + u: one
    // coalesced load per batch instead of 4 MV_U same-address loads, each of
    // which would move 64 lanes' worth of data; v_readlane broadcasts
    const int cl = lane / MV_U, ul = lane % MV_U;
    const double* my_col = colp[cl & 3];
    for (int kb = w * MV_U; kb < kend; kb += MV_W * MV_U) {
      double lv[MV_U];
#pragma unroll
      for (int u = 0; u < MV_U; ++u) {
        const int k = min(kb + u, kend - 1);
        lv[u] = gload(Wr + (size_t)k * ld);
      }
      const double sv = (lane < 4 * MV_U) ? gload(my_col + min(kb + ul, kend - 1)) : 0.0;
#pragma unroll
      for (int u = 0; u < MV_U; ++u)
        if (kb + u < kend) {
#pragma unroll
          for (int c = 0; c < 4; ++c) acc[c] = fma(lv[u], readlane_d(sv, c * MV_U + u), acc[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) red[w][c][lane] = acc[c];
    __syncthreads();
    if (w == 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (c >= nc) continue;
        // rows above the diagonal read W's upper part of the diagonal block,
        // which is exactly zero; rows >= n are identity rows times zero input.
        double v = red[0][c][lane];
#pragma unroll
        for (int q = 1; q < MV_W; ++q) v += red[q][c][lane];
        dst[(size_t)(c0 + c) * ldd + i0 + lane] = v;
        ssq = fma(v, v, ssq);
      }
    }
    __syncthreads();
  }
  (void)nsrc;
  if (part_sq != nullptr && w == 0) {
    const double t = wave_sum(ssq);
    if (lane == 0) part_sq[rb] = t;
  }
}

// sum of a per-thread value over the 256 threads of a tile workgroup, in a
// fixed order (wave butterflies, then the 4 wave sums in index order)
__device__ inline double block_sum4(double v, double* sh4, int lane, int w) {
  v = wave_sum(v);
  __syncthreads();
  if (lane == 0) sh4[w] = v;
  __syncthreads();
  return (sh4[0] + sh4[1]) + (sh4[2] + sh4[3]);
}

}  // namespace lk

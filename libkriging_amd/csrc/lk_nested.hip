// NestedKriging optimal ("NK") aggregation on gfx950, FP64.
//
// Reference: NestedKriging::predict_nk (src/lib/NestedKriging.cpp:353-465),
// after Rulliere, Durrande, Bachoc & Chevalier (Stat. Comput. 2018).  For a
// prediction point x and submodels g = 1..p (common prior sigma2 k(.,.;theta),
// constant trend beta0):
//   k_g(x)   = k(X_g, x)                       (n_g)
//   M_g(x)   = beta0 + k_g' R_g^-1 (y_g - beta0)
//   KM_gg    = k_g' R_g^-1 k_g + jitter
//   KM_ij    = (R_i^-1 k_i)' R_ij (R_j^-1 k_j)  (i != j, R_ij = k(X_i, X_j))
//   w        = KM^-1 diag(KM)
//   mean     = beta0 + w' (M - beta0),  var = sigma2 max(0, 1 - w' diag(KM))
// where R_g = k(X_g, X_g) + jitter I (NestedKriging.cpp:337-346).
//
// The dominant term is sum_{i<j} n_i n_j per point (N^2 q / 2 multiply-adds
// for q points): k_pair forms R_ij on the fly from X in LDS (never stored)
// and contracts it on the FP64 MFMA tiles against W_j = R_j^-1 k_j, then
// against W_i in the epilogue.  One workgroup owns one group pair and 64
// points and loops over all the pair's row tiles, so the per-pair sums are
// reduced in registers in a fixed order (deterministic, no atomics).
#include <hip/hip_runtime.h>

#include <type_traits>

#include <math.h>

#include "lk_device.hpp"
#include "lk_kernels.h"

namespace lk {

// ---------------------------------------------------------------------------
// (1) C(a, t) = k(x_a - x_t; theta) for the group-ordered design rows a and
// the chunk's points t; zero on pad rows / pad points.  grid (qpad/64, Npad/64)
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(TPB) void k_nk_cross(NkArgs a) {
  const int t0 = blockIdx.x * NB, i0 = blockIdx.y * NB, tid = threadIdx.x;
  const NkGroup G = a.groups[a.tile_group[blockIdx.y]];
  const int d = a.d, qc = a.qc, ld = a.Npad;
  const int nvalid = G.off + G.n;  // rows >= nvalid are the group's padding
  __shared__ double sXo[DT * NB];
  __shared__ double sXn[DT * NB];
  for (int q = tid; q < d * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    sXo[q] = a.Xg[(size_t)k * ld + i0 + r];
    sXn[q] = (t0 + r < qc) ? a.Xc[(size_t)k * qc + t0 + r] : 0.0;
  }
  double it[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) it[k] = (k < d) ? 1.0 / a.theta[k] : 1.0;
  __syncthreads();
  const int ii = tid & 63, gi = i0 + ii;
  double dx[DT];
  for (int r = 0; r < NB / 4; ++r) {
    const int jj = (tid >> 6) + 4 * r, gt = t0 + jj;
    double v = 0.0;
    if (gi < nvalid && gt < qc) {
#pragma unroll
      for (int k = 0; k < DT; ++k) dx[k] = (k < d) ? sXo[k * NB + ii] - sXn[k * NB + jj] : 0.0;
      v = kern_rho<DT>(a.kernel, dx, it);
    }
    a.C[(size_t)gt * ld + gi] = v;
  }
}

// (2) U_g = L_g^-1 C_g: tile (row tile of group g, 64 points), K over the
// group's blocks up to the diagonal.  grid (qpad/64, Npad/64)
__global__ __launch_bounds__(TPB) void k_nk_trmm(NkArgs a) {
  const int t0 = blockIdx.x * NB, i0 = blockIdx.y * NB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const NkGroup G = a.groups[a.tile_group[blockIdx.y]];
  const int il = i0 - G.off, ld = a.Npad;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  tile_mainloop<false, true, false>(a.Linv + G.loff, G.npad, il, a.C + G.off, ld, t0, 0, il + NB, sbuf, acc, tid,
                                    lane, wr, wc);
  acc_store(acc, a.U + (size_t)t0 * ld + i0, ld, lane, wr, wc);
}

// (3) W_g = L_g^-T U_g = R_g^-1 k_g, written over C (after k_nk_stats has
// used C): K over the group's blocks from the diagonal down.
__global__ __launch_bounds__(TPB) void k_nk_trmm_t(NkArgs a) {
  const int t0 = blockIdx.x * NB, i0 = blockIdx.y * NB;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const NkGroup G = a.groups[a.tile_group[blockIdx.y]];
  const int il = i0 - G.off, ld = a.Npad;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  tile_mainloop<true, true, false>(a.Linv + G.loff, G.npad, il, a.U + G.off, ld, t0, il, G.npad, sbuf, acc, tid,
                                   lane, wr, wc);
  acc_store(acc, a.C + (size_t)t0 * ld + i0, ld, lane, wr, wc);
}

// (4) per (point, group): M = beta0 + C_g' alpha_g, Kd = |U_g|^2.
// One wave per point; grid (ceil(qc / 4), p)
__global__ __launch_bounds__(TPB) void k_nk_stats(NkArgs a) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int t = blockIdx.x * 4 + w, g = blockIdx.y;
  if (t >= a.qc) return;
  const NkGroup G = a.groups[g];
  const double* c = a.C + (size_t)t * a.Npad + G.off;
  const double* u = a.U + (size_t)t * a.Npad + G.off;
  const double* al = a.alpha + G.off;
  double m = 0.0, s = 0.0;
  for (int i = lane; i < G.n; i += 64) {
    m = fma(c[i], al[i], m);
    s = fma(u[i], u[i], s);
  }
  m = wave_sum(m);
  s = wave_sum(s);
  if (lane == 0) {
    a.M[(size_t)t * a.p + g] = a.beta0 + m;
    a.Kd[(size_t)t * a.p + g] = s;
  }
}

// (5) alpha_g = L_g^-T r_g (r_g = L_g^-1 (y_g - beta0), group-ordered): one
// wave per row, reading the row's column of L_g^-1 below the diagonal.
// grid (Npad / 4)
__global__ __launch_bounds__(TPB) void k_nk_alpha(NkArgs a, const double* __restrict__ r) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = blockIdx.x * 4 + w;
  if (row >= a.Npad) return;
  const NkGroup G = a.groups[a.tile_group[row / NB]];
  const int il = row - G.off;
  double s = 0.0;
  if (il < G.n) {
    const double* col = a.Linv + G.loff + (size_t)il * G.npad;
    for (int k = il + lane; k < G.n; k += 64) s = fma(col[k], r[G.off + k], s);
  }
  s = wave_sum(s);
  if (lane == 0) a.alpha[row] = s;
}

// ---------------------------------------------------------------------------
// (6) KM_ij(t) = sum_a W_i(a, t) sum_b R(x_a, x_b) W_j(b, t) for i < j.
// grid (npairs, qpad / NK_QT), 256 threads; wave w owns the NK_QW points
// t0 + NK_QW w .. of the workgroup's NK_QT.  Per 64-row tile of group i and
// chunk of NK_PKC rows of group j:
//   R(a, b) generated once into LDS (4 kernel values per thread), then
//   reused by all NK_QT points on the MFMA
//   acc(a, t) += sum_b R(a, b) W_j(b, t)   (64 x NK_QW per wave)
// and per row tile  part(t) += sum_a W_i(a, t) acc(a, t)  in registers.
// Every sum over rows stays inside one wave (fixed order, no atomics).
// Three waves per SIMD (168 registers, a few spilled outside the MFMA loop):
// one workgroup's kernel-value generation runs beside the others' MFMAs
// (0.37 -> 0.41 of the FP64 peak against two waves per SIMD).
// ---------------------------------------------------------------------------
constexpr int NK_PKC = 16;      // rows of group j per chunk
constexpr int NK_QW = NK_QT / 4;  // points per wave
constexpr int NK_QB = NK_QW / 16;  // 16-point MFMA blocks per wave
constexpr int NK_SB = NK_QT + 1;  // W_j chunk row stride (odd: transposed stores conflict-free)

template <int DT, int KT>
__global__ __launch_bounds__(TPB, 3) void k_nk_pair(NkArgs a) {
  const int2 pr = a.pairs[blockIdx.x];
  const int t0 = blockIdx.y * NK_QT;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r16 = lane & 15, kq = lane >> 4;
  const NkGroup Gi = a.groups[pr.x], Gj = a.groups[pr.y];
  const int d = a.d, ld = a.Npad;
  constexpr int WV2 = NK_PKC * NK_QT / (2 * TPB);  // W_j 16-byte pairs per thread per chunk
  constexpr int XR = (DT * NK_PKC + TPB - 1) / TPB;  // X_j values per thread per chunk
  __shared__ double sA[NK_PKC * LDSS];   // R chunk: sA[kk * LDSS + ii] = R(a_ii, b_kk)
  __shared__ double sB[NK_PKC * NK_SB];  // W_j chunk: sB[kk * NK_SB + jj] = W_j(b_kk, t0 + jj)
  __shared__ double sXa[DT * NB];        // rows of the current a tile
  __shared__ double sXb[DT * NK_PKC];    // rows of the current b chunk
  const double* it = a.itheta;  // kernel arguments: scalar registers
  double part[NK_QB][4];
#pragma unroll
  for (int bj = 0; bj < NK_QB; ++bj)
#pragma unroll
    for (int e = 0; e < 4; ++e) part[bj][e] = 0.0;

  const double* Wi = a.C + (size_t)t0 * ld + Gi.off;
  const double* Wj = a.C + (size_t)t0 * ld + Gj.off;
  const int ii = tid & 63;
  // register stage of the next chunk (W_j pairs, X_j coordinates), loaded
  // while the current chunk is generated and contracted
  d2 rw[WV2];
  double rx[XR];
  auto load_chunk = [&](int b0) {
#pragma unroll
    for (int r = 0; r < WV2; ++r) {
      const int q = tid + r * TPB;
      const int bp = q % (NK_PKC / 2), t = q / (NK_PKC / 2);
      rw[r] = gload(reinterpret_cast<const d2*>(Wj + (size_t)t * ld + b0 + 2 * bp));
    }
#pragma unroll
    for (int r = 0; r < XR; ++r) {
      const int q = tid + r * TPB;
      rx[r] = (q < d * NK_PKC) ? a.Xg[(size_t)(q / NK_PKC) * ld + Gj.off + b0 + (q % NK_PKC)] : 0.0;
    }
  };
  for (int a0 = 0; a0 < Gi.npad; a0 += NB) {
    for (int q = tid; q < d * NB; q += TPB) {
      const int k = q / NB, r = q % NB;
      sXa[q] = a.Xg[(size_t)k * ld + Gi.off + a0 + r];
    }
    d4 acc[4][NK_QB];
#pragma unroll
    for (int bi = 0; bi < 4; ++bi)
#pragma unroll
      for (int bj = 0; bj < NK_QB; ++bj) acc[bi][bj] = d4{0.0, 0.0, 0.0, 0.0};
    load_chunk(0);
    for (int b0 = 0; b0 < Gj.npad; b0 += NK_PKC) {
#pragma unroll
      for (int r = 0; r < WV2; ++r) {
        const int q = tid + r * TPB;
        const int bp = q % (NK_PKC / 2), t = q / (NK_PKC / 2);
        sB[(2 * bp) * NK_SB + t] = rw[r].x;
        sB[(2 * bp + 1) * NK_SB + t] = rw[r].y;
      }
#pragma unroll
      for (int r = 0; r < XR; ++r) {
        const int q = tid + r * TPB;
        if (q < d * NK_PKC) sXb[q] = rx[r];
      }
      __syncthreads();
      if (b0 + NK_PKC < Gj.npad) load_chunk(b0 + NK_PKC);
      double dx[DT];
#pragma unroll 1
      for (int r = 0; r < (NK_PKC * NB) / TPB; ++r) {
        const int kk = (tid >> 6) + 4 * r;
#pragma unroll
        for (int k = 0; k < DT; ++k) dx[k] = (k < d) ? sXa[k * NB + ii] - sXb[k * NK_PKC + kk] : 0.0;
        // the padded coordinates (d .. DT - 1) are skipped
        sA[kk * LDSS + ii] = kern_rho<DT>(KT, dx, it, d);
      }
      __syncthreads();
#pragma unroll
      for (int k0 = 0; k0 < NK_PKC; k0 += 4) {
        const double* pa = sA + (k0 + kq) * LDSS + r16;
        const double* pb = sB + (k0 + kq) * NK_SB + NK_QW * w + r16;
        double av[4], bv[NK_QB];
#pragma unroll
        for (int q = 0; q < 4; ++q) av[q] = pa[16 * q];
#pragma unroll
        for (int q = 0; q < NK_QB; ++q) bv[q] = pb[16 * q];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
          for (int bj = 0; bj < NK_QB; ++bj)
            acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(bv[bj], av[bi], acc[bi][bj], 0, 0, 0);
      }
      __syncthreads();
    }
    // weight by W_i(a, t) and sum this tile's rows into part: lane (r16, kq),
    // register e of block (bi, bj) holds row 16 bi + r16, point
    // NK_QW w + 16 bj + kq + 4 e
#pragma unroll
    for (int bj = 0; bj < NK_QB; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const double* wcol = Wi + (size_t)(NK_QW * w + 16 * bj + kq + 4 * e) * ld + a0 + r16;
        double s = part[bj][e];
#pragma unroll
        for (int bi = 0; bi < 4; ++bi) s = fma(acc[bi][bj][e], gload(wcol + 16 * bi), s);
        part[bj][e] = s;
      }
  }
  // reduce over the 16 lanes holding different rows of the same points
  const size_t pp = (size_t)a.p * a.p;
#pragma unroll
  for (int bj = 0; bj < NK_QB; ++bj)
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double v = part[bj][e];
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) v += __shfl_xor(v, off, 64);
      const int t = t0 + NK_QW * w + 16 * bj + kq + 4 * e;
      if (r16 == 0 && t < a.qc) {
        double* K = a.KM + (size_t)t * pp;
        K[(size_t)pr.y * a.p + pr.x] = v;
        K[(size_t)pr.x * a.p + pr.y] = v;
      }
    }
}

// ---------------------------------------------------------------------------
// (7) per point: KM = cross + diag(Kd + jitter); w = KM^-1 diag(KM) by
// Cholesky (arma::solve with likely_sympd); on a non-positive pivot, LU with
// partial pivoting on the same system (the general solver Armadillo falls
// back to).  One workgroup per point; the factor lives in KF (global, L2).
// ---------------------------------------------------------------------------
constexpr int NK_PMAX = 2048;  // groups supported by the solve kernel (LDS vectors)

__global__ __launch_bounds__(TPB) void k_nk_solve(NkArgs a) {
  const int t = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int p = a.p;
  const size_t pp = (size_t)p * p;
  double* K = a.KM + (size_t)t * pp;
  double* F = a.KF + (size_t)t * pp;
  __shared__ double kM[NK_PMAX];
  __shared__ double z[NK_PMAX];
  __shared__ double sw[4];
  __shared__ int swi[4];
  __shared__ int spk;
  __shared__ int sflag;
  for (int i = tid; i < p; i += TPB) {
    const double v = a.Kd[(size_t)t * p + i] + a.jitter;
    K[(size_t)i * p + i] = v;
    kM[i] = v;
  }
  __syncthreads();
  for (size_t q = tid; q < pp; q += TPB) F[q] = K[q];
  if (tid == 0) sflag = 0;
  __syncthreads();
  // right-looking Cholesky, lower triangle of F (column-major)
  for (int k = 0; k < p; ++k) {
    const double dkk = F[(size_t)k * p + k];
    if (!(dkk > 0.0)) {  // uniform: every thread read the same value
      if (tid == 0) sflag = 1;
      break;
    }
    const double r = sqrt(dkk);
    __syncthreads();  // every thread has read F(k, k)
    for (int i = k + tid; i < p; i += TPB) F[(size_t)k * p + i] = (i == k) ? r : F[(size_t)k * p + i] / r;
    __syncthreads();
    for (int j = k + 1 + w; j < p; j += 4) {
      const double ljk = F[(size_t)k * p + j];
      for (int i = j + lane; i < p; i += 64) F[(size_t)j * p + i] = fma(-F[(size_t)k * p + i], ljk, F[(size_t)j * p + i]);
    }
    __syncthreads();
  }
  __syncthreads();
  const bool chol_ok = sflag == 0;
  if (chol_ok) {
    // L z = kM (column sweep), then L' w = z
    for (int i = tid; i < p; i += TPB) z[i] = kM[i];
    __syncthreads();
    for (int k = 0; k < p; ++k) {
      const double zk = z[k] / F[(size_t)k * p + k];
      __syncthreads();
      if (tid == 0) z[k] = zk;
      for (int i = k + 1 + tid; i < p; i += TPB) z[i] = fma(-F[(size_t)k * p + i], zk, z[i]);
      __syncthreads();
    }
    for (int k = p - 1; k >= 0; --k) {
      const double wk = z[k] / F[(size_t)k * p + k];
      __syncthreads();
      if (tid == 0) z[k] = wk;
      for (int i = tid; i < k; i += TPB) z[i] = fma(-F[(size_t)i * p + k], wk, z[i]);
      __syncthreads();
    }
  } else {
    // LU with partial pivoting on a fresh copy of KM (row-wise elimination)
    for (size_t q = tid; q < pp; q += TPB) F[q] = K[q];
    for (int i = tid; i < p; i += TPB) z[i] = kM[i];
    if (tid == 0) sflag = 1;
    __syncthreads();
    for (int k = 0; k < p; ++k) {
      // pivot: largest |F(i, k)|, i >= k, first index on ties
      double best = -1.0;
      int bi = k;
      for (int i = k + tid; i < p; i += TPB) {
        const double v = fabs(F[(size_t)k * p + i]);
        if (v > best) {
          best = v;
          bi = i;
        }
      }
      for (int off = 32; off > 0; off >>= 1) {
        const double ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (lane == 0) {
        sw[w] = best;
        swi[w] = bi;
      }
      __syncthreads();
      if (tid == 0) {
        double b = sw[0];
        int ix = swi[0];
        for (int q = 1; q < 4; ++q)
          if (sw[q] > b || (sw[q] == b && swi[q] < ix)) {
            b = sw[q];
            ix = swi[q];
          }
        spk = ix;
        if (!(b > 0.0)) sflag = 2;
      }
      __syncthreads();
      if (sflag == 2) break;
      const int pk = spk;
      if (pk != k) {
        for (int j = tid; j < p; j += TPB) {
          const double u = F[(size_t)j * p + k];
          F[(size_t)j * p + k] = F[(size_t)j * p + pk];
          F[(size_t)j * p + pk] = u;
        }
        if (tid == 0) {
          const double u = z[k];
          z[k] = z[pk];
          z[pk] = u;
        }
      }
      __syncthreads();
      const double dkk = F[(size_t)k * p + k];
      for (int i = k + 1 + tid; i < p; i += TPB) F[(size_t)k * p + i] /= dkk;
      __syncthreads();
      for (int j = k + 1 + w; j < p; j += 4) {
        const double ukj = F[(size_t)j * p + k];
        for (int i = k + 1 + lane; i < p; i += 64) F[(size_t)j * p + i] = fma(-F[(size_t)k * p + i], ukj, F[(size_t)j * p + i]);
      }
      if (tid == 0) {
        const double zk = z[k];
        for (int i = k + 1; i < p; ++i) z[i] = fma(-F[(size_t)k * p + i], zk, z[i]);
      }
      __syncthreads();
    }
    if (sflag != 2) {
      for (int k = p - 1; k >= 0; --k) {
        const double wk = z[k] / F[(size_t)k * p + k];
        __syncthreads();
        if (tid == 0) z[k] = wk;
        for (int i = tid; i < k; i += TPB) z[i] = fma(-F[(size_t)k * p + i], wk, z[i]);
        __syncthreads();
      }
    }
  }
  __syncthreads();
  // mean and variance: fixed-order reductions
  double sm = 0.0, sv = 0.0;
  const double* Mt = a.M + (size_t)t * p;
  for (int i = tid; i < p; i += TPB) {
    sm = fma(z[i], Mt[i] - a.beta0, sm);
    sv = fma(z[i], kM[i], sv);
  }
  sm = wave_sum(sm);
  sv = wave_sum(sv);
  __shared__ double rs[2][4];
  if (lane == 0) {
    rs[0][w] = sm;
    rs[1][w] = sv;
  }
  __syncthreads();
  if (tid == 0) {
    const int st = sflag;
    if (st == 2) {
      a.out[t] = nan("");
      a.out[a.qc + t] = nan("");
    } else {
      const double m = (rs[0][0] + rs[0][1]) + (rs[0][2] + rs[0][3]);
      const double v = (rs[1][0] + rs[1][1]) + (rs[1][2] + rs[1][3]);
      a.out[t] = a.beta0 + m;
      a.out[a.qc + t] = a.sigma2 * fmax(0.0, 1.0 - v);
    }
    a.info[t] = st;
  }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_nk_prepare(const NkArgs& a, hipStream_t stream) {
  const dim3 grid(a.qpad / NB, a.Npad / NB);
  if (a.d <= 4)
    hipLaunchKernelGGL(k_nk_cross<4>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 8)
    hipLaunchKernelGGL(k_nk_cross<8>, grid, dim3(TPB), 0, stream, a);
  else if (a.d <= 16)
    hipLaunchKernelGGL(k_nk_cross<16>, grid, dim3(TPB), 0, stream, a);
  else
    hipLaunchKernelGGL(k_nk_cross<32>, grid, dim3(TPB), 0, stream, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_nk_trmm, grid, dim3(TPB), 0, stream, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_nk_stats, dim3((a.qc + 3) / 4, a.p), dim3(TPB), 0, stream, a);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  hipLaunchKernelGGL(k_nk_trmm_t, grid, dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_nk_pairs(const NkArgs& a, hipStream_t stream) {
  if (a.npairs <= 0) return hipSuccess;
  if (a.qpad % NK_QT != 0) return hipErrorInvalidValue;
  const dim3 grid(a.npairs, a.qpad / NK_QT);
  // one instance per (dimension bound, covariance kernel): the kernel-value
  // generation runs one kernel's code, not a branch per value
  auto go = [&](auto dtag) {
    constexpr int DT = decltype(dtag)::value;
    switch (a.kernel) {
      case K_GAUSS: hipLaunchKernelGGL((k_nk_pair<DT, K_GAUSS>), grid, dim3(TPB), 0, stream, a); break;
      case K_EXP: hipLaunchKernelGGL((k_nk_pair<DT, K_EXP>), grid, dim3(TPB), 0, stream, a); break;
      case K_MATERN32: hipLaunchKernelGGL((k_nk_pair<DT, K_MATERN32>), grid, dim3(TPB), 0, stream, a); break;
      default: hipLaunchKernelGGL((k_nk_pair<DT, K_MATERN52>), grid, dim3(TPB), 0, stream, a); break;
    }
  };
  if (a.d <= 4)
    go(std::integral_constant<int, 4>{});
  else if (a.d <= 8)
    go(std::integral_constant<int, 8>{});
  else if (a.d <= 16)
    go(std::integral_constant<int, 16>{});
  else
    go(std::integral_constant<int, 32>{});
  return hipGetLastError();
}

hipError_t launch_nk_solve(const NkArgs& a, hipStream_t stream) {
  if (a.p > NK_PMAX) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_nk_solve, dim3(a.qc), dim3(TPB), 0, stream, a);
  return hipGetLastError();
}

hipError_t launch_nk_alpha(const NkArgs& a, const double* r, hipStream_t stream) {
  hipLaunchKernelGGL(k_nk_alpha, dim3(a.Npad / 4), dim3(TPB), 0, stream, a, r);
  return hipGetLastError();
}

}  // namespace lk

// Leave-one-out objective (Kriging::_leaveOneOut, Kriging.cpp:353-468), FP64,
// gfx950, run after the L^-1 path of the batched pipeline (L in A, L^-1 in
// W, F* / E* / x = L^-T E* in V).
//
// With Q* an orthonormal basis of F*, A = Q*' L^-1 and
// B = L^-T L^-1 - A'A (the reference's Q matrix), the reference computes
// sigma2LOO = 1 / diag(B), errors = sigma2LOO . (B y) with B y = x, and
// loo = |errors|^2 / n.  Here A'A = G'G with G = R*^-T F*' L^-1 (R* the upper
// Cholesky factor of F*'F*: F* R*^-1 spans the same space as Q*), so
//   diag(B)_j = |L^-1 e_j|^2 - |G e_j|^2
// comes from one pass over the columns of L^-1 (k_loo_g) without forming B.
//
// Gradient (Kriging.cpp:440-466).  With G_k = dR/dtheta_k (R_ij g_k(x_i - x_j)
// off the diagonal, zero on it), s = sigma2LOO, e = errors:
//   grad_k = 2/n sum_i e_i (s_i^2 diag(B G_k B)_i x_i - s_i (B G_k x)_i)
//          = 2/n sum_{j > l} G_k,jl (2 C_jl - (Bw)_j x_l - (Bw)_l x_j)
// where a = s^3 x^2, w = s^2 x and C = B diag(a) B = Bs Bs' with
// Bs = B diag(sqrt a) (a >= 0).  Bs is formed once (k_loo_bs, L^-1 tiles on
// MFMA, into A's storage), Bs Bs' is contracted tile by tile with the kernel
// derivatives recomputed from X (k_loo_adjoint), and
// B w = L^-T (L^-1 w) - G' (G w) takes two triangular matrix-vector passes.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <math.h>

#include "lk_device.hpp"
#include "lk_kernels.h"

namespace lk {

namespace {
__host__ __device__ inline int tri_n(int t) { return t * (t + 1) / 2; }
}  // namespace

// ---------------------------------------------------------------------------
// G' = (R*^-T F*' L^-1)' and diag(B).  grid (nblk_max, nslots), 256 threads;
// wave w of column block cb handles columns j = 64 cb + w, w + 4, ...; trend
// columns go four at a time through the rows i >= j of column j of L^-1.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_loo_g(const Slot* __restrict__ slots, const SlotParams* __restrict__ params,
                                               SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || ts.tile >= S.nblk) return;
  const int p = S.p, n = S.n, ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const double* __restrict__ rs = S.red + red_rstar_off(S.nblk);  // R* upper, rs[a + c p]
  double* __restrict__ Gt = S.V + (size_t)v_loo_g(p) * ld;
  double* __restrict__ bd = S.V + (size_t)v_loo_s2(p) * ld;
  __shared__ double hbuf[4][PMAX];
  constexpr int XU = 4;
  for (int cc = w; cc < NB; cc += 4) {
    const int j = ts.tile * NB + cc;
    if (j >= n) {  // padded column: L^-1 e_j = e_j, G e_j = 0
      if (lane == 0) {
        for (int a = 0; a < p; ++a) Gt[(size_t)a * ld + j] = 0.0;
        bd[j] = 1.0;
      }
      continue;
    }
    const double* __restrict__ Wc = S.W + (size_t)j * ld;
    double nrm = 0.0;
    for (int a0 = 0; a0 < max(p, 1); a0 += 4) {
      double h[4] = {0.0, 0.0, 0.0, 0.0};
      const double* fcol[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) fcol[c] = S.V + (size_t)min(a0 + c, max(p - 1, 0)) * ld;
      for (int b0 = j; b0 < n; b0 += 64 * XU) {
        double wv[XU], fv[XU][4];
#pragma unroll
        for (int u = 0; u < XU; ++u) {
          const int i = min(b0 + 64 * u + lane, n - 1);
          wv[u] = gload(Wc + i);
#pragma unroll
          for (int c = 0; c < 4; ++c) fv[u][c] = gload(fcol[c] + i);
        }
#pragma unroll
        for (int u = 0; u < XU; ++u)
          if (b0 + 64 * u + lane < n) {
            if (a0 == 0) nrm = fma(wv[u], wv[u], nrm);
#pragma unroll
            for (int c = 0; c < 4; ++c) h[c] = fma(fv[u][c], wv[u], h[c]);
          }
      }
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const double v = wave_sum(h[c]);
        if (lane == 0 && a0 + c < p) hbuf[w][a0 + c] = v;
      }
    }
    nrm = wave_sum(nrm);
    if (lane == 0) {
      // R*' g = h (forward substitution, R*' lower)
      double gg = 0.0;
      for (int a = 0; a < p; ++a) {
        double t = hbuf[w][a];
        for (int q = 0; q < a; ++q) t -= rs[q + a * p] * hbuf[w][q];
        const double g = t / rs[a + a * p];
        hbuf[w][a] = g;
        gg = fma(g, g, gg);
        Gt[(size_t)a * ld + j] = g;
      }
      bd[j] = nrm - gg;
    }
  }
}

// ---------------------------------------------------------------------------
// sigma2LOO = 1 / diag(B), errors, loo value; weights w = s^2 x and
// sqrt(a) = s^{3/2} |x| of the gradient.  One workgroup per slot.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_loo_value(const Slot* __restrict__ slots,
                                                   const SlotParams* __restrict__ params,
                                                   SlotOut* __restrict__ out) {
  const int b = blockIdx.x;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b])) return;
  const int p = S.p, n = S.n, ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const double* __restrict__ x = S.V + (size_t)(p + 3) * ld;
  double* s2v = S.V + (size_t)v_loo_s2(p) * ld;
  double* ev = S.V + (size_t)v_loo_err(p) * ld;
  double* wv = S.V + (size_t)v_loo_w(p) * ld;
  double* sav = S.V + (size_t)v_loo_sa(p) * ld;
  double acc = 0.0;
  for (int i = tid; i < ld; i += TPB) {
    if (i < n) {
      const double s2 = 1.0 / s2v[i];
      const double xi = x[i];
      const double e = s2 * xi;
      acc = fma(e, e, acc);
      s2v[i] = s2;
      ev[i] = e;
      wv[i] = s2 * e;
      sav[i] = s2 * sqrt(s2) * fabs(xi);
    } else {
      s2v[i] = 0.0;
      ev[i] = 0.0;
      wv[i] = 0.0;
      sav[i] = 0.0;
    }
  }
  __shared__ double sh[4];
  const double tot = block_sum4(acc, sh, lane, w);
  if (tid == 0) out[b].ll = tot / n;
}

// ---------------------------------------------------------------------------
// t = L^-1 w (row blocks, as the other lower matrix-vector products)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(MV_TPB) void k_loo_wv(const Slot* __restrict__ slots,
                                                   const SlotParams* __restrict__ params,
                                                   SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || !params[b].want_grad || ts.tile >= S.nblk) return;
  const int p = S.p, ld = S.npad;
  lower_matvec_block(S, ts.tile, S.V + (size_t)v_loo_w(p) * ld, ld, S.n, 1, S.V + (size_t)v_loo_t(p) * ld, ld,
                     nullptr);
}

// ---------------------------------------------------------------------------
// B w = L^-T t - G' (G w), column j of L^-1 against t; G w (p values) is
// formed by every workgroup (p n reads).  grid (nblk_max, nslots)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(MV_TPB) void k_loo_bw(const Slot* __restrict__ slots,
                                                   const SlotParams* __restrict__ params,
                                                   SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || !params[b].want_grad || ts.tile >= S.nblk) return;
  const int p = S.p, n = S.n, ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const double* __restrict__ Gt = S.V + (size_t)v_loo_g(p) * ld;
  const double* __restrict__ wv = S.V + (size_t)v_loo_w(p) * ld;
  const double* __restrict__ t = S.V + (size_t)v_loo_t(p) * ld;
  double* bw = S.V + (size_t)v_loo_bw(p) * ld;
  __shared__ double gw[PMAX];
  __shared__ double part[MV_W];
  for (int a = 0; a < p; ++a) {
    double s = 0.0;
    for (int m = tid; m < n; m += MV_TPB) s = fma(gload(Gt + (size_t)a * ld + m), gload(wv + m), s);
    s = wave_sum(s);
    if (lane == 0) part[w] = s;
    __syncthreads();
    if (tid == 0) {
      double v = part[0];
      for (int q = 1; q < MV_W; ++q) v += part[q];
      gw[a] = v;
    }
    __syncthreads();
  }
  constexpr int XU = 4;
  for (int cc = w; cc < NB; cc += MV_W) {
    const int j = ts.tile * NB + cc;
    const double* __restrict__ Wc = S.W + (size_t)j * ld;
    double s = 0.0;
    for (int b0 = j; b0 < ld; b0 += 64 * XU) {
      double wc[XU], tv[XU];
#pragma unroll
      for (int u = 0; u < XU; ++u) {
        const int i = min(b0 + 64 * u + lane, ld - 1);
        wc[u] = gload(Wc + i);
        tv[u] = gload(t + i);
      }
#pragma unroll
      for (int u = 0; u < XU; ++u)
        if (b0 + 64 * u + lane < ld) s = fma(wc[u], tv[u], s);
    }
    s = wave_sum(s);
    if (lane == 0) {
      double v = 0.0;
      if (j < n) {
        v = s;
        for (int a = 0; a < p; ++a) v -= Gt[(size_t)a * ld + j] * gw[a];
      }
      bw[j] = v;
    }
  }
}

// ---------------------------------------------------------------------------
// Bs = B diag(sqrt a), B = L^-T L^-1 - G'G, into A (full, column-major).
// Lower tile (i >= j) of L^-T L^-1 on MFMA as in the likelihood gradient,
// minus G'G in the epilogue; the tile and its mirror are stored with their
// column scales.  grid (nblk_max (nblk_max + 1) / 2, nslots)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_loo_bs(const Slot* __restrict__ slots, const SlotParams* __restrict__ params,
                                                SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b]) || !params[b].want_grad) return;
  if (ts.tile >= tri_n(S.nblk)) return;
  int it, jt;
  tri_decode(ts.tile, it, jt);
  const int p = S.p, ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int i0 = it * NB, j0 = jt * NB;
  // staging chunks, then the G rows of both blocks (2 PMAX NB doubles)
  constexpr int kBuf = 4 * KC * LDSS > 2 * PMAX * NB ? 4 * KC * LDSS : 2 * PMAX * NB;
  __shared__ double sbuf[kBuf];
  __shared__ double sai[NB], saj[NB];
  d4 acc[2][2];
  acc_zero(acc);
  tile_mainloop<true, true, false>(S.W, ld, i0, S.W, ld, j0, it * NB, S.nblk * NB, sbuf, acc, tid, lane, wr, wc);
  const double* __restrict__ Gt = S.V + (size_t)v_loo_g(p) * ld;
  const double* __restrict__ sa = S.V + (size_t)v_loo_sa(p) * ld;
  double* Gi = sbuf;
  double* Gj = sbuf + PMAX * NB;
  for (int q = tid; q < p * NB; q += TPB) {
    const int a = q / NB, r = q % NB;
    Gi[q] = Gt[(size_t)a * ld + i0 + r];
    Gj[q] = Gt[(size_t)a * ld + j0 + r];
  }
  if (tid < NB) {
    sai[tid] = sa[i0 + tid];
    saj[tid] = sa[j0 + tid];
  }
  __syncthreads();
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int r = acc_row(lane, wr, bi), c = acc_col(lane, wc, bj, e);
        double v = acc[bi][bj][e];
        for (int a = 0; a < p; ++a) v = fma(-Gi[a * NB + r], Gj[a * NB + c], v);
        S.A[(size_t)(j0 + c) * ld + i0 + r] = v * saj[c];
        if (it != jt) S.A[(size_t)(i0 + r) * ld + j0 + c] = v * sai[r];
      }
}

// ---------------------------------------------------------------------------
// C = Bs Bs' tile by tile (full K) and its contraction with dR/dtheta_k:
// partial_k = sum over the tile's pairs gi > gj of
//   (2 C_ij - (Bw)_i x_j - (Bw)_j x_i) R_ij g_k(x_j - x_i)
// into red[grad_off + tile * GRAD_STRIDE + k].  grid (tiles, nslots)
// ---------------------------------------------------------------------------
template <int DT>
__global__ __launch_bounds__(TPB) void k_loo_adjoint(const Slot* __restrict__ slots,
                                                     const SlotParams* __restrict__ params,
                                                     SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  if (!slot_live(P, out[b]) || !P.want_grad) return;
  if (ts.tile >= tri_n(S.nblk)) return;
  int it, jt;
  tri_decode(ts.tile, it, jt);
  const int p = S.p, n = S.n, d = S.d, ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int i0 = it * NB, j0 = jt * NB;
  __shared__ double sbuf[4 * KC * LDSS];
  static_assert(4 * KC * LDSS >= 2 * DT * NB, "X rows of both blocks fit the staging buffer");
  d4 acc[2][2];
  acc_zero(acc);
  tile_mainloop<false, false, false>(S.A, ld, i0, S.A, ld, j0, 0, S.nblk * NB, sbuf, acc, tid, lane, wr, wc);

  double* sXi = sbuf;
  double* sXj = sbuf + DT * NB;
  __shared__ double sxi[NB], sxj[NB], sbi[NB], sbj[NB];
  const double* __restrict__ xv = S.V + (size_t)(p + 3) * ld;
  const double* __restrict__ bw = S.V + (size_t)v_loo_bw(p) * ld;
  for (int q = tid; q < d * NB; q += TPB) {
    const int k = q / NB, r = q % NB;
    sXi[q] = (i0 + r < n) ? S.X[(size_t)k * n + i0 + r] : 0.0;
    sXj[q] = (j0 + r < n) ? S.X[(size_t)k * n + j0 + r] : 0.0;
  }
  if (tid < NB) {
    sxi[tid] = xv[i0 + tid];
    sxj[tid] = xv[j0 + tid];
    sbi[tid] = bw[i0 + tid];
    sbj[tid] = bw[j0 + tid];
  }
  __syncthreads();
  double it_[DT], it2[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) {
    it_[k] = (k < d) ? 1.0 / P.theta[k] : 1.0;
    it2[k] = it_[k] * it_[k];
  }
  const double factor = (S.noise == NOISE_NUGGET) ? P.alpha : 1.0;
  double s[DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) s[k] = 0.0;
  // one element (r, c) of the tile, Bs Bs' value v
  // specialised per covariance kernel: one kernel's code runs
  auto element = [&](auto ktag, int r, int c, double v) __attribute__((always_inline)) {
    constexpr int KT = decltype(ktag)::value;
    const int gi = i0 + r, gj = j0 + c;
    if (gi >= n || gj >= gi) return;
    const double m = 2.0 * v - sbi[r] * sxj[c] - sbj[c] * sxi[r];
    double dx[DT], g[DT];
#pragma unroll
    for (int k = 0; k < DT; ++k) dx[k] = (k < d) ? (sXj[k * NB + c] - sXi[k * NB + r]) : 0.0;
    const double mr = m * factor * kern_rho_g<DT>(KT, dx, it_, it2, g);
#pragma unroll
    for (int k = 0; k < DT; ++k) s[k] = fma(mr, g[k], s[k]);
  };
  auto tile = [&](auto ktag) __attribute__((always_inline)) {
    if constexpr (DT <= 8) {
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int e = 0; e < 4; ++e) element(ktag, acc_row(lane, wr, bi), acc_col(lane, wc, bj, e), acc[bi][bj][e]);
    } else {
      // beyond 8 dimensions the sixteen elements do not unroll: the four
      // 16 x 16 blocks in a loop, each picked from the accumulators by
      // selects (a loop index into acc would move it to scratch memory)
#pragma unroll 1
      for (int q = 0; q < 4; ++q) {
        const d4 blk = (q == 0) ? acc[0][0] : ((q == 1) ? acc[0][1] : ((q == 2) ? acc[1][0] : acc[1][1]));
#pragma unroll
        for (int e = 0; e < 4; ++e) element(ktag, acc_row(lane, wr, q >> 1), acc_col(lane, wc, q & 1, e), blk[e]);
      }
    }
  };
  switch (S.kernel) {
    case K_GAUSS: tile(std::integral_constant<int, K_GAUSS>{}); break;
    case K_EXP: tile(std::integral_constant<int, K_EXP>{}); break;
    case K_MATERN32: tile(std::integral_constant<int, K_MATERN32>{}); break;
    default: tile(std::integral_constant<int, K_MATERN52>{}); break;
  }
  __shared__ double sred[4][DT];
#pragma unroll
  for (int k = 0; k < DT; ++k) {
    const double v = wave_sum(s[k]);
    if (lane == 0) sred[w][k] = v;
  }
  __syncthreads();
  if (tid < DT) {
    double* dst = S.red + red_grad_off(S.nblk) + (size_t)ts.tile * GRAD_STRIDE;
    dst[tid] = (sred[0][tid] + sred[1][tid]) + (sred[2][tid] + sred[3][tid]);
  }
}

// ---------------------------------------------------------------------------
// grad_k = 2 / n * sum of the tile partials (fixed order).  One workgroup per
// slot.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_loo_finalize_grad(const Slot* __restrict__ slots,
                                                           const SlotParams* __restrict__ params,
                                                           SlotOut* __restrict__ out) {
  const int b = blockIdx.x;
  const Slot S = slots[b];
  const SlotParams& P = params[b];
  SlotOut& O = out[b];
  if (!slot_live(P, O) || !P.want_grad) return;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ntiles = tri_n(S.nblk);
  for (int k = w; k < S.d; k += TPB / 64) {
    const double* src = S.red + red_grad_off(S.nblk) + k;
    double s = 0.0;
    for (int t = lane; t < ntiles; t += 64) s += src[(size_t)t * GRAD_STRIDE];
    s = wave_sum(s);
    if (lane == 0) O.grad[k] = 2.0 * s / S.n;
  }
}

// ---------------------------------------------------------------------------
// R^-1 = L^-T L^-1 stored in full into A (the model file's "Rinv").  Lower
// tiles on MFMA, each stored with its mirror.  grid (tiles, nslots)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(TPB) void k_rinv_store(const Slot* __restrict__ slots,
                                                    const SlotParams* __restrict__ params,
                                                    SlotOut* __restrict__ out) {
  const TileSlot ts = tile_slot();
  const int b = ts.slot;
  const Slot S = slots[b];
  if (!slot_live(params[b], out[b])) return;
  if (ts.tile >= tri_n(S.nblk)) return;
  int it, jt;
  tri_decode(ts.tile, it, jt);
  const int ld = S.npad;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, wr = w >> 1, wc = w & 1;
  const int i0 = it * NB, j0 = jt * NB;
  __shared__ double sbuf[4 * KC * LDSS];
  d4 acc[2][2];
  acc_zero(acc);
  tile_mainloop<true, true, false>(S.W, ld, i0, S.W, ld, j0, it * NB, S.nblk * NB, sbuf, acc, tid, lane, wr, wc);
  acc_store(acc, S.A + (size_t)j0 * ld + i0, ld, lane, wr, wc);
  if (it != jt) acc_store_t(acc, S.A + (size_t)i0 * ld + j0, ld, lane, wr, wc);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
hipError_t launch_rinv_store(const Batch& B) {
  hipLaunchKernelGGL(k_rinv_store, dim3(tri_n(B.nblk_max), B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params,
                     B.out);
  return hipGetLastError();
}
hipError_t launch_loo_g(const Batch& B) {
  hipLaunchKernelGGL(k_loo_g, dim3(B.nblk_max, B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_loo_value(const Batch& B) {
  hipLaunchKernelGGL(k_loo_value, dim3(B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_loo_wv(const Batch& B) {
  hipLaunchKernelGGL(k_loo_wv, dim3(B.nblk_max, B.nslots), dim3(MV_TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_loo_bw(const Batch& B) {
  hipLaunchKernelGGL(k_loo_bw, dim3(B.nblk_max, B.nslots), dim3(MV_TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_loo_bs(const Batch& B) {
  hipLaunchKernelGGL(k_loo_bs, dim3(tri_n(B.nblk_max), B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params,
                     B.out);
  return hipGetLastError();
}
hipError_t launch_loo_adjoint(const Batch& B) {
  const dim3 grid(tri_n(B.nblk_max), B.nslots);
  if (B.dmax <= 4)
    hipLaunchKernelGGL(k_loo_adjoint<4>, grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else if (B.dmax <= 8)
    hipLaunchKernelGGL(k_loo_adjoint<8>, grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else if (B.dmax <= 16)
    hipLaunchKernelGGL(k_loo_adjoint<16>, grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  else
    hipLaunchKernelGGL(k_loo_adjoint<32>, grid, dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}
hipError_t launch_loo_finalize_grad(const Batch& B) {
  hipLaunchKernelGGL(k_loo_finalize_grad, dim3(B.nslots), dim3(TPB), 0, B.stream, B.slots, B.params, B.out);
  return hipGetLastError();
}

}  // namespace lk

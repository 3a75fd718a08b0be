// Device bodies of the FGMRES basis kernels (classical Gram-Schmidt with
// re-orthogonalisation, fused into three passes over the basis).  Shared by the
// single-case kernels (pp_core.hip) and the batched kernels (pp_batch.hip).
#pragma once

#include "pp_common.hpp"
#include "pp_stencil.hpp"

namespace pp {

// Per-block sums of a * V_l[c] for l < nv and of a * a (slot nv), written to
// partial[l * nblk + b].  Eight vectors at a time are loaded together, reduced
// inside each wave by shuffles and parked in LDS; one barrier, then thread l
// adds the four wave sums of slot l.  256-thread blocks, nv < 64; every thread
// of the block must call it.
__device__ inline void block_dots(bool in, size_t c, double a, const double* __restrict__ V, size_t ldv, int nv,
                                  double* __restrict__ partial, size_t nblk, size_t b, double (*lds)[64]) {
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int lane = t & 63, w = t >> 6;
  for (int l0 = 0; l0 <= nv; l0 += 8) {
    double p[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int l = l0 + q;
      double v = 0.;
      if (in && l < nv) v = a * V[(size_t)l * ldv + c];
      else if (in && l == nv) v = a * a;
      p[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) p[q] = wave_sum(p[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (l0 + q <= nv) lds[w][l0 + q] = p[q];
    }
  }
  __syncthreads();
  if (t <= nv) partial[(size_t)t * nblk + b] = (lds[0][t] + lds[1][t]) + (lds[2][t] + lds[3][t]);
  __syncthreads();  // lds is reused by the block's next tile
}

// x from a lane DPP_CTRL selects (0 where the source lane is out of its row or
// the lane is masked out by row_mask / bank_mask), a double moved as two dwords
template <int DPP_CTRL, int ROW_MASK, int BANK_MASK>
__device__ __forceinline__ double dpp_f64(double x) {
  union {
    double d;
    int i[2];
  } u, r;
  u.d = x;
  r.i[0] = __builtin_amdgcn_update_dpp(0, u.i[0], DPP_CTRL, ROW_MASK, BANK_MASK, true);
  r.i[1] = __builtin_amdgcn_update_dpp(0, u.i[1], DPP_CTRL, ROW_MASK, BANK_MASK, true);
  return r.d;
}

// Sum of v over the 64 lanes of a wave, left in lane 63 (other lanes hold partial
// sums), without LDS: row shifts by 1, 2, 4 and 8 leave each row's sum in its
// last lane, then row broadcasts 15 and 31 carry the row sums into lane 63.
// Every lane of the wave must be active.
__device__ __forceinline__ double wave_sum_to63(double v) {
  v += dpp_f64<0x111, 0xF, 0xF>(v);  // row_shr:1
  v += dpp_f64<0x112, 0xF, 0xF>(v);  // row_shr:2
  v += dpp_f64<0x114, 0xF, 0xE>(v);  // row_shr:4 into banks 1-3
  v += dpp_f64<0x118, 0xF, 0xC>(v);  // row_shr:8 into banks 2-3
  v += dpp_f64<0x142, 0xA, 0xF>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xC, 0xF>(v);  // row_bcast:31 into rows 2 and 3
  return v;
}

// V_l[c] for l < min(nv, 8), loaded before the value they multiply is computed so
// the loads overlap that computation
__device__ __forceinline__ void wave_dots_prefetch(bool in, size_t c, const double* __restrict__ V, size_t ldv,
                                                   int nv, double (&v8)[8]) {
#pragma unroll
  for (int q = 0; q < 8; q++) v8[q] = (in && q < nv) ? V[(size_t)q * ldv + c] : 0.;
}

// Per-wave sums of a * V_l[c] for l < nv and of a * a (slot nv), written by the
// wave's last lane to partial[l * nw + wv] (nw waves in the launch, wv this
// wave's index).  v8 holds V_l[c] for l < 8 (wave_dots_prefetch); the rest are
// loaded eight at a time.  Every lane of the wave must call it (cells outside the
// grid pass in = false); nv < 64.
__device__ inline void wave_dots(bool in, size_t c, double a, const double* __restrict__ V, size_t ldv, int nv,
                                 const double (&v8)[8], double* __restrict__ partial, size_t nw, size_t wv) {
  const int lane = (threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z)) & 63;
  for (int l0 = 0; l0 <= nv; l0 += 8) {
    double p[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int l = l0 + q;
      double v = 0.;
      if (in && l < nv) v = a * (l0 == 0 ? v8[q] : V[(size_t)l * ldv + c]);
      else if (in && l == nv) v = a * a;
      p[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) p[q] = wave_sum_to63(p[q]);
    if (lane == 63) {
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (l0 + q <= nv) partial[(size_t)(l0 + q) * nw + wv] = p[q];
    }
  }
}

// Per-block sums of per-thread accumulators acc[0..ns) (ns <= S <= 64), written to
// partial[l * nblk + b]: wave sums by shuffles, the four wave results of a
// 256-thread block meet in LDS.  Every thread of the block must call it.
template <int S>
__device__ inline void block_sums(const double (&acc)[S], int ns, double* __restrict__ partial, size_t nblk,
                                  size_t b, double (*lds)[64]) {
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int lane = t & 63, w = t >> 6;
#pragma unroll
  for (int l = 0; l < S; l++) {
    if (l < ns) {
      const double v = wave_sum(acc[l]);
      if (lane == 0) lds[w][l] = v;
    }
  }
  __syncthreads();
  if (t < ns) partial[(size_t)t * nblk + b] = (lds[0][t] + lds[1][t]) + (lds[2][t] + lds[3][t]);
  __syncthreads();  // lds is reused by the block's next call
}

// gs_sub followed by block_dots of the result at element t, with the basis
// values V_l[t] (l < nv <= S) loaded once into registers for both; the same
// arithmetic as the two calls.  Every thread of the block must call it.
template <int S>
__device__ inline void gs_sub_block_dots(bool in, double* __restrict__ w, const double* __restrict__ V, size_t ldv,
                                         int nv, const double* __restrict__ h, size_t t,
                                         double* __restrict__ partial, size_t nblk, size_t b, double (*lds)[64]) {
  const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  const int lane = tid & 63, wv = tid >> 6;
  double vv[S];
#pragma unroll
  for (int l = 0; l < S; l++) vv[l] = (in && l < nv) ? V[(size_t)l * ldv + t] : 0.;
  double w1 = 0.;
  if (in) {
    w1 = w[t];
#pragma unroll
    for (int l = 0; l < S; l++)
      if (l < nv) w1 -= h[l] * vv[l];
    w[t] = w1;
  }
#pragma unroll
  for (int l0 = 0; l0 <= S; l0 += 8) {
    if (l0 > nv) break;
    double p[8];
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int l = l0 + q;
      double v = 0.;
      if (in && l < nv && l < S) v = w1 * vv[l < S ? l : 0];
      else if (in && l == nv) v = w1 * w1;
      p[q] = v;
    }
#pragma unroll
    for (int q = 0; q < 8; q++) p[q] = wave_sum(p[q]);
    if (lane == 0) {
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (l0 + q <= nv) lds[wv][l0 + q] = p[q];
    }
  }
  __syncthreads();
  if (tid <= nv) partial[(size_t)tid * nblk + b] = (lds[0][tid] + lds[1][tid]) + (lds[2][tid] + lds[3][tid]);
  __syncthreads();  // lds is reused by the block's next tile
}

// Power of two near 1/||w|| from ||w||^2, the scale of the single-precision copy
// of an unnormalised basis vector (the host computes the same value).
__host__ __device__ inline double conv_scale(double a2) {
  if (!(a2 > 0.)) return 1.;
  const int e = ilogb(a2);
  return ldexp(1.0, -(int)floor(0.5 * e));
}

// w[c] = (A z)[c] in flux form on a row cell, 0 elsewhere; returns w[c]
template <typename TZ>
__device__ __forceinline__ double apply_cell(const Dims& d, const double* __restrict__ g,
                                             const uint32_t* __restrict__ code, const TZ* __restrict__ z,
                                             double* __restrict__ w, int i, int j, int k) {
  const size_t c = d.idx(i, j, k);
  const uint32_t cd = code[c];
  Nbr nb;
  neighbours(d, i, j, k, nb);
  double xv[19], ax, dg;
  load_x(nb, z, xv);
  flux_row(g, d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, ax, dg);
  const double wv = (cd & ROW_BIT) ? ax : 0.;
  w[c] = wv;
  return wv;
}

// w1 = w - sum_l h_l V_l at element t (in place); returns w1
__device__ __forceinline__ double gs_sub(double* __restrict__ w, const double* __restrict__ V, size_t ldv, int nv,
                                         const double* __restrict__ h, size_t t) {
  double w1 = w[t];
  for (int l = 0; l < nv; l++) w1 -= h[l] * V[(size_t)l * ldv + t];
  w[t] = w1;
  return w1;
}

// red[1] = ||w - V h|| = sqrt(w.w - h.h), red[2] = 1/red[1] with h = red[off..off+nv)
// and w.w = red[off+nv]; mirrored to host.  One thread.
__device__ __forceinline__ void gs_norm_one(double* __restrict__ red, int off, int nv, double* __restrict__ host) {
  const double* h = red + off;
  double hh = 0.;
  for (int l = 0; l < nv; l++) hh += h[l] * h[l];
  const double s = h[nv] - hh;  // h[nv] holds w.w
  const double nrm = s > 0. ? sqrt(s) : 0.;
  red[1] = nrm;
  red[2] = nrm > 0. ? 1. / nrm : 0.;
  if (host) { host[1] = nrm; host[2] = red[2]; }
}

// v = (w - sum_l h_l V_l) * inv at element t (in place over w); f = (TF)v
template <typename TF>
__device__ __forceinline__ void axpy_scale_one(double* __restrict__ w, const double* __restrict__ V, size_t ldv,
                                               int nv, const double* __restrict__ h, double inv, TF* __restrict__ f,
                                               size_t t) {
  double acc = w[t];
  for (int l = 0; l < nv; l++) acc -= h[l] * V[(size_t)l * ldv + t];
  const double v = acc * inv;
  w[t] = v;
  if (f) f[t] = (TF)v;
}

}  // namespace pp

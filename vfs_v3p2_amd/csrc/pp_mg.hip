// Geometric multigrid kernels: Chebyshev smoothers (point-Jacobi and y-line
// tridiagonal), fused residual+restriction, IB-masked trilinear prolongation and
// the dense coarsest-level solve.  Templated on the level precision (float for
// the mixed-precision preconditioner, double for the plain FP64 form).
//
// The transfers follow the reference's MG primitives (MyRestriction
// poisson.c:364-471, MyInterpolation poisson.c:473-669): restriction sums the
// fluid children of a coarse cell (the coarse operator is in the same
// integrated flux form as the fine one, so coarse residuals add up), zero on
// coarse IB cells; prolongation uses the (27,9,9,3,9,3,3,1)/64 weights, one-sided
// next to coarse IB cells and at non-periodic walls, periodic axes wrap.
#include <algorithm>

#include "pp_mg_dev.hpp"

namespace pp {

namespace {

const dim3 kBlk(64, 4, 1);
inline dim3 cell_grid(const Dims& d) { return dim3((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz); }
inline dim3 line_grid(const Dims& d) { return dim3((d.nx + 63) / 64, (d.nz + 3) / 4, 1); }

template <typename T>
__global__ __launch_bounds__(256) void k_dinv(LevelView<T> L, T* __restrict__ dinv) {
  CELL_IJK(i, j, k);
  if (i >= L.d.nx || j >= L.d.ny) return;
  dinv_cell(L, dinv, i, j, k);
}

template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_smooth_point(LevelView<T> L, const T* __restrict__ f,
                                                      const T* __restrict__ dinv, const T* __restrict__ x_old,
                                                      T* x_prev, T* x_new, T* __restrict__ dv, T c1, T c2, T c0,
                                                      const T* __restrict__ fd) {
  CELL_IJK(i, j, k);
  if (i >= L.d.nx || j >= L.d.ny) return;
  smooth_point_cell<T, MODE>(L, f, dinv, x_old, x_prev, x_new, dv, c1, c2, c0, i, j, k, fd);
}

template <typename T>
__global__ __launch_bounds__(256) void k_line_rows(LevelView<T> L, T* __restrict__ lfac) {
  CELL_IJK(i, j, k);
  if (i >= L.d.nx || j >= L.d.ny) return;
  line_rows_cell(L, lfac, i, j, k);
}

template <typename T>
__global__ __launch_bounds__(256) void k_line_factors(Dims d, T* __restrict__ lfac) {
  const int i = blockIdx.x * 64 + threadIdx.x, k = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.nx || k >= d.nz) return;
  line_factors_col(d, lfac, i, k);
}

template <typename T>
__global__ __launch_bounds__(256) void k_resid(LevelView<T> L, const T* __restrict__ f, const T* __restrict__ x,
                                               T* __restrict__ r) {
  CELL_IJK(i, j, k);
  if (i >= L.d.nx || j >= L.d.ny) return;
  resid_cell(L, f, x, r, i, j, k);
}

template <typename T, bool ZG>
__global__ __launch_bounds__(256) void k_line_solve(Dims d, const T* __restrict__ r, const T* __restrict__ lfac,
                                                    T* __restrict__ tmp, const T* __restrict__ x_old,
                                                    T* __restrict__ x_new, T* __restrict__ dv, T c1, T c2) {
  const int i = blockIdx.x * 64 + threadIdx.x, k = blockIdx.y * 4 + threadIdx.y;
  if (i >= d.nx || k >= d.nz) return;
  line_solve_col<T, ZG>(d, r, lfac, tmp, x_old, x_new, dv, c1, c2, i, k);
}

// line_lanes<T>() lines (i = line_lanes<T>() blockIdx.x + line_lane<T>(), k =
// blockIdx.y) per (64, 16) block
template <typename T, bool ZG>
__global__ __launch_bounds__(1024) void k_line_seg(Dims d, const T* __restrict__ r,
                                                             const T* __restrict__ lfac,
                                                             const T* __restrict__ x_old, T* __restrict__ x_new,
                                                             T* __restrict__ dv, T c1, T c2) {
  __shared__ LineSegLds<T> sh;
  const int i = blockIdx.x * line_lanes<T>() + line_lane<T>(), k = blockIdx.y;
  line_solve_seg<T, ZG>(d, i < d.nx, i, k, r, lfac, x_old, x_new, dv, c1, c2, sh);
}

// residual and line step in one pass, same block layout as k_line_seg
template <typename T>
__global__ __launch_bounds__(1024) void k_line_fused(LevelView<T> L, const T* __restrict__ f,
                                                               const T* __restrict__ lfac,
                                                               const T* __restrict__ x_old, T* __restrict__ x_new,
                                                               T* __restrict__ dv, T c1, T c2) {
  __shared__ LineSegLds<T> sh;
  const int i = blockIdx.x * line_lanes<T>() + line_lane<T>(), k = blockIdx.y;
  line_step_cell<T>(L, i < L.d.nx, i, k, f, lfac, x_old, x_new, dv, c1, c2, sh);
}

// HALF: the coefficients are FP16 (gh), else T (g)
template <typename T, bool HALF>
__global__ __launch_bounds__(256) void k_stream_floor(size_t n, const void* __restrict__ gv,
                                                      const uint8_t* __restrict__ cls, const T* __restrict__ f,
                                                      const T* __restrict__ x, const T* __restrict__ xp,
                                                      T* __restrict__ out) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  T acc = (T)cls[c] + f[c] + x[c] + xp[c];
#pragma unroll
  for (int q = 0; q < 9; q++) {
    if constexpr (HALF) acc += (T) static_cast<const _Float16*>(gv)[(size_t)q * n + c];
    else acc += static_cast<const T*>(gv)[(size_t)q * n + c];
  }
  out[c] = acc;
}

template <typename T>
__global__ __launch_bounds__(256) void k_resid_restrict(LevelView<T> F, const T* __restrict__ f,
                                                        const T* __restrict__ x, Dims dc,
                                                        const uint8_t* __restrict__ clsc, int cmask,
                                                        T* __restrict__ fc) {
  const int ic = blockIdx.x * 64 + threadIdx.x, jc = blockIdx.y * 4 + threadIdx.y, kc = blockIdx.z;
  if (ic >= dc.nx || jc >= dc.ny) return;
  resid_restrict_cell(F, f, x, dc, clsc, cmask, fc, ic, jc, kc);
}

template <typename T>
__global__ __launch_bounds__(256) void k_resid_restrict_blk(LevelView<T> F, const T* __restrict__ f,
                                                            const T* __restrict__ x, Dims dc,
                                                            const uint8_t* __restrict__ clsc, T* __restrict__ fc,
                                                            int keep_y) {
  __shared__ T lds[2][2][32];
  const uint3 xb = xcd_block();
  resid_restrict_blk(F, f, x, dc, clsc, fc, (int)xb.x, (int)xb.y, (int)xb.z, keep_y != 0, lds);
}

template <typename T>
__global__ __launch_bounds__(256) void k_prolong_add(LevelView<T> F, Dims dc, const uint8_t* __restrict__ clsc,
                                                     int cmask, const T* __restrict__ xc, T* __restrict__ xf) {
  CELL_IJK(i, j, k);
  if (i >= F.d.nx || j >= F.d.ny) return;
  prolong_add_cell(F, dc, clsc, cmask, xc, xf, i, j, k);
}

template <typename T>
__global__ __launch_bounds__(256) void k_dense_solve(int n, const T* __restrict__ cinv, const T* __restrict__ f,
                                                     T* __restrict__ x) {
  __shared__ T fs[1024];
  __shared__ T part[4][64];
  dense_rows(n, cinv, f, x, (int)blockIdx.x, fs, part);
}

// ------------------------------------------------------------- coarse tail
// The bottom of a V-cycle in one workgroup: the last smoothed level (point
// Chebyshev, at most 1024 cells, one per thread) with its iterates in LDS, the
// residual restricted to the dense level, the dense solve, the prolongation and
// the post-smoothing.  The launch loads the level's face coefficients, code words,
// class bytes, f and f D^-1 into LDS once, so the phases between barriers touch
// only LDS, instead of one launch with cold caches per phase.  Thread t smooths
// cell order[t] (tail_order): interior cells grouped into waves of their own, the
// waves spread evenly over the SIMDs.  Every step is the arithmetic of the
// separate kernels (smooth_point_body, resid_restrict_blk / resid_restrict_cell,
// dense_rows, prolong_add_cell) in the same order, so the result is the same bits.
// dense inverses up to this many entries (64 KB in FP32) are staged in LDS
constexpr int TAIL_CINV_LDS = 16384;
template <typename T, typename TG>
struct TailLds {
  T cinv[TAIL_CINV_LDS];                                          // the dense inverse, when it fits
  TG g[9][TAIL_MAX_CELLS];                                        // the level's face coefficients
  T xa[TAIL_MAX_CELLS], xb[TAIL_MAX_CELLS], fd[TAIL_MAX_CELLS];  // iterates; f D^-1, then the residual
  T fc[TAIL_MAX_CELLS], xc[TAIL_MAX_CELLS];                        // dense level: right-hand side and solution
  T fv[TAIL_MAX_CELLS];                                            // the level's f
  uint32_t code[TAIL_MAX_CELLS];                                   // its code words
  uint8_t cls[TAIL_MAX_CELLS];                                     // its class bytes
  T part[4][4][64];                                               // dense_rows partial sums, 4 row blocks
  T c1[TAIL_MAX_DEGREE], c2[TAIL_MAX_DEGREE];                     // Chebyshev coefficients per step
  uint8_t clsc[TAIL_MAX_CELLS];                                   // dense level's class bytes
};

// (A x)_c, the diagonal and whether the cell has a row, as flux_by_class and
// smooth_point_body form them (coefficients already in registers)
template <typename T, typename TG>
__device__ __forceinline__ void tail_flux(const RowCoef<TG>& rc, uint32_t code, bool interior, bool scaled, T gs,
                                          const T (&xv)[19], T& ax, T& diag, bool& row) {
  row = true;
  if (__all(interior)) {
    flux_row_v<X_CENTRAL>(rc, CODE_INTERIOR, xv, ax, diag);
  } else {
    flux_row_v<-1>(rc, code, xv, ax, diag);
    row = (code & ROW_BIT) != 0;
  }
  if (scaled) {
    ax *= gs;
    diag *= gs;
  }
}

// the cell's stencil indices and its row's 18 coefficients (load_coef's faces) from LDS
template <typename T, typename TG>
__device__ __forceinline__ void tail_row(const TailLds<T, TG>& sh, const Dims& d, int i, int j, int k, Nbr& nb,
                                         RowCoef<TG>& rc) {
  neighbours(d, i, j, k, nb);
  const uint32_t own[6] = {nb.s[CP], nb.s[WP], nb.s[CP], nb.s[SP], nb.s[CP], nb.s[BP]};
#pragma unroll
  for (int f = 0; f < 6; f++)
#pragma unroll
    for (int e = 0; e < 3; e++) rc.v[f][e] = sh.g[3 * (f / 2) + e][own[f]];
}

// one three-term Chebyshev step from cur (neighbours) and prev (own cell): x_new
template <typename T, typename TG>
__device__ __forceinline__ T tail_step(const TailLds<T, TG>& sh, const Dims& d, int i, int j, int k, uint32_t code,
                                       bool interior, bool scaled, T gs, T fv, const T* cur, const T* prev,
                                       uint32_t c, T c1, T c2) {
  const T xp = (c1 != T(0)) ? prev[c] : T(0);
  Nbr nb;
  RowCoef<TG> rc;
  tail_row(sh, d, i, j, k, nb, rc);
  T xv[19];
#pragma unroll
  for (int s = 0; s < 19; s++) xv[s] = cur[nb.s[s]];
  T ax, dg;
  bool row;
  tail_flux(rc, code, interior, scaled, gs, xv, ax, dg, row);
  const bool live = row && dg != T(0);
  const T z = live ? (fv - ax) / (live ? dg : T(1)) : T(0);
  const T xc = xv[CP];
  const T dold = c1 == T(0) ? T(0) : xc - xp;
  return row ? xc + (c1 * dold + c2 * z) : T(0);
}

template <typename T, typename TG>
__device__ __forceinline__ void coarse_tail_body(const TailArgs<T>& A, const TG* __restrict__ g, bool scaled,
                                                 TailLds<T, TG>& sh) {
  const LevelView<T>& L = A.L;
  const Dims& d = L.d;
  const int n = (int)d.n(), nc = (int)A.dc.n();
  const int t = (int)threadIdx.x;
  // Phase 0 loads cell t's data into LDS; from then on thread t owns cell A.order[t]
  // (issued with the other loads, so it adds no round trip to memory).
  const bool own_t = t < n;
  const uint16_t ord = A.order ? A.order[t] : (own_t ? (uint16_t)t : TAIL_NO_CELL);
  // Every global load of the launch is issued before the first LDS store, so the
  // launch waits for memory once (a load-store loop over the dense inverse would
  // wait once per round).
  constexpr int CINV_PER_THREAD = TAIL_CINV_LDS / 1024;
  const int ncc = nc * nc;
  const bool cinv_lds = ncc <= TAIL_CINV_LDS;
  T cv[CINV_PER_THREAD];
#pragma unroll
  for (int u = 0; u < CINV_PER_THREAD; u++) {
    const int q = t + 1024 * u;
    cv[u] = (cinv_lds && q < ncc) ? A.cinv[q] : T(0);
  }
  TG gv[9];
  uint32_t code_t = 0;
  uint8_t cl_t = 0, clc = 0;
  T f_t = T(0), dv = T(0), cf1 = T(0), cf2 = T(0);
  if (own_t) {
#pragma unroll
    for (int q = 0; q < 9; q++) gv[q] = ld_at(g + (size_t)q * d.n(), (uint32_t)t);
    code_t = L.code[t];
    cl_t = L.cls[t];
    f_t = A.f[t];
    dv = A.dinv[t];
  }
  if (t < nc) clc = A.clsc[t];  // nc <= 1024: one dense cell per thread at most
  if (t < TAIL_MAX_DEGREE) {
    cf1 = A.c1[t];
    cf2 = A.c2[t];
  }
  if (own_t) {
#pragma unroll
    for (int q = 0; q < 9; q++) sh.g[q][t] = gv[q];
    sh.fd[t] = f_t * dv;
    sh.code[t] = code_t;
    sh.cls[t] = cl_t;
    sh.fv[t] = f_t;
  }
  if (t < nc) sh.clsc[t] = clc;
#pragma unroll
  for (int u = 0; u < CINV_PER_THREAD; u++) {
    const int q = t + 1024 * u;
    if (cinv_lds && q < ncc) sh.cinv[q] = cv[u];
  }
  if (t < TAIL_MAX_DEGREE) {
    sh.c1[t] = cf1;
    sh.c2[t] = cf2;
  }
  __syncthreads();
  const bool mine = ord < n;
  const uint32_t c = mine ? (uint32_t)ord : 0u;
  const int i = (int)c % d.nx, j = ((int)c / d.nx) % d.ny, k = (int)c / (d.nx * d.ny);
  const uint32_t code = mine ? sh.code[c] : 0u;
  const uint8_t cl = mine ? sh.cls[c] : (uint8_t)0;
  const T fv = mine ? sh.fv[c] : T(0);
  const bool interior = (cl & CLS_INTERIOR) != 0;
  const T gs = L.gs;

  // pre-smoothing from zero: the fused first two steps (SM_FIRST2), or one step
  T* cur = sh.xb;
  T* prev = sh.xa;
  int k0;
  if (A.first2) {
    if (mine) {
      const T c0 = sh.c2[0], c1 = sh.c1[1], c2 = sh.c2[1];
      Nbr nb;
      RowCoef<TG> rc;
      tail_row(sh, d, i, j, k, nb, rc);
      T xv[19];
#pragma unroll
      for (int s = 0; s < 19; s++) xv[s] = c0 * sh.fd[nb.s[s]];
      T ax, dg;
      bool row;
      tail_flux(rc, code, interior, scaled, gs, xv, ax, dg, row);
      const bool live = row && dg != T(0);
      const T z = live ? (fv - ax) / (live ? dg : T(1)) : T(0);
      const T xc = xv[CP];
      sh.xa[c] = row ? xc : T(0);
      sh.xb[c] = row ? xc + (c1 * xc + c2 * z) : T(0);
    }
    k0 = 2;
  } else {
    if (mine) {
      sh.xb[c] = sh.c2[0] * sh.fd[c];
      sh.xa[c] = T(0);
    }
    k0 = 1;
  }
  __syncthreads();
  for (int s = k0; s < A.degree; s++) {
    const T xn = mine ? tail_step(sh, d, i, j, k, code, interior, scaled, gs, fv, cur, prev, c, sh.c1[s], sh.c2[s])
                      : T(0);
    if (mine) prev[c] = xn;  // the step's x_prev slot takes x_new
    T* tmp = cur;
    cur = prev;
    prev = tmp;
    __syncthreads();
  }

  // residual of the fine rows, then the restriction (the form the launch path takes)
  if (mine) {
    Nbr nb;
    RowCoef<TG> rc;
    tail_row(sh, d, i, j, k, nb, rc);
    T xv[19];
#pragma unroll
    for (int s = 0; s < 19; s++) xv[s] = cur[nb.s[s]];
    T ax, dg;
    if (__all((code & ~DIVOK_BIT) == CODE_INTERIOR))
      flux_row_v<X_CENTRAL>(rc, code, xv, ax, dg);
    else
      flux_row_v<-1>(rc, code, xv, ax, dg);
    if (scaled) ax *= gs;
    sh.fd[c] = (code & ROW_BIT) ? fv - ax : T(0);
  }
  __syncthreads();
  const Dims& dc = A.dc;
  const int cm = (dc.nx != d.nx ? 1 : 0) | (dc.ny != d.ny ? 2 : 0) | (dc.nz != d.nz ? 4 : 0);
  for (int q = t; q < nc; q += 1024) {
    const int ic = q % dc.nx, jc = (q / dc.nx) % dc.ny, kc = q / (dc.nx * dc.ny);
    T v = T(0);
    if (cm == 7 || cm == 5) {
      // resid_restrict_blk: pairs in i, then j and k pairs (j kept when cm == 5)
      auto r_at = [&](int ii, int jj, int kk) -> T {
        return (ii < d.nx && jj < d.ny && kk < d.nz) ? sh.fd[d.idx(ii, jj, kk)] : T(0);
      };
      auto pair = [&](int jj, int kk) -> T { return r_at(2 * ic, jj, kk) + r_at(2 * ic + 1, jj, kk); };
      if (cm == 7)
        v = (pair(2 * jc, 2 * kc) + pair(2 * jc + 1, 2 * kc)) + (pair(2 * jc, 2 * kc + 1) + pair(2 * jc + 1, 2 * kc + 1));
      else
        v = pair(jc, 2 * kc) + pair(jc, 2 * kc + 1);
    } else {
      // resid_restrict_cell: children in (k, j, i) order, rows only
      const int i0 = (cm & 1) ? 2 * ic : ic, i1 = (cm & 1) ? min(2 * ic + 1, d.nx - 1) : ic;
      const int j0 = (cm & 2) ? 2 * jc : jc, j1 = (cm & 2) ? min(2 * jc + 1, d.ny - 1) : jc;
      const int kb = (cm & 4) ? 2 * kc : kc, k1 = (cm & 4) ? min(2 * kc + 1, d.nz - 1) : kc;
      for (int kk = kb; kk <= k1; kk++)
        for (int jj = j0; jj <= j1; jj++)
          for (int ii = i0; ii <= i1; ii++) {
            const size_t cf = d.idx(ii, jj, kk);
            if (L.code[cf] & ROW_BIT) v += sh.fd[cf];
          }
    }
    sh.fc[q] = (sh.clsc[q] & CLS_NONFLUID) ? T(0) : v;
  }
  __syncthreads();

  // dense solve (dense_rows): 16 waves = 4 row blocks of 64 rows x 4 column quarters
  {
    const int lane = t & 63, w = t >> 6, rb = w >> 2, wq = w & 3;
    for (int r0 = 0; r0 < nc; r0 += 256) {
      const int r = r0 + rb * 64 + lane;
      T a0 = T(0), a1 = T(0);
      if (r < nc && cinv_lds) {
        int cc = wq;
        for (; cc + 4 < nc; cc += 8) {
          a0 += sh.cinv[cc * nc + r] * sh.fc[cc];
          a1 += sh.cinv[(cc + 4) * nc + r] * sh.fc[cc + 4];
        }
        for (; cc < nc; cc += 4) a0 += sh.cinv[cc * nc + r] * sh.fc[cc];
      } else if (r < nc) {
        int cc = wq;
        for (; cc + 4 < nc; cc += 8) {
          a0 += A.cinv[(size_t)cc * nc + r] * sh.fc[cc];
          a1 += A.cinv[(size_t)(cc + 4) * nc + r] * sh.fc[cc + 4];
        }
        for (; cc < nc; cc += 4) a0 += A.cinv[(size_t)cc * nc + r] * sh.fc[cc];
      }
      sh.part[rb][wq][lane] = a0 + a1;
      __syncthreads();
      if (wq == 0 && r < nc)
        sh.xc[r] = (sh.part[rb][0][lane] + sh.part[rb][1][lane]) + (sh.part[rb][2][lane] + sh.part[rb][3][lane]);
      __syncthreads();
    }
  }

  // prolongation into cur (prolong_add_cell)
  if (mine) {
    int ic, jc, kc;
    const int ia = axis_pick(i, d.nx, d.per[0], cm & 1, ic);
    const int ja = axis_pick(j, d.ny, d.per[1], cm & 2, jc);
    const int ka = axis_pick(k, d.nz, d.per[2], cm & 4, kc);
    const int in = wrapc(ic + ia, dc.nx), jn = wrapc(jc + ja, dc.ny), kn = wrapc(kc + ka, dc.nz);
    const uint8_t mk = sh.clsc[dc.idx(ic, jc, kn)], mj = sh.clsc[dc.idx(ic, jn, kc)],
                  mi = sh.clsc[dc.idx(in, jc, kc)];
    const T x000 = sh.xc[dc.idx(ic, jc, kc)], x010 = sh.xc[dc.idx(ic, jn, kc)];
    const T x100 = sh.xc[dc.idx(in, jc, kc)], x110 = sh.xc[dc.idx(in, jn, kc)];
    const T x001 = sh.xc[dc.idx(ic, jc, kn)], x011 = sh.xc[dc.idx(ic, jn, kn)];
    const T x101 = sh.xc[dc.idx(in, jc, kn)], x111 = sh.xc[dc.idx(in, jn, kn)];
    const T xo = cur[c];
    const bool ui = ia != 0 && !(mi & CLS_NONFLUID);
    const bool uj = ja != 0 && !(mj & CLS_NONFLUID);
    const bool uk = ka != 0 && !(mk & CLS_NONFLUID);
    const T y0 = uj ? x010 : x000, y1 = uj ? x110 : x100;
    const T z0 = uj ? x011 : x001, z1 = uj ? x111 : x101;
    const T a_c = x000, a_j = y0, a_i = ui ? x100 : x000, a_ij = ui ? y1 : y0;
    const T b_c = uk ? x001 : a_c, b_j = uk ? z0 : a_j;
    const T b_i = uk ? (ui ? x101 : x001) : a_i, b_ij = uk ? (ui ? z1 : z0) : a_ij;
    const T v = (a_c * T(9) + a_j * T(3) + a_i * T(3) + a_ij) * T(3. / 64.) +
                (b_c * T(9) + b_j * T(3) + b_i * T(3) + b_ij) * T(1. / 64.);
    if (!(cl & CLS_NONFLUID)) cur[c] = xo + v;
  }
  __syncthreads();

  // post-smoothing; the last step goes to the level's x in memory
  for (int s = 0; s < A.degree; s++) {
    const T c1 = s == 0 ? T(0) : sh.c1[s];
    const T xn = mine ? tail_step(sh, d, i, j, k, code, interior, scaled, gs, fv, cur, prev, c, c1, sh.c2[s]) : T(0);
    if (s + 1 == A.degree) {
      if (mine) A.xout[c] = xn;
    } else {
      if (mine) prev[c] = xn;
      T* tmp = cur;
      cur = prev;
      prev = tmp;
      __syncthreads();
    }
  }
}

// HALF: the level's coefficients are FP16 (gh), else T (g)
template <typename T, bool HALF>
__global__ __launch_bounds__(1024) void k_coarse_tail(TailArgs<T> A) {
  if constexpr (HALF) {
    __shared__ TailLds<T, _Float16> sh;
    coarse_tail_body(A, A.L.gh, true, sh);
  } else {
    __shared__ TailLds<T, T> sh;
    coarse_tail_body(A, A.L.g, false, sh);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_normsq(size_t n, const T* __restrict__ a, double* __restrict__ partial,
                                                int nblk) {
  __shared__ double lds[4];
  double s = 0.;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)nblk * 256) {
    const double v = (double)a[t];
    s += v * v;
  }
  s = block_sum(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

template <typename T>
__global__ __launch_bounds__(256) void k_scale_by(size_t n, T* __restrict__ a, const double* __restrict__ nrm2,
                                                  const T* __restrict__ src) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const double s = nrm2[0] > 0. ? 1. / sqrt(nrm2[0]) : 0.;
  a[t] = (T)((double)(src ? src[t] : a[t]) * s);
}

// ---------------------------------------------------------- reference primitives
__device__ inline bool dmda_bnd(const Dmda& m, int i, int j, int k) {
  return i == 0 || j == 0 || k == 0 || i == m.mx - 1 || j == m.my - 1 || k == m.mz - 1;
}

__global__ __launch_bounds__(256) void k_ref_restrict(Dmda f, Dmda c, const double* __restrict__ nvf,
                                                      const double* __restrict__ nvc, const double* __restrict__ x,
                                                      double* __restrict__ fc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= c.mx || j >= c.my) return;
  const size_t p = c.idx(i, j, k);
  if (dmda_bnd(c, i, j, k)) { fc[p] = 0.; return; }
  const int ih = 2 * i, jh = 2 * j, kh = 2 * k;
  auto W = [&](int a, int b, int cc) {
    const size_t q = f.idx(a, b, cc);
    return x[q] * fmax(0., 1 - nvf[q]);
  };
  double v = 0.125 * (W(ih, jh, kh) + W(ih - 1, jh, kh) + W(ih, jh - 1, kh) + W(ih, jh, kh - 1) +
                      W(ih - 1, jh - 1, kh) + W(ih, jh - 1, kh - 1) + W(ih - 1, jh, kh - 1) +
                      W(ih - 1, jh - 1, kh - 1));
  if (nvc[p] > 0.1) v = 0.;
  fc[p] = v;
}

__global__ __launch_bounds__(256) void k_ref_prolong(Dmda f, Dmda c, const double* __restrict__ nvf,
                                                     const double* __restrict__ nvc, const double* __restrict__ x,
                                                     double* __restrict__ ff) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= f.mx || j >= f.my) return;
  const size_t p = f.idx(i, j, k);
  if (dmda_bnd(f, i, j, k) || nvf[p] > 0.1) { ff[p] = 0.; return; }
  const int ic = (i + 1) / 2, jc = (j + 1) / 2, kc = (k + 1) / 2;
  int ia = (i - 2 * ic) == 0 ? 1 : -1;
  int ja = (j - 2 * jc) == 0 ? 1 : -1;
  int ka = (k - 2 * kc) == 0 ? 1 : -1;
  if (i == 1 || i == f.mx - 2) ia = 0;
  if (j == 1 || j == f.my - 2) ja = 0;
  if (k == 1 || k == f.mz - 2) ka = 0;
  if (ka == -1 && nvc[c.idx(ic, jc, kc - 1)] > 0.1) ka = 0;
  else if (ka == 1 && nvc[c.idx(ic, jc, kc + 1)] > 0.1) ka = 0;
  if (ja == -1 && nvc[c.idx(ic, jc - 1, kc)] > 0.1) ja = 0;
  else if (ja == 1 && nvc[c.idx(ic, jc + 1, kc)] > 0.1) ja = 0;
  if (ia == -1 && nvc[c.idx(ic - 1, jc, kc)] > 0.1) ia = 0;
  else if (ia == 1 && nvc[c.idx(ic + 1, jc, kc)] > 0.1) ia = 0;
  auto X = [&](int a, int b, int cc) { return x[c.idx(a, b, cc)]; };
  ff[p] = (X(ic, jc, kc) * 9 + X(ic, jc + ja, kc) * 3 + X(ic + ia, jc, kc) * 3 + X(ic + ia, jc + ja, kc)) * 3. / 64. +
          (X(ic, jc, kc + ka) * 9 + X(ic, jc + ja, kc + ka) * 3 + X(ic + ia, jc, kc + ka) * 3 +
           X(ic + ia, jc + ja, kc + ka)) / 64.;
}

__global__ __launch_bounds__(256) void k_ref_nvert_product(Dmda f, Dmda c, const double* __restrict__ nvf,
                                                           double* __restrict__ nvc) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= c.mx || j >= c.my) return;
  const size_t p = c.idx(i, j, k);
  double v = 0.;
  if (!dmda_bnd(c, i, j, k)) {
    const int ih = 2 * i, jh = 2 * j, kh = 2 * k;
    auto V = [&](int a, int b, int cc) { return nvf[f.idx(a, b, cc)]; };
    if (V(ih, jh, kh) * V(ih - 1, jh, kh) * V(ih, jh - 1, kh) * V(ih, jh, kh - 1) * V(ih - 1, jh - 1, kh) *
            V(ih - 1, jh, kh - 1) * V(ih, jh - 1, kh - 1) * V(ih - 1, jh - 1, kh - 1) >
        0.1)
      v = 1.;  // PetscMax(1., 0.)
  }
  nvc[p] = v;
}

// isolated coarse fluid cells -> 1 (poisson.c:2083-2097).  A cell only flips when
// all six neighbours are already 1, so no decision depends on another flip and
// the parallel sweep equals the reference's sequential one.
__global__ __launch_bounds__(256) void k_ref_nvert_isolated(Dmda c, const double* __restrict__ pre,
                                                            double* __restrict__ nvc) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= c.mx || j >= c.my) return;
  if (dmda_bnd(c, i, j, k)) return;
  const size_t p = c.idx(i, j, k), sy = c.mx, sz = (size_t)c.mx * c.my;
  if (pre[p] < 0.1 && pre[p + 1] + pre[p - 1] > 1.1 && pre[p + sy] + pre[p - sy] > 1.1 &&
      pre[p + sz] + pre[p - sz] > 1.1)
    nvc[p] = 1.;
}

inline unsigned nb1(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

template <typename T>
void launch_smooth(hipStream_t s, LevelView<T> L, const T* lfac, int mode, const T* f, const T* dinv,
                   const T* x_old, T* x_prev, T* x_new, T* dv, T* scr, T c1, T c2, T c0, const T* fd) {
  if (lfac) {
    // y-line smoothing: residual over all cells, then one thread per line
    const size_t n = L.d.n();
    if (mode == SM_FIRST2) throw StatusError(1, "the line smoother has no fused first-two-steps mode");
    // lines of up to 256 rows are split over the waves of a block
    const bool seg = line_seg_fits<T>(L.d);
    const dim3 sgrid((L.d.nx + line_lanes<T>() - 1) / line_lanes<T>(), L.d.nz), sblk(64, 16);
    if (mode == SM_ZERO) {
      if (seg) k_line_seg<T, true><<<sgrid, sblk, 0, s>>>(L.d, f, lfac, x_old, x_new, dv, c1, c2);
      else k_line_solve<T, true><<<line_grid(L.d), kBlk, 0, s>>>(L.d, f, lfac, scr + n, x_old, x_new, dv, c1, c2);
    } else if (seg) {
      k_line_fused<T><<<sgrid, sblk, 0, s>>>(L, f, lfac, x_old, x_new, dv, c1, c2);
    } else {
      k_resid<T><<<cell_grid(L.d), kBlk, 0, s>>>(L, f, x_old, scr);
      HIPCHECK(hipGetLastError());
      k_line_solve<T, false><<<line_grid(L.d), kBlk, 0, s>>>(L.d, scr, lfac, scr + n, x_old, x_new, dv, c1, c2);
    }
  } else {
    switch (mode) {
      case SM_ZERO:
        k_smooth_point<T, SM_ZERO><<<cell_grid(L.d), kBlk, 0, s>>>(L, f, dinv, x_old, x_prev, x_new, dv, c1, c2, c0, fd);
        break;
      case SM_FIRST2:
        k_smooth_point<T, SM_FIRST2>
            <<<cell_grid(L.d), kBlk, 0, s>>>(L, f, dinv, x_old, x_prev, x_new, dv, c1, c2, c0, fd);
        break;
      case SM_POWER:
        k_smooth_point<T, SM_POWER><<<cell_grid(L.d), kBlk, 0, s>>>(L, f, dinv, x_old, x_prev, x_new, dv, c1, c2, c0, fd);
        break;
      default:
        k_smooth_point<T, SM_NEXT><<<cell_grid(L.d), kBlk, 0, s>>>(L, f, dinv, x_old, x_prev, x_new, dv, c1, c2, c0, fd);
    }
  }
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_stream_floor(hipStream_t s, LevelView<T> L, const T* f, const T* x, const T* x_prev, T* out) {
  const size_t n = L.d.n();
  if constexpr (std::is_same<T, float>::value) {
    if (L.gh) {
      k_stream_floor<T, true><<<nb1(n), 256, 0, s>>>(n, L.gh, L.cls, f, x, x_prev, out);
      HIPCHECK(hipGetLastError());
      return;
    }
  }
  k_stream_floor<T, false><<<nb1(n), 256, 0, s>>>(n, L.g, L.cls, f, x, x_prev, out);
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_line_factors(hipStream_t s, LevelView<T> L, T* lfac) {
  // every row's (a, b, c) in parallel, then the recurrence along each line
  k_line_rows<T><<<cell_grid(L.d), kBlk, 0, s>>>(L, lfac);
  HIPCHECK(hipGetLastError());
  k_line_factors<T><<<line_grid(L.d), kBlk, 0, s>>>(L.d, lfac);
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_dinv(hipStream_t s, LevelView<T> L, T* dinv) {
  k_dinv<T><<<cell_grid(L.d), kBlk, 0, s>>>(L, dinv);
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_resid_restrict(hipStream_t s, LevelView<T> F, const T* f, const T* x, Dims dc, const uint8_t* cls_c,
                           T* fc) {
  const int cm = (dc.nx != F.d.nx ? 1 : 0) | (dc.ny != F.d.ny ? 2 : 0) | (dc.nz != F.d.nz ? 4 : 0);
  if (cm == 7 || cm == 5) {  // 2x2x2 children, or 2x1x2 (i and k coarsened, j kept)
    const dim3 gr((F.d.nx + 63) / 64, (F.d.ny + 1) / 2, (F.d.nz + 1) / 2);
    k_resid_restrict_blk<T><<<gr, dim3(64, 2, 2), 0, s>>>(F, f, x, dc, cls_c, fc, cm == 5 ? 1 : 0);
  } else {
    k_resid_restrict<T><<<cell_grid(dc), kBlk, 0, s>>>(F, f, x, dc, cls_c, cm, fc);
  }
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_prolong_add(hipStream_t s, LevelView<T> F, Dims dc, const uint8_t* cls_c, const T* xc, T* xf) {
  const int cm = (dc.nx != F.d.nx ? 1 : 0) | (dc.ny != F.d.ny ? 2 : 0) | (dc.nz != F.d.nz ? 4 : 0);
  k_prolong_add<T><<<cell_grid(F.d), kBlk, 0, s>>>(F, dc, cls_c, cm, xc, xf);
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_dense_solve(hipStream_t s, int n, const T* cinv, const T* f, T* x) {
  if (n > 1024) throw StatusError(1, "coarsest level larger than 1024 cells");
  k_dense_solve<T><<<(n + 63) / 64, 256, 0, s>>>(n, cinv, f, x);
  HIPCHECK(hipGetLastError());
}

std::vector<uint16_t> tail_order(size_t n, const uint32_t* code) {
  if (n > (size_t)TAIL_MAX_CELLS) throw StatusError(1, "coarse tail: level larger than 1024 cells");
  // cells in two groups, each in index order: interior code words, then the rest
  std::vector<uint16_t> cells;
  cells.reserve(n);
  for (int pass = 0; pass < 2; pass++)
    for (size_t c = 0; c < n; c++)
      if (((code[c] & ~DIVOK_BIT) == CODE_INTERIOR) == (pass == 0)) cells.push_back((uint16_t)c);
  size_t ni = 0;
  while (ni < n && (code[cells[ni]] & ~DIVOK_BIT) == CODE_INTERIOR) ni++;
  // 64-cell waves, an all-interior one costed at 3 and any other at 5 (their
  // instruction counts relative to each other); the counts of each kind per SIMD
  // (at most four waves each) minimise the largest SIMD cost
  const int nw = (int)((n + 63) / 64), nf = (int)(ni / 64), ns = nw - nf;
  int pairs[15][2], np = 0;  // (slow, fast) waves one SIMD can hold
  for (int sl = 0; sl <= 4; sl++)
    for (int fa = 0; sl + fa <= 4; fa++) {
      pairs[np][0] = sl;
      pairs[np++][1] = fa;
    }
  int best_cost = 1 << 30, best[4] = {0, 0, 0, 0};
  for (int p0 = 0; p0 < np; p0++)
    for (int p1 = 0; p1 < np; p1++)
      for (int p2 = 0; p2 < np; p2++)
        for (int p3 = 0; p3 < np; p3++) {
          const int pk[4] = {p0, p1, p2, p3};
          int cs = 0, cf = 0, worst = 0;
          for (int q = 0; q < 4; q++) {
            cs += pairs[pk[q]][0];
            cf += pairs[pk[q]][1];
            worst = std::max(worst, 5 * pairs[pk[q]][0] + 3 * pairs[pk[q]][1]);
          }
          if (cs == ns && cf == nf && worst < best_cost) {
            best_cost = worst;
            for (int q = 0; q < 4; q++) best[q] = pk[q];
          }
        }
  // wave slots q, q + 4, q + 8, q + 12 run on SIMD q: the interior waves (the first
  // nf) take the fast places, the rest the slow ones
  std::vector<int> slot_of(nw, -1);
  int wf = 0, wsl = nf;
  for (int q = 0; q < 4; q++) {
    int u = 0;
    for (int e = 0; e < pairs[best[q]][1]; e++) slot_of[wf++] = q + 4 * u++;
    for (int e = 0; e < pairs[best[q]][0]; e++) slot_of[wsl++] = q + 4 * u++;
  }
  std::vector<uint16_t> order(1024, TAIL_NO_CELL);
  for (size_t p = 0; p < n; p++) order[(size_t)slot_of[p / 64] * 64 + p % 64] = cells[p];
  return order;
}

template <typename T>
void launch_coarse_tail(hipStream_t s, const TailArgs<T>& a) {
  if (a.L.d.n() > (size_t)TAIL_MAX_CELLS || a.dc.n() > (size_t)TAIL_MAX_CELLS || a.degree < 1 ||
      a.degree > TAIL_MAX_DEGREE)
    throw StatusError(1, "coarse tail: level or degree out of range");
  if constexpr (std::is_same<T, float>::value) {
    if (a.L.gh) k_coarse_tail<T, true><<<1, 1024, 0, s>>>(a);
    else k_coarse_tail<T, false><<<1, 1024, 0, s>>>(a);
  } else {
    // an FP64 step needs more than the 128 registers a 1024-thread workgroup allows
    throw StatusError(1, "coarse tail: FP32 levels only");
  }
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_normsq(hipStream_t s, size_t n, const T* a, double* partial, int* nblk) {
  const int nb = partial_blocks(n);
  *nblk = nb;
  k_normsq<T><<<nb, 256, 0, s>>>(n, a, partial, nb);
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_scale_by(hipStream_t s, size_t n, T* a, const double* norm2, const T* src) {
  k_scale_by<T><<<nb1(n), 256, 0, s>>>(n, a, norm2, src);
  HIPCHECK(hipGetLastError());
}

#define INST(T)                                                                                                     \
  template void launch_smooth<T>(hipStream_t, LevelView<T>, const T*, int, const T*, const T*, const T*, T*, T*, T*, \
                                 T*, T, T, T, const T*);                                                          \
  template void launch_line_factors<T>(hipStream_t, LevelView<T>, T*);                                            \
  template void launch_stream_floor<T>(hipStream_t, LevelView<T>, const T*, const T*, const T*, T*);             \
  template void launch_dinv<T>(hipStream_t, LevelView<T>, T*);                                                     \
  template void launch_resid_restrict<T>(hipStream_t, LevelView<T>, const T*, const T*, Dims, const uint8_t*, T*); \
  template void launch_prolong_add<T>(hipStream_t, LevelView<T>, Dims, const uint8_t*, const T*, T*);              \
  template void launch_dense_solve<T>(hipStream_t, int, const T*, const T*, T*);                                   \
  template void launch_coarse_tail<T>(hipStream_t, const TailArgs<T>&);                                            \
  template void launch_normsq<T>(hipStream_t, size_t, const T*, double*, int*);                                    \
  template void launch_scale_by<T>(hipStream_t, size_t, T*, const double*, const T*);
INST(float)
INST(double)
#undef INST

void launch_ref_restrict(hipStream_t s, Dmda f, Dmda c, const double* nvf, const double* nvc, const double* xf,
                         double* fc) {
  k_ref_restrict<<<dim3((c.mx + 63) / 64, (c.my + 3) / 4, c.mz), kBlk, 0, s>>>(f, c, nvf, nvc, xf, fc);
  HIPCHECK(hipGetLastError());
}

void launch_ref_prolong(hipStream_t s, Dmda f, Dmda c, const double* nvf, const double* nvc, const double* xc,
                        double* ff) {
  k_ref_prolong<<<dim3((f.mx + 63) / 64, (f.my + 3) / 4, f.mz), kBlk, 0, s>>>(f, c, nvf, nvc, xc, ff);
  HIPCHECK(hipGetLastError());
}

void launch_ref_nvert_restrict(hipStream_t s, Dmda f, Dmda c, const double* nvf, double* nvc, double* tmp) {
  const dim3 gr((c.mx + 63) / 64, (c.my + 3) / 4, c.mz);
  k_ref_nvert_product<<<gr, kBlk, 0, s>>>(f, c, nvf, tmp);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(nvc, tmp, c.n() * sizeof(double), hipMemcpyDeviceToDevice, s));
  k_ref_nvert_isolated<<<gr, kBlk, 0, s>>>(c, tmp, nvc);
  HIPCHECK(hipGetLastError());
}

}  // namespace pp

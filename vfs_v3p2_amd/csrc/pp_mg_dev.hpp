// Device bodies of the multigrid kernels: Chebyshev smoothers, fused residual +
// restriction, IB-masked prolongation and the dense coarsest solve.  Each body
// takes explicit cell (or tile) coordinates, so the single-case kernels
// (pp_mg.hip) and the batched persistent kernels (pp_batch.hip) share the
// arithmetic.
//
// The transfers follow the reference's MG primitives (MyRestriction
// poisson.c:364-471, MyInterpolation poisson.c:473-669): restriction sums the
// fluid children of a coarse cell (the coarse operator is in the same integrated
// flux form as the fine one, so coarse residuals add up), zero on coarse IB
// cells; prolongation uses the (27,9,9,3,9,3,3,1)/64 weights, one-sided next to
// coarse IB cells and at non-periodic walls, periodic axes wrap.
#pragma once

#include <type_traits>

#include "pp_kernels.hpp"
#include "pp_stencil.hpp"

namespace pp {

// (A x)_c and the diagonal on a level, from its coefficients in either storage
template <typename T>
__device__ __forceinline__ void level_flux(const LevelView<T>& L, uint32_t cd, size_t c, const Nbr& nb,
                                           const T (&xv)[19], T& ax, T& diag) {
  if constexpr (std::is_same<T, float>::value) {
    if (L.gh) {
      flux_row(L.gh, L.d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, ax, diag);
      ax *= L.gs;
      diag *= L.gs;
      return;
    }
  }
  flux_row(L.g, L.d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, ax, diag);
}

// D^-1 of the level operator, 0 on cells that carry no unknown.
template <typename T>
__device__ __forceinline__ void dinv_cell(const LevelView<T>& L, T* __restrict__ dinv, int i, int j, int k) {
  const size_t c = L.d.idx(i, j, k);
  const uint32_t cd = L.code[c];
  Nbr nb;
  neighbours(L.d, i, j, k, nb);
  T xv[19];
#pragma unroll
  for (int s = 0; s < 19; s++) xv[s] = T(0);
  T ax, diag;
  level_flux(L, cd, c, nb, xv, ax, diag);
  dinv[c] = ((cd & ROW_BIT) && diag != T(0)) ? T(1) / diag : T(0);
}

// (A x)_c, the diagonal and whether the cell has a row, with the level's class
// bytes picking the row form: a wave whose cells are all interior (class flag)
// takes the compile-time-weight form and reads no code word.  The coefficient
// loads are issued before the class test, together with the caller's stencil
// loads.
template <typename T, typename TG>
__device__ __forceinline__ void flux_by_class(const TG* __restrict__ g, size_t n, const uint32_t* __restrict__ code,
                                              size_t c, const Nbr& nb, const T (&xv)[19], bool interior, T& ax,
                                              T& diag, bool& row) {
  RowCoef<TG> rc;
  load_coef(g, n, (uint32_t)c, nb.s[WP], nb.s[SP], nb.s[BP], rc);
  if (__all(interior)) {
    flux_row_v<X_CENTRAL>(rc, CODE_INTERIOR, xv, ax, diag);
  } else {
    const uint32_t cd = code[c];
    flux_row_v<-1>(rc, cd, xv, ax, diag);
    row = (cd & ROW_BIT) != 0;
  }
}

// Chebyshev smoothing preconditioned by the diagonal, in three-term form: one step is
//   z = D^-1 (f - A x);  x_new = x + c1 (x - x_prev) + c2 z
// (the direction d = x - x_prev is not stored; c1 == 0 starts a new sequence and
// does not read x_prev).  x_new may be x_prev's buffer.  Modes:
//   SM_NEXT    one step from x_old;
//   SM_ZERO    the first step from x = 0:  x_new = c2 D^-1 f, x_prev = 0;
//   SM_FIRST2  the first two steps from x = 0 in one pass: the first step's
//              x1 = c0 D^-1 f is formed on the fly at the 19 neighbours from f and
//              D^-1 (or from fd = f D^-1 when the writer of f stored it), then one
//              step from x1 with (c1, c2); x_new = x2, x_prev = x1;
//   SM_POWER   dv = D^-1 (f - A x_old) only.
// The level's class bytes pick the row form: a wave whose cells are all interior
// takes the compile-time-weight form and reads no code word.  g holds the
// level's coefficients in their storage type (FP16 scaled by gs, or T); every
// load of the cell -- class, f, x_prev, the 19 stencil values, the 18
// coefficients -- is issued before any of them is used.
template <typename T, int MODE, typename TG>
__device__ __forceinline__ void smooth_point_body(const LevelView<T>& L, const TG* __restrict__ g, bool scaled,
                                                  const T* __restrict__ f, const T* __restrict__ dinv,
                                                  const T* __restrict__ x_old, T* x_prev, T* x_new,
                                                  T* __restrict__ dv, T c1, T c2, T c0, int i, int j, int k,
                                                  const T* __restrict__ fd) {
  const size_t c = L.d.idx(i, j, k);
  const uint8_t cl = L.cls[c];
  const T fc = f[c];
  T xp = T(0);
  if constexpr (MODE == SM_NEXT) {
    if (c1 != T(0)) xp = x_prev[c];
  }
  Nbr nb;
  neighbours(L.d, i, j, k, nb);
  T xv[19];
  if constexpr (MODE == SM_FIRST2) {
    if (fd) {
#pragma unroll
      for (int s = 0; s < 19; s++) xv[s] = c0 * fd[nb.s[s]];
    } else {
#pragma unroll
      for (int s = 0; s < 19; s++) xv[s] = c0 * (f[nb.s[s]] * dinv[nb.s[s]]);
    }
  } else {
    load_x(nb, x_old, xv);
  }
  T ax, diag;
  bool row = true;
  flux_by_class(g, L.d.n(), L.code, c, nb, xv, (cl & CLS_INTERIOR) != 0, ax, diag, row);
  if (scaled) {
    ax *= L.gs;
    diag *= L.gs;
  }
  // straight-line division with a safe divisor: a branch around it would let the
  // compiler move the load of f into the branch, after the stencil
  const bool live = row && diag != T(0);
  const T z = live ? (fc - ax) / (live ? diag : T(1)) : T(0);
  if constexpr (MODE == SM_POWER) {
    dv[c] = z;
    return;
  }
  const T xc = xv[CP];
  const T dold = (MODE == SM_FIRST2) ? xc : (c1 == T(0) ? T(0) : xc - xp);
  if constexpr (MODE == SM_FIRST2) x_prev[c] = row ? xc : T(0);
  x_new[c] = row ? xc + (c1 * dold + c2 * z) : T(0);
}

template <typename T, int MODE>
__device__ __forceinline__ void smooth_point_cell(const LevelView<T>& L, const T* __restrict__ f,
                                                  const T* __restrict__ dinv, const T* __restrict__ x_old,
                                                  T* x_prev, T* x_new, T* __restrict__ dv, T c1, T c2, T c0, int i,
                                                  int j, int k, const T* __restrict__ fd = nullptr) {
  if constexpr (MODE == SM_ZERO) {
    const size_t c = L.d.idx(i, j, k);
    x_new[c] = c2 * (f[c] * dinv[c]);
    x_prev[c] = T(0);
    return;
  } else {
    if constexpr (std::is_same<T, float>::value) {
      if (L.gh) {
        smooth_point_body<T, MODE>(L, L.gh, true, f, dinv, x_old, x_prev, x_new, dv, c1, c2, c0, i, j, k, fd);
        return;
      }
    }
    smooth_point_body<T, MODE>(L, L.g, false, f, dinv, x_old, x_prev, x_new, dv, c1, c2, c0, i, j, k, fd);
  }
}

// Thomas factors of the y-line tridiagonal part of A on the line (i, k), from the
// rows' SP / CP / NP coefficients (cells without an unknown get the identity):
// lfac = [a | cp | inv], each n long; den_j = b_j - a_j cp_{j-1}, inv_j = 1/den_j,
// cp_j = c_j inv_j.  First pass, one thread per cell: the row's (a, b, c) into
// lfac's three parts.
template <typename T>
__device__ __forceinline__ void line_rows_cell(const LevelView<T>& L, T* __restrict__ lfac, int i, int j, int k) {
  const size_t n = L.d.n();
  const size_t c = L.d.idx(i, j, k);
  const uint32_t cd = L.code[c];
  T a = T(0), b = T(1), cc = T(0);
  if (cd & ROW_BIT) {
    Nbr nb;
    neighbours(L.d, i, j, k, nb);
    T vol[19];
    if (L.gh) {
      build_row(L.gh, n, cd, c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
      a = vol[SP] * L.gs;
      b = vol[CP] * L.gs;
      cc = vol[NP] * L.gs;
    } else {
      build_row(L.g, n, cd, c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
      a = vol[SP];
      b = vol[CP];
      cc = vol[NP];
    }
  }
  lfac[c] = a;
  lfac[n + c] = b;
  lfac[2 * n + c] = cc;
}

// The factorisation along the line (i, k), in place over line_rows_cell's rows:
// den_j = b_j - a_j cp_{j-1}, inv_j = 1 / den_j (0 when den_j is 0), cp_j = c_j inv_j.
// The rows' loads do not depend on the recurrence, so eight rows are loaded at a time.
template <typename T>
__device__ __forceinline__ void line_factors_col(const Dims& d, T* __restrict__ lfac, int i, int k) {
  const size_t n = d.n();
  const size_t sy = (size_t)d.nx;
  const size_t c0 = d.idx(i, 0, k);
  T cprev = T(0);
  constexpr int LC = 8;
  for (int j0 = 0; j0 < d.ny; j0 += LC) {
    T b[LC], cc[LC], a[LC];
#pragma unroll
    for (int q = 0; q < LC; q++) {
      a[q] = b[q] = cc[q] = T(0);
      if (j0 + q < d.ny) {
        const size_t c = c0 + (size_t)(j0 + q) * sy;
        a[q] = lfac[c];
        b[q] = lfac[n + c];
        cc[q] = lfac[2 * n + c];
      }
    }
#pragma unroll
    for (int q = 0; q < LC; q++) {
      if (j0 + q < d.ny) {
        const size_t c = c0 + (size_t)(j0 + q) * sy;
        const T den = b[q] - a[q] * cprev;
        const T inv = den != T(0) ? T(1) / den : T(0);
        cprev = cc[q] * inv;
        lfac[n + c] = cprev;
        lfac[2 * n + c] = inv;
      }
    }
  }
}

// r = f - A x on cells with an unknown, 0 elsewhere.  The class byte picks the
// row form (interior waves read no code word); every load is issued before the
// vote.
template <typename T, typename TG>
__device__ __forceinline__ void resid_body(const LevelView<T>& L, const TG* __restrict__ g, bool scaled,
                                           const T* __restrict__ f, const T* __restrict__ x, T* __restrict__ r,
                                           int i, int j, int k) {
  const size_t c = L.d.idx(i, j, k);
  const uint8_t cl = L.cls[c];
  const T fc = f[c];
  Nbr nb;
  neighbours(L.d, i, j, k, nb);
  T xv[19], ax, dg;
  load_x(nb, x, xv);
  bool row = true;
  flux_by_class(g, L.d.n(), L.code, c, nb, xv, (cl & CLS_INTERIOR) != 0, ax, dg, row);
  if (scaled) ax *= L.gs;
  r[c] = row ? fc - ax : T(0);
}

template <typename T>
__device__ __forceinline__ void resid_cell(const LevelView<T>& L, const T* __restrict__ f, const T* __restrict__ x,
                                           T* __restrict__ r, int i, int j, int k) {
  if constexpr (std::is_same<T, float>::value) {
    if (L.gh) {
      resid_body(L, L.gh, true, f, x, r, i, j, k);
      return;
    }
  }
  resid_body(L, L.g, false, f, x, r, i, j, k);
}

// One Chebyshev step preconditioned by the y-line tridiagonal part T of A on the
// line (i, k): z = T^-1 r with the precomputed factors (forward sweep parked in
// tmp), d = c1 d + c2 z, x_new = x_old + d.  ZG: from x = 0 with r = f.
// The line is walked in chunks of LC rows whose inputs are loaded into registers
// first (independent loads in flight together), then the recurrence runs on them:
// the sequential sweep is otherwise bound by one load latency per row.
template <typename T, bool ZG>
__device__ __forceinline__ void line_solve_col(const Dims& d, const T* __restrict__ r, const T* __restrict__ lfac,
                                               T* __restrict__ tmp, const T* __restrict__ x_old,
                                               T* __restrict__ x_new, T* __restrict__ dv, T c1, T c2, int i,
                                               int k) {
  constexpr int LC = 16;
  const size_t n = d.n();
  const T* __restrict__ la = lfac;
  const T* __restrict__ lcp = lfac + n;
  const T* __restrict__ linv = lfac + 2 * n;
  const size_t sy = (size_t)d.nx;
  const size_t c0 = d.idx(i, 0, k);
  const int ny = d.ny;
  T dprev = T(0);
  for (int j0 = 0; j0 < ny; j0 += LC) {
    T rr[LC], aa[LC], vv[LC];
#pragma unroll
    for (int q = 0; q < LC; q++) {
      const int j = j0 + q < ny ? j0 + q : ny - 1;  // clamped rows are loaded but not used
      const size_t c = c0 + (size_t)j * sy;
      rr[q] = r[c];
      aa[q] = la[c];
      vv[q] = linv[c];
    }
#pragma unroll
    for (int q = 0; q < LC; q++) {
      if (j0 + q < ny) {
        dprev = (rr[q] - aa[q] * dprev) * vv[q];
        tmp[c0 + (size_t)(j0 + q) * sy] = dprev;
      }
    }
  }
  T znext = T(0);
  const bool rd = !ZG && c1 != T(0);
  for (int j1 = ny - 1; j1 >= 0; j1 -= LC) {
    T tt[LC], pp[LC], dd[LC], xo[LC];
#pragma unroll
    for (int q = 0; q < LC; q++) {
      const int j = j1 - q >= 0 ? j1 - q : 0;
      const size_t c = c0 + (size_t)j * sy;
      tt[q] = tmp[c];
      pp[q] = lcp[c];
      dd[q] = rd ? dv[c] : T(0);
      xo[q] = ZG ? T(0) : x_old[c];
    }
#pragma unroll
    for (int q = 0; q < LC; q++) {
      if (j1 - q >= 0) {
        const size_t c = c0 + (size_t)(j1 - q) * sy;
        const T z = tt[q] - pp[q] * znext;
        znext = z;
        const T dn = c1 * dd[q] + c2 * z;
        dv[c] = dn;
        x_new[c] = xo[q] + dn;
      }
    }
  }
}

// The same Chebyshev step with each line split into segments of a few rows held
// in registers.  A (64, 16) block of 1024 threads holds line_lanes<T>() lines:
// in FP32 64 lines of 16 segments of up to 16 rows (wave w takes segment w of
// every line), in FP64 32 lines of 32 segments of up to 8 rows (lanes 0-31 of
// wave w take segment 2w, lanes 32-63 segment 2w + 1), so lines of up to 256 rows
// fit either way within the registers a 1024-thread block allows.  Lines of up
// to 256 rows keep 16 or 32 times more lanes busy than one thread per line, and
// the forward sweep stays in registers instead of going through memory.  Both
// sweeps are affine recurrences, d_j = (r_j - a_j d_{j-1}) inv_j and
// z_j = d_j - cp_j z_{j+1}; each segment first runs its sweep from a zero carry,
// keeping the product of its multipliers, the carries into every segment follow
// from one pass over the segments of a line (LDS), and each segment then reruns
// the recurrence exactly from its carry.  Every thread of the block must call it
// (block barriers); lanes without a line pass live = false.
// (line_lanes, line_segs, line_seg_rows and line_seg_fits: pp_kernels.hpp)
// this thread's line within the block and its segment
template <typename T>
__device__ __forceinline__ int line_lane() {
  return (int)(threadIdx.x & (line_lanes<T>() - 1));
}
template <typename T>
__device__ __forceinline__ int line_seg() {
  return (int)(threadIdx.y * (64 / line_lanes<T>()) + threadIdx.x / line_lanes<T>());
}
template <typename T>
struct LineSegLds {
  T part[2][line_segs<T>()][line_lanes<T>()];  // a segment's zero-carry end value and multiplier product
  T carry[line_segs<T>()][line_lanes<T>()];    // carry into each segment
};

// The two segmented sweeps of one Chebyshev step: rr holds the segment's nr rows
// of r (rows j0 .. j0 + nr - 1 of the line whose row j0 is cell c0).
template <typename T, bool ZG>
__device__ __forceinline__ void line_seg_sweeps(const Dims& d, size_t c0, int nr,
                                                const T (&rr)[line_seg_rows<T>()], const T* __restrict__ lfac,
                                                const T* __restrict__ x_old, T* __restrict__ x_new,
                                                T* __restrict__ dv, T c1, T c2, LineSegLds<T>& sh) {
  constexpr int LM = line_seg_rows<T>();
  const int lane = line_lane<T>(), s = line_seg<T>();
  const size_t n = d.n();
  const size_t sy = (size_t)d.nx;
  const T* __restrict__ la = lfac;
  const T* __restrict__ lcp = lfac + n;
  const T* __restrict__ linv = lfac + 2 * n;
  T aa[LM], vv[LM];
#pragma unroll
  for (int q = 0; q < LM; q++) {
    aa[q] = vv[q] = T(0);
    if (q < nr) {
      const size_t c = c0 + (size_t)q * sy;
      aa[q] = la[c];
      vv[q] = linv[c];
    }
  }
  // forward sweep from a zero carry: end value and product of the multipliers
  T dl = T(0), prod = T(1);
#pragma unroll
  for (int q = 0; q < LM; q++)
    if (q < nr) {
      dl = (rr[q] - aa[q] * dl) * vv[q];
      prod = prod * -(aa[q] * vv[q]);
    }
  sh.part[0][s][lane] = dl;
  sh.part[1][s][lane] = prod;
  __syncthreads();
  if (s == 0) {
    T c = T(0);
    for (int t = 0; t < line_segs<T>(); t++) {
      sh.carry[t][lane] = c;
      c = sh.part[0][t][lane] + sh.part[1][t][lane] * c;
    }
  }
  __syncthreads();
  T dd[LM];
  T dprev = sh.carry[s][lane];
#pragma unroll
  for (int q = 0; q < LM; q++) {
    dd[q] = T(0);
    if (q < nr) {
      dprev = (rr[q] - aa[q] * dprev) * vv[q];
      dd[q] = dprev;
    }
  }
  // backward sweep
  const bool rd = !ZG && c1 != T(0);
  T pp[LM], dvo[LM], xo[LM];
#pragma unroll
  for (int q = 0; q < LM; q++) {
    pp[q] = dvo[q] = xo[q] = T(0);
    if (q < nr) {
      const size_t c = c0 + (size_t)q * sy;
      pp[q] = lcp[c];
      if (rd) dvo[q] = dv[c];
      if (!ZG) xo[q] = x_old[c];
    }
  }
  T zl = T(0);
  prod = T(1);
#pragma unroll
  for (int q = LM - 1; q >= 0; q--)
    if (q < nr) {
      zl = dd[q] - pp[q] * zl;
      prod = prod * -pp[q];
    }
  __syncthreads();  // every thread has read its forward carry
  sh.part[0][s][lane] = zl;
  sh.part[1][s][lane] = prod;
  __syncthreads();
  if (s == 0) {
    T c = T(0);
    for (int t = line_segs<T>() - 1; t >= 0; t--) {
      sh.carry[t][lane] = c;
      c = sh.part[0][t][lane] + sh.part[1][t][lane] * c;
    }
  }
  __syncthreads();
  T znext = sh.carry[s][lane];
#pragma unroll
  for (int q = LM - 1; q >= 0; q--)
    if (q < nr) {
      const size_t c = c0 + (size_t)q * sy;
      const T z = dd[q] - pp[q] * znext;
      znext = z;
      const T dn = c1 * dvo[q] + c2 * z;
      dv[c] = dn;
      x_new[c] = xo[q] + dn;
    }
  __syncthreads();  // the LDS is free for the block's next tile
}

// rows of this thread's segment: first row j0, count nr (0 for a lane without a line)
template <typename T>
__device__ __forceinline__ void line_seg_rows_of(const Dims& d, bool live, int& j0, int& nr) {
  const int L = (d.ny + line_segs<T>() - 1) / line_segs<T>();
  j0 = line_seg<T>() * L;
  nr = live ? max(0, min(L, d.ny - j0)) : 0;
}

template <typename T, bool ZG>
__device__ __forceinline__ void line_solve_seg(const Dims& d, bool live, int i, int k, const T* __restrict__ r,
                                               const T* __restrict__ lfac, const T* __restrict__ x_old,
                                               T* __restrict__ x_new, T* __restrict__ dv, T c1, T c2,
                                               LineSegLds<T>& sh) {
  constexpr int LM = line_seg_rows<T>();
  int j0, nr;
  line_seg_rows_of<T>(d, live, j0, nr);
  const size_t sy = (size_t)d.nx;
  const size_t c0 = live ? d.idx(i, j0 < d.ny ? j0 : 0, k) : 0;
  T rr[LM];
#pragma unroll
  for (int q = 0; q < LM; q++) rr[q] = q < nr ? r[c0 + (size_t)q * sy] : T(0);
  line_seg_sweeps<T, ZG>(d, c0, nr, rr, lfac, x_old, x_new, dv, c1, c2, sh);
}

// One Chebyshev step (from a non-zero x) with the y-line preconditioner and the
// residual r = f - A x computed in the same pass: each thread walks its segment's
// rows, keeping the stencil's values of x in the five columns (i, k), (i +- 1, k),
// (i, k +- 1) for the rows below, at and above the current one, so a row loads
// five values of x for the row above plus the four (i +- 1, k +- 1) values of its
// own row (9 instead of 19), and its lower j-face coefficients are the row below's
// upper ones.  The row's stencil values and coefficients are the ones the residual
// kernel reads, and its form is voted over the wave's 64 lines at that row.
template <typename T, typename TG>
__device__ __forceinline__ void line_step_fused(const LevelView<T>& L, const TG* __restrict__ g, bool scaled, bool live,
                                                int i, int k, const T* __restrict__ f, const T* __restrict__ lfac,
                                                const T* __restrict__ x_old, T* __restrict__ x_new,
                                                T* __restrict__ dv, T c1, T c2, LineSegLds<T>& sh) {
  constexpr int LM = line_seg_rows<T>();
  const Dims& d = L.d;
  int j0, nr;
  line_seg_rows_of<T>(d, live, j0, nr);
  const uint32_t nx = (uint32_t)d.nx, plane = (uint32_t)d.nx * (uint32_t)d.ny;
  const size_t n = d.n();
  const int ii = live ? i : 0, kk = live ? k : 0;
  const uint32_t ip = (uint32_t)wrap_p(ii, d.nx, d.per[0]), im = (uint32_t)wrap_m(ii, d.nx, d.per[0]);
  const uint32_t kp = (uint32_t)wrap_p(kk, d.nz, d.per[2]) * plane, km = (uint32_t)wrap_m(kk, d.nz, d.per[2]) * plane;
  const uint32_t ic = (uint32_t)ii, kc = (uint32_t)kk * plane;
  // column bases (cell index of row 0) of the marched columns C, E, W, T, B and of
  // the four corner columns read at the row itself
  const uint32_t col[5] = {ic + kc, ip + kc, im + kc, ic + kp, ic + km};
  const uint32_t cor[4] = {ip + kp, ip + km, im + kp, im + km};  // TE, BE, TW, BW
  const int jr = j0 < d.ny ? j0 : 0;
  const uint32_t rowm0 = (uint32_t)wrap_m(jr, d.ny, d.per[1]) * nx;
  T xm[5], x0[5];
#pragma unroll
  for (int a = 0; a < 5; a++) {
    xm[a] = nr > 0 ? (T)ld_at(x_old, col[a] + rowm0) : T(0);
    x0[a] = nr > 0 ? (T)ld_at(x_old, col[a] + (uint32_t)jr * nx) : T(0);
  }
  // upper j-face coefficients of the row below the segment
  TG gs_prev[3];
#pragma unroll
  for (int e = 0; e < 3; e++) gs_prev[e] = nr > 0 ? ld_at(g + (size_t)(3 + e) * n, ic + kc + rowm0) : TG(0);
  T rr[LM];
#pragma unroll
  for (int q = 0; q < LM; q++) {
    rr[q] = T(0);
    if (q < nr) {
      const int j = j0 + q;
      const uint32_t rj = (uint32_t)j * nx, rp = (uint32_t)wrap_p(j, d.ny, d.per[1]) * nx;
      const uint32_t c = ic + kc + rj;
      const uint8_t cl = L.cls[c];
      const T fc = f[c];
      T xp_[5];
#pragma unroll
      for (int a = 0; a < 5; a++) xp_[a] = (T)ld_at(x_old, col[a] + rp);
      T xv[19];
      xv[CP] = x0[0]; xv[NP] = xp_[0]; xv[SP] = xm[0];
      xv[EP] = x0[1]; xv[NE] = xp_[1]; xv[SE] = xm[1];
      xv[WP] = x0[2]; xv[NW] = xp_[2]; xv[SW] = xm[2];
      xv[TP] = x0[3]; xv[TN] = xp_[3]; xv[TS] = xm[3];
      xv[BP] = x0[4]; xv[BN] = xp_[4]; xv[BS] = xm[4];
      xv[TE] = (T)ld_at(x_old, cor[0] + rj);
      xv[BE] = (T)ld_at(x_old, cor[1] + rj);
      xv[TW] = (T)ld_at(x_old, cor[2] + rj);
      xv[BW] = (T)ld_at(x_old, cor[3] + rj);
      RowCoef<TG> rc;
#pragma unroll
      for (int e = 0; e < 3; e++) {
        rc.v[0][e] = ld_at(g + (size_t)e * n, c);
        rc.v[1][e] = ld_at(g + (size_t)e * n, im + kc + rj);
        rc.v[2][e] = ld_at(g + (size_t)(3 + e) * n, c);
        rc.v[3][e] = gs_prev[e];
        rc.v[4][e] = ld_at(g + (size_t)(6 + e) * n, c);
        rc.v[5][e] = ld_at(g + (size_t)(6 + e) * n, ic + km + rj);
      }
      T ax, dg;
      bool row = true;
      if (__all((cl & CLS_INTERIOR) != 0)) {
        flux_row_v<X_CENTRAL>(rc, CODE_INTERIOR, xv, ax, dg);
      } else {
        const uint32_t cd = L.code[c];
        flux_row_v<-1>(rc, cd, xv, ax, dg);
        row = (cd & ROW_BIT) != 0;
      }
      if (scaled) ax *= L.gs;
      rr[q] = row ? fc - ax : T(0);
#pragma unroll
      for (int a = 0; a < 5; a++) {
        xm[a] = x0[a];
        x0[a] = xp_[a];
      }
#pragma unroll
      for (int e = 0; e < 3; e++) gs_prev[e] = rc.v[2][e];
    }
  }
  const size_t c0 = nr > 0 ? d.idx(ii, j0, kk) : 0;
  line_seg_sweeps<T, false>(d, c0, nr, rr, lfac, x_old, x_new, dv, c1, c2, sh);
}

template <typename T>
__device__ __forceinline__ void line_step_cell(const LevelView<T>& L, bool live, int i, int k,
                                               const T* __restrict__ f, const T* __restrict__ lfac,
                                               const T* __restrict__ x_old, T* __restrict__ x_new,
                                               T* __restrict__ dv, T c1, T c2, LineSegLds<T>& sh) {
  if constexpr (std::is_same<T, float>::value) {
    if (L.gh) {
      line_step_fused(L, L.gh, true, live, i, k, f, lfac, x_old, x_new, dv, c1, c2, sh);
      return;
    }
  }
  line_step_fused(L, L.g, false, live, i, k, f, lfac, x_old, x_new, dv, c1, c2, sh);
}

// fc[C] = sum over fluid children of (f - A x); zero on coarse IB cells.  Any
// coarsening mask; one thread per coarse cell.
template <typename T>
__device__ __forceinline__ void resid_restrict_cell(const LevelView<T>& F, const T* __restrict__ f,
                                                    const T* __restrict__ x, const Dims& dc,
                                                    const uint8_t* __restrict__ clsc, int cmask, T* __restrict__ fc,
                                                    int ic, int jc, int kc) {
  const size_t C = dc.idx(ic, jc, kc);
  if (clsc[C] & CLS_NONFLUID) { fc[C] = T(0); return; }
  const int i0 = (cmask & 1) ? 2 * ic : ic, i1 = (cmask & 1) ? min(2 * ic + 1, F.d.nx - 1) : ic;
  const int j0 = (cmask & 2) ? 2 * jc : jc, j1 = (cmask & 2) ? min(2 * jc + 1, F.d.ny - 1) : jc;
  const int k0 = (cmask & 4) ? 2 * kc : kc, k1 = (cmask & 4) ? min(2 * kc + 1, F.d.nz - 1) : kc;
  T sum = T(0);
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        const size_t c = F.d.idx(i, j, k);
        const uint32_t cd = F.code[c];
        if (!(cd & ROW_BIT)) continue;
        Nbr nb;
        neighbours(F.d, i, j, k, nb);
        T xv[19], ax, dg;
        load_x(nb, x, xv);
        level_flux(F, cd, c, nb, xv, ax, dg);
        sum += f[c] - ax;
      }
  fc[C] = sum;
}

// fc[C] from the fine residual r (f - A x on the rows, as resid_cell writes it): the
// sum resid_restrict_cell forms, over the same children in the same (k, j, i) order,
// rows only, zero on coarse IB cells -- the same value to the last bit, with the
// children's residuals computed one thread per fine cell beforehand.
template <typename T>
__device__ __forceinline__ void restrict_sum_cell(const LevelView<T>& F, const T* __restrict__ r, const Dims& dc,
                                                  const uint8_t* __restrict__ clsc, int cmask, T* __restrict__ fc,
                                                  int ic, int jc, int kc) {
  const size_t C = dc.idx(ic, jc, kc);
  if (clsc[C] & CLS_NONFLUID) { fc[C] = T(0); return; }
  const int i0 = (cmask & 1) ? 2 * ic : ic, i1 = (cmask & 1) ? min(2 * ic + 1, F.d.nx - 1) : ic;
  const int j0 = (cmask & 2) ? 2 * jc : jc, j1 = (cmask & 2) ? min(2 * jc + 1, F.d.ny - 1) : jc;
  const int k0 = (cmask & 4) ? 2 * kc : kc, k1 = (cmask & 4) ? min(2 * kc + 1, F.d.nz - 1) : kc;
  T sum = T(0);
  for (int k = k0; k <= k1; k++)
    for (int j = j0; j <= j1; j++)
      for (int i = i0; i <= i1; i++) {
        const size_t c = F.d.idx(i, j, k);
        if (F.code[c] & ROW_BIT) sum += r[c];
      }
  fc[C] = sum;
}

// Residual + restriction when all three axes are coarsened, for the 64x2x2 fine
// tile (bx, by, bz) of a (64,2,2) block: one thread per fine cell, pair sums in i
// by lane shuffle, the j and k pairs through LDS; the tile covers coarse cells
// (32 bx .. 32 bx + 31, by, bz).  Every thread of the block must call it.
template <typename T>
__device__ __forceinline__ void resid_restrict_blk(const LevelView<T>& F, const T* __restrict__ f,
                                                   const T* __restrict__ x, const Dims& dc,
                                                   const uint8_t* __restrict__ clsc, T* __restrict__ fc, int bx,
                                                   int by, int bz, bool keep_y, T (*lds)[2][32]) {
  const int tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;
  const int i = bx * 64 + tx, j = by * 2 + ty, k = bz * 2 + tz;
  T r = T(0);
  if (i < F.d.nx && j < F.d.ny && k < F.d.nz) {
    const size_t c = F.d.idx(i, j, k);
    const uint32_t cd = F.code[c];
    Nbr nb;
    neighbours(F.d, i, j, k, nb);
    T xv[19], ax, dg;
    load_x(nb, x, xv);
    level_flux(F, cd, c, nb, xv, ax, dg);
    if (cd & ROW_BIT) r = f[c] - ax;
  }
  r += __shfl_xor(r, 1, 64);
  if ((tx & 1) == 0) lds[tz][ty][tx >> 1] = r;
  __syncthreads();
  if (keep_y) {
    // i and k coarsened only: the two j rows of the tile are two coarse cells
    if (tz == 0 && tx < 32) {
      const int ic = bx * 32 + tx, jc = by * 2 + ty, kc = bz;
      if (ic < dc.nx && jc < dc.ny && kc < dc.nz) {
        const size_t C = dc.idx(ic, jc, kc);
        const T v = lds[0][ty][tx] + lds[1][ty][tx];
        fc[C] = (clsc[C] & CLS_NONFLUID) ? T(0) : v;
      }
    }
  } else if (ty == 0 && tz == 0 && tx < 32) {
    const int ic = bx * 32 + tx, jc = by, kc = bz;
    if (ic < dc.nx && jc < dc.ny && kc < dc.nz) {
      const size_t C = dc.idx(ic, jc, kc);
      const T v = (lds[0][0][tx] + lds[0][1][tx]) + (lds[1][0][tx] + lds[1][1][tx]);
      fc[C] = (clsc[C] & CLS_NONFLUID) ? T(0) : v;
    }
  }
  __syncthreads();  // the LDS tile is reused by the block's next tile
}

__device__ __forceinline__ int axis_pick(int f, int n, int per, bool coarsened, int& ic) {
  if (!coarsened) { ic = f; return 0; }
  ic = f >> 1;
  int ia = (f & 1) ? 1 : -1;
  if (!per && (f == 0 || f == n - 1)) ia = 0;
  return ia;
}
__device__ __forceinline__ int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// xf += P xc at fine cell (i, j, k): trilinear (27,9,9,3,9,3,3,1)/64, one-sided
// next to coarse IB cells (poisson.c:500-505; the three neighbour tests are
// independent).  Every operand is loaded at once - the fine mask, the three
// coarse neighbours' masks and the coarse values at both choices on each axis
// (cached) - and the one-sided choice is a register select, so the cell costs
// one round trip to memory instead of three.
template <typename T>
__device__ __forceinline__ void prolong_add_cell(const LevelView<T>& F, const Dims& dc,
                                                 const uint8_t* __restrict__ clsc, int cmask,
                                                 const T* __restrict__ xc, T* __restrict__ xf, int i, int j, int k) {
  const size_t c = F.d.idx(i, j, k);
  int ic, jc, kc;
  const int ia = axis_pick(i, F.d.nx, F.d.per[0], cmask & 1, ic);
  const int ja = axis_pick(j, F.d.ny, F.d.per[1], cmask & 2, jc);
  const int ka = axis_pick(k, F.d.nz, F.d.per[2], cmask & 4, kc);
  const int in = wrapc(ic + ia, dc.nx), jn = wrapc(jc + ja, dc.ny), kn = wrapc(kc + ka, dc.nz);
  const uint8_t fm = F.cls[c];
  const uint8_t mk = clsc[dc.idx(ic, jc, kn)], mj = clsc[dc.idx(ic, jn, kc)], mi = clsc[dc.idx(in, jc, kc)];
  const T x000 = xc[dc.idx(ic, jc, kc)], x010 = xc[dc.idx(ic, jn, kc)];
  const T x100 = xc[dc.idx(in, jc, kc)], x110 = xc[dc.idx(in, jn, kc)];
  const T x001 = xc[dc.idx(ic, jc, kn)], x011 = xc[dc.idx(ic, jn, kn)];
  const T x101 = xc[dc.idx(in, jc, kn)], x111 = xc[dc.idx(in, jn, kn)];
  const T xo = xf[c];
  // neighbour used on each axis: the other coarse cell unless the axis is not
  // coarsened, the cell sits on a non-periodic wall, or that coarse cell is IB
  const bool ui = ia != 0 && !(mi & CLS_NONFLUID);
  const bool uj = ja != 0 && !(mj & CLS_NONFLUID);
  const bool uk = ka != 0 && !(mk & CLS_NONFLUID);
  const T y0 = uj ? x010 : x000, y1 = uj ? x110 : x100;  // (ic, j2, kc), (in, j2, kc)
  const T z0 = uj ? x011 : x001, z1 = uj ? x111 : x101;  // (ic, j2, kn), (in, j2, kn)
  const T a_c = x000, a_j = y0, a_i = ui ? x100 : x000, a_ij = ui ? y1 : y0;
  const T b_c = uk ? x001 : a_c, b_j = uk ? z0 : a_j;
  const T b_i = uk ? (ui ? x101 : x001) : a_i, b_ij = uk ? (ui ? z1 : z0) : a_ij;
  const T v = (a_c * T(9) + a_j * T(3) + a_i * T(3) + a_ij) * T(3. / 64.) +
              (b_c * T(9) + b_j * T(3) + b_i * T(3) + b_ij) * T(1. / 64.);
  if (!(fm & CLS_NONFLUID)) xf[c] = xo + v;
}

// x = Cinv f for rows 64 rb .. 64 rb + 63, Cinv column-major n x n (n <= 1024):
// four waves take interleaved quarters of the columns, partial sums meet in
// LDS.  256-thread block; every thread must call it.
template <typename T>
__device__ __forceinline__ void dense_rows(int n, const T* __restrict__ cinv, const T* __restrict__ f,
                                           T* __restrict__ x, int rb, T* fs, T (*part)[64]) {
  for (int t = threadIdx.x; t < n && t < 1024; t += 256) fs[t] = f[t];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = rb * 64 + lane;
  T a0 = T(0), a1 = T(0);
  if (r < n) {
    int c = w;
    for (; c + 4 < n; c += 8) {
      a0 += cinv[(size_t)c * n + r] * fs[c];
      a1 += cinv[(size_t)(c + 4) * n + r] * fs[c + 4];
    }
    for (; c < n; c += 4) a0 += cinv[(size_t)c * n + r] * fs[c];
  }
  part[w][lane] = a0 + a1;
  __syncthreads();
  if (w == 0 && r < n) x[r] = (part[0][lane] + part[1][lane]) + (part[2][lane] + part[3][lane]);
  __syncthreads();  // fs and part are reused by the block's next tile
}

}  // namespace pp

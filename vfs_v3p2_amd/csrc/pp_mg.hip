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
#include "pp_kernels.hpp"
#include "pp_stencil.hpp"

namespace pp {

namespace {

const dim3 kBlk(64, 4, 1);
inline dim3 cell_grid(const Dims& d) { return dim3((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz); }
inline dim3 line_grid(const Dims& d) { return dim3((d.nx + 63) / 64, (d.nz + 3) / 4, 1); }

template <typename T>
__device__ __forceinline__ T row_times(const T (&vol)[19], const Nbr& nb, const T* __restrict__ x) {
  T acc = T(0);
#pragma unroll
  for (int s = 0; s < 19; s++)
    if (vol[s] != T(0)) acc += vol[s] * x[nb.s[s]];
  return acc;
}

// D^-1 of the level operator, 0 on cells that carry no unknown.
template <typename T>
__global__ __launch_bounds__(256) void k_dinv(LevelView<T> L, T* __restrict__ dinv) {
  CELL_IJK(i, j, k);
  if (i >= L.d.nx || j >= L.d.ny) return;
  const size_t c = L.d.idx(i, j, k);
  const uint32_t cd = L.code[c];
  Nbr nb;
  neighbours(L.d, i, j, k, nb);
  T xv[19];
#pragma unroll
  for (int s = 0; s < 19; s++) xv[s] = T(0);
  T ax, diag;
  flux_row(L.g, L.d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, ax, diag);
  dinv[c] = ((cd & ROW_BIT) && diag != T(0)) ? T(1) / diag : T(0);
}

// Chebyshev smoothing preconditioned by the diagonal.  One step is
//   z = D^-1 (f - A x_old);  d = c1 d + c2 z;  x_new = x_old + d
// (c1 == 0 starts a new sequence and does not read d).  Modes:
//   SM_NEXT    one step from x_old;
//   SM_ZERO    the first step from x = 0:  x_new = d = c2 D^-1 f;
//   SM_FIRST2  the first two steps from x = 0 in one pass: the first step's
//              x1 = d1 = c0 D^-1 f is formed on the fly at the 19 neighbours from
//              f and D^-1, then one step from x1 with (c1, c2).
template <typename T, int MODE>
__global__ __launch_bounds__(256) void k_smooth_point(LevelView<T> L, const T* __restrict__ f,
                                                      const T* __restrict__ dinv, const T* __restrict__ x_old,
                                                      T* __restrict__ x_new, T* __restrict__ dv, T c1, T c2, T c0) {
  CELL_IJK(i, j, k);
  if (i >= L.d.nx || j >= L.d.ny) return;
  const size_t c = L.d.idx(i, j, k);
  if constexpr (MODE == SM_ZERO) {
    const T v = c2 * (f[c] * dinv[c]);
    x_new[c] = v;
    dv[c] = v;
    return;
  }
  const uint32_t cd = L.code[c];
  Nbr nb;
  neighbours(L.d, i, j, k, nb);
  T xv[19];
  if constexpr (MODE == SM_FIRST2) {
#pragma unroll
    for (int s = 0; s < 19; s++) xv[s] = c0 * (f[nb.s[s]] * dinv[nb.s[s]]);
  } else {
    load_x(nb, x_old, xv);
  }
  T ax, diag;
  flux_row(L.g, L.d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, ax, diag);
  const T fc = f[c];
  const T dold = (MODE == SM_FIRST2) ? xv[CP] : (c1 == T(0) ? T(0) : dv[c]);
  if (!(cd & ROW_BIT)) { x_new[c] = T(0); dv[c] = T(0); return; }
  const T z = diag != T(0) ? (fc - ax) / diag : T(0);
  const T dn = c1 * dold + c2 * z;
  dv[c] = dn;
  x_new[c] = xv[CP] + dn;
}

// One Chebyshev step preconditioned by the y-line tridiagonal part T of A
// (block Jacobi over (i,k) lines); one thread per line, Thomas algorithm with
// the forward sweep's factors parked in scr.
template <typename T, bool ZG>
__global__ __launch_bounds__(256) void k_smooth_line(LevelView<T> L, const T* __restrict__ f,
                                                     const T* __restrict__ x_old, T* __restrict__ x_new,
                                                     T* __restrict__ dv, T* __restrict__ scr, T c1, T c2) {
  const int i = blockIdx.x * 64 + threadIdx.x, k = blockIdx.y * 4 + threadIdx.y;
  if (i >= L.d.nx || k >= L.d.nz) return;
  const size_t n = L.d.n();
  T* __restrict__ cpv = scr;
  T* __restrict__ dpv = scr + n;
  T cprev = T(0), dprev = T(0);
  for (int j = 0; j < L.d.ny; j++) {
    const size_t c = L.d.idx(i, j, k);
    const uint32_t cd = L.code[c];
    T a = T(0), b = T(1), cc = T(0), r = T(0);
    if (cd & ROW_BIT) {
      Nbr nb;
      neighbours(L.d, i, j, k, nb);
      T vol[19];
      build_row(L.g, n, cd, c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
      r = f[c];
      if (!ZG) r -= row_times(vol, nb, x_old);
      a = vol[SP];
      b = vol[CP];
      cc = vol[NP];
    }
    const T den = b - a * cprev;
    const T inv = den != T(0) ? T(1) / den : T(0);
    cprev = cc * inv;
    dprev = (r - a * dprev) * inv;
    cpv[c] = cprev;
    dpv[c] = dprev;
  }
  T znext = T(0);
  for (int j = L.d.ny - 1; j >= 0; j--) {
    const size_t c = L.d.idx(i, j, k);
    const T z = dpv[c] - cpv[c] * znext;
    znext = z;
    const T dold = (ZG || c1 == T(0)) ? T(0) : dv[c];
    const T dn = c1 * dold + c2 * z;
    dv[c] = dn;
    x_new[c] = (ZG ? T(0) : x_old[c]) + dn;
  }
}

// fc[C] = sum over fluid children of (f - A x); zero on coarse IB cells
template <typename T>
__global__ __launch_bounds__(256) void k_resid_restrict(LevelView<T> F, const T* __restrict__ f,
                                                        const T* __restrict__ x, Dims dc,
                                                        const uint8_t* __restrict__ clsc, int cmask,
                                                        T* __restrict__ fc) {
  const int ic = blockIdx.x * 64 + threadIdx.x, jc = blockIdx.y * 4 + threadIdx.y, kc = blockIdx.z;
  if (ic >= dc.nx || jc >= dc.ny) return;
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
        flux_row(F.g, F.d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, ax, dg);
        sum += f[c] - ax;
      }
  fc[C] = sum;
}

// Residual + restriction when all three axes are coarsened: one thread per fine
// cell (block 64x2x2 covers 32x1x1 coarse cells), pair sums in i by lane shuffle,
// the j and k pairs through LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_resid_restrict_blk(LevelView<T> F, const T* __restrict__ f,
                                                            const T* __restrict__ x, Dims dc,
                                                            const uint8_t* __restrict__ clsc, T* __restrict__ fc) {
  __shared__ T lds[2][2][32];
  const int tx = threadIdx.x, ty = threadIdx.y, tz = threadIdx.z;
  const uint3 xb = xcd_block();
  const int i = (int)xb.x * 64 + tx, j = (int)xb.y * 2 + ty, k = (int)xb.z * 2 + tz;
  T r = T(0);
  if (i < F.d.nx && j < F.d.ny && k < F.d.nz) {
    const size_t c = F.d.idx(i, j, k);
    const uint32_t cd = F.code[c];
    Nbr nb;
    neighbours(F.d, i, j, k, nb);
    T xv[19], ax, dg;
    load_x(nb, x, xv);
    flux_row(F.g, F.d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, ax, dg);
    if (cd & ROW_BIT) r = f[c] - ax;
  }
  r += __shfl_xor(r, 1, 64);
  if ((tx & 1) == 0) lds[tz][ty][tx >> 1] = r;
  __syncthreads();
  if (ty == 0 && tz == 0 && tx < 32) {
    const int ic = (int)xb.x * 32 + tx, jc = (int)xb.y, kc = (int)xb.z;
    if (ic < dc.nx && jc < dc.ny && kc < dc.nz) {
      const size_t C = dc.idx(ic, jc, kc);
      const T v = (lds[0][0][tx] + lds[0][1][tx]) + (lds[1][0][tx] + lds[1][1][tx]);
      fc[C] = (clsc[C] & CLS_NONFLUID) ? T(0) : v;
    }
  }
}

__device__ inline int axis_pick(int f, int n, int per, bool coarsened, int ncoarse, int& ic) {
  if (!coarsened) { ic = f; return 0; }
  ic = f >> 1;
  int ia = (f & 1) ? 1 : -1;
  if (!per && (f == 0 || f == n - 1)) ia = 0;
  (void)ncoarse;
  return ia;
}
__device__ inline int wrapc(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

template <typename T>
__global__ __launch_bounds__(256) void k_prolong_add(LevelView<T> F, Dims dc, const uint8_t* __restrict__ clsc,
                                                     int cmask, const T* __restrict__ xc, T* __restrict__ xf) {
  CELL_IJK(i, j, k);
  if (i >= F.d.nx || j >= F.d.ny) return;
  const size_t c = F.d.idx(i, j, k);
  if (F.cls[c] & CLS_NONFLUID) return;
  int ic, jc, kc;
  int ia = axis_pick(i, F.d.nx, F.d.per[0], cmask & 1, dc.nx, ic);
  int ja = axis_pick(j, F.d.ny, F.d.per[1], cmask & 2, dc.ny, jc);
  int ka = axis_pick(k, F.d.nz, F.d.per[2], cmask & 4, dc.nz, kc);
  const int in = wrapc(ic + ia, dc.nx), jn = wrapc(jc + ja, dc.ny), kn = wrapc(kc + ka, dc.nz);
  // one-sided next to coarse IB cells, in the reference's order k, j, i (poisson.c:500-505)
  if (ka != 0 && (clsc[dc.idx(ic, jc, kn)] & CLS_NONFLUID)) ka = 0;
  if (ja != 0 && (clsc[dc.idx(ic, jn, kc)] & CLS_NONFLUID)) ja = 0;
  if (ia != 0 && (clsc[dc.idx(in, jc, kc)] & CLS_NONFLUID)) ia = 0;
  const int i2 = ia ? in : ic, j2 = ja ? jn : jc, k2 = ka ? kn : kc;
  const T v = (xc[dc.idx(ic, jc, kc)] * T(9) + xc[dc.idx(ic, j2, kc)] * T(3) + xc[dc.idx(i2, jc, kc)] * T(3) +
               xc[dc.idx(i2, j2, kc)]) * T(3. / 64.) +
              (xc[dc.idx(ic, jc, k2)] * T(9) + xc[dc.idx(ic, j2, k2)] * T(3) + xc[dc.idx(i2, jc, k2)] * T(3) +
               xc[dc.idx(i2, j2, k2)]) * T(1. / 64.);
  xf[c] += v;
}

// x = Cinv f with Cinv stored column-major: each block owns 64 rows, its four
// waves take interleaved quarters of the columns, partial sums meet in LDS.
template <typename T>
__global__ __launch_bounds__(256) void k_dense_solve(int n, const T* __restrict__ cinv, const T* __restrict__ f,
                                                     T* __restrict__ x) {
  __shared__ T fs[1024];
  __shared__ T part[4][64];
  for (int t = threadIdx.x; t < n && t < 1024; t += 256) fs[t] = f[t];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = blockIdx.x * 64 + lane;
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
__global__ __launch_bounds__(256) void k_scale_by(size_t n, T* __restrict__ a, const double* __restrict__ nrm2) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const double s = nrm2[0] > 0. ? 1. / sqrt(nrm2[0]) : 0.;
  a[t] = (T)((double)a[t] * s);
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
void launch_smooth(hipStream_t s, LevelView<T> L, int line, int mode, const T* f, const T* dinv, const T* x_old,
                   T* x_new, T* dv, T* scr, T c1, T c2, T c0) {
  if (line) {
    if (mode == SM_FIRST2) throw StatusError(1, "the line smoother has no fused first-two-steps mode");
    if (mode == SM_ZERO) k_smooth_line<T, true><<<line_grid(L.d), kBlk, 0, s>>>(L, f, x_old, x_new, dv, scr, c1, c2);
    else k_smooth_line<T, false><<<line_grid(L.d), kBlk, 0, s>>>(L, f, x_old, x_new, dv, scr, c1, c2);
  } else {
    switch (mode) {
      case SM_ZERO:
        k_smooth_point<T, SM_ZERO><<<cell_grid(L.d), kBlk, 0, s>>>(L, f, dinv, x_old, x_new, dv, c1, c2, c0);
        break;
      case SM_FIRST2:
        k_smooth_point<T, SM_FIRST2><<<cell_grid(L.d), kBlk, 0, s>>>(L, f, dinv, x_old, x_new, dv, c1, c2, c0);
        break;
      default:
        k_smooth_point<T, SM_NEXT><<<cell_grid(L.d), kBlk, 0, s>>>(L, f, dinv, x_old, x_new, dv, c1, c2, c0);
    }
  }
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
  if (cm == 7) {
    const dim3 gr((F.d.nx + 63) / 64, (F.d.ny + 1) / 2, (F.d.nz + 1) / 2);
    k_resid_restrict_blk<T><<<gr, dim3(64, 2, 2), 0, s>>>(F, f, x, dc, cls_c, fc);
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

template <typename T>
void launch_normsq(hipStream_t s, size_t n, const T* a, double* partial, int* nblk) {
  const int nb = partial_blocks(n);
  *nblk = nb;
  k_normsq<T><<<nb, 256, 0, s>>>(n, a, partial, nb);
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_scale_by(hipStream_t s, size_t n, T* a, const double* norm2) {
  k_scale_by<T><<<nb1(n), 256, 0, s>>>(n, a, norm2);
  HIPCHECK(hipGetLastError());
}

#define INST(T)                                                                                                     \
  template void launch_smooth<T>(hipStream_t, LevelView<T>, int, int, const T*, const T*, const T*, T*, T*, T*, T, T, \
                                 T);                                                                              \
  template void launch_dinv<T>(hipStream_t, LevelView<T>, T*);                                                     \
  template void launch_resid_restrict<T>(hipStream_t, LevelView<T>, const T*, const T*, Dims, const uint8_t*, T*); \
  template void launch_prolong_add<T>(hipStream_t, LevelView<T>, Dims, const uint8_t*, const T*, T*);              \
  template void launch_dense_solve<T>(hipStream_t, int, const T*, const T*, T*);                                   \
  template void launch_normsq<T>(hipStream_t, size_t, const T*, double*, int*);                                    \
  template void launch_scale_by<T>(hipStream_t, size_t, T*, const double*);
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

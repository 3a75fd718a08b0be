// Per-element bodies of the pressure-path kernels around the solver (RHS,
// VolumeFlux, true residual, null space, Convert_Phi2_Phi, UpdatePressure,
// Projection, Contra2Cart, Divergence).  The single-case kernels (pp_core.hip)
// and the batched phase kernels (pp_batch.hip) both call these, so the two
// forms do the same arithmetic on every element.
#pragma once

#include "pp_kernels.hpp"
#include "pp_stencil.hpp"

namespace pp {

// periodic image of DMDA point (poisson.c:153-174)
__device__ inline bool image_of(const Dmda& m, int i, int j, int k, int& a, int& b, int& c, int fl[3]) {
  a = i; b = j; c = k;
  fl[0] = fl[1] = fl[2] = 0;
  if (m.per[0] && i == 0) { a = m.mx - 2; fl[0] = 1; }
  else if (m.per[0] && i == m.mx - 1) { a = 1; fl[0] = 1; }
  if (m.per[1] && j == 0) { b = m.my - 2; fl[1] = 1; }
  else if (m.per[1] && j == m.my - 1) { b = 1; fl[1] = 1; }
  if (m.per[2] && k == 0) { c = m.mz - 2; fl[2] = 1; }
  else if (m.per[2] && k == m.mz - 1) { c = 1; fl[2] = 1; }
  return fl[0] | fl[1] | fl[2];
}

__device__ __forceinline__ bool on_boundary(const Dmda& m, int i, int j, int k) {
  return i == 0 || j == 0 || k == 0 || i == m.mx - 1 || j == m.my - 1 || k == m.mz - 1;
}

// Number of boundary points of the DMDA box and the e-th of them: the two k
// planes, then the two j planes without their k-boundary rows, then the two i
// planes without their j- and k-boundary rows.
HD inline uint32_t boundary_count(const Dmda& m) {
  const uint32_t mx = m.mx, my = m.my, mz = m.mz;
  return 2u * mx * my + 2u * mx * (mz - 2u) + 2u * (my - 2u) * (mz - 2u);
}
HD inline bool boundary_point(const Dmda& m, uint32_t e, int& i, int& j, int& k) {
  const uint32_t mx = m.mx, my = m.my, mz = m.mz;
  const uint32_t pxy = mx * my;
  if (e < 2u * pxy) {
    const uint32_t r = e < pxy ? e : e - pxy;
    k = e < pxy ? 0 : (int)mz - 1;
    j = (int)(r / mx);
    i = (int)(r - (uint32_t)j * mx);
    return true;
  }
  e -= 2u * pxy;
  const uint32_t kin = mz - 2u, pxz = mx * kin;
  if (e < 2u * pxz) {
    const uint32_t r = e < pxz ? e : e - pxz;
    j = e < pxz ? 0 : (int)my - 1;
    const uint32_t kk = r / mx;
    k = 1 + (int)kk;
    i = (int)(r - kk * mx);
    return true;
  }
  e -= 2u * pxz;
  const uint32_t jin = my - 2u, pyz = jin * kin;
  if (e < 2u * pyz) {
    const uint32_t r = e < pyz ? e : e - pyz;
    i = e < pyz ? 0 : (int)mx - 1;
    const uint32_t kk = r / jin;
    k = 1 + (int)kk;
    j = 1 + (int)(r - kk * jin);
    return true;
  }
  return false;
}

// ------------------------------------------------------------------------- RHS
// b[c] of interior cell (i, j, k) (poisson_hypre.c:346-376); returns b[c]^2
__device__ inline double rhs_cell(const Dmda& m, const Dims& d, const double* __restrict__ u,
                                  const uint8_t* __restrict__ cls, double fac, double* __restrict__ b, int i, int j,
                                  int k) {
  const size_t c = d.idx(i, j, k);
  const int I = i + 1, J = j + 1, K = k + 1;
  const size_t p = m.idx(I, J, K);
  double val = 0.;
  if (cls[c] & CLS_LT) {
    val -= u[3 * p + 0];
    val += (I == 1 && m.per[0]) ? u[3 * m.idx(m.mx - 2, J, K) + 0] : u[3 * (p - 1) + 0];
    val -= u[3 * p + 1];
    val += (J == 1 && m.per[1]) ? u[3 * m.idx(I, m.my - 2, K) + 1] : u[3 * (p - m.mx) + 1];
    val -= u[3 * p + 2];
    val += (K == 1 && m.per[2]) ? u[3 * m.idx(I, J, m.mz - 2) + 2] : u[3 * (p - (size_t)m.mx * m.my) + 2];
    val *= fac;
  }
  b[c] = val;
  return val * val;
}

// ------------------------------------------------------------------ VolumeFlux
// one face of the sums: poisson.c:2938-2957 zeroes and skips |u| <= 1e-10 faces
// when zero_small
__device__ __forceinline__ void ibflux_sum_face(const IbFace& F, double* __restrict__ u, int zero_small, double& fl,
                                                double& ar) {
  const double v = u[F.pos];
  if (zero_small) {
    if (fabs(v) > 1.e-10) { fl += F.sign * v; ar += F.area; }
    else u[F.pos] = 0.;
  } else {
    fl += F.sign * v;
    ar += F.area;
  }
}
// correction of one face from sums = [flux, area]; immersed == 0 sets it to zero
// (poisson.c:2715-2750)
__device__ __forceinline__ void ibflux_correct_face(const IbFace& F, double* __restrict__ u, int immersed,
                                                    const double* __restrict__ sums) {
  const double area = sums[1];
  const double corr = area > 1.e-15 ? sums[0] / area : 0.;
  if (immersed == 0) {
    u[F.pos] = 0.;
    return;
  }
  const double v = u[F.pos];
  if (fabs(v) > 1.e-10) u[F.pos] = v - F.sign * (F.area * corr);
}

// ----------------------------------------------------------- FP64 true residual
// r[c] = b[c] - (A x)[c] on row cells, 0 elsewhere; returns r[c]^2
template <typename TX>
__device__ inline double residual_cell(const Dims& d, const double* __restrict__ g, const uint32_t* __restrict__ code,
                                       const double* __restrict__ b, const TX* __restrict__ x, double* __restrict__ r,
                                       int i, int j, int k) {
  const size_t c = d.idx(i, j, k);
  const uint32_t cd = code[c];
  Nbr nb;
  neighbours(d, i, j, k, nb);
  double xv[19], ax, dg;
  load_x(nb, x, xv);
  flux_row(g, d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, ax, dg);
  const double res = (cd & ROW_BIT) ? b[c] - ax : 0.;
  r[c] = res;
  return res * res;
}

// |aj r| on Divergence-counted cells: with r = b - A phi this is the max |aj div u|
// the projection of phi leaves (div u_new = dt/coeff (b - A phi) for st = 1)
__device__ __forceinline__ double divest_elem(const uint32_t* __restrict__ code, const double* __restrict__ aj,
                                              const double* __restrict__ r, size_t t) {
  return (code[t] & DIVOK_BIT) ? fabs(aj[t] * r[t]) : 0.;
}

// ------------------------------------------------------------ null space, scatter
// Convert_Phi2_Phi (poisson.c:131-139) and, when p is given, UpdatePressure's
// P += Phi (poisson.c:2548) for interior cell (i, j, k)
__device__ __forceinline__ void scatter_cell(const Dmda& m, const Dims& d, const double* __restrict__ x,
                                             const uint8_t* __restrict__ cls, double* __restrict__ phi,
                                             double* __restrict__ p, int i, int j, int k) {
  const size_t c = d.idx(i, j, k), q = m.idx(i + 1, j + 1, k + 1);
  const uint8_t b = cls[c];
  const double ph = (b & CLS_INTGT) ? 0. : x[c];
  phi[q] = ph;
  if (p && (b & CLS_LT)) p[q] += ph;
}

// Initial guess from the caller's Phi: the inverse of Convert_Phi2_Phi.  The
// reference keeps its HYPRE solution vector between time steps and GMRES starts
// from it (poisson_hypre.c:565-574, 722); that vector equals Phi on every cell
// that carries an unknown.
__device__ __forceinline__ void gather_phi_cell(const Dmda& m, const Dims& d, const double* __restrict__ phi,
                                                const uint8_t* __restrict__ cls, double* __restrict__ x, int i, int j,
                                                int k) {
  const size_t c = d.idx(i, j, k);
  x[c] = (cls[c] & CLS_INTGT) ? 0. : phi[m.idx(i + 1, j + 1, k + 1)];
}

// UpdatePressure: P += Phi on interior cells with nvert < 0.1 (poisson.c:2548)
__device__ __forceinline__ void add_phi_cell(const Dmda& m, const Dims& d, const uint8_t* __restrict__ cls,
                                             const double* __restrict__ phi, double* __restrict__ p, int i, int j,
                                             int k) {
  if (!(cls[d.idx(i, j, k)] & CLS_LT)) return;
  const size_t q = m.idx(i + 1, j + 1, k + 1);
  p[q] += phi[q];
}

// Ghost point (i, j, k) of a DMDA scalar: its periodic image; other boundary
// points are zeroed when zero_other (Convert_Phi2_Phi starts from VecSet(Phi,0)).
// The image of a ghost point with one periodic flag may itself be a non-periodic
// boundary point; it is resolved the same way.  Images are never ghost points
// along a periodic axis, so no image is written by the same pass.
__device__ __forceinline__ void periodic_point(const Dmda& m, double* __restrict__ a, int zero_other, int i, int j,
                                               int k) {
  int aa, bb, cc, fl[3];
  if (image_of(m, i, j, k, aa, bb, cc, fl)) {
    if (on_boundary(m, aa, bb, cc) && zero_other) a[m.idx(i, j, k)] = 0.;
    else a[m.idx(i, j, k)] = a[m.idx(aa, bb, cc)];
  } else if (zero_other) {
    a[m.idx(i, j, k)] = 0.;
  }
}

// ------------------------------------------------------------------ Projection
// Cross difference of phi for the face (p, q) from its case code (poisson.c:2234-2257).
__device__ inline double proj_cross(int cc, const double* __restrict__ phi, size_t p, size_t q, size_t se) {
  switch (cc) {
    case X_LOWER: return (phi[p] + phi[q] - phi[p - se] - phi[q - se]) * 0.5;
    case X_UPPER: return (phi[p + se] + phi[q + se] - phi[p] - phi[q]) * 0.5;
    case X_CENTRAL: return (phi[p + se] + phi[q + se] - phi[p - se] - phi[q - se]) * 0.25;
    default: return 0.;
  }
}

// The three upper-face fluxes of interior cell (i, j, k), read, corrected and
// stored together so each cache line of the AoS array is written once.
__device__ inline void projection_cell(const Dmda& m, const Dims& d, const double* __restrict__ g,
                                       const uint32_t* __restrict__ code, const uint8_t* __restrict__ cls,
                                       const double* __restrict__ phi, double* __restrict__ u, double fac, int i,
                                       int j, int k) {
  const size_t n = d.n(), c = d.idx(i, j, k), p = m.idx(i + 1, j + 1, k + 1);
  const uint8_t b = cls[c];
  if (!(b & (CLS_PROJ_I | CLS_PROJ_J | CLS_PROJ_K))) return;
  const uint32_t cd = code[c];
  const size_t sx = 1, sy = m.mx, sz = (size_t)m.mx * m.my;
  double u0 = u[3 * p + 0], u1 = u[3 * p + 1], u2 = u[3 * p + 2];
  if (b & CLS_PROJ_I) {  // poisson.c:2221-2265; face i+ is face 0, cross j then k
    const size_t q = p + sx;
    const double dc = phi[q] - phi[p];
    const double de = proj_cross((cd >> 6) & 3u, phi, p, q, sy);
    const double dz = proj_cross((cd >> 8) & 3u, phi, p, q, sz);
    u0 -= (dc * g[0 * n + c] + de * g[1 * n + c] + dz * g[2 * n + c]) * fac;
  }
  if (b & CLS_PROJ_J) {  // poisson.c:2267-2314; face j+ is face 2, cross i then k
    const size_t q = p + sy;
    const double dc = proj_cross((cd >> 14) & 3u, phi, p, q, sx);
    const double de = phi[q] - phi[p];
    const double dz = proj_cross((cd >> 16) & 3u, phi, p, q, sz);
    u1 -= (dc * g[3 * n + c] + de * g[4 * n + c] + dz * g[5 * n + c]) * fac;
  }
  if (b & CLS_PROJ_K) {  // poisson.c:2316-2363; face k+ is face 4, cross i then j
    const size_t q = p + sz;
    const double dc = proj_cross((cd >> 22) & 3u, phi, p, q, sx);
    const double de = proj_cross((cd >> 24) & 3u, phi, p, q, sy);
    const double dz = phi[q] - phi[p];
    u2 -= (dc * g[6 * n + c] + de * g[7 * n + c] + dz * g[8 * n + c]) * fac;
  }
  u[3 * p + 0] = u0;
  u[3 * p + 1] = u1;
  u[3 * p + 2] = u2;
}

// per-component periodic ghost faces of Ucont (poisson.c:2401-2441)
__device__ __forceinline__ void ucont_ghost_point(const Dmda& m, double* __restrict__ u, int i, int j, int k) {
  int a, b, c, fl[3];
  if (!image_of(m, i, j, k, a, b, c, fl)) return;
  const size_t p = m.idx(i, j, k), s = m.idx(a, b, c);
  if (fl[0]) u[3 * p + 0] = u[3 * s + 0];
  if (fl[1]) u[3 * p + 1] = u[3 * s + 1];
  if (fl[2]) u[3 * p + 2] = u[3 * s + 2];
}

// ----------------------------------------------------------------- Contra2Cart
// Cartesian velocity of interior cell (i, j, k) from the averaged contravariant
// fluxes by Cramer's rule (rhs.c:166-189)
__device__ inline void c2c_core_cell(const Dmda& m, const Dims& d, const double* __restrict__ met,
                                     const uint8_t* __restrict__ cls, const double* __restrict__ u,
                                     double* __restrict__ ucat, int i, int j, int k) {
#pragma clang fp contract(off)
  const size_t n = d.n(), c = d.idx(i, j, k);
  if (!(cls[c] & CLS_LT)) return;
  const size_t p = m.idx(i + 1, j + 1, k + 1);
  const double m00 = met[0 * n + c], m01 = met[1 * n + c], m02 = met[2 * n + c];
  const double m10 = met[3 * n + c], m11 = met[4 * n + c], m12 = met[5 * n + c];
  const double m20 = met[6 * n + c], m21 = met[7 * n + c], m22 = met[8 * n + c];
  const double q0 = 0.5 * (u[3 * (p - 1) + 0] + u[3 * p + 0]);
  const double q1 = 0.5 * (u[3 * (p - m.mx) + 1] + u[3 * p + 1]);
  const double q2 = 0.5 * (u[3 * (p - (size_t)m.mx * m.my) + 2] + u[3 * p + 2]);
  const double det = m00 * (m11 * m22 - m12 * m21) - m01 * (m10 * m22 - m12 * m20) + m02 * (m10 * m21 - m11 * m20);
  const double det0 = q0 * (m11 * m22 - m12 * m21) - q1 * (m01 * m22 - m02 * m21) + q2 * (m01 * m12 - m02 * m11);
  const double det1 = -q0 * (m10 * m22 - m12 * m20) + q1 * (m00 * m22 - m02 * m20) - q2 * (m00 * m12 - m02 * m10);
  const double det2 = q0 * (m10 * m21 - m11 * m20) - q1 * (m00 * m21 - m01 * m20) + q2 * (m00 * m11 - m01 * m10);
  ucat[3 * p + 0] = det0 / det;
  ucat[3 * p + 1] = det1 / det;
  ucat[3 * p + 2] = det2 / det;
}

// whole-vector periodic copy at boundary point (i, j, k) (rhs.c:198-241,
// 645-673); when snap is given the point is also copied there (the snapshot the
// boundary rules read)
__device__ __forceinline__ void vec_ghost_point(const Dmda& m, double* __restrict__ a, double* __restrict__ snap,
                                                int i, int j, int k) {
  const size_t p = m.idx(i, j, k);
  int aa, bb, cc, fl[3];
  double v0 = a[3 * p], v1 = a[3 * p + 1], v2 = a[3 * p + 2];
  if (image_of(m, i, j, k, aa, bb, cc, fl)) {
    const size_t s = m.idx(aa, bb, cc);
    v0 = a[3 * s]; v1 = a[3 * s + 1]; v2 = a[3 * s + 2];
    a[3 * p] = v0; a[3 * p + 1] = v1; a[3 * p + 2] = v2;
  }
  if (snap) { snap[3 * p] = v0; snap[3 * p + 1] = v1; snap[3 * p + 2] = v2; }
}

// boundary rules of Contra2Cart at boundary point (i, j, k) (rhs.c:257-260,
// 491-515); sources on boundary points come from the snapshot, interior sources
// are untouched until the solid-cell pass runs
__device__ inline void c2c_bc_point(const Dmda& m, const double* __restrict__ nv, const double* __restrict__ snap,
                                    double* __restrict__ ucat, const int bc[6], int i, int j, int k) {
  const size_t p = m.idx(i, j, k);
  if ((int)(nv[p] + 0.1) == 3) { ucat[3 * p] = ucat[3 * p + 1] = ucat[3 * p + 2] = 0.; return; }
  auto src = [&](int a, int b, int c) -> const double* {
    return (on_boundary(m, a, b, c) ? snap : ucat) + 3 * m.idx(a, b, c);
  };
  const double* s = nullptr;
  if (i == 0 && bc[0] == 1 && (j != 0 && k != 0)) s = src(i + 1, j, k);
  if (i == m.mx - 1 && bc[1] == 1 && (j != 0 && k != 0)) s = src(i - 1, j, k);
  if (j == 0 && bc[2] == 1 && (i != 0 && k != 0)) s = src(i, j + 1, k);
  if (j == m.my - 1 && bc[3] == 1 && (i != 0 && k != 0)) s = src(i, j - 1, k);
  if (k == 0 && bc[4] == 1 && (i != 0 && j != 0)) s = src(i, j, k + 1);
  if (k == m.mz - 1 && bc[5] == 1 && (i != 0 && j != 0)) s = src(i, j, k - 1);
  if (s) {
    const double v0 = -s[0], v1 = -s[1], v2 = -s[2];
    ucat[3 * p] = v0; ucat[3 * p + 1] = v1; ucat[3 * p + 2] = v2;
  }
}

// Ucat = 0 on solid interior cells (rhs.c:257)
__device__ __forceinline__ void c2c_solid_cell(const Dmda& m, const Dims& d, const uint8_t* __restrict__ cls,
                                               double* __restrict__ ucat, int i, int j, int k) {
  if (!(cls[d.idx(i, j, k)] & CLS_SOLID)) return;
  const size_t p = m.idx(i + 1, j + 1, k + 1);
  ucat[3 * p] = ucat[3 * p + 1] = ucat[3 * p + 2] = 0.;
}

// ------------------------------------------------------------------- Divergence
// |aj div u| of interior cell (i, j, k) if Divergence counts it, else 0 (main.c:1283
// operand order)
__device__ inline double divergence_cell(const Dmda& m, const Dims& d, const uint32_t* __restrict__ code,
                                         const double* __restrict__ aj, const double* __restrict__ u, int i, int j,
                                         int k) {
#pragma clang fp contract(off)
  const size_t c = d.idx(i, j, k);
  if (!(code[c] & DIVOK_BIT)) return 0.;
  const size_t p = m.idx(i + 1, j + 1, k + 1);
  return fabs((u[3 * p] - u[3 * (p - 1)] + u[3 * p + 1] - u[3 * (p - m.mx) + 1] + u[3 * p + 2] -
               u[3 * (p - (size_t)m.mx * m.my) + 2]) *
              aj[c]);
}

// ------------------------------------------------------------------ vectors
// x[t] += sum_v y_v Z_v[t] (the Krylov solution update)
template <typename TZ>
__device__ __forceinline__ void update_x_elem(double* __restrict__ x, const TZ* __restrict__ Z, size_t ldz, int nz,
                                              const double* __restrict__ y, size_t t) {
  double acc = x[t];
  for (int v = 0; v < nz; v++) acc += y[v] * (double)Z[(size_t)v * ldz + t];
  x[t] = acc;
}

// v[t] = a[t] / sqrt(norm2); f[t] = (TF)v[t]
template <typename TF>
__device__ __forceinline__ void scale_conv_elem(const double* __restrict__ a, double norm2, double* __restrict__ v,
                                                TF* __restrict__ f, size_t t) {
  const double s = norm2 > 0. ? 1. / sqrt(norm2) : 0.;
  const double x = a[t] * s;
  v[t] = x;
  if (f) f[t] = (TF)x;
}

}  // namespace pp

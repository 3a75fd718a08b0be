// The explicit momentum predictor on the device: Formfunction_2
// (momentum.c:529-2034) and the four-stage RungeKutta update around it
// (rhs.c:1511-1631), the step that produces the Ucont* the projection corrects.
//
// Scope is the reference's default branch: no LES/RANS, no level set, no
// skew-symmetric form, fourth-order central convection (second_order,
// inviscid, laplacian and clark off), no o-cylinder force bookkeeping
// (bctype 11).  The inflow/outflow boundary (InflowFlux, FormBCS) and the IB
// velocity interpolation (ibm_interpolation_advanced) stay with the caller,
// whose Ucat carries their results in IB-adjacent cells (Contra2Cart leaves
// those cells as they are).
//
// One thread per DMDA point in every kernel.  The face fluxes of the three
// families are written to DMDA arrays (convective and viscous apart, since the
// cell step takes a fourth-order difference of the convective ones), their
// periodic ghosts copied, then the cell divergence Fp and the contravariant
// right-hand side.  The arithmetic keeps the reference's operand order without
// contraction, so the results equal a CPU restatement bit for bit.
#include <limits>

#include "pp_engine.hpp"

namespace pp {

namespace {

#pragma clang fp contract(off)

struct MomIn {
  const double* fm[3][3];  // face metrics: [family i/j/k][csi, eta, zet]
  const double* faj[3];    // iaj, jaj, kaj
  const double* cm[3];     // cell csi, eta, zet
  const double* nv;
  int bc5;       // bctype[5] (the outflow special case of the k faces)
  int immersed;  // the reference's `immersed`, as given
};

struct V3 {
  double x[3];
};

__device__ __forceinline__ bool dmda_point(const Dmda& m, int& i, int& j, int& k) {
  i = blockIdx.x * 64 + threadIdx.x;
  j = blockIdx.y * 4 + threadIdx.y;
  k = blockIdx.z;
  return i < m.mx && j < m.my;
}

// the periodic image of a ghost point (the `if(periodic)` copies,
// momentum.c:713-740, 1556-1582, 1724-1748); false for a point that is none
__device__ __forceinline__ bool ghost_image(const Dmda& m, int i, int j, int k, size_t& src) {
  int a = i, b = j, c = k;
  bool f = false;
  if (m.per[0] && i == 0) a = m.mx - 2, f = true;
  else if (m.per[0] && i == m.mx - 1) a = 1, f = true;
  if (m.per[1] && j == 0) b = m.my - 2, f = true;
  else if (m.per[1] && j == m.my - 1) b = 1, f = true;
  if (m.per[2] && k == 0) c = m.mz - 2, f = true;
  else if (m.per[2] && k == m.mz - 1) c = 1, f = true;
  src = m.idx(a, b, c);
  return f;
}

// lUcont with its whole-vector periodic ghosts (momentum.c:713-740); the images
// are never ghosts themselves, so one pass suffices
__global__ __launch_bounds__(256) void k_mm_local_ucont(Dmda m, const double* __restrict__ u, double* __restrict__ lu) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  const size_t p = m.idx(i, j, k);
  size_t s;
  if (!ghost_image(m, i, j, k, s)) s = p;
  lu[3 * p] = u[3 * s];
  lu[3 * p + 1] = u[3 * s + 1];
  lu[3 * p + 2] = u[3 * s + 2];
}

// Compute_du_i / _j / _k (inline.h:406-611): du[c][e] = d U_c / d xi_e at face D
template <int D>
__device__ __forceinline__ void face_du(const size_t (&st)[3], size_t p, const double* __restrict__ u,
                                        const double* __restrict__ nv, double (&du)[3][3]) {
  const size_t sd = st[D];
#pragma unroll
  for (int c = 0; c < 3; c++) du[c][D] = u[3 * (p + sd) + c] - u[3 * p + c];
#pragma unroll
  for (int e = 0; e < 3; e++) {
    if (e == D) continue;
    const size_t se = st[e];
    if (nv[p + se] > 0.1 || nv[p + sd + se] > 0.1) {
#pragma unroll
      for (int c = 0; c < 3; c++)
        du[c][e] = (u[3 * (p + sd) + c] + u[3 * p + c] - u[3 * (p + sd - se) + c] - u[3 * (p - se) + c]) * 0.5;
    } else if (nv[p - se] > 0.1 || nv[p + sd - se] > 0.1) {
#pragma unroll
      for (int c = 0; c < 3; c++)
        du[c][e] = (u[3 * (p + sd + se) + c] + u[3 * (p + se) + c] - u[3 * (p + sd) + c] - u[3 * p + c]) * 0.5;
    } else {
#pragma unroll
      for (int c = 0; c < 3; c++)
        du[c][e] =
            (u[3 * (p + sd + se) + c] + u[3 * (p + se) + c] - u[3 * (p + sd - se) + c] - u[3 * (p - se) + c]) * 0.25;
    }
  }
}

// Convective and viscous flux of face D of DMDA point (i, j, k): the i, j and k
// loops of momentum.c:743-1010, 1011-1250, 1251-1521
template <int D>
__global__ __launch_bounds__(256) void k_mm_faces(Dmda m, MomIn in, const double* __restrict__ uc,
                                                  const double* __restrict__ ucat, double nu,
                                                  double* __restrict__ div, double* __restrict__ visc) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  const int pos[3] = {i, j, k}, mm[3] = {m.mx, m.my, m.mz};
  if (i == m.mx - 1 || j == m.my - 1 || k == m.mz - 1) return;
  if ((D != 0 && i == 0) || (D != 1 && j == 0) || (D != 2 && k == 0)) return;
  const size_t st[3] = {1, (size_t)m.mx, (size_t)m.mx * m.my};
  const size_t p = m.idx(i, j, k), sd = st[D];
  const double* __restrict__ nv = in.nv;
  const double ajc = in.faj[D][p];
  double mt[3][3];
#pragma unroll
  for (int e = 0; e < 3; e++)
#pragma unroll
    for (int x = 0; x < 3; x++) mt[e][x] = in.fm[D][e][3 * p + x];
  double du[3][3];
  face_du<D>(st, p, ucat, nv, du);
  double g[3];
#pragma unroll
  for (int e = 0; e < 3; e++) g[e] = mt[e][0] * mt[D][0] + mt[e][1] * mt[D][1] + mt[e][2] * mt[D][2];
  double r[3][3];
#pragma unroll
  for (int c = 0; c < 3; c++)
#pragma unroll
    for (int x = 0; x < 3; x++) r[c][x] = du[c][0] * mt[0][x] + du[c][1] * mt[1][x] + du[c][2] * mt[2][x];
  const int pd = pos[D], md = mm[D];
  size_t L = p - sd, R = p + 2 * sd;
  if (pd == 0 || pd == md - 2) {
    if (m.per[D]) {
      L = (size_t)((long long)p + (long long)(md - 3 - pd) * (long long)sd);
      R = (size_t)((long long)p + (long long)(2 - pd) * (long long)sd);
    } else {
      L = p;
      R = p + sd;
    }
  } else if (nv[L] + nv[R] > 0.1) {
    L = p;
    R = p + sd;
  }
  double ucon = uc[3 * p + D];
  if (m.per[D] && (pd == 0 || pd == md - 2)) ucon = uc[3 * (size_t)((long long)p + (long long)(md - 2 - pd) * (long long)sd) + D];
  if (D == 2 && k == m.mz - 2 && in.bc5 == 4 && (int)nv[p] == 0) ucon = uc[3 * (p - sd) + 2];
  const double up = -0.5 * (ucon + fabs(ucon));
  const double um = -0.5 * (ucon - fabs(ucon));
  double dv[3];
  if (in.immersed && pd != md - 2 && nv[p] > 0.1) {
#pragma unroll
    for (int c = 0; c < 3; c++)
      dv[c] = um * (0.125 * (-ucat[3 * (p + 2 * sd) + c] - 2. * ucat[3 * (p + sd) + c] + 3. * ucat[3 * p + c]) +
                    ucat[3 * (p + sd) + c]) +
              up * (0.125 * (-ucat[3 * p + c] - 2. * ucat[3 * p + c] + 3. * ucat[3 * (p + sd) + c]) + ucat[3 * p + c]);
  } else if (in.immersed && pd != 0 && nv[p + sd] > 0.1) {
#pragma unroll
    for (int c = 0; c < 3; c++)
      dv[c] = um * (0.125 * (-ucat[3 * (p + sd) + c] - 2. * ucat[3 * (p + sd) + c] + 3. * ucat[3 * p + c]) +
                    ucat[3 * (p + sd) + c]) +
              up * (0.125 * (-ucat[3 * (p - sd) + c] - 2. * ucat[3 * p + c] + 3. * ucat[3 * (p + sd) + c]) +
                    ucat[3 * p + c]);
  } else {
    const double fl = uc[3 * p + D];
#pragma unroll
    for (int c = 0; c < 3; c++)
      dv[c] = -fl * 0.0625 * (-ucat[3 * L + c] + 9. * ucat[3 * p + c] + 9. * ucat[3 * (p + sd) + c] - ucat[3 * R + c]);
  }
  if (nv[p] + nv[p + sd] > 0.1 && in.immersed == 3) dv[0] = dv[1] = dv[2] = 0.;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    div[3 * p + c] = dv[c];
    double v = 0.;
    v += (g[0] * du[c][0] + g[1] * du[c][1] + g[2] * du[c][2] + r[0][c] * mt[D][0] + r[1][c] * mt[D][1] +
          r[2][c] * mt[D][2]) *
         ajc * (nu);
    visc[3 * p + c] = v;
  }
}

// whole-vector periodic ghost copies of n3 DMDA vector arrays, in place
__global__ __launch_bounds__(256) void k_mm_periodic(Dmda m, double* a0, double* a1, double* a2, double* a3,
                                                     double* a4, double* a5, int narr) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  size_t s;
  if (!ghost_image(m, i, j, k, s)) return;
  const size_t p = m.idx(i, j, k);
  double* arr[6] = {a0, a1, a2, a3, a4, a5};
  for (int q = 0; q < narr; q++) {
    double* a = arr[q];
    a[3 * p] = a[3 * s];
    a[3 * p + 1] = a[3 * s + 1];
    a[3 * p + 2] = a[3 * s + 2];
  }
}

// fourth-order stencil ends of the cell difference along one axis
// (momentum.c:1605-1662): returns 1/denom and the R, L plane offsets
__device__ __forceinline__ double cell_ends(int per, int pos, int mq, const double* __restrict__ nv, size_t p,
                                            size_t s, long long& R, long long& Lo) {
  R = 1;
  Lo = -2;
  double denom = 3.;
  if (pos == 1) {
    if (per) Lo = (long long)(mq - 3) - pos;
    else R = 0, Lo = -1, denom = 1.;
  } else if (pos == 2 || pos == mq - 3) {
    if (!per) R = 0, Lo = -1, denom = 1.;
  } else if (pos == mq - 2) {
    if (per) R = 1 - pos;
    else R = 0, Lo = -1, denom = 1.;
  }
  if (nv[p - s] + nv[p] + nv[p + s] > 0.1) R = 0, Lo = -1, denom = 1.;
  return 1. / denom;
}

// Fp at the interior cells (momentum.c:1584-1718); other points left as they are
__global__ __launch_bounds__(256) void k_mm_fp(Dmda m, const double* __restrict__ nv, const double* __restrict__ d0,
                                               const double* __restrict__ d1, const double* __restrict__ d2,
                                               const double* __restrict__ v0, const double* __restrict__ v1,
                                               const double* __restrict__ v2, double* __restrict__ fp) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  if (i == 0 || j == 0 || k == 0 || i == m.mx - 1 || j == m.my - 1 || k == m.mz - 1) return;
  const size_t st[3] = {1, (size_t)m.mx, (size_t)m.mx * m.my};
  const size_t p = m.idx(i, j, k);
  const double* dd[3] = {d0, d1, d2};
  const int pos[3] = {i, j, k}, mq[3] = {m.mx, m.my, m.mz};
  double dv[3];
#pragma unroll
  for (int c = 0; c < 3; c++)
    dv[c] = (d0[3 * p + c] - d0[3 * (p - 1) + c] + d1[3 * p + c] - d1[3 * (p - st[1]) + c] + d2[3 * p + c] -
             d2[3 * (p - st[2]) + c]);
  double d4[3] = {0., 0., 0.};
#pragma unroll
  for (int a = 0; a < 3; a++) {
    long long R, Lo;
    const double inv = cell_ends(m.per[a], pos[a], mq[a], nv, p, st[a], R, Lo);
    const size_t pr = (size_t)((long long)p + R * (long long)st[a]);
    const size_t pl = (size_t)((long long)p + Lo * (long long)st[a]);
#pragma unroll
    for (int c = 0; c < 3; c++) d4[c] += (dd[a][3 * pr + c] - dd[a][3 * pl + c]) * inv;
  }
  const double rr = 1.0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    double f = 9. / 8. * dv[c] + -1. / 8. * d4[c];
    f += (v0[3 * p + c] - v0[3 * (p - 1) + c] + v1[3 * p + c] - v1[3 * (p - st[1]) + c] + v2[3 * p + c] -
          v2[3 * (p - st[2]) + c]) /
         rr;
    fp[3 * p + c] = f;
  }
}

// the contravariant right-hand side (momentum.c:1750-1965) at every DMDA point;
// acc = 0: Rhs starts at zero (the caller's VecSet), acc = 1: the terms add to what
// Rhs holds (FormFunction_SNES adds them to its time terms); boundary-plane points
// end at zero either way
__global__ __launch_bounds__(256) void k_mm_rhs(Dmda m, MomIn in, const double* __restrict__ fp, double scale,
                                                int acc, double* __restrict__ rhs) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  const size_t st[3] = {1, (size_t)m.mx, (size_t)m.mx * m.my};
  const size_t p = m.idx(i, j, k);
  const int pos[3] = {i, j, k}, mq[3] = {m.mx, m.my, m.mz};
  double out[3] = {0., 0., 0.};
  const bool interior = i > 0 && j > 0 && k > 0 && i < m.mx - 1 && j < m.my - 1 && k < m.mz - 1;
  if (interior && acc) {
    out[0] = rhs[3 * p];
    out[1] = rhs[3 * p + 1];
    out[2] = rhs[3 * p + 2];
  }
  if (interior) {
    const double* nv = in.nv;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double* a0 = in.cm[d] + 3 * p;
      const double* a1 = in.cm[d] + 3 * (p + st[d]);
      const double* f0 = fp + 3 * p;
      const double* f1 = fp + 3 * (p + st[d]);
      out[d] += scale * (0.5 * (a0[0] * f0[0] + a0[1] * f0[1] + a0[2] * f0[2]) +
                         0.5 * (a1[0] * f1[0] + a1[1] * f1[1] + a1[2] * f1[2])) *
                in.faj[d][p];
      if (nv[p] + nv[p + st[d]] > 0.1 || (!m.per[d] && pos[d] == mq[d] - 2)) out[d] = 0.;
    }
  } else {
#pragma unroll
    for (int a = 0; a < 3; a++)
      if (pos[a] == mq[a] - 2 && !m.per[a]) out[a] = 0.;
  }
  rhs[3 * p] = out[0];
  rhs[3 * p + 1] = out[1];
  rhs[3 * p + 2] = out[2];
}

// one Runge-Kutta stage: Rhs -= dP (VecAXPY), Ucont = a Rhs + pUcont (VecWAXPY)
__global__ __launch_bounds__(256) void k_mm_stage(size_t n3, const double* __restrict__ rhs,
                                                  const double* __restrict__ dp, const double* __restrict__ pu,
                                                  double a, double* __restrict__ u) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n3) return;
  const double r = rhs[q] + -1. * dp[q];
  u[q] = a * r + pu[q];
}

__global__ __launch_bounds__(256) void k_mm_copy(size_t n, const double* __restrict__ a, double* __restrict__ b) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n) b[q] = a[q];
}

// ------------------------------------------------------------ implicit form
struct Bc6 {
  int b[6];
};

// U with the wall components FormFunction_SNES zeroes (momentum.c:2323-2356):
// bctype 1 / -1 walls zero the normal flux on their wall face (i = 0 and i = mx - 2
// for the two i walls, likewise j and k), cavity (2) and couette (12) walls the j
// faces, slip (10) walls the i or j faces away from the boundary rows of the other
// axes.  Each point zeroes its own components (the rules written at i = mx - 1
// land on i = mx - 2).
__global__ __launch_bounds__(256) void k_mm_bc_zero(Dmda m, Bc6 bc, const double* __restrict__ u,
                                                    double* __restrict__ uz) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  const size_t p = m.idx(i, j, k);
  const int* b = bc.b;
  const bool ik = j != 0 && j != m.my - 1 && k != 0 && k != m.mz - 1;
  const bool jk = i != 0 && i != m.mx - 1 && k != 0 && k != m.mz - 1;
  const bool zx = (i == 0 && (b[0] == 1 || b[0] == -1)) || (i == m.mx - 2 && (b[1] == 1 || b[1] == -1)) ||
                  (b[0] == 10 && i == 0 && ik) || (b[1] == 10 && i == m.mx - 2 && ik);
  const bool zy = (j == 0 && (b[2] == 1 || b[2] == -1 || b[2] == 12)) ||
                  (j == m.my - 2 && (b[3] == 1 || b[3] == -1 || b[3] == 2 || b[3] == 12)) ||
                  (b[2] == 10 && j == 0 && jk) || ((b[3] == 10 || b[3] == -10) && j == m.my - 2 && jk);
  const bool zz = (k == 0 && (b[4] == 1 || b[4] == -1)) || (k == m.mz - 2 && (b[5] == 1 || b[5] == -1));
  uz[3 * p] = zx ? 0. : u[3 * p];
  uz[3 * p + 1] = zy ? 0. : u[3 * p + 1];
  uz[3 * p + 2] = zz ? 0. : u[3 * p + 2];
}

// IB_BC (momentum.c:2037-2220) on the Ucont that Formfunction_2 reads, for a body
// at rest: with immersed == 3 the faces between interior cells whose nvert sum
// exceeds 1.1 are zeroed (momentum.c:2158-2162), and the slip (bctype 10) faces
// next to the first and last interior cells on every row, ghost rows included
// (momentum.c:2190-2196).  The normal components' periodic ghost copies
// (momentum.c:2198-2213) are left out: Formfunction_2 reads Ucont through its own
// whole-vector periodic ghosts, which overwrite them.  A slip face on a periodic
// axis (where the reference's point order would matter) is refused by the caller.
__global__ __launch_bounds__(256) void k_mm_ib_bc(Dmda m, Bc6 bc, int immersed, const double* __restrict__ nv,
                                                  const double* __restrict__ u, double* __restrict__ ul) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  const size_t p = m.idx(i, j, k);
  const size_t sy = (size_t)m.mx, sz = (size_t)m.mx * m.my;
  const int* b = bc.b;
  const bool in = i > 0 && i < m.mx - 1 && j > 0 && j < m.my - 1 && k > 0 && k < m.mz - 1;
  const bool ib = immersed == 3 && in;
  const bool zx = (ib && nv[p + 1] + nv[p] > 1.1) || (b[0] == 10 && i == 0) || (b[1] == 10 && i == m.mx - 2);
  const bool zy = (ib && nv[p + sy] + nv[p] > 1.1) || (b[2] == 10 && j == 0) || (b[3] == 10 && j == m.my - 2);
  const bool zz = (ib && nv[p + sz] + nv[p] > 1.1) || (b[4] == 10 && k == 0) || (b[5] == 10 && k == m.mz - 2);
  ul[3 * p] = zx ? 0. : u[3 * p];
  ul[3 * p + 1] = zy ? 0. : u[3 * p + 1];
  ul[3 * p + 2] = zz ? 0. : u[3 * p + 2];
}

// the time terms, summed as the VecAXPY sequence from a zero vector does:
// -(Uz - Uo)/dt (coeff 1) or -(1.5 Uz - 2 Uo + 0.5 Urm1)/dt
__global__ __launch_bounds__(256) void k_mm_time(size_t n3, const double* __restrict__ uz,
                                                 const double* __restrict__ uo, const double* __restrict__ urm1,
                                                 int bdf1, double dt, double* __restrict__ f) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n3) return;
  double r = 0.;
  if (bdf1) {
    r = r + -1. / dt * uz[q];
    r = r + 1. / dt * uo[q];
  } else {
    r = r + -1.5 / dt * uz[q];
    r = r + 2. / dt * uo[q];
    r = r + -0.5 / dt * urm1[q];
  }
  f[q] = r;
}

// F += 0.5 RHS_o (coeff 1), F += -1 dP
__global__ __launch_bounds__(256) void k_mm_tail(size_t n3, const double* __restrict__ rhs_o,
                                                 const double* __restrict__ dp, double* __restrict__ f) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= n3) return;
  double r = f[q];
  if (rhs_o) r = r + 0.5 * rhs_o[q];
  f[q] = r + -1. * dp[q];
}

// w = (a - b) / h
__global__ __launch_bounds__(256) void k_mm_fd(size_t n3, const double* __restrict__ a, const double* __restrict__ b,
                                               double h, double* __restrict__ w) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n3) w[q] = (a[q] - b[q]) / h;
}

// y = x + a v
__global__ __launch_bounds__(256) void k_mm_axpby(size_t n3, const double* __restrict__ x,
                                                  const double* __restrict__ v, double a, double* __restrict__ y) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n3) y[q] = x[q] + a * v[q];
}

// y = a x
__global__ __launch_bounds__(256) void k_mm_scale(size_t n3, const double* __restrict__ x, double a,
                                                  double* __restrict__ y) {
  const size_t q = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (q < n3) y[q] = a * x[q];
}

// the wall components Implicit_MatrixFree zeroes in the solution
// (implicitsolver.c:408-415): bctype 1, -1, -2 and 10 on every face, also 2 on the
// j top, at i = 0 / mx - 2, j = 0 / my - 2, k = 0 / mz - 2
__global__ __launch_bounds__(256) void k_mm_bc_post(Dmda m, Bc6 bc, double* __restrict__ u) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  const size_t p = m.idx(i, j, k);
  auto wall = [](int b) { return b == 1 || b == -1 || b == -2 || b == 10; };
  const int* b = bc.b;
  if ((wall(b[0]) && i == 0) || (wall(b[1]) && i == m.mx - 2)) u[3 * p] = 0.;
  if ((wall(b[2]) && j == 0) || ((wall(b[3]) || b[3] == 2) && j == m.my - 2)) u[3 * p + 1] = 0.;
  if ((wall(b[4]) && k == 0) || (wall(b[5]) && k == m.mz - 2)) u[3 * p + 2] = 0.;
}

dim3 box_grid(const Dmda& m) { return dim3((m.mx + 63) / 64, (m.my + 3) / 4, m.mz); }

// Formfunction_2 on device arrays; ws holds 7 DMDA vector arrays (3 M doubles each)
void formfunction2(hipStream_t s, const Dmda& m, const MomIn& in, const double* ucont, const double* ucat,
                   double ren, double scale, double* ws, double* rhs, bool acc = false) {
  const size_t M3 = 3 * m.n();
  double* lu = ws;
  double* dv[3] = {ws + M3, ws + 2 * M3, ws + 3 * M3};
  double* vs[3] = {ws + 4 * M3, ws + 5 * M3, ws + 6 * M3};
  double* fp = ws + 7 * M3;
  const dim3 gr = box_grid(m), bl(64, 4);
  const double nu = 1. / ren;
  k_mm_local_ucont<<<gr, bl, 0, s>>>(m, ucont, lu);
  k_mm_faces<0><<<gr, bl, 0, s>>>(m, in, lu, ucat, nu, dv[0], vs[0]);
  k_mm_faces<1><<<gr, bl, 0, s>>>(m, in, lu, ucat, nu, dv[1], vs[1]);
  k_mm_faces<2><<<gr, bl, 0, s>>>(m, in, lu, ucat, nu, dv[2], vs[2]);
  if (m.per[0] || m.per[1] || m.per[2])
    k_mm_periodic<<<gr, bl, 0, s>>>(m, dv[0], dv[1], dv[2], vs[0], vs[1], vs[2], 6);
  k_mm_fp<<<gr, bl, 0, s>>>(m, in.nv, dv[0], dv[1], dv[2], vs[0], vs[1], vs[2], fp);
  if (m.per[0] || m.per[1] || m.per[2])
    k_mm_periodic<<<gr, bl, 0, s>>>(m, fp, nullptr, nullptr, nullptr, nullptr, nullptr, 1);
  k_mm_rhs<<<gr, bl, 0, s>>>(m, in, fp, scale, acc ? 1 : 0, rhs);
  HIPCHECK(hipGetLastError());
}

MomIn make_in(const vfs_grid* g, const double* const met[15], const double* nvert) {
  MomIn in{};
  for (int f = 0; f < 3; f++)
    for (int e = 0; e < 3; e++) in.fm[f][e] = met[3 * f + e];
  for (int f = 0; f < 3; f++) in.faj[f] = met[9 + f];
  for (int f = 0; f < 3; f++) in.cm[f] = met[12 + f];
  in.nv = nvert;
  in.bc5 = g->bctype[5];
  in.immersed = g->immersed;
  return in;
}

// The implicit momentum equation of one block on the device: its residual
// F(U) (FormFunction_SNES, momentum.c:2285-2440, immersed == 0) and a
// Newton-Krylov solve of F(U) = 0 in the shape of Implicit_MatrixFree
// (implicitsolver.c:281-390): Jacobian-free products by forward differences,
// unpreconditioned restarted GMRES(30) with an Eisenstat-Walker forcing term,
// relative tolerance rtol on ||F||.  The reference globalises with PETSc's trust
// region (SNESTR); this solver backtracks along the Newton step instead, so the
// iterates differ while the converged U satisfies the same test.
struct ImplicitMomentum {
  hipStream_t s;
  Dmda m;
  Dims d;
  MomIn in;
  const vfs_pp_handle* h;
  Bc6 bc;
  const double *dp, *uo, *urm1, *rhs_o;
  double ren, dt, coeff;
  size_t n3;
  Arena ar;
  double *ws = nullptr, *uz = nullptr, *ul = nullptr, *partial = nullptr, *red = nullptr;
  bool ib_bc = false;  // IB_BC changes the Ucont Formfunction_2 reads (immersed 3 or slip faces)
  double* hpin = nullptr;  // pinned host mirror of red
  double* hpin_dev = nullptr;
  int evals = 0;

  unsigned nb() const { return (unsigned)((n3 + 255) / 256); }

  void init() {
    n3 = 3 * m.n();
    ws = ar.alloc<double>(8 * n3);
    uz = ar.alloc<double>(n3);
    for (int f = 0; f < 6; f++) ib_bc = ib_bc || bc.b[f] == 10;
    ib_bc = ib_bc || in.immersed == 3;
    ul = ib_bc ? ar.alloc<double>(n3) : uz;
    partial = ar.alloc<double>((size_t)64 * 1024);
    red = ar.alloc<double>(128);
    HIPCHECK(hipHostMalloc(&hpin, 128 * sizeof(double)));
    HIPCHECK(hipHostGetDevicePointer((void**)&hpin_dev, hpin, 0));
  }
  ~ImplicitMomentum() {
    if (hpin) (void)hipHostFree(hpin);
  }

  // F(U); Uz and Ucat of U land in uz and ucat
  void residual(const double* u, double* ucat, double* f) {
    const dim3 grd = box_grid(m), blk(64, 4);
    k_mm_bc_zero<<<grd, blk, 0, s>>>(m, bc, u, uz);
    k_mm_local_ucont<<<grd, blk, 0, s>>>(m, uz, ws);
    HIPCHECK(hipGetLastError());
    launch_c2c(s, m, d, h->met9, h->cls, h->nvert, ws, ucat, h->snap, h->grid.bctype);
    const bool bdf1 = coeff > 0.9 && coeff < 1.1;
    k_mm_time<<<nb(), 256, 0, s>>>(n3, uz, uo, urm1 ? urm1 : uo, bdf1 ? 1 : 0, dt, f);
    HIPCHECK(hipGetLastError());
    if (ib_bc) {
      k_mm_ib_bc<<<grd, blk, 0, s>>>(m, bc, in.immersed, in.nv, uz, ul);
      HIPCHECK(hipGetLastError());
    }
    formfunction2(s, m, in, ul, ucat, ren, bdf1 ? 0.5 : 1.0, ws, f, true);
    k_mm_tail<<<nb(), 256, 0, s>>>(n3, bdf1 ? rhs_o : nullptr, dp, f);
    HIPCHECK(hipGetLastError());
    evals++;
  }

  double norm(const double* a) {
    int nblk = 0;
    launch_norm2(s, n3, a, partial, &nblk);
    launch_finalize(s, partial, 1, nblk, red, 0, hpin_dev);
    HIPCHECK(hipStreamSynchronize(s));
    return std::sqrt(hpin[0]);
  }

  // h = V[0..nv)^T w into red[0..nv) and hpin[0..nv)
  void dots(const double* w, const double* V, int nv) {
    int nblk = 0;
    launch_multidot(s, n3, w, V, n3, nv, partial, &nblk);
    launch_finalize(s, partial, nv, nblk, red, 0, hpin_dev);
    HIPCHECK(hipStreamSynchronize(s));
  }

  // Newton-Krylov solve in place on u; returns the stats
  vfs_pp_newton_stats solve(double* u, double* ucat, double rtol, int max_newton) {
    constexpr int M = 30;           // GMRES restart
    constexpr int MAX_KRYLOV = 1000;  // per linear solve
    const double err = std::sqrt(std::numeric_limits<double>::epsilon());
    const double t0 = now_ms();
    vfs_pp_newton_stats st{};
    double* V = ar.alloc<double>((size_t)(M + 1) * n3);
    double* F = ar.alloc<double>(n3);
    double* Ft = ar.alloc<double>(n3);
    double* Ut = ar.alloc<double>(n3);
    double* S = ar.alloc<double>(n3);
    double* ydev = ar.alloc<double>(M);
    std::vector<double> H((size_t)(M + 1) * M), cs(M), sn(M), gv(M + 1), y(M);
    residual(u, ucat, F);
    double fnorm = norm(F);
    st.fnorm0 = fnorm;
    if (!std::isfinite(fnorm)) throw StatusError(VFS_PP_ENONFINITE, "non-finite momentum residual");
    double eta_prev = 0.3, fprev = fnorm;
    for (int it = 0; it < max_newton; it++) {
      if (fnorm <= rtol * st.fnorm0 || fnorm == 0.) { st.converged = 1; break; }
      // Eisenstat-Walker forcing (version 3: gamma 1, alpha (1 + sqrt 5)/2, safeguard
      // above 0.1, at most 0.9; 0.3 at the first step)
      double eta = 0.3;
      if (it > 0) {
        const double alpha = 0.5 * (1. + std::sqrt(5.));
        eta = std::pow(fnorm / fprev, alpha);
        const double guard = std::pow(eta_prev, alpha);
        if (guard > 0.1) eta = std::max(eta, guard);
        eta = std::min(eta, 0.9);
      }
      eta_prev = eta;
      const double unorm = norm(u);
      // GMRES(M) on J s = -F from s = 0
      launch_fill(s, n3, S, 0.);
      double rnorm = fnorm;
      const double target = eta * fnorm;
      int kits = 0;
      bool first = true;
      while (kits < MAX_KRYLOV) {
        if (first) {
          k_mm_scale<<<nb(), 256, 0, s>>>(n3, F, -1. / rnorm, V);  // r0 = -F
          first = false;
        } else {
          // true residual r = -F - J s
          const double sn2 = norm(S);
          const double hfd = err * std::sqrt(1. + unorm) / sn2;
          k_mm_axpby<<<nb(), 256, 0, s>>>(n3, u, S, hfd, Ut);
          residual(Ut, ucat, Ft);
          k_mm_fd<<<nb(), 256, 0, s>>>(n3, Ft, F, hfd, V);  // J s
          k_mm_axpby<<<nb(), 256, 0, s>>>(n3, F, V, 1., V);  // F + J s
          rnorm = norm(V);
          if (rnorm <= target) break;
          k_mm_scale<<<nb(), 256, 0, s>>>(n3, V, -1. / rnorm, V);
        }
        HIPCHECK(hipGetLastError());
        std::fill(gv.begin(), gv.end(), 0.);
        gv[0] = rnorm;
        int j = 0;
        bool conv = false;
        for (; j < M && kits < MAX_KRYLOV; j++, kits++) {
          double* vj = V + (size_t)j * n3;
          double* w = V + (size_t)(j + 1) * n3;
          const double hfd = err * std::sqrt(1. + unorm);  // |v_j| = 1
          k_mm_axpby<<<nb(), 256, 0, s>>>(n3, u, vj, hfd, Ut);
          residual(Ut, ucat, Ft);
          k_mm_fd<<<nb(), 256, 0, s>>>(n3, Ft, F, hfd, w);
          HIPCHECK(hipGetLastError());
          for (int l = 0; l <= j; l++) H[(size_t)l * M + j] = 0.;
          for (int pass = 0; pass < 2; pass++) {  // classical Gram-Schmidt, twice
            dots(w, V, j + 1);
            for (int l = 0; l <= j; l++) H[(size_t)l * M + j] += hpin[l];
            launch_axpy_multi(s, n3, w, V, n3, j + 1, red, -1.);
          }
          const double hn = norm(w);
          H[(size_t)(j + 1) * M + j] = hn;
          if (hn > 0.) k_mm_scale<<<nb(), 256, 0, s>>>(n3, w, 1. / hn, w);
          for (int l = 0; l < j; l++) {
            const double a = H[(size_t)l * M + j], b = H[(size_t)(l + 1) * M + j];
            H[(size_t)l * M + j] = cs[l] * a + sn[l] * b;
            H[(size_t)(l + 1) * M + j] = -sn[l] * a + cs[l] * b;
          }
          const double a = H[(size_t)j * M + j], b = H[(size_t)(j + 1) * M + j];
          const double r = std::sqrt(a * a + b * b);
          cs[j] = r == 0. ? 1. : a / r;
          sn[j] = r == 0. ? 0. : b / r;
          H[(size_t)j * M + j] = r;
          H[(size_t)(j + 1) * M + j] = 0.;
          gv[j + 1] = -sn[j] * gv[j];
          gv[j] = cs[j] * gv[j];
          if (std::fabs(gv[j + 1]) <= target || hn == 0.) {
            j++;
            kits++;
            conv = true;
            break;
          }
        }
        for (int l = j - 1; l >= 0; l--) {
          double acc = gv[l];
          for (int q = l + 1; q < j; q++) acc -= H[(size_t)l * M + q] * y[q];
          y[l] = acc / H[(size_t)l * M + l];
        }
        if (j > 0) {
          HIPCHECK(hipMemcpyAsync(ydev, y.data(), j * sizeof(double), hipMemcpyHostToDevice, s));
          launch_axpy_multi(s, n3, S, V, n3, j, ydev, 1.);
        }
        if (conv) break;
      }
      st.krylov_iterations += kits;
      // backtracking along the Newton step
      double lam = 1., ft = 0.;
      for (int ls = 0; ls < 12; ls++) {
        k_mm_axpby<<<nb(), 256, 0, s>>>(n3, u, S, lam, Ut);
        residual(Ut, ucat, Ft);
        ft = norm(Ft);
        if (std::isfinite(ft) && ft <= (1. - 1e-4 * lam) * fnorm) break;
        lam *= 0.5;
      }
      if (!std::isfinite(ft)) throw StatusError(VFS_PP_ENONFINITE, "non-finite momentum residual");
      const double snorm = lam * norm(S);
      launch_copy(s, n3, Ut, u);
      launch_copy(s, n3, Ft, F);
      fprev = fnorm;
      fnorm = ft;
      st.newton_iterations++;
      if (fnorm <= rtol * st.fnorm0 || fnorm == 0.) { st.converged = 1; break; }
      if (snorm < 1e-8 * norm(u)) { st.converged = 1; break; }  // step below stol
    }
    st.fnorm = fnorm;
    // the solution's wall faces (implicitsolver.c:408-415) and its Ucat
    k_mm_bc_post<<<box_grid(m), dim3(64, 4), 0, s>>>(m, bc, u);
    k_mm_local_ucont<<<box_grid(m), dim3(64, 4), 0, s>>>(m, u, ws);
    HIPCHECK(hipGetLastError());
    launch_c2c(s, m, d, h->met9, h->cls, h->nvert, ws, ucat, h->snap, h->grid.bctype);
    HIPCHECK(hipStreamSynchronize(s));
    st.residual_evals = evals;
    st.ms_total = now_ms() - t0;
    return st;
  }
};

// checks shared by the implicit entry points
void implicit_setup(vfs_pp_handle* h, const double* const met[15], const double* dp, const double* uo,
                    const double* urm1, double ren, double dt, double coeff, ImplicitMomentum& P) {
  ensure_unbound(h);
  ensure_geom(h);
  for (int f = 0; f < 6; f++) {
    const int b = h->grid.bctype[f];
    if (!(b == 0 || b == 1 || b == -1 || b == 2 || b == 10 || b == -10 || b == 12))
      throw StatusError(VFS_PP_EINVAL, "unsupported bctype for the implicit momentum form");
    if ((b == 10 || b == -10) && h->grid.periodic[f / 2])
      throw StatusError(VFS_PP_EINVAL, "a slip boundary on a periodic axis");
  }
  if (!(coeff > 0.9 && coeff < 1.1) && !urm1)
    throw StatusError(VFS_PP_EINVAL, "coeff 1.5 needs Ucont of the step before the previous one");
  P.s = h->s;
  P.m = h->m;
  P.d = h->d;
  vfs_grid gr = h->grid;
  gr.immersed = h->immersed_ref;
  P.in = make_in(&gr, met, h->nvert);
  P.h = h;
  for (int f = 0; f < 6; f++) P.bc.b[f] = h->grid.bctype[f];
  P.dp = dp;
  P.uo = uo;
  P.urm1 = urm1;
  P.ren = ren;
  P.dt = dt;
  P.coeff = coeff;
  P.init();
}

}  // namespace
}  // namespace pp

using namespace pp;

int vfs_pp_formfunction2_device(const vfs_grid* g, const double* icsi, const double* ieta, const double* izet,
                                const double* jcsi, const double* jeta, const double* jzet, const double* kcsi,
                                const double* keta, const double* kzet, const double* iaj, const double* jaj,
                                const double* kaj, const double* csi, const double* eta, const double* zet,
                                const double* nvert, const double* ucont, const double* ucat, double ren,
                                double scale, double* rhs) {
  ABI_BEGIN
  if (!g || !nvert || !ucont || !ucat || !rhs || !(ren > 0.)) return VFS_PP_EINVAL;
  const double* met[15] = {icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, csi, eta, zet};
  for (auto p : met)
    if (!p) return VFS_PP_EINVAL;
  if (g->nx < 2 || g->ny < 2 || g->nz < 2) return VFS_PP_EINVAL;
  for (int a = 0; a < 3; a++)
    if (g->periodic[a] && (a == 0 ? g->nx : a == 1 ? g->ny : g->nz) < 3) return VFS_PP_EINVAL;
  const Dmda m{g->nx + 2, g->ny + 2, g->nz + 2, {g->periodic[0] ? 1 : 0, g->periodic[1] ? 1 : 0, g->periodic[2] ? 1 : 0}};
  Arena tmp;
  double* ws = tmp.alloc<double>(8 * 3 * m.n());
  formfunction2(nullptr, m, make_in(g, met, nvert), ucont, ucat, ren, scale, ws, rhs);
  HIPCHECK(hipStreamSynchronize(nullptr));
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_runge_kutta_device(vfs_pp_handle* h, const double* icsi, const double* ieta, const double* izet,
                              const double* jcsi, const double* jeta, const double* jzet, const double* kcsi,
                              const double* keta, const double* kzet, const double* iaj, const double* jaj,
                              const double* kaj, const double* csi, const double* eta, const double* zet,
                              const double* dp, double* ucont, double* ucat, double ren, double dt, double st) {
  ABI_BEGIN
  if (!h || !dp || !ucont || !ucat || !(ren > 0.) || !(dt > 0.)) return VFS_PP_EINVAL;
  const double* met[15] = {icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, csi, eta, zet};
  for (auto p : met)
    if (!p) return VFS_PP_EINVAL;
  ensure_unbound(h);
  ensure_geom(h);
  hipStream_t s = h->s;
  const Dmda m = h->m;
  const size_t M3 = 3 * m.n();
  vfs_grid gr = h->grid;
  gr.immersed = h->immersed_ref;
  const MomIn in = make_in(&gr, met, h->nvert);
  Arena tmp;
  double* ws = tmp.alloc<double>(8 * M3);
  double* pu = tmp.alloc<double>(M3);
  double* rhs = tmp.alloc<double>(M3);
  const unsigned nb = (unsigned)((M3 + 255) / 256);
  const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
  // Contra2Cart reads Ucont with fresh whole-vector periodic ghosts (rhs.c:115-142)
  // in a local copy; the caller's ghost faces keep their values, as the
  // reference's global Ucont does
  double* lu = ws;
  const dim3 grd = box_grid(m), blk(64, 4);
  auto contra2cart = [&]() {
    k_mm_local_ucont<<<grd, blk, 0, s>>>(m, ucont, lu);
    HIPCHECK(hipGetLastError());
    launch_c2c(s, m, h->d, h->met9, h->cls, h->nvert, lu, ucat, h->snap, h->grid.bctype);
  };
  k_mm_copy<<<nb, 256, 0, s>>>(M3, ucont, pu);
  for (int q = 0; q < 4; q++) {
    contra2cart();
    formfunction2(s, m, in, ucont, ucat, ren, 1.0, ws, rhs);
    const double a = alfa[q] * dt * st;
    k_mm_stage<<<nb, 256, 0, s>>>(M3, rhs, dp, pu, a, ucont);
    HIPCHECK(hipGetLastError());
    contra2cart();
  }
  HIPCHECK(hipStreamSynchronize(s));
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_momentum_residual_device(vfs_pp_handle* h, const double* icsi, const double* ieta, const double* izet,
                                    const double* jcsi, const double* jeta, const double* jzet, const double* kcsi,
                                    const double* keta, const double* kzet, const double* iaj, const double* jaj,
                                    const double* kaj, const double* csi, const double* eta, const double* zet,
                                    const double* dp, const double* ucont_o, const double* ucont_rm1,
                                    const double* rhs_o, const double* ucont, double ren, double dt, double coeff,
                                    double* ucat, double* f) {
  ABI_BEGIN
  if (!h || !dp || !ucont_o || !ucont || !ucat || !f || !(ren > 0.) || !(dt > 0.)) return VFS_PP_EINVAL;
  const double* met[15] = {icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, csi, eta, zet};
  for (auto p : met)
    if (!p) return VFS_PP_EINVAL;
  ImplicitMomentum P;
  implicit_setup(h, met, dp, ucont_o, ucont_rm1, ren, dt, coeff, P);
  P.rhs_o = rhs_o;
  P.residual(ucont, ucat, f);
  HIPCHECK(hipStreamSynchronize(P.s));
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_implicit_momentum_device(vfs_pp_handle* h, const double* icsi, const double* ieta, const double* izet,
                                    const double* jcsi, const double* jeta, const double* jzet, const double* kcsi,
                                    const double* keta, const double* kzet, const double* iaj, const double* jaj,
                                    const double* kaj, const double* csi, const double* eta, const double* zet,
                                    const double* dp, const double* ucont_o, const double* ucont_rm1,
                                    const double* rhs_o, double* ucont, double* ucat, double ren, double dt,
                                    double coeff, double rtol, int max_newton, vfs_pp_newton_stats* stats) {
  ABI_BEGIN
  if (!h || !dp || !ucont_o || !ucont || !ucat || !(ren > 0.) || !(dt > 0.) || !(rtol > 0.) || max_newton < 1)
    return VFS_PP_EINVAL;
  const double* met[15] = {icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, csi, eta, zet};
  for (auto p : met)
    if (!p) return VFS_PP_EINVAL;
  ImplicitMomentum P;
  implicit_setup(h, met, dp, ucont_o, ucont_rm1, ren, dt, coeff, P);
  P.rhs_o = rhs_o;
  const vfs_pp_newton_stats st = P.solve(ucont, ucat, rtol, max_newton);
  if (stats) *stats = st;
  return st.converged ? VFS_PP_OK : VFS_PP_ENOCONV;
  ABI_END
}

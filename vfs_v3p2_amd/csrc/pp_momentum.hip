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

// the contravariant right-hand side (momentum.c:1750-1965) at every DMDA point
__global__ __launch_bounds__(256) void k_mm_rhs(Dmda m, MomIn in, const double* __restrict__ fp, double scale,
                                                double* __restrict__ rhs) {
  int i, j, k;
  if (!dmda_point(m, i, j, k)) return;
  const size_t st[3] = {1, (size_t)m.mx, (size_t)m.mx * m.my};
  const size_t p = m.idx(i, j, k);
  const int pos[3] = {i, j, k}, mq[3] = {m.mx, m.my, m.mz};
  double out[3] = {0., 0., 0.};  // VecSet(Rhs, 0)
  const bool interior = i > 0 && j > 0 && k > 0 && i < m.mx - 1 && j < m.my - 1 && k < m.mz - 1;
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

dim3 box_grid(const Dmda& m) { return dim3((m.mx + 63) / 64, (m.my + 3) / 4, m.mz); }

// Formfunction_2 on device arrays; ws holds 7 DMDA vector arrays (3 M doubles each)
void formfunction2(hipStream_t s, const Dmda& m, const MomIn& in, const double* ucont, const double* ucat,
                   double ren, double scale, double* ws, double* rhs) {
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
  k_mm_rhs<<<gr, bl, 0, s>>>(m, in, fp, scale, rhs);
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

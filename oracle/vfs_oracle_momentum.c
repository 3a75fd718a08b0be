/*
 * CPU oracle, momentum predictor -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C FP64 restatement of the explicit momentum step of VFS_v3p2 that
 * produces the intermediate Ucont* the pressure projection then corrects:
 *   Formfunction_2 (momentum.c:529-2034) -- convective and viscous right-hand
 *     side in contravariant form, the reference's default branch: no LES/RANS,
 *     no level set, no skew-symmetric form, fourth-order central convection
 *     (second_order = 0, inviscid = 0, laplacian = 0, clark = 0), no o-cylinder
 *     force bookkeeping (bctype 11);
 *   RungeKutta (rhs.c:1511-1631) -- the four-stage explicit update with
 *     alfa = 1/4, 1/3, 1/2, 1 around Contra2Cart (rhs.c:50), without the
 *     inflow/outflow boundary (InflowFlux, FormBCS) and the IB velocity
 *     interpolation (ibm_interpolation_advanced), whose results the caller's
 *     Ucat carries.
 * Operand order follows the reference so that a device restatement compiled
 * without contraction agrees bit for bit.  Same DMDA layout as vfs_oracle.c.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "vfs_oracle.h"

#define IXM(g, i, j, k) ((((size_t)(k)) * (size_t)(g)->my + (size_t)(j)) * (size_t)(g)->mx + (size_t)(i))

/* periodic image of a ghost point (the `if(periodic)` copies, momentum.c:713-740,
   1556-1582, 1724-1748): plane 0 takes plane m-2 and plane m-1 takes plane 1
   along every periodic axis; returns 0 for a point that is no ghost */
static int mm_image(const or_grid *g, int i, int j, int k, size_t *src) {
  int a = i, b = j, c = k, f = 0;
  if (g->per[0] && i == 0) a = g->mx - 2, f = 1;
  else if (g->per[0] && i == g->mx - 1) a = 1, f = 1;
  if (g->per[1] && j == 0) b = g->my - 2, f = 1;
  else if (g->per[1] && j == g->my - 1) b = 1, f = 1;
  if (g->per[2] && k == 0) c = g->mz - 2, f = 1;
  else if (g->per[2] && k == g->mz - 1) c = 1, f = 1;
  *src = IXM(g, a, b, c);
  return f;
}

/* whole-vector copies into the ghost points of a DMDA vector field */
static void mm_periodic_vec(const or_grid *g, double *v) {
  if (!(g->per[0] || g->per[1] || g->per[2])) return;
  for (int k = 0; k < g->mz; k++)
    for (int j = 0; j < g->my; j++)
      for (int i = 0; i < g->mx; i++) {
        size_t s, p = IXM(g, i, j, k);
        if (!mm_image(g, i, j, k, &s)) continue;
        v[3 * p] = v[3 * s];
        v[3 * p + 1] = v[3 * s + 1];
        v[3 * p + 2] = v[3 * s + 2];
      }
}

/* Compute_du_i / _j / _k (inline.h:406-611) for the face d+1/2 of point p:
   du[c][e] = d U_c / d xi_e (c: velocity component, e: computational axis).
   Along d the plain difference; along another axis e the difference of the
   two-point face averages, one-sided (x 0.5) away from a blocked side (upper
   tested first) and central (x 0.25) otherwise. */
static void mm_du(const size_t st[3], int d, size_t p, const double *u, const double *nv, double du[3][3]) {
  const size_t sd = st[d];
  for (int c = 0; c < 3; c++) du[c][d] = u[3 * (p + sd) + c] - u[3 * p + c];
  for (int e = 0; e < 3; e++) {
    if (e == d) continue;
    const size_t se = st[e];
    if (nv[p + se] > 0.1 || nv[p + sd + se] > 0.1) {
      for (int c = 0; c < 3; c++)
        du[c][e] = (u[3 * (p + sd) + c] + u[3 * p + c] - u[3 * (p + sd - se) + c] - u[3 * (p - se) + c]) * 0.5;
    } else if (nv[p - se] > 0.1 || nv[p + sd - se] > 0.1) {
      for (int c = 0; c < 3; c++)
        du[c][e] =
            (u[3 * (p + sd + se) + c] + u[3 * (p + se) + c] - u[3 * (p + sd) + c] - u[3 * p + c]) * 0.5;
    } else {
      for (int c = 0; c < 3; c++)
        du[c][e] = (u[3 * (p + sd + se) + c] + u[3 * (p + se) + c] - u[3 * (p + sd - se) + c] -
                    u[3 * (p - se) + c]) * 0.25;
    }
  }
}

/* The convective and viscous fluxes of the faces d+1/2 (the i, j and k loops of
   momentum.c:743-1010, 1011-1250, 1251-1521). fm[0..2]: the face family's csi,
   eta, zet vectors; faj its Jacobian. */
static void mm_faces(const or_grid *g, int d, const int bctype[6], int immersed, const double *const fm[3],
                     const double *faj, const double *nv, const double *uc, const double *ucat, double nu,
                     double *div, double *visc) {
  const int m[3] = {g->mx, g->my, g->mz};
  const size_t st[3] = {1, (size_t)g->mx, (size_t)g->mx * g->my};
  for (int k = 0; k < g->mz; k++)
    for (int j = 0; j < g->my; j++)
      for (int i = 0; i < g->mx; i++) {
        const int pos[3] = {i, j, k};
        if (i == g->mx - 1 || j == g->my - 1 || k == g->mz - 1) continue;
        int skip = 0;
        for (int a = 0; a < 3; a++)
          if (a != d && pos[a] == 0) skip = 1;
        if (skip) continue;
        const size_t p = IXM(g, i, j, k), sd = st[d];
        const double ajc = faj[p];
        const double *m0 = &fm[0][3 * p], *m1 = &fm[1][3 * p], *m2 = &fm[2][3 * p];
        const double *nrm = d == 0 ? m0 : (d == 1 ? m1 : m2);
        double du[3][3];
        mm_du(st, d, p, ucat, nv, du);
        /* g11, g21, g31: (csi, eta, zet) . n with the metric's component first */
        const double g1 = m0[0] * nrm[0] + m0[1] * nrm[1] + m0[2] * nrm[2];
        const double g2 = m1[0] * nrm[0] + m1[1] * nrm[1] + m1[2] * nrm[2];
        const double g3 = m2[0] * nrm[0] + m2[1] * nrm[1] + m2[2] * nrm[2];
        /* r[c][x] = du_c/dc * csi_x + du_c/de * eta_x + du_c/dz * zet_x (r11 = r[0][0], r21 = r[1][0], ...) */
        double r[3][3];
        for (int c = 0; c < 3; c++)
          for (int x = 0; x < 3; x++) r[c][x] = du[c][0] * m0[x] + du[c][1] * m1[x] + du[c][2] * m2[x];
        /* fourth-order stencil ends, moved in at walls and beside IB points */
        const int pd = pos[d], md = m[d];
        size_t L = p - sd, R = p + 2 * sd;
        if (pd == 0 || pd == md - 2) {
          if (g->per[d]) {
            L = (size_t)((long)p + (long)(md - 3 - pd) * (long)sd); /* plane m-3 */
            R = (size_t)((long)p + (long)(2 - pd) * (long)sd);      /* plane 2 */
          } else {
            L = p;
            R = p + sd;
          }
        } else if (nv[L] + nv[R] > 0.1) {
          L = p;
          R = p + sd;
        }
        double ucon = uc[3 * p + d];
        if (g->per[d] && (pd == 0 || pd == md - 2)) ucon = uc[3 * (p + (size_t)(md - 2 - pd) * sd) + d];
        if (d == 2 && k == g->mz - 2 && bctype[5] == 4 && (int)nv[p] == 0) ucon = uc[3 * (p - sd) + 2];
        const double up = -0.5 * (ucon + fabs(ucon));
        const double um = -0.5 * (ucon - fabs(ucon));
        double *dv = &div[3 * p];
        if (immersed && pd != md - 2 && nv[p] > 0.1) {
          for (int c = 0; c < 3; c++)
            dv[c] = um * (0.125 * (-ucat[3 * (p + 2 * sd) + c] - 2. * ucat[3 * (p + sd) + c] + 3. * ucat[3 * p + c]) +
                          ucat[3 * (p + sd) + c]) +
                    up * (0.125 * (-ucat[3 * p + c] - 2. * ucat[3 * p + c] + 3. * ucat[3 * (p + sd) + c]) +
                          ucat[3 * p + c]);
        } else if (immersed && pd != 0 && nv[p + sd] > 0.1) {
          for (int c = 0; c < 3; c++)
            dv[c] = um * (0.125 * (-ucat[3 * (p + sd) + c] - 2. * ucat[3 * (p + sd) + c] + 3. * ucat[3 * p + c]) +
                          ucat[3 * (p + sd) + c]) +
                    up * (0.125 * (-ucat[3 * (p - sd) + c] - 2. * ucat[3 * p + c] + 3. * ucat[3 * (p + sd) + c]) +
                          ucat[3 * p + c]);
        } else {
          for (int c = 0; c < 3; c++)
            dv[c] = -uc[3 * p + d] * 0.0625 *
                    (-ucat[3 * L + c] + 9. * ucat[3 * p + c] + 9. * ucat[3 * (p + sd) + c] - ucat[3 * R + c]);
        }
        if (nv[p] + nv[p + sd] > 0.1 && immersed == 3) dv[0] = dv[1] = dv[2] = 0;
        /* visc = 0, then += the laminar flux (les, rans, clark off) */
        double *vs = &visc[3 * p];
        for (int c = 0; c < 3; c++) {
          double v = 0;
          v += (g1 * du[c][0] + g2 * du[c][1] + g3 * du[c][2] + r[0][c] * nrm[0] + r[1][c] * nrm[1] +
                r[2][c] * nrm[2]) *
               ajc * (nu);
          vs[c] = v;
        }
      }
}

/* fourth-order stencil ends of the cell difference along one axis
   (momentum.c:1605-1662): returns 1/denom, sets the R and L offsets in planes */
static double mm_cell_ends(int per, int pos, int m, const double *nv, size_t p, size_t s, long *rr, long *ll) {
  long R = 1, Lo = -2;
  double denom = 3.;
  if (pos == 1) {
    if (per) Lo = (long)(m - 3) - pos;
    else R = 0, Lo = -1, denom = 1.;
  } else if (pos == 2 || pos == m - 3) {
    if (!per) R = 0, Lo = -1, denom = 1.;
  } else if (pos == m - 2) {
    if (per) R = 1 - pos;
    else R = 0, Lo = -1, denom = 1.;
  }
  if (nv[p - s] + nv[p] + nv[p + s] > 0.1) R = 0, Lo = -1, denom = 1.;
  *rr = R;
  *ll = Lo;
  return 1. / denom;
}

/* acc = 0: Rhs starts at zero (the caller's VecSet, rhs.c:1552); acc = 1: the
   terms add to what Rhs holds (FormFunction_SNES, momentum.c:2410-2413) */
static void mm_formfunction2(const or_grid *g, const int bctype[6], int immersed, const double *icsi,
                             const double *ieta, const double *izet, const double *jcsi, const double *jeta,
                             const double *jzet, const double *kcsi, const double *keta, const double *kzet,
                             const double *iaj, const double *jaj, const double *kaj, const double *csi,
                             const double *eta, const double *zet, const double *nvert, const double *ucont,
                             const double *ucat, double ren, double scale, double *rhs, int acc) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  const size_t n = (size_t)mx * my * mz;
  const size_t st[3] = {1, (size_t)mx, (size_t)mx * my};
  /* lUcont with its whole-vector periodic ghosts (momentum.c:713-740) */
  double *uc = (double *)malloc(3 * n * sizeof(double));
  memcpy(uc, ucont, 3 * n * sizeof(double));
  mm_periodic_vec(g, uc);
  double *div[3], *visc[3];
  for (int d = 0; d < 3; d++) {
    div[d] = (double *)calloc(3 * n, sizeof(double));
    visc[d] = (double *)calloc(3 * n, sizeof(double));
  }
  double *fp = (double *)calloc(3 * n, sizeof(double));
  const double nu = 1. / ren;
  const double *fmi[3] = {icsi, ieta, izet}, *fmj[3] = {jcsi, jeta, jzet}, *fmk[3] = {kcsi, keta, kzet};
  mm_faces(g, 0, bctype, immersed, fmi, iaj, nvert, uc, ucat, nu, div[0], visc[0]);
  mm_faces(g, 1, bctype, immersed, fmj, jaj, nvert, uc, ucat, nu, div[1], visc[1]);
  mm_faces(g, 2, bctype, immersed, fmk, kaj, nvert, uc, ucat, nu, div[2], visc[2]);
  for (int d = 0; d < 3; d++) {
    mm_periodic_vec(g, div[d]);
    mm_periodic_vec(g, visc[d]);
  }
  /* Fp at the interior cells, momentum.c:1584-1718 */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        const size_t p = IXM(g, i, j, k);
        const int pos[3] = {i, j, k}, m[3] = {mx, my, mz};
        double dv[3];
        for (int c = 0; c < 3; c++)
          dv[c] = (div[0][3 * p + c] - div[0][3 * (p - 1) + c] + div[1][3 * p + c] - div[1][3 * (p - st[1]) + c] +
                   div[2][3 * p + c] - div[2][3 * (p - st[2]) + c]);
        long R[3], L[3];
        double inv[3];
        for (int a = 0; a < 3; a++) inv[a] = mm_cell_ends(g->per[a], pos[a], m[a], nvert, p, st[a], &R[a], &L[a]);
        double d4[3] = {0, 0, 0};
        for (int a = 0; a < 3; a++) {
          const size_t pr = (size_t)((long)p + R[a] * (long)st[a]), pl = (size_t)((long)p + L[a] * (long)st[a]);
          for (int c = 0; c < 3; c++) d4[c] += (div[a][3 * pr + c] - div[a][3 * pl + c]) * inv[a];
        }
        for (int c = 0; c < 3; c++) fp[3 * p + c] = 9. / 8. * dv[c] + -1. / 8. * d4[c];
        const double rr = 1.0;
        for (int c = 0; c < 3; c++)
          fp[3 * p + c] += (visc[0][3 * p + c] - visc[0][3 * (p - 1) + c] + visc[1][3 * p + c] -
                            visc[1][3 * (p - st[1]) + c] + visc[2][3 * p + c] - visc[2][3 * (p - st[2]) + c]) /
                           rr;
      }
  mm_periodic_vec(g, fp);
  /* Rhs (momentum.c:1750-1866): += on what the caller left there */
  if (!acc) memset(rhs, 0, 3 * n * sizeof(double));
  const double *cm[3] = {csi, eta, zet}, *caj[3] = {iaj, jaj, kaj};
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        const size_t p = IXM(g, i, j, k);
        const int pos[3] = {i, j, k}, m[3] = {mx, my, mz};
        for (int d = 0; d < 3; d++) {
          const double *a0 = &cm[d][3 * p], *a1 = &cm[d][3 * (p + st[d])];
          const double *f0 = &fp[3 * p], *f1 = &fp[3 * (p + st[d])];
          rhs[3 * p + d] += scale * (0.5 * (a0[0] * f0[0] + a0[1] * f0[1] + a0[2] * f0[2]) +
                                     0.5 * (a1[0] * f1[0] + a1[1] * f1[1] + a1[2] * f1[2])) *
                            caj[d][p];
        }
        for (int d = 0; d < 3; d++)
          if (nvert[p] + nvert[p + st[d]] > 0.1 || (!g->per[d] && pos[d] == m[d] - 2)) rhs[3 * p + d] = 0;
      }
  /* boundary planes (momentum.c:1883-1965) */
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        const size_t p = IXM(g, i, j, k);
        const int pos[3] = {i, j, k}, m[3] = {mx, my, mz};
        for (int a = 0; a < 3; a++) {
          if (pos[a] == 0 || pos[a] == m[a] - 1) rhs[3 * p] = rhs[3 * p + 1] = rhs[3 * p + 2] = 0;
          if (pos[a] == m[a] - 2 && !g->per[a]) rhs[3 * p + a] = 0;
        }
      }
  free(uc);
  for (int d = 0; d < 3; d++) {
    free(div[d]);
    free(visc[d]);
  }
  free(fp);
}

void or_formfunction2(const or_grid *g, const int bctype[6], int immersed, const double *icsi,
                      const double *ieta, const double *izet, const double *jcsi, const double *jeta,
                      const double *jzet, const double *kcsi, const double *keta, const double *kzet,
                      const double *iaj, const double *jaj, const double *kaj, const double *csi,
                      const double *eta, const double *zet, const double *nvert, const double *ucont,
                      const double *ucat, double ren, double scale, double *rhs) {
  mm_formfunction2(g, bctype, immersed, icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, csi,
                   eta, zet, nvert, ucont, ucat, ren, scale, rhs, 0);
}

void or_runge_kutta(const or_grid *g, const int bctype[6], int immersed, const double *icsi, const double *ieta,
                    const double *izet, const double *jcsi, const double *jeta, const double *jzet,
                    const double *kcsi, const double *keta, const double *kzet, const double *iaj,
                    const double *jaj, const double *kaj, const double *csi, const double *eta, const double *zet,
                    const double *nvert, const double *dp, double *ucont, double *ucat, double ren, double dt,
                    double st) {
  const size_t n3 = 3 * (size_t)g->mx * g->my * g->mz;
  const double alfa[4] = {0.25, 1 / 3., 0.5, 1.};
  double *pu = (double *)malloc(n3 * sizeof(double));
  double *rhs = (double *)malloc(n3 * sizeof(double));
  memcpy(pu, ucont, n3 * sizeof(double));
  for (int s = 0; s < 4; s++) {
    or_contra2cart(g, bctype, csi, eta, zet, nvert, ucont, ucat);
    or_formfunction2(g, bctype, immersed, icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj,
                     csi, eta, zet, nvert, ucont, ucat, ren, 1.0, rhs);
    for (size_t q = 0; q < n3; q++) rhs[q] = rhs[q] + -1. * dp[q]; /* VecAXPY(Rhs, -1, dP) */
    const double a = alfa[s] * dt * st;
    for (size_t q = 0; q < n3; q++) ucont[q] = a * rhs[q] + pu[q]; /* VecWAXPY */
    or_contra2cart(g, bctype, csi, eta, zet, nvert, ucont, ucat);
  }
  free(pu);
  free(rhs);
}

/* ------------------------------------------------------- implicit momentum
 * The nonlinear residual the reference's matrix-free SNES drives to zero
 * (FormFunction_SNES, momentum.c:2285-2440, called by Implicit_MatrixFree,
 * implicitsolver.c:281), for a body at rest (movefsi = rotatefsi = 0), no wall
 * model, forcing or level set:
 *   Uz  = U with the wall components zeroed (momentum.c:2323-2356: bctype 1 and
 *         -1 on every face, cavity 2 and couette 12 on j, slip 10 on i and j),
 *   Ucat = Contra2Cart(Uz) (IB and solid cells keep the Ucat passed in),
 *   F   = -(Uz - Uo)/dt (coeff 1) or -(1.5 Uz - 2 Uo + 0.5 Urm1)/dt (otherwise),
 *         summed as the VecAXPY sequence does,
 *   Ul  = IB_BC applied to Uz (or_momentum_ib_bc; the local Ucont that
 *         Formfunction_2 reads, while the time terms use the global Uz),
 *   F  += Formfunction_2(Ul, Ucat) scaled by 0.5 (coeff 1; plus 0.5 RHS_o) or 1,
 *   F  += -1 dP.
 * uz and ucat receive Uz and its Ucat. */

/* IB_BC (momentum.c:2037-2220) on a Ucont vector field, for a body at rest:
 * with immersed == 3, the faces between interior cells whose nvert sum exceeds
 * 1.1 are zeroed (momentum.c:2158-2162); the slip (bctype 10) faces next to the
 * first and last interior cells are zeroed on every row, ghost rows included
 * (momentum.c:2190-2196); the periodic ghosts of each axis's normal component
 * copy their images (momentum.c:2198-2213), in the reference's point order.  Its
 * periodic Ucat copy (momentum.c:2075-2103) repeats what Contra2Cart leaves. */
void or_momentum_ib_bc(const or_grid *g, const int bc[6], int immersed, const double *nvert, double *u) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  if (immersed == 3)
    for (int k = 1; k < mz - 1; k++)
      for (int j = 1; j < my - 1; j++)
        for (int i = 1; i < mx - 1; i++) {
          const size_t p = IXM(g, i, j, k);
          if (nvert[IXM(g, i + 1, j, k)] + nvert[p] > 1.1) u[3 * p] = 0;
          if (nvert[IXM(g, i, j + 1, k)] + nvert[p] > 1.1) u[3 * p + 1] = 0;
          if (nvert[IXM(g, i, j, k + 1)] + nvert[p] > 1.1) u[3 * p + 2] = 0;
        }
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        if (bc[0] == 10 && i == 1) u[3 * IXM(g, 0, j, k)] = 0;
        if (bc[1] == 10 && i == mx - 2) u[3 * IXM(g, mx - 2, j, k)] = 0;
        if (bc[2] == 10 && j == 1) u[3 * IXM(g, i, 0, k) + 1] = 0;
        if (bc[3] == 10 && j == my - 2) u[3 * IXM(g, i, my - 2, k) + 1] = 0;
        if (bc[4] == 10 && k == 1) u[3 * IXM(g, i, j, 0) + 2] = 0;
        if (bc[5] == 10 && k == mz - 2) u[3 * IXM(g, i, j, mz - 2) + 2] = 0;
        if (g->per[0] && i == 0) u[3 * IXM(g, 0, j, k)] = u[3 * IXM(g, mx - 2, j, k)];
        if (g->per[0] && i == mx - 1) u[3 * IXM(g, mx - 1, j, k)] = u[3 * IXM(g, 1, j, k)];
        if (g->per[1] && j == 0) u[3 * IXM(g, i, 0, k) + 1] = u[3 * IXM(g, i, my - 2, k) + 1];
        if (g->per[1] && j == my - 1) u[3 * IXM(g, i, my - 1, k) + 1] = u[3 * IXM(g, i, 1, k) + 1];
        if (g->per[2] && k == 0) u[3 * IXM(g, i, j, 0) + 2] = u[3 * IXM(g, i, j, mz - 2) + 2];
        if (g->per[2] && k == mz - 1) u[3 * IXM(g, i, j, mz - 1) + 2] = u[3 * IXM(g, i, j, 1) + 2];
      }
}
void or_momentum_bc_zero(const or_grid *g, const int bc[6], double *u) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        const int ik = j != 0 && j != my - 1 && k != 0 && k != mz - 1;  /* slip-rule interior in j, k */
        const int jk = i != 0 && i != mx - 1 && k != 0 && k != mz - 1;  /* slip-rule interior in i, k */
        if (i == 0 && (bc[0] == 1 || bc[0] == -1)) u[3 * IXM(g, i, j, k)] = 0;
        if (i == mx - 1 && (bc[1] == 1 || bc[1] == -1)) u[3 * IXM(g, i - 1, j, k)] = 0;
        if (j == 0 && (bc[2] == 1 || bc[2] == -1 || bc[2] == 12)) u[3 * IXM(g, i, j, k) + 1] = 0;
        if (j == my - 1 && (bc[3] == 1 || bc[3] == -1 || bc[3] == 2 || bc[3] == 12)) u[3 * IXM(g, i, j - 1, k) + 1] = 0;
        if (k == 0 && (bc[4] == 1 || bc[4] == -1)) u[3 * IXM(g, i, j, k) + 2] = 0;
        if (k == mz - 1 && (bc[5] == 1 || bc[5] == -1)) u[3 * IXM(g, i, j, k - 1) + 2] = 0;
        if (bc[0] == 10 && i == 0 && ik) u[3 * IXM(g, i, j, k)] = 0;
        if (bc[1] == 10 && i == mx - 1 && ik) u[3 * IXM(g, i - 1, j, k)] = 0;
        if (bc[2] == 10 && j == 0 && jk) u[3 * IXM(g, i, j, k) + 1] = 0;
        if ((bc[3] == 10 || bc[3] == -10) && j == my - 1 && jk) u[3 * IXM(g, i, j - 1, k) + 1] = 0;
      }
}

void or_momentum_residual(const or_grid *g, const int bctype[6], int immersed, const double *icsi,
                          const double *ieta, const double *izet, const double *jcsi, const double *jeta,
                          const double *jzet, const double *kcsi, const double *keta, const double *kzet,
                          const double *iaj, const double *jaj, const double *kaj, const double *csi,
                          const double *eta, const double *zet, const double *nvert, const double *dp,
                          const double *uo, const double *urm1, const double *rhs_o, const double *u, double ren,
                          double dt, double coeff, double *uz, double *ucat, double *f) {
  const size_t n3 = 3 * (size_t)g->mx * g->my * g->mz;
  memcpy(uz, u, n3 * sizeof(double));
  or_momentum_bc_zero(g, bctype, uz);
  or_contra2cart(g, bctype, csi, eta, zet, nvert, uz, ucat);
  const int bdf1 = coeff > 0.9 && coeff < 1.1;
  for (size_t q = 0; q < n3; q++) {
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
  double *ul = (double *)malloc(n3 * sizeof(double));
  memcpy(ul, uz, n3 * sizeof(double));
  or_momentum_ib_bc(g, bctype, immersed, nvert, ul);
  mm_formfunction2(g, bctype, immersed, icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, csi,
                   eta, zet, nvert, ul, ucat, ren, bdf1 ? 0.5 : 1.0, f, 1);
  free(ul);
  if (bdf1)
    for (size_t q = 0; q < n3; q++) f[q] = f[q] + 0.5 * rhs_o[q];
  for (size_t q = 0; q < n3; q++) f[q] = f[q] + -1. * dp[q];
}

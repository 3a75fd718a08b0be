/*
 * CPU oracle for the VFS pressure-projection path -- TEST INFRASTRUCTURE ONLY.
 * See vfs_oracle.h for the layout conventions.  Each function restates the
 * arithmetic of the cited reference lines (trunglendsu/VFS_v3p2) in FP64,
 * keeping the reference's branch structure, operand order and accumulation
 * order wherever the result is meant to be compared bit for bit.
 */
#include "vfs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define TH 0.1 /* poisson_threshold, poisson.c:16 */

#define IX(g, i, j, k) ((((size_t)(k)) * (size_t)(g)->my + (size_t)(j)) * (size_t)(g)->mx + (size_t)(i))

double or_threshold(void) { return TH; }

int or_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

static int is_bnd(const or_grid *g, int i, int j, int k) {
  return i == 0 || i == g->mx - 1 || j == 0 || j == g->my - 1 || k == 0 || k == g->mz - 1;
}

/* periodic image of a DMDA point (poisson.c:153-174); returns 1 if any axis was mapped */
static int periodic_image(const or_grid *g, int i, int j, int k, int *a, int *b, int *c, int flag[3]) {
  *a = i; *b = j; *c = k;
  flag[0] = flag[1] = flag[2] = 0;
  if (g->per[0] && i == 0) { *a = g->mx - 2; flag[0] = 1; }
  else if (g->per[0] && i == g->mx - 1) { *a = 1; flag[0] = 1; }
  if (g->per[1] && j == 0) { *b = g->my - 2; flag[1] = 1; }
  else if (g->per[1] && j == g->my - 1) { *b = 1; flag[1] = 1; }
  if (g->per[2] && k == 0) { *c = g->mz - 2; flag[2] = 1; }
  else if (g->per[2] && k == g->mz - 1) { *c = 1; flag[2] = 1; }
  return flag[0] | flag[1] | flag[2];
}

static void periodic_copy_scalar(const or_grid *g, double *f) {
  if (!(g->per[0] || g->per[1] || g->per[2])) return;
  /* an image has every periodic index interior, so no image is itself written */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < g->mz; k++)
    for (int j = 0; j < g->my; j++)
      for (int i = 0; i < g->mx; i++) {
        int a, b, c, fl[3];
        if (periodic_image(g, i, j, k, &a, &b, &c, fl)) f[IX(g, i, j, k)] = f[IX(g, a, b, c)];
      }
}

/* ---------------------------------------------------------------- setup_lidx2 */
int or_setup_gid(const or_grid *g, const double *nvert, double *gid) {
  size_t n = (size_t)g->mx * g->my * g->mz;
  for (size_t p = 0; p < n; p++) gid[p] = -1.0;
  int cnt = 0;
  for (int k = 0; k < g->mz; k++)
    for (int j = 0; j < g->my; j++)
      for (int i = 0; i < g->mx; i++) {
        if (is_bnd(g, i, j, k)) continue;
        size_t p = IX(g, i, j, k);
        if ((int)nvert[p] > TH) continue; /* poisson.c:239 casts to int */
        gid[p] = (double)cnt++;
      }
  periodic_copy_scalar(g, gid); /* poisson.c:298-331 */
  return cnt;
}

/* ---------------------------------------------------------- face coefficients */
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

void or_face_coeffs(const or_grid *g, const double *icsi, const double *ieta, const double *izet,
                    const double *jcsi, const double *jeta, const double *jzet,
                    const double *kcsi, const double *keta, const double *kzet,
                    const double *iaj, const double *jaj, const double *kaj, double *g9) {
  size_t n = (size_t)g->mx * g->my * g->mz;
  memset(g9, 0, 9 * n * sizeof(double));
  for (int k = 0; k < g->mz; k++)
    for (int j = 0; j < g->my; j++)
      for (int i = 0; i < g->mx; i++) {
        int a, b, c, fl[3];
        periodic_image(g, i, j, k, &a, &b, &c, fl);
        size_t s = IX(g, a, b, c), p = IX(g, i, j, k);
        /* operand order of poisson.c:1132-1140: (v.x*w.x + v.y*w.y + v.z*w.z) * aj */
        g9[0 * n + p] = dot3(&icsi[3 * s], &icsi[3 * s]) * iaj[s];
        g9[1 * n + p] = dot3(&ieta[3 * s], &icsi[3 * s]) * iaj[s];
        g9[2 * n + p] = dot3(&izet[3 * s], &icsi[3 * s]) * iaj[s];
        g9[3 * n + p] = dot3(&jcsi[3 * s], &jeta[3 * s]) * jaj[s];
        g9[4 * n + p] = dot3(&jeta[3 * s], &jeta[3 * s]) * jaj[s];
        g9[5 * n + p] = dot3(&jzet[3 * s], &jeta[3 * s]) * jaj[s];
        g9[6 * n + p] = dot3(&kcsi[3 * s], &kzet[3 * s]) * kaj[s];
        g9[7 * n + p] = dot3(&keta[3 * s], &kzet[3 * s]) * kaj[s];
        g9[8 * n + p] = dot3(&kzet[3 * s], &kzet[3 * s]) * kaj[s];
      }
}

/* ------------------------------------------------------------ operator rows */
/* slot index of offset (di,dj,dk), poisson.c:956-974 */
static int slot_of(int di, int dj, int dk) {
  static const int tab[3][3][3] = {
      /* dk=-1 */ {{-1, OR_BS, -1}, {OR_BW, OR_BP, OR_BE}, {-1, OR_BN, -1}},
      /* dk= 0 */ {{OR_SW, OR_SP, OR_SE}, {OR_WP, OR_CP, OR_EP}, {OR_NW, OR_NP, OR_NE}},
      /* dk=+1 */ {{-1, OR_TS, -1}, {OR_TW, OR_TP, OR_TE}, {-1, OR_TN, -1}}};
  return tab[dk + 1][dj + 1][di + 1];
}

/* Cross-difference case for the face between cells p and q (q = p +/- e_d), taken
 * in direction e (poisson.c:1182-1230 and the 11 rotated copies). */
static int cross_case(const or_grid *g, const double *nv, size_t p, size_t q, size_t se, int pos_e,
                      int m_e, int per_e) {
  double sum_hi = nv[p + se] + nv[q + se];
  double sum_lo = nv[p - se] + nv[q - se];
  if ((pos_e == m_e - 2 && !per_e) || sum_hi > TH) {
    if (sum_lo < TH && pos_e != 1) return OR_X_LOWER;
    return OR_X_ZERO;
  } else if ((pos_e == 1 && !per_e) || sum_lo > TH) {
    if (sum_hi < TH) return OR_X_UPPER;
    return OR_X_ZERO;
  }
  (void)g;
  return OR_X_CENTRAL;
}

void or_stencil(const or_grid *g, const double *g9, const double *nvert, double *vol19,
                uint8_t *rowflag, uint8_t *cases) {
  const size_t n = (size_t)g->mx * g->my * g->mz;
  const size_t stride[3] = {1, (size_t)g->mx, (size_t)g->mx * g->my};
  const int m[3] = {g->mx, g->my, g->mz};
  memset(vol19, 0, 19 * n * sizeof(double));
  memset(rowflag, 0, n);
  memset(cases, 0, 18 * n);
#pragma omp parallel for schedule(static)
  for (int k = 1; k < g->mz - 1; k++)
    for (int j = 1; j < g->my - 1; j++)
      for (int i = 1; i < g->mx - 1; i++) {
        const size_t c = IX(g, i, j, k);
        /* rows exist only for fluid cells: poisson.c:1158-1169 */
        if (nvert[c] > 0.1) continue;
        rowflag[c] = 1;
        const int pos[3] = {i, j, k};
        double vol[19];
        for (int s = 0; s < 19; s++) vol[s] = 0.;
        /* faces in the reference's order: i+, i-, j+, j-, k+, k- */
        for (int d = 0; d < 3; d++)
          for (int sgn = 1; sgn >= -1; sgn -= 2) {
            const int face = 2 * d + (sgn > 0 ? 0 : 1);
            const size_t q = sgn > 0 ? c + stride[d] : c - stride[d];
            const int edge = sgn > 0 ? (pos[d] != m[d] - 2) : (pos[d] != 1);
            if (!(nvert[q] < TH && (edge || g->per[d]))) continue;
            cases[18 * c + 3 * face] = 1;
            const size_t owner = sgn > 0 ? c : q; /* face coefficients live on the lower cell */
            int xi = 0;
            for (int e = 0; e < 3; e++) {
              const double gv = g9[(size_t)(3 * d + e) * n + owner];
              int off_q[3] = {0, 0, 0};
              off_q[d] = sgn;
              if (e == d) {
                /* g_dd (phi_q - phi_c) term, e.g. poisson.c:1183-1184 */
                vol[OR_CP] -= gv;
                vol[slot_of(off_q[0], off_q[1], off_q[2])] += gv;
                continue;
              }
              const int cc = cross_case(g, nvert, c, q, stride[e], pos[e], m[e], g->per[e]);
              cases[18 * c + 3 * face + 1 + xi] = (uint8_t)cc;
              xi++;
              if (cc == OR_X_ZERO) continue;
              const double w = (cc == OR_X_CENTRAL) ? 0.25 : 0.5;
              const double t = w * gv; /* exact: w is a power of two */
              int hi = (cc == OR_X_LOWER) ? 0 : 1;  /* upper row offset in e */
              int lo = (cc == OR_X_UPPER) ? 0 : -1; /* lower row offset in e */
              int o1[3] = {0, 0, 0}, o2[3], o3[3] = {0, 0, 0}, o4[3];
              /* the row gets sgn * F; F's cross part is t * (phi[c+hi]+phi[q+hi]-phi[c+lo]-phi[q+lo]) */
              o1[e] = hi;
              memcpy(o2, off_q, sizeof o2); o2[e] += hi;
              o3[e] = lo;
              memcpy(o4, off_q, sizeof o4); o4[e] += lo;
              if (sgn > 0) {
                vol[slot_of(o1[0], o1[1], o1[2])] += t;
                vol[slot_of(o2[0], o2[1], o2[2])] += t;
                vol[slot_of(o3[0], o3[1], o3[2])] -= t;
                vol[slot_of(o4[0], o4[1], o4[2])] -= t;
              } else {
                vol[slot_of(o1[0], o1[1], o1[2])] -= t;
                vol[slot_of(o2[0], o2[1], o2[2])] -= t;
                vol[slot_of(o3[0], o3[1], o3[2])] += t;
                vol[slot_of(o4[0], o4[1], o4[2])] += t;
              }
            }
          }
        for (int s = 0; s < 19; s++) vol19[19 * c + s] = vol[s];
      }
}

static const int SLOT_OFF[19][3] = {{0, 0, 0},  {1, 0, 0},   {-1, 0, 0}, {0, 1, 0},  {0, -1, 0},
                                    {0, 0, 1},  {0, 0, -1},  {1, 1, 0},  {1, -1, 0}, {-1, 1, 0},
                                    {-1, -1, 0}, {0, 1, 1},  {0, 1, -1}, {0, -1, 1}, {0, -1, -1},
                                    {1, 0, 1},  {1, 0, -1},  {-1, 0, 1}, {-1, 0, -1}};

int or_assemble_csr(const or_grid *g, const double *vol19, const uint8_t *rowflag,
                    const double *gid, int drop, int *rowptr, int *col, double *val) {
  int nnz = 0, row = 0;
  rowptr[0] = 0;
  for (int k = 1; k < g->mz - 1; k++)
    for (int j = 1; j < g->my - 1; j++)
      for (int i = 1; i < g->mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if (!rowflag[c]) continue;
        int r = (int)gid[c];
        if (r != row) return -1; /* rows are numbered in this loop order */
        int start = nnz;
        for (int s = 0; s < 19; s++) {
          size_t q = IX(g, i + SLOT_OFF[s][0], j + SLOT_OFF[s][1], k + SLOT_OFF[s][2]);
          int idx = (int)gid[q];
          double v = vol19[19 * c + s];
          /* poisson.c:1608 (drop=1); drop=0 keeps every nonzero coefficient */
          const double cut = drop ? 1.e-10 : 0.;
          if ((fabs(v) > cut && idx >= 0) || s == OR_CP) {
            int dup = -1;
            for (int t = start; t < nnz; t++)
              if (col[t] == idx) dup = t;
            if (dup >= 0) { val[dup] = v; continue; } /* SetValues overwrites */
            col[nnz] = idx;
            val[nnz] = v;
            nnz++;
          }
        }
        row++;
        rowptr[row] = nnz;
      }
  return nnz;
}

void or_csr_matvec(int nrow, const int *rowptr, const int *col, const double *val,
                   const double *x, double *y) {
#pragma omp parallel for schedule(static)
  for (int r = 0; r < nrow; r++) {
    double s = 0.;
    for (int t = rowptr[r]; t < rowptr[r + 1]; t++) s += val[t] * x[col[t]];
    y[r] = s;
  }
}

void or_stencil_apply(const or_grid *g, const double *vol19, const uint8_t *rowflag,
                      const double *phi, double *y) {
  size_t n = (size_t)g->mx * g->my * g->mz;
  memset(y, 0, n * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int k = 1; k < g->mz - 1; k++)
    for (int j = 1; j < g->my - 1; j++)
      for (int i = 1; i < g->mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if (!rowflag[c]) continue;
        double s = 0.;
        for (int t = 0; t < 19; t++)
          s += vol19[19 * c + t] * phi[IX(g, i + SLOT_OFF[t][0], j + SLOT_OFF[t][1], k + SLOT_OFF[t][2])];
        y[c] = s;
      }
}

/* ----------------------------------------------------------------------- RHS */
void or_rhs(const or_grid *g, const double *ucont, const double *nvert, double dt, double st,
            double coeff, double *b) {
  size_t n = (size_t)g->mx * g->my * g->mz;
  memset(b, 0, n * sizeof(double));
  const int mx = g->mx, my = g->my, mz = g->mz;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if (nvert[c] >= 0.1) continue; /* poisson_hypre.c:346 */
        double val = 0;
        val -= ucont[3 * c + 0];
        if (i == 1 && g->per[0]) val += ucont[3 * IX(g, mx - 2, j, k) + 0];
        else val += ucont[3 * IX(g, i - 1, j, k) + 0];
        val -= ucont[3 * c + 1];
        if (j == 1 && g->per[1]) val += ucont[3 * IX(g, i, my - 2, k) + 1];
        else val += ucont[3 * IX(g, i, j - 1, k) + 1];
        val -= ucont[3 * c + 2];
        if (k == 1 && g->per[2]) val += ucont[3 * IX(g, i, j, mz - 2) + 2];
        else val += ucont[3 * IX(g, i, j, k - 1) + 2];
        val *= -1.0 / dt * st * coeff; /* poisson_hypre.c:376 */
        b[c] = val;
      }
}

/* ---------------------------------------------------------------- VolumeFlux */
static double norm3(const double *v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* Sums over the interior are taken per k plane (in i, j order) and the plane
 * partials are added in k order, so the result does not depend on the thread
 * count; the planes themselves run in parallel (each writes only its own faces). */
static double sum_planes(const double *part, int mz) {
  double s = 0;
  for (int k = 1; k < mz - 1; k++) s += part[k];
  return s;
}

void or_volume_flux(const or_grid *g, int immersed, double *ucont, const double *csi,
                    const double *eta, const double *zet, const double *nvert, double out[4]) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  const size_t st[3] = {1, (size_t)mx, (size_t)mx * my};
  const double *met[3] = {csi, eta, zet};
  const int mm[3] = {mx, my, mz};
  double *pf = (double *)calloc(2 * (size_t)mz, sizeof(double)), *pa = pf + mz;
  if (immersed == 0 || immersed == 3) {
    /* poisson.c:2634-2828 (poisson_hypre.c:667 takes it for every value but 1 and
       2; the corrected faces are zeroed for 0 and 3, poisson.c:2728) */
    for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel for schedule(static)
      for (int k = 1; k < mz - 1; k++) {
        double flux = 0, area = 0;
        for (int j = 1; j < my - 1; j++)
          for (int i = 1; i < mx - 1; i++) {
            size_t c = IX(g, i, j, k);
            int pos[3] = {i, j, k};
            for (int d = 0; d < 3; d++) {
              double nq = nvert[c + st[d]];
              if (pos[d] > mm[d] - 2) continue;
              if (nvert[c] < TH && nq > TH) { flux += ucont[3 * c + d]; area += norm3(&met[d][3 * c]); }
              if (nvert[c] > TH && nq < TH) { flux -= ucont[3 * c + d]; area += norm3(&met[d][3 * c]); }
            }
          }
        pf[k] = flux;
        pa[k] = area;
      }
      double flux = sum_planes(pf, mz), area = sum_planes(pa, mz);
      out[2 * pass] = flux;
      out[2 * pass + 1] = area;
      if (pass == 1) break;
      double corr = (area > 1.e-15) ? flux / area : 0.;
#pragma omp parallel for collapse(2) schedule(static)
      for (int k = 1; k < mz - 1; k++)
        for (int j = 1; j < my - 1; j++)
          for (int i = 1; i < mx - 1; i++) {
            size_t c = IX(g, i, j, k);
            int pos[3] = {i, j, k};
            for (int d = 0; d < 3; d++) {
              double nq = nvert[c + st[d]];
              if (pos[d] > mm[d] - 2) continue;
              if (nvert[c] < TH && nq > TH) { ucont[3 * c + d] -= norm3(&met[d][3 * c]) * corr; ucont[3 * c + d] = 0; }
              if (nvert[c] > TH && nq < TH) { ucont[3 * c + d] += norm3(&met[d][3 * c]) * corr; ucont[3 * c + d] = 0; }
            }
          }
    }
    free(pf);
    return;
  }
  /* immersed == 1: poisson.c:2898-3240 with flg = 0 */
  const double eps = 1.e-10, ibmval = 1.1;
#pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++) {
    double flux = 0, area = 0;
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        int pos[3] = {i, j, k};
        for (int d = 0; d < 3; d++) {
          double nq = nvert[c + st[d]];
          if (!(pos[d] < mm[d] - 2)) continue;
          if (nvert[c] < 0.1 && nq > 0.1 && nq < ibmval) {
            if (fabs(ucont[3 * c + d]) > eps) { flux += ucont[3 * c + d]; area += norm3(&met[d][3 * c]); }
            else ucont[3 * c + d] = 0.;
          }
        }
        if (nvert[c] > 0.1 && nvert[c] < ibmval)
          for (int d = 0; d < 3; d++) {
            double nq = nvert[c + st[d]];
            if (!(pos[d] < mm[d] - 2)) continue;
            if (nq < 0.1) {
              if (fabs(ucont[3 * c + d]) > eps) { flux -= ucont[3 * c + d]; area += norm3(&met[d][3 * c]); }
              else ucont[3 * c + d] = 0.;
            }
          }
      }
    pf[k] = flux;
    pa[k] = area;
  }
  double flux = sum_planes(pf, mz), area = sum_planes(pa, mz);
  out[0] = flux;
  out[1] = area;
  double corr = (area > 1.e-15) ? flux / area : 0.;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        int pos[3] = {i, j, k};
        for (int d = 0; d < 3; d++) {
          double nq = nvert[c + st[d]];
          if (!(pos[d] < mm[d] - 2)) continue;
          if (nvert[c] < 0.1 && nq > 0.1 && nq < ibmval && fabs(ucont[3 * c + d]) > eps)
            ucont[3 * c + d] -= norm3(&met[d][3 * c]) * corr;
        }
        if (nvert[c] > 0.1 && nvert[c] < 1.1)
          for (int d = 0; d < 3; d++) {
            double nq = nvert[c + st[d]];
            if (!(pos[d] < mm[d] - 2)) continue;
            if (nq < 0.1 && fabs(ucont[3 * c + d]) > eps) ucont[3 * c + d] += norm3(&met[d][3 * c]) * corr;
          }
      }
#pragma omp parallel for schedule(static)
  for (int k = 1; k < mz - 1; k++) {
    double fl = 0, ar = 0;
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        int pos[3] = {i, j, k};
        for (int d = 0; d < 3; d++) {
          double nq = nvert[c + st[d]];
          if (!(pos[d] < mm[d] - 2)) continue;
          if (nvert[c] < 0.1 && nq > 0.1 && nq < ibmval) { fl += ucont[3 * c + d]; ar += norm3(&met[d][3 * c]); }
        }
        if (nvert[c] > 0.1 && nvert[c] < ibmval)
          for (int d = 0; d < 3; d++) {
            double nq = nvert[c + st[d]];
            if (!(pos[d] < mm[d] - 2)) continue;
            if (nq < 0.1) { fl -= ucont[3 * c + d]; ar += norm3(&met[d][3 * c]); }
          }
      }
    pf[k] = fl;
    pa[k] = ar;
  }
  out[2] = sum_planes(pf, mz);
  out[3] = sum_planes(pa, mz);
  free(pf);
}

/* ------------------------------------------------------ null space, scatter */
void or_remove_nullspace(const or_grid *g, const double *nvert, const double *gid, double *x) {
  const int mz = g->mz;
  double *ps = (double *)calloc((size_t)mz, sizeof(double));
  int cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
  for (int k = 1; k < mz - 1; k++) {
    double sum = 0;
    for (int j = 1; j < g->my - 1; j++)
      for (int i = 1; i < g->mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if (nvert[c] < 0.1) { sum += x[(int)gid[c]]; cnt++; }
      }
    ps[k] = sum;
  }
  double val = -sum_planes(ps, mz) / (double)cnt;
  free(ps);
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < g->mz - 1; k++)
    for (int j = 1; j < g->my - 1; j++)
      for (int i = 1; i < g->mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if (nvert[c] < 0.1) x[(int)gid[c]] += val;
      }
}

/* The same mean removal and scatter on an interior array (nz, ny, nx) as the CPU
 * port of the engine's solver holds it: x -= mean over fluid cells, then phi =
 * x on cells whose int(nvert) is 0 (Convert_Phi2_Phi, poisson.c:1086-1125),
 * zero elsewhere, with periodic ghost images. */
void or_interior_to_phi(const or_grid *g, const double *nvert, double *x, double *phi) {
  const int mx = g->mx, my = g->my, mz = g->mz, nx = mx - 2, ny = my - 2;
  double *ps = (double *)calloc((size_t)mz, sizeof(double));
  int cnt = 0;
#pragma omp parallel for reduction(+ : cnt) schedule(static)
  for (int k = 1; k < mz - 1; k++) {
    double sum = 0;
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++)
        if (nvert[IX(g, i, j, k)] < 0.1) { sum += x[((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1)]; cnt++; }
    ps[k] = sum;
  }
  double mean = cnt ? sum_planes(ps, mz) / (double)cnt : 0.;
  free(ps);
  memset(phi, 0, (size_t)mx * my * mz * sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k), q = ((size_t)(k - 1) * ny + (j - 1)) * nx + (i - 1);
        if (nvert[c] < 0.1) x[q] -= mean;
        phi[c] = (int)nvert[c] > TH ? 0. : x[q];
      }
  periodic_copy_scalar(g, phi);
}

void or_convert_phi(const or_grid *g, const double *nvert, const double *phi2, double *phi) {
  size_t n = (size_t)g->mx * g->my * g->mz;
  memset(phi, 0, n * sizeof(double));
  int pos = 0;
  for (int k = 1; k < g->mz - 1; k++)
    for (int j = 1; j < g->my - 1; j++)
      for (int i = 1; i < g->mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if ((int)nvert[c] > TH) phi[c] = 0;
        else phi[c] = phi2[pos++];
      }
  periodic_copy_scalar(g, phi);
}

void or_update_pressure(const or_grid *g, const double *nvert, double *p, double *phi) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < g->mz - 1; k++)
    for (int j = 1; j < g->my - 1; j++)
      for (int i = 1; i < g->mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if (nvert[c] < 0.1) p[c] += phi[c];
      }
  periodic_copy_scalar(g, p);
  periodic_copy_scalar(g, phi);
}

/* ---------------------------------------------------------------- Projection */
/* cross difference of phi for the face (p,q) in direction e, poisson.c:2234-2257 */
static double proj_cross(const or_grid *g, const double *nv, const double *phi, size_t p, size_t q,
                         size_t se, int pos_e, int m_e, int per_e) {
  int cc = cross_case(g, nv, p, q, se, pos_e, m_e, per_e);
  switch (cc) {
    case OR_X_LOWER: return (phi[p] + phi[q] - phi[p - se] - phi[q - se]) * 0.5;
    case OR_X_UPPER: return (phi[p + se] + phi[q + se] - phi[p] - phi[q]) * 0.5;
    case OR_X_CENTRAL: return (phi[p + se] + phi[q + se] - phi[p - se] - phi[q - se]) * 0.25;
    default: return 0.;
  }
}

void or_projection(const or_grid *g, const double *icsi, const double *ieta, const double *izet,
                   const double *jcsi, const double *jeta, const double *jzet,
                   const double *kcsi, const double *keta, const double *kzet,
                   const double *iaj, const double *jaj, const double *kaj,
                   const double *nvert, const double *phi, double *ucont, double dt, double st,
                   double coeff) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  const size_t sx = 1, sy = (size_t)mx, sz = (size_t)mx * my;
  const int pi = g->per[0], pj = g->per[1], pk = g->per[2];
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        const size_t c = IX(g, i, j, k);
        /* i face, poisson.c:2221-2265 */
        if (j > 0 && j < my - 1 && k > 0 && k < mz - 1)
          if ((i > 0 && i < mx - 2) || (pi && i == mx - 2)) {
            const size_t q = c + sx;
            double dpdc = phi[q] - phi[c];
            if (i == mx - 2 && pi) dpdc = phi[IX(g, 1, j, k)] - phi[c];
            double dpde = proj_cross(g, nvert, phi, c, q, sy, j, my, pj);
            double dpdz = proj_cross(g, nvert, phi, c, q, sz, k, mz, pk);
            if ((nvert[c] + nvert[q]) < TH) {
              const double *a = &icsi[3 * c], *b = &ieta[3 * c], *z = &izet[3 * c];
              ucont[3 * c + 0] -= (dpdc * (a[0] * a[0] + a[1] * a[1] + a[2] * a[2]) * iaj[c] +
                                   dpde * (b[0] * a[0] + b[1] * a[1] + b[2] * a[2]) * iaj[c] +
                                   dpdz * (z[0] * a[0] + z[1] * a[1] + z[2] * a[2]) * iaj[c]) *
                                  dt * st / coeff / 1.0;
            }
          }
        /* j face, poisson.c:2267-2314 */
        if (i > 0 && i < mx - 1 && k > 0 && k < mz - 1)
          if ((j > 0 && j < my - 2) || (pj && j == my - 2)) {
            const size_t q = c + sy;
            double dpdc = proj_cross(g, nvert, phi, c, q, sx, i, mx, pi);
            double dpde;
            if (j == my - 2 && pj) dpde = phi[IX(g, i, 1, k)] - phi[c];
            else dpde = phi[q] - phi[c];
            double dpdz = proj_cross(g, nvert, phi, c, q, sz, k, mz, pk);
            if ((nvert[c] + nvert[q]) < TH) {
              const double *a = &jcsi[3 * c], *b = &jeta[3 * c], *z = &jzet[3 * c];
              ucont[3 * c + 1] -= (dpdc * (a[0] * b[0] + a[1] * b[1] + a[2] * b[2]) * jaj[c] +
                                   dpde * (b[0] * b[0] + b[1] * b[1] + b[2] * b[2]) * jaj[c] +
                                   dpdz * (z[0] * b[0] + z[1] * b[1] + z[2] * b[2]) * jaj[c]) *
                                  dt * st / coeff / 1.0;
            }
          }
        /* k face, poisson.c:2316-2363 */
        if (i > 0 && i < mx - 1 && j > 0 && j < my - 1)
          if ((k > 0 && k < mz - 2) || (pk && k == mz - 2)) {
            const size_t q = c + sz;
            double dpdc = proj_cross(g, nvert, phi, c, q, sx, i, mx, pi);
            double dpde = proj_cross(g, nvert, phi, c, q, sy, j, my, pj);
            double dpdz;
            if (k == mz - 2 && pk) dpdz = phi[IX(g, i, j, 1)] - phi[c];
            else dpdz = phi[q] - phi[c];
            if ((nvert[c] + nvert[q]) < TH) {
              const double *a = &kcsi[3 * c], *b = &keta[3 * c], *z = &kzet[3 * c];
              ucont[3 * c + 2] -= (dpdc * (a[0] * z[0] + a[1] * z[1] + a[2] * z[2]) * kaj[c] +
                                   dpde * (b[0] * z[0] + b[1] * z[1] + b[2] * z[2]) * kaj[c] +
                                   dpdz * (z[0] * z[0] + z[1] * z[1] + z[2] * z[2]) * kaj[c]) *
                                  dt * st / coeff / 1.0;
            }
          }
      }
  /* periodic ghost faces, poisson.c:2401-2441 */
  if (pi || pj || pk)
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          int a, b, cc, fl[3];
          if (!periodic_image(g, i, j, k, &a, &b, &cc, fl)) continue;
          size_t p = IX(g, i, j, k), s = IX(g, a, b, cc);
          if (fl[0]) ucont[3 * p + 0] = ucont[3 * s + 0];
          if (fl[1]) ucont[3 * p + 1] = ucont[3 * s + 1];
          if (fl[2]) ucont[3 * p + 2] = ucont[3 * s + 2];
        }
}

/* --------------------------------------------------------------- Contra2Cart */
void or_contra2cart(const or_grid *g, const int bctype[6], const double *csi, const double *eta,
                    const double *zet, const double *nvert, double *ucont_in, double *ucat) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  const size_t n = (size_t)mx * my * mz;
  const int anyper = g->per[0] || g->per[1] || g->per[2];
  /* local copy of Ucont with whole-struct periodic ghost copies, rhs.c:115-142 */
  double *lu = (double *)malloc(3 * n * sizeof(double));
  memcpy(lu, ucont_in, 3 * n * sizeof(double));
  if (anyper)
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          int a, b, c, fl[3];
          if (!periodic_image(g, i, j, k, &a, &b, &c, fl)) continue;
          size_t p = IX(g, i, j, k), s = IX(g, a, b, c);
          lu[3 * p + 0] = lu[3 * s + 0]; lu[3 * p + 1] = lu[3 * s + 1]; lu[3 * p + 2] = lu[3 * s + 2];
        }
  /* core Cramer solve, rhs.c:144-189 */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if (!(nvert[c] < 0.1)) continue;
        const double *m0 = &csi[3 * c], *m1 = &eta[3 * c], *m2 = &zet[3 * c];
        double q0 = 0.5 * (lu[3 * IX(g, i - 1, j, k) + 0] + lu[3 * c + 0]);
        double q1 = 0.5 * (lu[3 * IX(g, i, j - 1, k) + 1] + lu[3 * c + 1]);
        double q2 = 0.5 * (lu[3 * IX(g, i, j, k - 1) + 2] + lu[3 * c + 2]);
        double det = m0[0] * (m1[1] * m2[2] - m1[2] * m2[1]) - m0[1] * (m1[0] * m2[2] - m1[2] * m2[0]) +
                     m0[2] * (m1[0] * m2[1] - m1[1] * m2[0]);
        double det0 = q0 * (m1[1] * m2[2] - m1[2] * m2[1]) - q1 * (m0[1] * m2[2] - m0[2] * m2[1]) +
                      q2 * (m0[1] * m1[2] - m0[2] * m1[1]);
        double det1 = -q0 * (m1[0] * m2[2] - m1[2] * m2[0]) + q1 * (m0[0] * m2[2] - m0[2] * m2[0]) -
                      q2 * (m0[0] * m1[2] - m0[2] * m1[0]);
        double det2 = q0 * (m1[0] * m2[1] - m1[1] * m2[0]) - q1 * (m0[0] * m2[1] - m0[1] * m2[0]) +
                      q2 * (m0[0] * m1[1] - m0[1] * m1[0]);
        ucat[3 * c + 0] = det0 / det;
        ucat[3 * c + 1] = det1 / det;
        ucat[3 * c + 2] = det2 / det;
      }
  /* periodic copy, rhs.c:198-241 */
  if (anyper)
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          int a, b, c, fl[3];
          if (!periodic_image(g, i, j, k, &a, &b, &c, fl)) continue;
          size_t p = IX(g, i, j, k), s = IX(g, a, b, c);
          for (int d = 0; d < 3; d++) ucat[3 * p + d] = ucat[3 * s + d];
        }
  /* boundary values: solid cells and no-slip walls, rhs.c:257-260, 491-515 */
  double *lc = lu; /* reuse the buffer as the snapshot of Ucat */
  memcpy(lc, ucat, 3 * n * sizeof(double));
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        size_t c = IX(g, i, j, k);
        if ((int)(nvert[c] + 0.1) == 3) {
          ucat[3 * c] = ucat[3 * c + 1] = ucat[3 * c + 2] = 0.;
          continue;
        }
        size_t src = (size_t)-1;
        if (i == 0 && bctype[0] == 1 && (j != 0 && k != 0)) src = IX(g, i + 1, j, k);
        if (i == mx - 1 && bctype[1] == 1 && (j != 0 && k != 0)) src = IX(g, i - 1, j, k);
        if (src != (size_t)-1) for (int d = 0; d < 3; d++) ucat[3 * c + d] = -lc[3 * src + d];
        src = (size_t)-1;
        if (j == 0 && bctype[2] == 1 && (i != 0 && k != 0)) src = IX(g, i, j + 1, k);
        if (j == my - 1 && bctype[3] == 1 && (i != 0 && k != 0)) src = IX(g, i, j - 1, k);
        if (src != (size_t)-1) for (int d = 0; d < 3; d++) ucat[3 * c + d] = -lc[3 * src + d];
        src = (size_t)-1;
        if (k == 0 && bctype[4] == 1 && (i != 0 && j != 0)) src = IX(g, i, j, k + 1);
        if (k == mz - 1 && bctype[5] == 1 && (i != 0 && j != 0)) src = IX(g, i, j, k - 1);
        if (src != (size_t)-1) for (int d = 0; d < 3; d++) ucat[3 * c + d] = -lc[3 * src + d];
      }
  /* final periodic copy, rhs.c:645-673 */
  if (anyper)
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          int a, b, c, fl[3];
          if (!periodic_image(g, i, j, k, &a, &b, &c, fl)) continue;
          size_t p = IX(g, i, j, k), s = IX(g, a, b, c);
          for (int d = 0; d < 3; d++) ucat[3 * p + d] = ucat[3 * s + d];
        }
  free(lu);
}

/* ---------------------------------------------------------------- Divergence */
double or_divergence(const or_grid *g, const double *u, const double *aj, const double *nvert,
                     double *div) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  const size_t sy = (size_t)mx, sz = (size_t)mx * my;
  double mx_div = 0.;
  if (div) memset(div, 0, (size_t)mx * my * mz * sizeof(double));
#pragma omp parallel for collapse(2) reduction(max : mx_div) schedule(static)
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        double d = fabs((u[3 * c] - u[3 * (c - 1)] + u[3 * c + 1] - u[3 * (c - sy) + 1] + u[3 * c + 2] -
                         u[3 * (c - sz) + 2]) *
                        aj[c]);
        if (nvert[c] + nvert[c + sz] + nvert[c - sz] + nvert[c + sy] + nvert[c - sy] + nvert[c + 1] +
                nvert[c - 1] >
            0.1)
          d = 0.;
        if (div) div[c] = d;
        if (d > mx_div) mx_div = d;
      }
  return mx_div;
}

/* ------------------------------------------------------------ nvert encoding */
void or_nvert_classify(const or_grid *g, const int bctype[6], double *nv) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  /* dilation, ibm.c:331-353 (order independent: it only ever writes 2) */
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        size_t c = IX(g, i, j, k);
        if (nv[c] < 0) nv[c] = 0;
        int ip = i < mx - 1 ? i + 1 : i, im = i > 0 ? i - 1 : i;
        int jp = j < my - 1 ? j + 1 : j, jm = j > 0 ? j - 1 : j;
        int kp = k < mz - 1 ? k + 1 : k, km = k > 0 ? k - 1 : k;
        if ((int)(nv[c] + 0.5) != 4)
          for (int kk = km; kk < kp + 1; kk++)
            for (int jj = jm; jj < jp + 1; jj++)
              for (int ii = im; ii < ip + 1; ii++)
                if ((int)(nv[IX(g, ii, jj, kk)] + 0.5) == 4) nv[c] = nv[c] > 2 ? nv[c] : 2;
      }
  /* remap 2->1, 4->3 on the interior, ibm.c:540-547 */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        if ((int)(nv[c] + 0.5) == 2) nv[c] = 1;
        if ((int)(nv[c] + 0.5) == 4) nv[c] = 3;
      }
  /* wall-model face layers, ibm.c:550-568 */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        size_t c = IX(g, i, j, k);
        int pos[6] = {i == 1, i == mx - 2, j == 1, j == my - 2, k == 1, k == mz - 2};
        for (int f = 0; f < 6; f++)
          if ((bctype[f] == -1 || bctype[f] == -2) && pos[f]) nv[c] = 1;
      }
}

/* ------------------------------------------------------------- MG primitives */
void or_mg_restrict(const or_grid *gf, const or_grid *gc, const double *nvf, const double *nvc,
                    const double *x, double *f) {
  for (int k = 0; k < gc->mz; k++)
    for (int j = 0; j < gc->my; j++)
      for (int i = 0; i < gc->mx; i++) {
        size_t c = IX(gc, i, j, k);
        if (is_bnd(gc, i, j, k)) { f[c] = 0.; continue; }
        int ih = 2 * i, jh = 2 * j, kh = 2 * k;
#define W(a, b, cc) (x[IX(gf, a, b, cc)] * fmax(0., 1 - nvf[IX(gf, a, b, cc)]))
        f[c] = 0.125 * (W(ih, jh, kh) + W(ih - 1, jh, kh) + W(ih, jh - 1, kh) + W(ih, jh, kh - 1) +
                        W(ih - 1, jh - 1, kh) + W(ih, jh - 1, kh - 1) + W(ih - 1, jh, kh - 1) +
                        W(ih - 1, jh - 1, kh - 1));
#undef W
        if (nvc[c] > 0.1) f[c] = 0.;
      }
}

void or_mg_prolong(const or_grid *gf, const or_grid *gc, const double *nvf, const double *nvc,
                   const double *x, double *f) {
  const int mx = gf->mx, my = gf->my, mz = gf->mz;
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        int ic = (i + 1) / 2, jc = (j + 1) / 2, kc = (k + 1) / 2;
        int ia = (i - 2 * ic) == 0 ? 1 : -1;
        int ja = (j - 2 * jc) == 0 ? 1 : -1;
        int ka = (k - 2 * kc) == 0 ? 1 : -1;
        if (i == 1 || i == mx - 2) ia = 0;
        if (j == 1 || j == my - 2) ja = 0;
        if (k == 1 || k == mz - 2) ka = 0;
        if (ka == -1 && nvc[IX(gc, ic, jc, kc - 1)] > 0.1) ka = 0;
        else if (ka == 1 && nvc[IX(gc, ic, jc, kc + 1)] > 0.1) ka = 0;
        if (ja == -1 && nvc[IX(gc, ic, jc - 1, kc)] > 0.1) ja = 0;
        else if (ja == 1 && nvc[IX(gc, ic, jc + 1, kc)] > 0.1) ja = 0;
        if (ia == -1 && nvc[IX(gc, ic - 1, jc, kc)] > 0.1) ia = 0;
        else if (ia == 1 && nvc[IX(gc, ic + 1, jc, kc)] > 0.1) ia = 0;
#define X(a, b, cc) x[IX(gc, a, b, cc)]
        f[IX(gf, i, j, k)] = (X(ic, jc, kc) * 9 + X(ic, jc + ja, kc) * 3 + X(ic + ia, jc, kc) * 3 +
                              X(ic + ia, jc + ja, kc)) * 3. / 64. +
                             (X(ic, jc, kc + ka) * 9 + X(ic, jc + ja, kc + ka) * 3 + X(ic + ia, jc, kc + ka) * 3 +
                              X(ic + ia, jc + ja, kc + ka)) / 64.;
#undef X
      }
  for (int k = 0; k < mz; k++)
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) {
        size_t c = IX(gf, i, j, k);
        if (is_bnd(gf, i, j, k)) f[c] = 0.;
        if (nvf[c] > 0.1) f[c] = 0.;
      }
}

void or_mg_nvert_restrict(const or_grid *gf, const or_grid *gc, const double *nvf, double *nvc) {
  size_t n = (size_t)gc->mx * gc->my * gc->mz;
  for (size_t p = 0; p < n; p++) nvc[p] = 0.;
  for (int k = 1; k < gc->mz - 1; k++)
    for (int j = 1; j < gc->my - 1; j++)
      for (int i = 1; i < gc->mx - 1; i++) {
        int ih = 2 * i, jh = 2 * j, kh = 2 * k;
#define V(a, b, cc) nvf[IX(gf, a, b, cc)]
        if (V(ih, jh, kh) * V(ih - 1, jh, kh) * V(ih, jh - 1, kh) * V(ih, jh, kh - 1) * V(ih - 1, jh - 1, kh) *
                V(ih - 1, jh, kh - 1) * V(ih, jh - 1, kh - 1) * V(ih - 1, jh - 1, kh - 1) >
            0.1) {
          size_t c = IX(gc, i, j, k);
          nvc[c] = nvc[c] > 1. ? nvc[c] : 1.;
        }
#undef V
      }
  /* isolated coarse fluid cells, poisson.c:2083-2097 (sequential in-place sweep) */
  const size_t sy = (size_t)gc->mx, sz = (size_t)gc->mx * gc->my;
  for (int k = 1; k < gc->mz - 1; k++)
    for (int j = 1; j < gc->my - 1; j++)
      for (int i = 1; i < gc->mx - 1; i++) {
        size_t c = IX(gc, i, j, k);
        if (nvc[c] < 0.1 && nvc[c + 1] + nvc[c - 1] > 1.1 && nvc[c + sy] + nvc[c - sy] > 1.1 &&
            nvc[c + sz] + nvc[c - sz] > 1.1)
          nvc[c] = 1.;
      }
}

/* ---------------------------------------------------------------- GMRES(m) */
/* Dot product over a fixed split into DOT_BLOCKS contiguous blocks, each summed
 * in order by one thread, the block sums then added in block order: the result
 * is the same for every thread count and every run. */
#define DOT_BLOCKS 256
static double det_dot(size_t n, const double *a, const double *b) {
  double part[DOT_BLOCKS];
#pragma omp parallel for schedule(static)
  for (int blk = 0; blk < DOT_BLOCKS; blk++) {
    size_t lo = n * (size_t)blk / DOT_BLOCKS, hi = n * (size_t)(blk + 1) / DOT_BLOCKS;
    double s = 0.;
    for (size_t i = lo; i < hi; i++) s += a[i] * b[i];
    part[blk] = s;
  }
  double s = 0.;
  for (int blk = 0; blk < DOT_BLOCKS; blk++) s += part[blk];
  return s;
}

static double vdot(int n, const double *a, const double *b) { return det_dot((size_t)n, a, b); }

int or_gmres(int n, const int *rowptr, const int *col, const double *val, const double *b,
             double *x, int m, double rtol, int maxit, int precond, double *rel) {
  double *V = (double *)malloc((size_t)(m + 1) * n * sizeof(double));
  double *Z = (double *)malloc((size_t)m * n * sizeof(double));
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double *)malloc(m * sizeof(double)), *sn = (double *)malloc(m * sizeof(double));
  double *gv = (double *)malloc((m + 1) * sizeof(double)), *y = (double *)malloc(m * sizeof(double));
  double *dinv = (double *)malloc(n * sizeof(double));
  double *w = (double *)malloc(n * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int r = 0; r < n; r++) {
    double d = 1.;
    if (precond == 1)
      for (int t = rowptr[r]; t < rowptr[r + 1]; t++)
        if (col[t] == r && val[t] != 0.) d = 1. / val[t];
    dinv[r] = d;
  }
  double bn = sqrt(vdot(n, b, b));
  if (bn == 0.) bn = 1.;
  int it = 0;
  double res = 0.;
  for (;;) {
    or_csr_matvec(n, rowptr, col, val, x, w);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) V[i] = b[i] - w[i];
    double beta = sqrt(vdot(n, V, V));
    res = beta / bn;
    if (res <= rtol || it >= maxit) break;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) V[i] /= beta;
    for (int i = 0; i <= m; i++) gv[i] = 0.;
    gv[0] = beta;
    int j;
    for (j = 0; j < m && it < maxit; j++, it++) {
      double *zj = Z + (size_t)j * n, *vj = V + (size_t)j * n, *vn = V + (size_t)(j + 1) * n;
#pragma omp parallel for schedule(static)
      for (int i = 0; i < n; i++) zj[i] = dinv[i] * vj[i];
      or_csr_matvec(n, rowptr, col, val, zj, vn);
      for (int l = 0; l <= j; l++) { /* modified Gram-Schmidt */
        double h = vdot(n, vn, V + (size_t)l * n);
        H[l * m + j] = h;
        const double *vl = V + (size_t)l * n;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) vn[i] -= h * vl[i];
      }
      double hn = sqrt(vdot(n, vn, vn));
      H[(j + 1) * m + j] = hn;
      if (hn != 0.) {
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) vn[i] /= hn;
      }
      for (int l = 0; l < j; l++) {
        double t1 = cs[l] * H[l * m + j] + sn[l] * H[(l + 1) * m + j];
        double t2 = -sn[l] * H[l * m + j] + cs[l] * H[(l + 1) * m + j];
        H[l * m + j] = t1; H[(l + 1) * m + j] = t2;
      }
      double a = H[j * m + j], bb = H[(j + 1) * m + j], r = sqrt(a * a + bb * bb);
      cs[j] = r == 0. ? 1. : a / r; sn[j] = r == 0. ? 0. : bb / r;
      H[j * m + j] = r; H[(j + 1) * m + j] = 0.;
      gv[j + 1] = -sn[j] * gv[j]; gv[j] = cs[j] * gv[j];
      if (fabs(gv[j + 1]) / bn <= rtol) { j++; it++; break; }
    }
    for (int l = j - 1; l >= 0; l--) {
      double s = gv[l];
      for (int q = l + 1; q < j; q++) s -= H[l * m + q] * y[q];
      y[l] = s / H[l * m + l];
    }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
      double xi = x[i];
      for (int l = 0; l < j; l++) xi += y[l] * Z[(size_t)l * n + i];
      x[i] = xi;
    }
  }
  if (rel) *rel = res;
  free(V); free(Z); free(H); free(cs); free(sn); free(gv); free(y); free(dinv); free(w);
  return it;
}

/* ===================================================================== CPU MG
 * CPU port of the engine's solver, for the CPU baseline and for checking the
 * device multigrid: geometric multigrid (coarse face coefficients summed over
 * the active fine faces and divided by the coarsening factor of the difference
 * direction, coarse nvert by the reference's MyNvertRestriction rule, Chebyshev
 * smoothing preconditioned by the diagonal, summed restriction of fluid-child
 * residuals, trilinear prolongation with the (27,9,9,3,9,3,3,1)/64 weights,
 * bordered dense solve on the coarsest level) inside FP64 FGMRES(m) with CGS2.
 * Vectors are interior arrays (nx*ny*nz, i fastest) with periodic wrap. */

typedef struct {
  int nx, ny, nz, per[3];
  size_t n;
  or_grid g;        /* DMDA view of this level */
  double *g9;       /* DMDA face coefficients */
  double *nv;       /* DMDA nvert */
  double *vol;      /* DMDA row coefficients, 19 per point */
  uint8_t *row, *cases;
  double *f, *x, *xt, *dv, *r;
  double c1[16], c2[16], lam;
  int cmask;
} mg_lev;

struct or_mg {
  int nl, degree;
  mg_lev L[16];
  double *cinv; /* coarsest bordered inverse, n x n row-major */
};

static size_t lidx(const mg_lev *L, int i, int j, int k) { return ((size_t)k * L->ny + j) * L->nx + i; }
static size_t ldm(const mg_lev *L, int i, int j, int k) { return IX(&L->g, i + 1, j + 1, k + 1); }
static int wr(int i, int n, int per) { return i < 0 ? (per ? i + n : 0) : (i >= n ? (per ? i - n : n - 1) : i); }

/* y = A x on interior vectors using the level's row coefficients */
static double row_apply(const mg_lev *L, int i, int j, int k, const double *x) {
  size_t p = ldm(L, i, j, k);
  if (!L->row[p]) return 0.;
  double s = 0.;
  for (int t = 0; t < 19; t++) {
    double v = L->vol[19 * p + t];
    if (v == 0.) continue;
    int a = wr(i + SLOT_OFF[t][0], L->nx, L->per[0]);
    int b = wr(j + SLOT_OFF[t][1], L->ny, L->per[1]);
    int c = wr(k + SLOT_OFF[t][2], L->nz, L->per[2]);
    s += v * x[lidx(L, a, b, c)];
  }
  return s;
}

static void lev_rows(mg_lev *L) {
  size_t M = (size_t)L->g.mx * L->g.my * L->g.mz;
  L->vol = (double *)malloc(19 * M * sizeof(double));
  L->row = (uint8_t *)malloc(M);
  L->cases = (uint8_t *)malloc(18 * M);
  or_stencil(&L->g, L->g9, L->nv, L->vol, L->row, L->cases);
}

static void lev_alloc(mg_lev *L) {
  L->n = (size_t)L->nx * L->ny * L->nz;
  L->f = (double *)calloc(L->n, sizeof(double));
  L->x = (double *)calloc(L->n, sizeof(double));
  L->xt = (double *)calloc(L->n, sizeof(double));
  L->dv = (double *)calloc(L->n, sizeof(double));
  L->r = (double *)calloc(L->n, sizeof(double));
}

static void dm_periodic(const or_grid *g, double *a) { periodic_copy_scalar(g, a); }

/* axes coarsened from a level of nx x ny x nz cells: every axis with at least 3
 * cells, except j when keep_y (semi-coarsening in i and k) while i or k can be */
static int coarse_mask(int nx, int ny, int nz, int keep_y) {
  int cm = (nx >= 3 ? 1 : 0) | (ny >= 3 ? 2 : 0) | (nz >= 3 ? 4 : 0);
  if (keep_y && (cm & 5)) cm &= ~2;
  return cm;
}

static void coarsen_level(mg_lev *F, mg_lev *C, int keep_y) {
  int cm = coarse_mask(F->nx, F->ny, F->nz, keep_y);
  F->cmask = cm;
  C->nx = (cm & 1) ? (F->nx + 1) / 2 : F->nx;
  C->ny = (cm & 2) ? (F->ny + 1) / 2 : F->ny;
  C->nz = (cm & 4) ? (F->nz + 1) / 2 : F->nz;
  for (int a = 0; a < 3; a++) C->per[a] = F->per[a];
  C->g.mx = C->nx + 2; C->g.my = C->ny + 2; C->g.mz = C->nz + 2;
  for (int a = 0; a < 3; a++) C->g.per[a] = C->per[a];
  size_t M = (size_t)C->g.mx * C->g.my * C->g.mz;
  C->nv = (double *)calloc(M, sizeof(double));
  C->g9 = (double *)calloc(9 * M, sizeof(double));
  const int fn[3] = {F->nx, F->ny, F->nz};
  for (int k = 0; k < C->nz; k++)
    for (int j = 0; j < C->ny; j++)
      for (int i = 0; i < C->nx; i++) {
        int lo[3], hi[3], ci[3] = {i, j, k};
        for (int a = 0; a < 3; a++) {
          if (cm & (1 << a)) { lo[a] = 2 * ci[a]; hi[a] = 2 * ci[a] + 1 < fn[a] ? 2 * ci[a] + 1 : fn[a] - 1; }
          else lo[a] = hi[a] = ci[a];
        }
        double prod = 1.;
        for (int kk = lo[2]; kk <= hi[2]; kk++)
          for (int jj = lo[1]; jj <= hi[1]; jj++)
            for (int ii = lo[0]; ii <= hi[0]; ii++) prod *= F->nv[ldm(F, ii, jj, kk)];
        if (prod > 0.1) C->nv[ldm(C, i, j, k)] = 1.;
        for (int d = 0; d < 3; d++) {
          int l2[3] = {lo[0], lo[1], lo[2]};
          l2[d] = hi[d];
          double acc[3] = {0, 0, 0};
          for (int kk = l2[2]; kk <= hi[2]; kk++)
            for (int jj = l2[1]; jj <= hi[1]; jj++)
              for (int ii = l2[0]; ii <= hi[0]; ii++) {
                size_t p = ldm(F, ii, jj, kk);
                if (!F->cases[18 * p + 3 * (2 * d)]) continue; /* active upper face */
                for (int e = 0; e < 3; e++) acc[e] += F->g9[(size_t)(3 * d + e) * ((size_t)F->g.mx * F->g.my * F->g.mz) + p];
              }
          for (int e = 0; e < 3; e++) {
            const double fac = (cm & (1 << e)) ? 2. : 1.;
            C->g9[(size_t)(3 * d + e) * M + ldm(C, i, j, k)] = acc[e] / fac;
          }
        }
      }
  dm_periodic(&C->g, C->nv);
  {
    double *pre = (double *)malloc(M * sizeof(double));
    memcpy(pre, C->nv, M * sizeof(double));
    const size_t sy = C->g.mx, sz = (size_t)C->g.mx * C->g.my;
    for (int k = 1; k < C->g.mz - 1; k++)
      for (int j = 1; j < C->g.my - 1; j++)
        for (int i = 1; i < C->g.mx - 1; i++) {
          size_t p = IX(&C->g, i, j, k);
          if (pre[p] < 0.1 && pre[p + 1] + pre[p - 1] > 1.1 && pre[p + sy] + pre[p - sy] > 1.1 &&
              pre[p + sz] + pre[p - sz] > 1.1)
            C->nv[p] = 1.;
        }
    free(pre);
    dm_periodic(&C->g, C->nv);
  }
  for (int e = 0; e < 9; e++) dm_periodic(&C->g, C->g9 + (size_t)e * M);
  lev_rows(C);
  /* rows with no coupling become solid */
  for (int it = 0; it < 3; it++) {
    int changed = 0;
    for (int k = 0; k < C->nz; k++)
      for (int j = 0; j < C->ny; j++)
        for (int i = 0; i < C->nx; i++) {
          size_t p = ldm(C, i, j, k);
          if (C->row[p] && C->vol[19 * p] == 0.) { C->nv[p] = 1.; changed = 1; }
        }
    if (!changed) break;
    dm_periodic(&C->g, C->nv);
    free(C->vol); free(C->row); free(C->cases);
    lev_rows(C);
  }
}

/* one Chebyshev step: z = D^-1 (f - A x), d = c1 d + c2 z, x_new = x + d */
static void cheb_step(mg_lev *L, const double *xin, double *xout, double c1, double c2, int zero_guess) {
#pragma omp parallel for schedule(static)
  for (int k = 0; k < L->nz; k++)
    for (int j = 0; j < L->ny; j++)
      for (int i = 0; i < L->nx; i++) {
        size_t c = lidx(L, i, j, k), p = ldm(L, i, j, k);
        if (!L->row[p]) { xout[c] = 0.; L->dv[c] = 0.; continue; }
        double diag = L->vol[19 * p];
        double r = L->f[c] - (zero_guess ? 0. : row_apply(L, i, j, k, xin));
        double z = diag != 0. ? r / diag : 0.;
        double dn = (zero_guess || c1 == 0.) ? c2 * z : c1 * L->dv[c] + c2 * z;
        L->dv[c] = dn;
        xout[c] = (zero_guess ? 0. : xin[c]) + dn;
      }
}

static double vnorm(size_t n, const double *a) { return sqrt(det_dot(n, a, a)); }

static double power_lambda(mg_lev *L) {
  unsigned s = 1234u;
  for (size_t c = 0; c < L->n; c++) {
    s = s * 1664525u + 1013904223u;
    L->x[c] = ((double)(s >> 8) / (double)(1u << 24)) * 2. - 1.;
  }
  memset(L->f, 0, L->n * sizeof(double));
  double best = 0.;
  for (int it = 0; it < 12; it++) {
    double nx = vnorm(L->n, L->x);
    for (size_t c = 0; c < L->n; c++) L->x[c] /= nx;
    cheb_step(L, L->x, L->xt, 0., 1., 0);
    double lam = vnorm(L->n, L->dv);
    if (it >= 6 && lam > best) best = lam;
    memcpy(L->x, L->dv, L->n * sizeof(double));
  }
  return best > 0. ? best : 2.;
}

static void cheb_coeffs(mg_lev *L, int deg, double lo, double hi) {
  double a = lo * L->lam, b = hi * L->lam;
  double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta, rho = 1. / sigma;
  L->c1[0] = 0.; L->c2[0] = 1. / theta;
  for (int k = 1; k < deg; k++) {
    double rn = 1. / (2. * sigma - rho);
    L->c1[k] = rn * rho; L->c2[k] = 2. * rn / delta; rho = rn;
  }
}

static double *bordered_inverse(mg_lev *L) {
  int n = (int)L->n;
  int *comp = (int *)malloc(n * sizeof(int));
  for (int c = 0; c < n; c++) comp[c] = -1;
  int *stack = (int *)malloc(n * sizeof(int)), nc = 0;
  for (int k0 = 0; k0 < L->nz; k0++)
    for (int j0 = 0; j0 < L->ny; j0++)
      for (int i0 = 0; i0 < L->nx; i0++) {
        int c0 = (int)lidx(L, i0, j0, k0);
        if (!L->row[ldm(L, i0, j0, k0)] || comp[c0] >= 0) continue;
        int sp = 0;
        stack[sp++] = c0; comp[c0] = nc;
        while (sp) {
          int u = stack[--sp];
          int i = u % L->nx, j = (u / L->nx) % L->ny, k = u / (L->nx * L->ny);
          size_t p = ldm(L, i, j, k);
          for (int t = 1; t < 19; t++) {
            if (L->vol[19 * p + t] == 0.) continue;
            int a = wr(i + SLOT_OFF[t][0], L->nx, L->per[0]), b = wr(j + SLOT_OFF[t][1], L->ny, L->per[1]),
                cc = wr(k + SLOT_OFF[t][2], L->nz, L->per[2]);
            int v = (int)lidx(L, a, b, cc);
            if (L->row[ldm(L, a, b, cc)] && comp[v] < 0) { comp[v] = nc; stack[sp++] = v; }
          }
        }
        nc++;
      }
  int N = n + nc;
  double *B = (double *)calloc((size_t)N * N, sizeof(double)), *I = (double *)calloc((size_t)N * N, sizeof(double));
  for (int k = 0; k < L->nz; k++)
    for (int j = 0; j < L->ny; j++)
      for (int i = 0; i < L->nx; i++) {
        int c = (int)lidx(L, i, j, k);
        size_t p = ldm(L, i, j, k);
        if (!L->row[p]) { B[(size_t)c * N + c] = 1.; continue; }
        for (int t = 0; t < 19; t++) {
          double v = L->vol[19 * p + t];
          if (v == 0.) continue;
          int a = wr(i + SLOT_OFF[t][0], L->nx, L->per[0]), b = wr(j + SLOT_OFF[t][1], L->ny, L->per[1]),
              cc = wr(k + SLOT_OFF[t][2], L->nz, L->per[2]);
          B[(size_t)c * N + lidx(L, a, b, cc)] += v;
        }
        B[(size_t)c * N + n + comp[c]] = 1.;
        B[(size_t)(n + comp[c]) * N + c] = 1.;
      }
  for (int r = 0; r < N; r++) I[(size_t)r * N + r] = 1.;
  for (int col = 0; col < N; col++) {
    int piv = col;
    double best = fabs(B[(size_t)col * N + col]);
    for (int r = col + 1; r < N; r++)
      if (fabs(B[(size_t)r * N + col]) > best) { best = fabs(B[(size_t)r * N + col]); piv = r; }
    if (piv != col)
      for (int t = 0; t < N; t++) {
        double tmp = B[(size_t)piv * N + t]; B[(size_t)piv * N + t] = B[(size_t)col * N + t]; B[(size_t)col * N + t] = tmp;
        tmp = I[(size_t)piv * N + t]; I[(size_t)piv * N + t] = I[(size_t)col * N + t]; I[(size_t)col * N + t] = tmp;
      }
    double inv = 1. / B[(size_t)col * N + col];
    for (int t = 0; t < N; t++) { B[(size_t)col * N + t] *= inv; I[(size_t)col * N + t] *= inv; }
    for (int r = 0; r < N; r++) {
      if (r == col) continue;
      double f = B[(size_t)r * N + col];
      if (f == 0.) continue;
      for (int t = 0; t < N; t++) { B[(size_t)r * N + t] -= f * B[(size_t)col * N + t]; I[(size_t)r * N + t] -= f * I[(size_t)col * N + t]; }
    }
  }
  double *out = (double *)malloc((size_t)n * n * sizeof(double));
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) out[(size_t)r * n + c] = I[(size_t)r * N + c];
  free(B); free(I); free(comp); free(stack);
  return out;
}

/* Coupling of each fluid row along j against i and k: yc = |a_N| + |a_S|,
 * oc = |a_E| + |a_W| + |a_T| + |a_B| from the assembled row; the largest yc / oc and
 * the fraction of rows with yc < oc / 10 (rows with oc = 0 are skipped). */
void or_level_coupling(const or_grid *g, const double *g9, const double *nvert, double *worst, double *weak_frac) {
  mg_lev L;
  memset(&L, 0, sizeof L);
  L.nx = g->mx - 2; L.ny = g->my - 2; L.nz = g->mz - 2;
  L.g = *g;
  L.g9 = (double *)g9;
  L.nv = (double *)nvert;
  lev_rows(&L);
  double w = 0.;
  size_t rows = 0, weak = 0;
  for (int k = 0; k < L.nz; k++)
    for (int j = 0; j < L.ny; j++)
      for (int i = 0; i < L.nx; i++) {
        size_t p = ldm(&L, i, j, k);
        if (!L.row[p]) continue;
        const double *v = L.vol + 19 * p;
        double yc = fabs(v[3]) + fabs(v[4]);             /* (0,+1,0), (0,-1,0) */
        double oc = fabs(v[1]) + fabs(v[2]) + fabs(v[5]) + fabs(v[6]);
        if (!(oc > 0.)) continue;
        if (yc / oc > w) w = yc / oc;
        rows++;
        if (yc < 0.1 * oc) weak++;
      }
  *worst = w;
  *weak_frac = rows ? (double)weak / (double)rows : 0.;
  free(L.vol); free(L.row); free(L.cases);
}

int or_level_coarsen(const or_grid *gf, const double *g9f, const double *nvf, int keep_y, or_grid *gc, double *g9c,
                     double *nvc) {
  const int nx = gf->mx - 2, ny = gf->my - 2, nz = gf->mz - 2;
  const int cm = coarse_mask(nx, ny, nz, keep_y);
  gc->mx = ((cm & 1) ? (nx + 1) / 2 : nx) + 2;
  gc->my = ((cm & 2) ? (ny + 1) / 2 : ny) + 2;
  gc->mz = ((cm & 4) ? (nz + 1) / 2 : nz) + 2;
  for (int a = 0; a < 3; a++) gc->per[a] = gf->per[a];
  if (!g9c || !nvc) return cm;
  mg_lev F, C;
  memset(&F, 0, sizeof F);
  memset(&C, 0, sizeof C);
  F.nx = nx; F.ny = ny; F.nz = nz;
  for (int a = 0; a < 3; a++) F.per[a] = gf->per[a];
  F.g = *gf;
  F.g9 = (double *)g9f;
  F.nv = (double *)nvf;
  lev_rows(&F);
  coarsen_level(&F, &C, keep_y);
  const size_t M = (size_t)C.g.mx * C.g.my * C.g.mz;
  memcpy(g9c, C.g9, 9 * M * sizeof(double));
  memcpy(nvc, C.nv, M * sizeof(double));
  free(F.vol); free(F.row); free(F.cases);
  free(C.g9); free(C.nv); free(C.vol); free(C.row); free(C.cases);
  return cm;
}

or_mg *or_mg_create(const or_grid *g, const double *g9, const double *nvert, int degree, double lo, double hi,
                    int coarse_max) {
  or_mg *mg = (or_mg *)calloc(1, sizeof(or_mg));
  mg->degree = degree;
  mg_lev *L0 = &mg->L[0];
  L0->nx = g->mx - 2; L0->ny = g->my - 2; L0->nz = g->mz - 2;
  for (int a = 0; a < 3; a++) L0->per[a] = g->per[a];
  L0->g = *g;
  size_t M = (size_t)g->mx * g->my * g->mz;
  L0->g9 = (double *)malloc(9 * M * sizeof(double));
  memcpy(L0->g9, g9, 9 * M * sizeof(double));
  L0->nv = (double *)malloc(M * sizeof(double));
  memcpy(L0->nv, nvert, M * sizeof(double));
  lev_rows(L0);
  lev_alloc(L0);
  mg->nl = 1;
  while (mg->nl < 16) {
    mg_lev *F = &mg->L[mg->nl - 1];
    int can = F->nx >= 3 || F->ny >= 3 || F->nz >= 3;
    if ((long)F->n <= coarse_max || !can) break;
    mg_lev *C = &mg->L[mg->nl];
    coarsen_level(F, C, 0);
    lev_alloc(C);
    mg->nl++;
  }
  for (int l = 0; l + 1 < mg->nl; l++) {
    mg->L[l].lam = power_lambda(&mg->L[l]);
    cheb_coeffs(&mg->L[l], degree, lo, hi);
  }
  mg->cinv = bordered_inverse(&mg->L[mg->nl - 1]);
  return mg;
}

void or_mg_destroy(or_mg *mg) {
  if (!mg) return;
  for (int l = 0; l < mg->nl; l++) {
    mg_lev *L = &mg->L[l];
    free(L->g9); free(L->nv); free(L->vol); free(L->row); free(L->cases);
    free(L->f); free(L->x); free(L->xt); free(L->dv); free(L->r);
  }
  free(mg->cinv);
  free(mg);
}

static void mg_cycle(or_mg *mg, int l, double *out) {
  mg_lev *L = &mg->L[l];
  if (l == mg->nl - 1) {
    int n = (int)L->n;
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n; r++) {
      double s = 0.;
      for (int c = 0; c < n; c++) s += mg->cinv[(size_t)r * n + c] * L->f[c];
      out[r] = s;
    }
    return;
  }
  double *x = L->x, *y = L->xt;
  for (int k = 0; k < mg->degree; k++) {
    cheb_step(L, x, y, L->c1[k], L->c2[k], k == 0);
    double *t = x; x = y; y = t;
  }
  mg_lev *C = &mg->L[l + 1];
  const int cm = L->cmask;
  /* restriction of fluid-child residuals */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < C->nz; k++)
    for (int j = 0; j < C->ny; j++)
      for (int i = 0; i < C->nx; i++) {
        size_t cc = lidx(C, i, j, k);
        if (C->nv[ldm(C, i, j, k)] > 0.1) { C->f[cc] = 0.; continue; }
        int ci[3] = {i, j, k}, fn[3] = {L->nx, L->ny, L->nz}, lo[3], hi[3];
        for (int a = 0; a < 3; a++) {
          if (cm & (1 << a)) { lo[a] = 2 * ci[a]; hi[a] = 2 * ci[a] + 1 < fn[a] ? 2 * ci[a] + 1 : fn[a] - 1; }
          else lo[a] = hi[a] = ci[a];
        }
        double s = 0.;
        for (int kk = lo[2]; kk <= hi[2]; kk++)
          for (int jj = lo[1]; jj <= hi[1]; jj++)
            for (int ii = lo[0]; ii <= hi[0]; ii++)
              if (L->row[ldm(L, ii, jj, kk)]) s += L->f[lidx(L, ii, jj, kk)] - row_apply(L, ii, jj, kk, x);
        C->f[cc] = s;
      }
  double *xc = (double *)malloc(C->n * sizeof(double));
  mg_cycle(mg, l + 1, xc);
  /* prolongation and correction */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < L->nz; k++)
    for (int j = 0; j < L->ny; j++)
      for (int i = 0; i < L->nx; i++) {
        if (L->nv[ldm(L, i, j, k)] > 0.1) continue;
        int f[3] = {i, j, k}, fn[3] = {L->nx, L->ny, L->nz}, cn[3] = {C->nx, C->ny, C->nz}, c0[3], o[3], nb[3];
        for (int a = 0; a < 3; a++) {
          if (!(cm & (1 << a))) { c0[a] = f[a]; o[a] = 0; }
          else {
            c0[a] = f[a] / 2;
            o[a] = (f[a] & 1) ? 1 : -1;
            if (!L->per[a] && (f[a] == 0 || f[a] == fn[a] - 1)) o[a] = 0;
          }
          int q = c0[a] + o[a];
          nb[a] = q < 0 ? q + cn[a] : (q >= cn[a] ? q - cn[a] : q);
        }
        if (o[2] && C->nv[ldm(C, c0[0], c0[1], nb[2])] > 0.1) o[2] = 0;
        if (o[1] && C->nv[ldm(C, c0[0], nb[1], c0[2])] > 0.1) o[1] = 0;
        if (o[0] && C->nv[ldm(C, nb[0], c0[1], c0[2])] > 0.1) o[0] = 0;
        int i2 = o[0] ? nb[0] : c0[0], j2 = o[1] ? nb[1] : c0[1], k2 = o[2] ? nb[2] : c0[2];
#define XC(a, b, cc) xc[lidx(C, a, b, cc)]
        double v = (XC(c0[0], c0[1], c0[2]) * 9 + XC(c0[0], j2, c0[2]) * 3 + XC(i2, c0[1], c0[2]) * 3 + XC(i2, j2, c0[2])) * (3. / 64.) +
                   (XC(c0[0], c0[1], k2) * 9 + XC(c0[0], j2, k2) * 3 + XC(i2, c0[1], k2) * 3 + XC(i2, j2, k2)) * (1. / 64.);
#undef XC
        x[lidx(L, i, j, k)] += v;
      }
  free(xc);
  for (int k = 0; k < mg->degree; k++) {
    cheb_step(L, x, y, L->c1[k], L->c2[k], 0);
    double *t = x; x = y; y = t;
  }
  memcpy(out, x, L->n * sizeof(double));
}

void or_mg_vcycle(or_mg *mg, const double *r, double *z) {
  memcpy(mg->L[0].f, r, mg->L[0].n * sizeof(double));
  mg_cycle(mg, 0, z);
}

size_t or_mg_n(const or_mg *mg) { return mg->L[0].n; }
int or_mg_levels(const or_mg *mg) { return mg->nl; }

/* FGMRES(m) + CGS2 on interior vectors, right preconditioned by one V-cycle.
 * b, x interior (x holds the initial guess).  Returns iterations. */
int or_mg_fgmres(or_mg *mg, const double *b, double *x, int m, double rtol, int maxit, double *rel, int *vcyc) {
  mg_lev *L0 = &mg->L[0];
  size_t n = L0->n;
  double *V = (double *)malloc((size_t)(m + 1) * n * sizeof(double));
  double *Z = (double *)malloc((size_t)m * n * sizeof(double));
  double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
  double *cs = (double *)malloc(m * sizeof(double)), *sn = (double *)malloc(m * sizeof(double));
  double *gv = (double *)malloc((m + 1) * sizeof(double)), *y = (double *)malloc(m * sizeof(double));
  double bn = vnorm(n, b);
  int it = 0;
  double res = 1.;
  *vcyc = 0;
  for (;;) {
#pragma omp parallel for schedule(static)
    for (int k = 0; k < L0->nz; k++)
      for (int j = 0; j < L0->ny; j++)
        for (int i = 0; i < L0->nx; i++) {
          size_t c = lidx(L0, i, j, k);
          V[c] = L0->row[ldm(L0, i, j, k)] ? b[c] - row_apply(L0, i, j, k, x) : 0.;
        }
    double beta = vnorm(n, V);
    res = bn > 0 ? beta / bn : 0.;
    if (beta <= rtol * bn || it >= maxit) break;
    for (size_t i = 0; i < n; i++) V[i] /= beta;
    for (int i = 0; i <= m; i++) gv[i] = 0.;
    gv[0] = beta;
    int j;
    for (j = 0; j < m && it < maxit; j++) {
      double *vj = V + (size_t)j * n, *zj = Z + (size_t)j * n, *w = V + (size_t)(j + 1) * n;
      or_mg_vcycle(mg, vj, zj);
      (*vcyc)++;
#pragma omp parallel for schedule(static)
      for (int k = 0; k < L0->nz; k++)
        for (int jj = 0; jj < L0->ny; jj++)
          for (int i = 0; i < L0->nx; i++) w[lidx(L0, i, jj, k)] = row_apply(L0, i, jj, k, zj);
      for (int l = 0; l <= j; l++) H[l * m + j] = 0.;
      for (int pass = 0; pass < 2; pass++)
        for (int l = 0; l <= j; l++) {
          const double *vl = V + (size_t)l * n;
          double h = det_dot(n, w, vl);
#pragma omp parallel for schedule(static)
          for (size_t i = 0; i < n; i++) w[i] -= h * vl[i];
          H[l * m + j] += h;
        }
      double hn = vnorm(n, w);
      H[(j + 1) * m + j] = hn;
      if (hn != 0.) for (size_t i = 0; i < n; i++) w[i] /= hn;
      for (int l = 0; l < j; l++) {
        double t1 = cs[l] * H[l * m + j] + sn[l] * H[(l + 1) * m + j];
        double t2 = -sn[l] * H[l * m + j] + cs[l] * H[(l + 1) * m + j];
        H[l * m + j] = t1; H[(l + 1) * m + j] = t2;
      }
      double a = H[j * m + j], bb = H[(j + 1) * m + j], r = sqrt(a * a + bb * bb);
      cs[j] = r == 0. ? 1. : a / r; sn[j] = r == 0. ? 0. : bb / r;
      H[j * m + j] = r; H[(j + 1) * m + j] = 0.;
      gv[j + 1] = -sn[j] * gv[j]; gv[j] = cs[j] * gv[j];
      it++;
      if (fabs(gv[j + 1]) <= rtol * bn) { j++; break; }
    }
    for (int l = j - 1; l >= 0; l--) {
      double s = gv[l];
      for (int q = l + 1; q < j; q++) s -= H[l * m + q] * y[q];
      y[l] = s / H[l * m + l];
    }
    for (int l = 0; l < j; l++) {
      const double *zl = Z + (size_t)l * n;
#pragma omp parallel for schedule(static)
      for (size_t i = 0; i < n; i++) x[i] += y[l] * zl[i];
    }
  }
  if (rel) *rel = res;
  free(V); free(Z); free(H); free(cs); free(sn); free(gv); free(y);
  return it;
}

/* ---------------------------------------------------------------- FormMetrics */
typedef struct { double x, y, z; } or_v3;

static or_v3 v3_at(const double *a, size_t p) { or_v3 v = {a[3 * p], a[3 * p + 1], a[3 * p + 2]}; return v; }
static void v3_put(double *a, size_t p, or_v3 v) { a[3 * p] = v.x; a[3 * p + 1] = v.y; a[3 * p + 2] = v.z; }

/* metric triple and inverse Jacobian from the derivatives along xi, eta, zeta
 * (metrics.c:700-724 and the same block in the j and k sections) */
static void metric_triple(or_v3 c, or_v3 e, or_v3 z, or_v3 *csi, or_v3 *eta, or_v3 *zet, double *aj) {
  csi->x = e.y * z.z - e.z * z.y;
  csi->y = -e.x * z.z + e.z * z.x;
  csi->z = e.x * z.y - e.y * z.x;
  eta->x = z.y * c.z - z.z * c.y;
  eta->y = -z.x * c.z + z.z * c.x;
  eta->z = z.x * c.y - z.y * c.x;
  zet->x = c.y * e.z - c.z * e.y;
  zet->y = -c.x * e.z + c.z * e.x;
  zet->z = c.x * e.y - c.y * e.x;
  double det = c.x * (e.y * z.z - e.z * z.y) - c.y * (e.x * z.z - e.z * z.x) + c.z * (e.x * z.y - e.y * z.x);
  *aj = 1. / det;
}

/* one-sided at the ends of [lo, hi], central inside, the periodic seam average at
 * hi when `seam` (metrics.c:628-700): f(t) is the face-centre value at index t */
static or_v3 dcent(const double *cent, const or_grid *g, int ax, int lo, int hi, int seam, int i, int j, int k) {
  int t = ax == 0 ? i : (ax == 1 ? j : k);
#define CAT(s) v3_at(cent, ax == 0 ? IX(g, (s), j, k) : (ax == 1 ? IX(g, i, (s), k) : IX(g, i, j, (s))))
  or_v3 r;
  if (t == lo) {
    or_v3 a = CAT(t + 1), b = CAT(t);
    r.x = a.x - b.x; r.y = a.y - b.y; r.z = a.z - b.z;
  } else if (t == hi) {
    if (seam) {
      or_v3 a = CAT(1), b = CAT(0), c = CAT(t), d = CAT(t - 1);
      r.x = (a.x - b.x + c.x - d.x) * 0.5; r.y = (a.y - b.y + c.y - d.y) * 0.5; r.z = (a.z - b.z + c.z - d.z) * 0.5;
    } else {
      or_v3 a = CAT(t), b = CAT(t - 1);
      r.x = a.x - b.x; r.y = a.y - b.y; r.z = a.z - b.z;
    }
  } else {
    or_v3 a = CAT(t + 1), b = CAT(t - 1);
    r.x = (a.x - b.x) * 0.5; r.y = (a.y - b.y) * 0.5; r.z = (a.z - b.z) * 0.5;
  }
#undef CAT
  return r;
}

/* 0.5 * (a + b - c - d) on nodes given by index */
static or_v3 half4(const double *co, size_t a, size_t b, size_t c, size_t d) {
  or_v3 r;
  r.x = 0.5 * (co[3 * a] + co[3 * b] - co[3 * c] - co[3 * d]);
  r.y = 0.5 * (co[3 * a + 1] + co[3 * b + 1] - co[3 * c + 1] - co[3 * d + 1]);
  r.z = 0.5 * (co[3 * a + 2] + co[3 * b + 2] - co[3 * c + 2] - co[3 * d + 2]);
  return r;
}

void or_form_metrics(const or_grid *g, const double *co, double *const out[16]) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  const size_t n = (size_t)mx * my * mz;
  double *icsi = out[0], *ieta = out[1], *izet = out[2], *jcsi = out[3], *jeta = out[4], *jzet = out[5];
  double *kcsi = out[6], *keta = out[7], *kzet = out[8], *iaj = out[9], *jaj = out[10], *kaj = out[11];
  double *aj = out[12], *csi = out[13], *eta = out[14], *zet = out[15];
  for (int q = 0; q < 16; q++) memset(out[q], 0, ((q < 9 || q >= 13) ? 3 : 1) * n * sizeof(double));
  double *cent = (double *)calloc(3 * n, sizeof(double));

  /* face centres and face metrics, one face family at a time (metrics.c:613-1004) */
  for (int fam = 0; fam < 3; fam++) {
    memset(cent, 0, 3 * n * sizeof(double));
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) {
          size_t a, b, c, d;
          if (fam == 0) { /* centx: corners over j and k (613-622) */
            if (k < 1 || j < 1) continue;
            a = IX(g, i, j, k); b = IX(g, i, j, k - 1); c = IX(g, i, j - 1, k); d = IX(g, i, j - 1, k - 1);
          } else if (fam == 1) { /* centy: corners over i and k (739-748) */
            if (k < 1 || i < 1) continue;
            a = IX(g, i, j, k); b = IX(g, i, j, k - 1); c = IX(g, i - 1, j, k); d = IX(g, i - 1, j, k - 1);
          } else { /* centz: corners over i and j (859-868) */
            if (j < 1 || i < 1) continue;
            a = IX(g, i, j, k); b = IX(g, i, j - 1, k); c = IX(g, i - 1, j, k); d = IX(g, i - 1, j - 1, k);
          }
          const size_t p = IX(g, i, j, k);
          for (int v = 0; v < 3; v++)
            cent[3 * p + v] = (co[3 * a + v] + co[3 * b + v] + co[3 * c + v] + co[3 * d + v]) * 0.25;
        }
    /* derivative ranges: the face-normal axis runs 0..m-2 (seam when periodic),
     * the two tangential axes 1..m-2 */
    int lo[3] = {1, 1, 1}, seam[3] = {0, 0, 0};
    lo[fam] = 0;
    seam[fam] = g->per[fam];
    double *o1 = fam == 0 ? icsi : (fam == 1 ? jcsi : kcsi);
    double *o2 = fam == 0 ? ieta : (fam == 1 ? jeta : keta);
    double *o3 = fam == 0 ? izet : (fam == 1 ? jzet : kzet);
    double *oj = fam == 0 ? iaj : (fam == 1 ? jaj : kaj);
    for (int k = lo[2]; k <= mz - 2; k++)
      for (int j = lo[1]; j <= my - 2; j++)
        for (int i = lo[0]; i <= mx - 2; i++) {
          or_v3 dc = dcent(cent, g, 0, lo[0], mx - 2, seam[0], i, j, k);
          or_v3 de = dcent(cent, g, 1, lo[1], my - 2, seam[1], i, j, k);
          or_v3 dz = dcent(cent, g, 2, lo[2], mz - 2, seam[2], i, j, k);
          or_v3 a, b, c;
          double J;
          metric_triple(dc, de, dz, &a, &b, &c, &J);
          const size_t p = IX(g, i, j, k);
          v3_put(o1, p, a); v3_put(o2, p, b); v3_put(o3, p, c);
          oj[p] = J;
        }
  }

  /* NEWMETRIC node-plane area vectors (metrics.c:185-270) */
  double *ct = (double *)calloc(3 * n, sizeof(double));
  double *et = (double *)calloc(3 * n, sizeof(double));
  double *zt = (double *)calloc(3 * n, sizeof(double));
  for (int k = 1; k <= mz - 2; k++)
    for (int j = 1; j <= my - 2; j++)
      for (int i = 0; i <= mx - 2; i++) {
        or_v3 e = half4(co, IX(g, i, j, k), IX(g, i, j, k - 1), IX(g, i, j - 1, k), IX(g, i, j - 1, k - 1));
        or_v3 z = half4(co, IX(g, i, j - 1, k), IX(g, i, j, k), IX(g, i, j - 1, k - 1), IX(g, i, j, k - 1));
        or_v3 r = {e.y * z.z - e.z * z.y, -e.x * z.z + e.z * z.x, e.x * z.y - e.y * z.x};
        v3_put(ct, IX(g, i, j, k), r);
      }
  for (int k = 1; k <= mz - 2; k++)
    for (int j = 0; j <= my - 2; j++)
      for (int i = 1; i <= mx - 2; i++) {
        or_v3 c = half4(co, IX(g, i, j, k), IX(g, i, j, k - 1), IX(g, i - 1, j, k), IX(g, i - 1, j, k - 1));
        or_v3 z = half4(co, IX(g, i, j, k), IX(g, i - 1, j, k), IX(g, i, j, k - 1), IX(g, i - 1, j, k - 1));
        or_v3 r = {z.y * c.z - z.z * c.y, -z.x * c.z + z.z * c.x, z.x * c.y - z.y * c.x};
        v3_put(et, IX(g, i, j, k), r);
      }
  for (int k = 0; k <= mz - 2; k++)
    for (int j = 1; j <= my - 2; j++)
      for (int i = 1; i <= mx - 2; i++) {
        or_v3 c = half4(co, IX(g, i, j, k), IX(g, i, j - 1, k), IX(g, i - 1, j, k), IX(g, i - 1, j - 1, k));
        or_v3 e = half4(co, IX(g, i, j, k), IX(g, i - 1, j, k), IX(g, i, j - 1, k), IX(g, i - 1, j - 1, k));
        or_v3 r = {c.y * e.z - c.z * e.y, -c.x * e.z + c.z * e.x, c.x * e.y - c.y * e.x};
        v3_put(zt, IX(g, i, j, k), r);
      }
  /* cell metrics as AxByC(0.5, ., 0.5, .) of the two planes (metrics.c:286-296)
   * and the cell Jacobian from the 8 corners (metrics.c:302-351) */
  for (int k = 1; k <= mz - 2; k++)
    for (int j = 1; j <= my - 2; j++)
      for (int i = 1; i <= mx - 2; i++) {
        const size_t p = IX(g, i, j, k);
        const size_t pi = IX(g, i - 1, j, k), pj = IX(g, i, j - 1, k), pk = IX(g, i, j, k - 1);
        for (int v = 0; v < 3; v++) {
          csi[3 * p + v] = 0.5 * ct[3 * p + v] + 0.5 * ct[3 * pi + v];
          eta[3 * p + v] = 0.5 * et[3 * p + v] + 0.5 * et[3 * pj + v];
          zet[3 * p + v] = 0.5 * zt[3 * p + v] + 0.5 * zt[3 * pk + v];
        }
        /* corners: (i|i-1, j|j-1, k|k-1) */
        const size_t n0 = IX(g, i, j, k), n1 = IX(g, i, j - 1, k), n2 = IX(g, i, j, k - 1),
                     n3 = IX(g, i, j - 1, k - 1), n4 = IX(g, i - 1, j, k), n5 = IX(g, i - 1, j - 1, k),
                     n6 = IX(g, i - 1, j, k - 1), n7 = IX(g, i - 1, j - 1, k - 1);
        or_v3 d[3];
        for (int v = 0; v < 3; v++) {
          const double c0 = co[3 * n0 + v], c1 = co[3 * n1 + v], c2 = co[3 * n2 + v], c3 = co[3 * n3 + v],
                       c4 = co[3 * n4 + v], c5 = co[3 * n5 + v], c6 = co[3 * n6 + v], c7 = co[3 * n7 + v];
          const double xc = 0.25 * (c0 + c1 + c2 + c3 - c4 - c5 - c6 - c7);
          const double xe = 0.25 * (c0 + c4 + c2 + c6 - c1 - c5 - c3 - c7);
          const double xz = 0.25 * (c0 + c1 + c4 + c5 - c2 - c3 - c6 - c7);
          if (v == 0) { d[0].x = xc; d[1].x = xe; d[2].x = xz; }
          if (v == 1) { d[0].y = xc; d[1].y = xe; d[2].y = xz; }
          if (v == 2) { d[0].z = xc; d[1].z = xe; d[2].z = xz; }
        }
        or_v3 a, b, c;
        metric_triple(d[0], d[1], d[2], &a, &b, &c, &aj[p]);
      }
  free(ct); free(et); free(zt); free(cent);

  /* boundary planes copy the adjacent interior plane, i then j then k
   * (metrics.c:374-454), then the periodic images (metrics.c:477-509) */
  double *cell[4] = {csi, eta, zet, aj};
  for (int q = 0; q < 4; q++) {
    const int w = q < 3 ? 3 : 1;
    double *a = cell[q];
#define CP(dst, src) for (int v = 0; v < w; v++) a[w * (dst) + v] = a[w * (src) + v]
    for (int k = 0; k < mz; k++)
      for (int j = 0; j < my; j++) { CP(IX(g, 0, j, k), IX(g, 1, j, k)); CP(IX(g, mx - 1, j, k), IX(g, mx - 2, j, k)); }
    for (int k = 0; k < mz; k++)
      for (int i = 0; i < mx; i++) { CP(IX(g, i, 0, k), IX(g, i, 1, k)); CP(IX(g, i, my - 1, k), IX(g, i, my - 2, k)); }
    for (int j = 0; j < my; j++)
      for (int i = 0; i < mx; i++) { CP(IX(g, i, j, 0), IX(g, i, j, 1)); CP(IX(g, i, j, mz - 1), IX(g, i, j, mz - 2)); }
    if (g->per[0])
      for (int k = 0; k < mz; k++)
        for (int j = 0; j < my; j++) { CP(IX(g, 0, j, k), IX(g, mx - 2, j, k)); CP(IX(g, mx - 1, j, k), IX(g, 1, j, k)); }
    if (g->per[1])
      for (int k = 0; k < mz; k++)
        for (int i = 0; i < mx; i++) { CP(IX(g, i, 0, k), IX(g, i, my - 2, k)); CP(IX(g, i, my - 1, k), IX(g, i, 1, k)); }
    if (g->per[2])
      for (int j = 0; j < my; j++)
        for (int i = 0; i < mx; i++) { CP(IX(g, i, j, 0), IX(g, i, j, mz - 2)); CP(IX(g, i, j, mz - 1), IX(g, i, j, 1)); }
#undef CP
  }
}

/* ------------------------------------------------------- IB ray-parity search */
/* s of randomdirection (ibm.c:48-79): srand(seed), then rand() / (RAND_MAX + 1) - 0.5 */
double or_ray_s(int seed) {
  srand((unsigned)seed);
  return rand() / ((double)RAND_MAX + 1) - 0.5;
}

/* Moller-Trumbore as compgeom.c:17-60 writes it (EPSILON 1e-15) */
static int tri_hit(const double o[3], const double d[3], const double a[3], const double b[3], const double c[3],
                   double *t) {
  double e1[3], e2[3], tv[3], pv[3], qv[3];
  e1[0] = b[0] - a[0]; e1[1] = b[1] - a[1]; e1[2] = b[2] - a[2];
  e2[0] = c[0] - a[0]; e2[1] = c[1] - a[1]; e2[2] = c[2] - a[2];
  pv[0] = d[1] * e2[2] - d[2] * e2[1];
  pv[1] = d[2] * e2[0] - d[0] * e2[2];
  pv[2] = d[0] * e2[1] - d[1] * e2[0];
  double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
  if (det > -1.e-15 && det < 1.e-15) return 0;
  double inv = 1.0 / det;
  tv[0] = o[0] - a[0]; tv[1] = o[1] - a[1]; tv[2] = o[2] - a[2];
  double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
  if (u < 0.0 || u > 1.0) return 0;
  qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
  qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
  qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
  double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
  if (v < 0.0 || u + v > 1.0) return 0;
  *t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
  return 1;
}

typedef struct { int *v; int n, cap; } or_bin;

static void bin_insert(or_bin *b, int e) { /* insertnode, variables.c:7-27: once per bin */
  for (int q = 0; q < b->n; q++)
    if (b->v[q] == e) return;
  if (b->n == b->cap) {
    b->cap = b->cap ? 2 * b->cap : 8;
    b->v = (int *)realloc(b->v, (size_t)b->cap * sizeof(int));
  }
  b->v[b->n++] = e;
}

int or_ib_search(const or_grid *g, const int bctype[6], const double *co, int n_v, const double *xb,
                 const double *yb, const double *zb, int n_elmt, const int *nv1, const int *nv2, const int *nv3,
                 int body, int classify, double *nvert) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  const int ncx = 40, ncy = 40, ncz = 40; /* ibm.c:100 */
  /* element normals, ibm_io.c:249-266 */
  double *nf = (double *)malloc(3 * (size_t)(n_elmt > 0 ? n_elmt : 1) * sizeof(double));
  for (int e = 0; e < n_elmt; e++) {
    int a = nv1[e], b = nv2[e], c = nv3[e];
    double dx12 = xb[b] - xb[a], dy12 = yb[b] - yb[a], dz12 = zb[b] - zb[a];
    double dx13 = xb[c] - xb[a], dy13 = yb[c] - yb[a], dz13 = zb[c] - zb[a];
    double nx = dy12 * dz13 - dz12 * dy13, ny = -dx12 * dz13 + dz12 * dx13, nz = dx12 * dy13 - dy12 * dx13;
    double dr = sqrt(nx * nx + ny * ny + nz * nz);
    nf[3 * e] = nx / dr; nf[3 * e + 1] = ny / dr; nf[3 * e + 2] = nz / dr;
  }
  /* bounding box + 0.05 and the 40^3 bins, ibm.c:128-207 */
  double x0 = 1.e23, x1 = -1.e23, y0 = 1.e23, y1 = -1.e23, z0 = 1.e23, z1 = -1.e23;
  for (int q = 0; q < n_v; q++) {
    x0 = x0 < xb[q] ? x0 : xb[q]; x1 = x1 > xb[q] ? x1 : xb[q];
    y0 = y0 < yb[q] ? y0 : yb[q]; y1 = y1 > yb[q] ? y1 : yb[q];
    z0 = z0 < zb[q] ? z0 : zb[q]; z1 = z1 > zb[q] ? z1 : zb[q];
  }
  x0 -= 0.05; x1 += 0.05; y0 -= 0.05; y1 += 0.05; z0 -= 0.05; z1 += 0.05;
  const double dcx = (x1 - x0) / (ncx - 1.), dcy = (y1 - y0) / (ncy - 1.), dcz = (z1 - z0) / (ncz - 1.);
  or_bin *bins = (or_bin *)calloc((size_t)ncx * ncy * ncz, sizeof(or_bin));
  for (int e = 0; e < n_elmt; e++) {
    int a = nv1[e], b = nv2[e], c = nv3[e];
#define MN3(p) ((p[a] < p[b] ? p[a] : p[b]) < p[c] ? (p[a] < p[b] ? p[a] : p[b]) : p[c])
#define MX3(p) ((p[a] > p[b] ? p[a] : p[b]) > p[c] ? (p[a] > p[b] ? p[a] : p[b]) : p[c])
    int i0 = (int)floor((MN3(xb) - x0) / dcx), i1 = (int)floor((MX3(xb) - x0) / dcx) + 1;
    int j0 = (int)floor((MN3(yb) - y0) / dcy), j1 = (int)floor((MX3(yb) - y0) / dcy) + 1;
    int k0 = (int)floor((MN3(zb) - z0) / dcz), k1 = (int)floor((MX3(zb) - z0) / dcz) + 1;
#undef MN3
#undef MX3
    i0 = i0 < 0 ? 0 : i0; i1 = i1 > ncx ? ncx : i1;
    j0 = j0 < 0 ? 0 : j0; j1 = j1 > ncx ? ncy : j1;
    k0 = k0 < 0 ? 0 : k0; k1 = k1 > ncz ? ncz : k1;
    for (int k = k0; k < k1; k++)
      for (int j = j0; j < j1; j++)
        for (int i = i0; i < i1; i++) bin_insert(&bins[(k * ncy + j) * ncx + i], e);
  }
  unsigned char *seen = (unsigned char *)malloc((size_t)(n_elmt > 0 ? n_elmt : 1));
  double *ht = (double *)malloc(1000 * sizeof(double));
  int *he = (int *)malloc(1000 * sizeof(int));
  int status = 0;
  /* search over interior cell centres, ibm.c:245-270 and point_cell_advanced 1777-1902 */
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        const size_t c = IX(g, i, j, k);
        if (body == 0 && nvert[c] < 5.) nvert[c] = 0;
        double p[3];
        for (int v = 0; v < 3; v++)
          p[v] = 0.125 * (co[3 * IX(g, i, j, k) + v] + co[3 * IX(g, i, j - 1, k) + v] +
                          co[3 * IX(g, i, j, k - 1) + v] + co[3 * IX(g, i, j - 1, k - 1) + v] +
                          co[3 * IX(g, i - 1, j, k) + v] + co[3 * IX(g, i - 1, j - 1, k) + v] +
                          co[3 * IX(g, i - 1, j, k - 1) + v] + co[3 * IX(g, i - 1, j - 1, k - 1) + v]);
        if (!(p[0] > x0 && p[0] < x1 && p[1] > y0 && p[1] < y1 && p[2] > z0 && p[2] < z1)) continue;
        const int ic = (int)floor((p[0] - x0) / dcx), jc = (int)floor((p[1] - y0) / dcy),
                  kc = (int)floor((p[2] - z0) / dcz);
        double val = -1;
        for (int st = 1; st <= 1000 && val < 0; st++) {
          const double s = or_ray_s(st);
          double d[3];
          d[0] = (dcx * (ic + 0.5) + x0 + s * dcx) - p[0];
          d[1] = (dcy * (jc + 0.5) + y0 + s * dcy) - p[1];
          d[2] = (z1 + 0.2) - p[2];
          double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
          d[0] /= len; d[1] /= len; d[2] /= len;
          memset(seen, 0, (size_t)(n_elmt > 0 ? n_elmt : 1));
          int nh = 0, sing = 0;
          for (int kk = kc; kk < ncz && !sing; kk++) {
            const or_bin *b = &bins[(kk * ncy + jc) * ncx + ic];
            for (int q = 0; q < b->n && !sing; q++) {
              const int e = b->v[q];
              if (seen[e]) continue;
              seen[e] = 1;
              const double A[3] = {xb[nv1[e]], yb[nv1[e]], zb[nv1[e]]};
              const double B[3] = {xb[nv2[e]], yb[nv2[e]], zb[nv2[e]]};
              const double Cc[3] = {xb[nv3[e]], yb[nv3[e]], zb[nv3[e]]};
              double t;
              if (tri_hit(p, d, A, B, Cc, &t) && t > 0) {
                if (nh == 1000) { status = 1; sing = 1; break; }
                for (int r = 0; r < nh; r++) {
                  const double *n0 = &nf[3 * he[r]], *n1 = &nf[3 * e];
                  const double nd = n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2];
                  if (fabs(t - ht[r]) < 1.e-8 && nd > -0.95) sing = 1;
                }
                ht[nh] = t; he[nh] = e; nh++;
              }
            }
          }
          if (!sing) val = (nh % 2) ? 4 : 0;
        }
        if (val < 0) { status = 2; val = 0; }
        nvert[c] = nvert[c] < val ? val : nvert[c];
        if (nvert[c] < 0) nvert[c] = 0;
      }
  for (int q = 0; q < ncx * ncy * ncz; q++) free(bins[q].v);
  free(bins); free(seen); free(ht); free(he); free(nf);
  if (classify) or_nvert_classify(g, bctype, nvert); /* ibm.c:325-353, 540-568 */
  return status;
}

/* ---------------------------------------------------------- Pressure_Gradient */
/* momentum.c:223-452 with levelset off, no -dpdz forcing and no -flux inlet term
 * (their defaults, main.c:73 and 95): the contravariant pressure-gradient flux
 * dP on the i, j, k faces of every interior cell, with the one-sided tangential
 * differences next to nvert == 1 cells and non-periodic walls.  p gets its
 * periodic ghost copies first, as the reference does to lP. */
void or_pressure_gradient(const or_grid *g, const double *icsi, const double *ieta, const double *izet,
                          const double *jcsi, const double *jeta, const double *jzet, const double *kcsi,
                          const double *keta, const double *kzet, const double *iaj, const double *jaj,
                          const double *kaj, const double *nvert, double *p, double *dp) {
  const int mx = g->mx, my = g->my, mz = g->mz;
  const int pi = g->per[0], pj = g->per[1], pk = g->per[2];
  periodic_copy_scalar(g, p);
  memset(dp, 0, 3 * (size_t)mx * my * mz * sizeof(double));
#define P(k, j, i) p[IX(g, i, j, k)]
#define N1(k, j, i) ((int)(nvert[IX(g, i, j, k)] + 0.5) == 1)
#define DOT(a, b) (a[3 * c] * b[3 * c] + a[3 * c + 1] * b[3 * c + 1] + a[3 * c + 2] * b[3 * c + 2])
  for (int k = 1; k < mz - 1; k++)
    for (int j = 1; j < my - 1; j++)
      for (int i = 1; i < mx - 1; i++) {
        const size_t c = IX(g, i, j, k);
        const double g11 = DOT(icsi, icsi), g12 = DOT(ieta, icsi), g13 = DOT(izet, icsi);
        const double g21 = DOT(jcsi, jeta), g22 = DOT(jeta, jeta), g23 = DOT(jzet, jeta);
        const double g31 = DOT(kcsi, kzet), g32 = DOT(keta, kzet), g33 = DOT(kzet, kzet);
        double dpdc, dpde, dpdz;
        /* i face */
        if (i == mx - 2 && pi) dpdc = P(k, j, 1) - P(k, j, i);
        else dpdc = P(k, j, i + 1) - P(k, j, i);
        if (N1(k, j + 1, i) || N1(k, j + 1, i + 1) || (j == my - 2 && !pj))
          dpde = (P(k, j, i) - P(k, j - 1, i) + P(k, j, i + 1) - P(k, j - 1, i + 1)) * 0.5;
        else if (N1(k, j - 1, i) || N1(k, j - 1, i + 1) || (j == 1 && !pj))
          dpde = (P(k, j + 1, i) - P(k, j, i) + P(k, j + 1, i + 1) - P(k, j, i + 1)) * 0.5;
        else dpde = (P(k, j + 1, i) - P(k, j - 1, i) + P(k, j + 1, i + 1) - P(k, j - 1, i + 1)) * 0.25;
        if (N1(k + 1, j, i) || N1(k + 1, j, i + 1) || (k == mz - 2 && !pk))
          dpdz = (P(k, j, i) - P(k - 1, j, i) + P(k, j, i + 1) - P(k - 1, j, i + 1)) * 0.5;
        else if (N1(k - 1, j, i) || N1(k - 1, j, i + 1) || (k == 1 && !pk))
          dpdz = (P(k + 1, j, i) - P(k, j, i) + P(k + 1, j, i + 1) - P(k, j, i + 1)) * 0.5;
        else dpdz = (P(k + 1, j, i) - P(k - 1, j, i) + P(k + 1, j, i + 1) - P(k - 1, j, i + 1)) * 0.25;
        double dx = (dpdc * g11 + dpde * g12 + dpdz * g13) * iaj[c];
        /* j face */
        if (N1(k, j, i + 1) || N1(k, j + 1, i + 1) || (i == mx - 2 && !pi))
          dpdc = (P(k, j, i) - P(k, j, i - 1) + P(k, j + 1, i) - P(k, j + 1, i - 1)) * 0.5;
        else if (N1(k, j, i - 1) || N1(k, j + 1, i - 1) || (i == 1 && !pi))
          dpdc = (P(k, j, i + 1) - P(k, j, i) + P(k, j + 1, i + 1) - P(k, j + 1, i)) * 0.5;
        else dpdc = (P(k, j, i + 1) - P(k, j, i - 1) + P(k, j + 1, i + 1) - P(k, j + 1, i - 1)) * 0.25;
        if (j == my - 2 && pj) dpde = P(k, 1, i) - P(k, j, i);
        else dpde = P(k, j + 1, i) - P(k, j, i);
        if (N1(k + 1, j, i) || N1(k + 1, j + 1, i) || (k == mz - 2 && !pk))
          dpdz = (P(k, j, i) - P(k - 1, j, i) + P(k, j + 1, i) - P(k - 1, j + 1, i)) * 0.5;
        else if (N1(k - 1, j, i) || N1(k - 1, j + 1, i) || (k == 1 && !pk))
          dpdz = (P(k + 1, j, i) - P(k, j, i) + P(k + 1, j + 1, i) - P(k, j + 1, i)) * 0.5;
        else dpdz = (P(k + 1, j, i) - P(k - 1, j, i) + P(k + 1, j + 1, i) - P(k - 1, j + 1, i)) * 0.25;
        double dy = (dpdc * g21 + dpde * g22 + dpdz * g23) * jaj[c];
        /* k face */
        if (N1(k, j, i + 1) || N1(k + 1, j, i + 1) || (i == mx - 2 && !pi))
          dpdc = (P(k, j, i) - P(k, j, i - 1) + P(k + 1, j, i) - P(k + 1, j, i - 1)) * 0.5;
        else if (N1(k, j, i - 1) || N1(k + 1, j, i - 1) || (i == 1 && !pi))
          dpdc = (P(k, j, i + 1) - P(k, j, i) + P(k + 1, j, i + 1) - P(k + 1, j, i)) * 0.5;
        else dpdc = (P(k, j, i + 1) - P(k, j, i - 1) + P(k + 1, j, i + 1) - P(k + 1, j, i - 1)) * 0.25;
        if (N1(k, j + 1, i) || N1(k + 1, j + 1, i) || (j == my - 2 && !pj))
          dpde = (P(k, j, i) - P(k, j - 1, i) + P(k + 1, j, i) - P(k + 1, j - 1, i)) * 0.5;
        else if (N1(k, j - 1, i) || N1(k + 1, j - 1, i) || (j == 1 && !pj))
          dpde = (P(k, j + 1, i) - P(k, j, i) + P(k + 1, j + 1, i) - P(k + 1, j, i)) * 0.5;
        else dpde = (P(k, j + 1, i) - P(k, j - 1, i) + P(k + 1, j + 1, i) - P(k + 1, j - 1, i)) * 0.25;
        if (k == mz - 2 && pk) dpdz = P(1, j, i) - P(k, j, i);
        else dpdz = P(k + 1, j, i) - P(k, j, i);
        double dz = (dpdc * g31 + dpde * g32 + dpdz * g33) * kaj[c];
        /* faces without a fluid pair or on a non-periodic wall carry none (momentum.c:426-436) */
        if (nvert[c] + nvert[IX(g, i + 1, j, k)] > 0.1 || (!pi && i == mx - 2)) dx = 0;
        if (nvert[c] + nvert[IX(g, i, j + 1, k)] > 0.1 || (!pj && j == my - 2)) dy = 0;
        if (nvert[c] + nvert[IX(g, i, j, k + 1)] > 0.1 || (!pk && k == mz - 2)) dz = 0;
        dp[3 * c] = dx;
        dp[3 * c + 1] = dy;
        dp[3 * c + 2] = dz;
      }
#undef P
#undef N1
#undef DOT
}

// The reference's 19-point curvilinear operator (PoissonLHSNew, poisson.c:976-1691)
// in matrix-free form.
//
// Per cell the engine keeps the nine face coefficients of its three "upper"
// faces (g11,g12,g13 on i+1/2; g21,g22,g23 on j+1/2; g31,g32,g33 on k+1/2, the
// values poisson.c:1132-1140 computes) and one 32-bit code word that records
// every branch PoissonLHSNew takes for that row:
//   bits 0..5   face active, faces in the order i+, i-, j+, j-, k+, k-
//               (poisson.c:1171, 1235, 1293, 1361, 1420, 1491)
//   bits 6..29  two 2-bit cross-difference cases per face: 0 zero, 1 central
//               (0.25), 2 one-sided toward the lower neighbour (0.5), 3
//               one-sided toward the upper neighbour (0.5); cross directions in
//               increasing axis order (poisson.c:1182-1230 and rotations)
//   bit 30      the cell has a row (fluid, nvert <= 0.1, poisson.c:1158-1169)
//   bit 31      Divergence counts the cell (7-point nvert sum <= 0.1, main.c:1284)
// build_row() rebuilds the 19 row coefficients from these in the reference's
// accumulation order, so they are bit-identical to the assembled values.
#pragma once

#include "pp_common.hpp"

namespace pp {

enum Slot { CP = 0, EP, WP, NP, SP, TP, BP, NE, SE, NW, SW, TN, BN, TS, BS, TE, BE, TW, BW };

constexpr uint32_t ROW_BIT = 1u << 30;
constexpr uint32_t DIVOK_BIT = 1u << 31;

enum XCase { X_ZERO = 0, X_CENTRAL = 1, X_LOWER = 2, X_UPPER = 3 };

// offset -> slot (poisson.c:956-974); -1 for the 8 corners, which the stencil lacks
HD constexpr int slot_of(int di, int dj, int dk) {
  return (dk == 0)
             ? ((dj == 0) ? (di == 0 ? CP : (di > 0 ? EP : WP))
                          : (dj > 0 ? (di == 0 ? NP : (di > 0 ? NE : NW)) : (di == 0 ? SP : (di > 0 ? SE : SW))))
             : (dk > 0 ? ((dj == 0) ? (di == 0 ? TP : (di > 0 ? TE : TW)) : (dj > 0 ? (di == 0 ? TN : -1) : (di == 0 ? TS : -1)))
                       : ((dj == 0) ? (di == 0 ? BP : (di > 0 ? BE : BW)) : (dj > 0 ? (di == 0 ? BN : -1) : (di == 0 ? BS : -1))));
}

struct Off3 {
  int d[3];
};
HD constexpr Off3 slot_off(int s) {
  return s == CP   ? Off3{{0, 0, 0}}
         : s == EP ? Off3{{1, 0, 0}}
         : s == WP ? Off3{{-1, 0, 0}}
         : s == NP ? Off3{{0, 1, 0}}
         : s == SP ? Off3{{0, -1, 0}}
         : s == TP ? Off3{{0, 0, 1}}
         : s == BP ? Off3{{0, 0, -1}}
         : s == NE ? Off3{{1, 1, 0}}
         : s == SE ? Off3{{1, -1, 0}}
         : s == NW ? Off3{{-1, 1, 0}}
         : s == SW ? Off3{{-1, -1, 0}}
         : s == TN ? Off3{{0, 1, 1}}
         : s == BN ? Off3{{0, 1, -1}}
         : s == TS ? Off3{{0, -1, 1}}
         : s == BS ? Off3{{0, -1, -1}}
         : s == TE ? Off3{{1, 0, 1}}
         : s == BE ? Off3{{1, 0, -1}}
         : s == TW ? Off3{{-1, 0, 1}}
                   : Off3{{-1, 0, -1}};
}

// Cross-difference case of the face between DMDA points p and q = p +/- e_d in
// direction e (the branch structure of poisson.c:1182-1230).
template <typename NV>
HD inline int cross_case(const NV* nv, size_t p, size_t q, size_t se, int pos_e, int m_e, int per_e) {
  const double sum_hi = (double)nv[p + se] + (double)nv[q + se];
  const double sum_lo = (double)nv[p - se] + (double)nv[q - se];
  if ((pos_e == m_e - 2 && !per_e) || sum_hi > TH) {
    if (sum_lo < TH && pos_e != 1) return X_LOWER;
    return X_ZERO;
  } else if ((pos_e == 1 && !per_e) || sum_lo > TH) {
    if (sum_hi < TH) return X_UPPER;
    return X_ZERO;
  }
  return X_CENTRAL;
}

// Code word of DMDA interior point (I,J,K) from a DMDA nvert array.
template <typename NV>
HD inline uint32_t make_code(const NV* nv, int mx, int my, int mz, const int per[3], int I, int J, int K) {
  const size_t st[3] = {1, (size_t)mx, (size_t)mx * my};
  const int m[3] = {mx, my, mz};
  const int pos[3] = {I, J, K};
  const size_t c = (size_t)K * st[2] + (size_t)J * st[1] + (size_t)I;
  if ((double)nv[c] > 0.1) return 0u;
  uint32_t code = ROW_BIT;
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      const int f = 2 * d + s;
      const size_t q = s == 0 ? c + st[d] : c - st[d];
      const bool edge = s == 0 ? (pos[d] != m[d] - 2) : (pos[d] != 1);
      if (!((double)nv[q] < TH && (edge || per[d]))) continue;
      code |= 1u << f;
      int xi = 0;
      for (int e = 0; e < 3; e++) {
        if (e == d) continue;
        const int cc = cross_case(nv, c, q, st[e], pos[e], m[e], per[e]);
        code |= (uint32_t)cc << (6 + 4 * f + 2 * xi);
        xi++;
      }
    }
  // main.c:1284 operand order
  const double s7 = (double)nv[c] + (double)nv[c + st[2]] + (double)nv[c - st[2]] + (double)nv[c + st[1]] +
                    (double)nv[c - st[1]] + (double)nv[c + 1] + (double)nv[c - 1];
  if (!(s7 > 0.1)) code |= DIVOK_BIT;
  return code;
}

// ---------------------------------------------------------------------------
// Row rebuild.  Faces in the reference's order i+, i-, j+, j-, k+, k-; inside a
// face the terms g_d1, g_d2, g_d3 in that order (normal term where e == d).
template <int D, int SGN, int E, typename T>
HDI void face_term(T (&vol)[19], T gv, uint32_t code) {
  constexpr int F = 2 * D + (SGN > 0 ? 0 : 1);
  constexpr int qd[3] = {D == 0 ? SGN : 0, D == 1 ? SGN : 0, D == 2 ? SGN : 0};
  if constexpr (E == D) {
    constexpr int sQ = slot_of(qd[0], qd[1], qd[2]);
    vol[CP] -= gv;
    vol[sQ] += gv;
  } else {
    constexpr int XI = (E < D) ? E : E - 1;
    constexpr int ee[3] = {E == 0 ? 1 : 0, E == 1 ? 1 : 0, E == 2 ? 1 : 0};
    const int cc = (int)((code >> (6 + 4 * F + 2 * XI)) & 3u);
    constexpr int sC_hi = slot_of(ee[0], ee[1], ee[2]);
    constexpr int sQ_hi = slot_of(qd[0] + ee[0], qd[1] + ee[1], qd[2] + ee[2]);
    constexpr int sC_lo = slot_of(-ee[0], -ee[1], -ee[2]);
    constexpr int sQ_lo = slot_of(qd[0] - ee[0], qd[1] - ee[1], qd[2] - ee[2]);
    constexpr int sC = CP;
    constexpr int sQ = slot_of(qd[0], qd[1], qd[2]);
    // Branch-free form of the three cases (a store per case into different slots
    // would be merged into one dynamically indexed store and push vol to scratch):
    //   central: +0.25 g on the upper pair, -0.25 g on the lower pair
    //   lower:   +0.5 g on (c, q),          -0.5 g on the lower pair
    //   upper:   +0.5 g on the upper pair,  -0.5 g on (c, q)
    // Each non-zero addend equals the reference's w*g exactly and lands in the same
    // order; the remaining addends are exact zeros.
    const T whi = cc == X_CENTRAL ? T(0.25) : (cc == X_UPPER ? T(0.5) : T(0));
    const T wlo = cc == X_CENTRAL ? T(0.25) : (cc == X_LOWER ? T(0.5) : T(0));
    const T wmd = cc == X_LOWER ? T(0.5) : (cc == X_UPPER ? T(-0.5) : T(0));
    const T th = whi * gv, tl = wlo * gv, tm = wmd * gv;
    if constexpr (SGN > 0) {
      vol[sC_hi] += th; vol[sQ_hi] += th; vol[sC_lo] -= tl; vol[sQ_lo] -= tl; vol[sC] += tm; vol[sQ] += tm;
    } else {
      vol[sC_hi] -= th; vol[sQ_hi] -= th; vol[sC_lo] += tl; vol[sQ_lo] += tl; vol[sC] -= tm; vol[sQ] -= tm;
    }
  }
}

template <int D, int SGN, typename T, typename TG>
HDI void face_terms(T (&vol)[19], const TG* __restrict__ g, size_t n, size_t owner, uint32_t code) {
  constexpr int F = 2 * D + (SGN > 0 ? 0 : 1);
  if (!((code >> F) & 1u)) return;
  const T g0 = (T)g[(size_t)(3 * D + 0) * n + owner];
  const T g1 = (T)g[(size_t)(3 * D + 1) * n + owner];
  const T g2 = (T)g[(size_t)(3 * D + 2) * n + owner];
  face_term<D, SGN, 0>(vol, g0, code);
  face_term<D, SGN, 1>(vol, g1, code);
  face_term<D, SGN, 2>(vol, g2, code);
}

// 19 row coefficients of cell c (interior index); cw/cs/cb are the interior
// indices of its i-1, j-1, k-1 neighbours (wrapped), which own the lower faces.
template <typename T, typename TG>
HDI void build_row(const TG* __restrict__ g, size_t n, uint32_t code, size_t c, size_t cw, size_t cs, size_t cb,
                         T (&vol)[19]) {
#pragma unroll
  for (int s = 0; s < 19; s++) vol[s] = T(0);
  face_terms<0, +1>(vol, g, n, c, code);
  face_terms<0, -1>(vol, g, n, cw, code);
  face_terms<1, +1>(vol, g, n, c, code);
  face_terms<1, -1>(vol, g, n, cs, code);
  face_terms<2, +1>(vol, g, n, c, code);
  face_terms<2, -1>(vol, g, n, cb, code);
}

// ---------------------------------------------------------------------------
// Flux-form evaluation of (A x)_c and the diagonal a_cc, used by the solver's
// hot kernels.  Same operator as build_row (same masks, same face
// coefficients), summed face by face instead of coefficient by coefficient, so
// it differs from the assembled row only by rounding.  xv holds x at the 19
// stencil points in slot order, loaded up front so all loads are in flight at
// once.
//
// KC < 0 decodes each cross-difference case from the code word; KC == X_CENTRAL
// is the row of an interior cell (every face active, every cross difference
// central), with the weights known at compile time.
template <int D, int SGN, int E, int KC, typename T>
HDI void flux_term(T& nrm, T& crs, T& diag, T gv, uint32_t code, const T (&xv)[19]) {
  constexpr int F = 2 * D + (SGN > 0 ? 0 : 1);
  constexpr int qd[3] = {D == 0 ? SGN : 0, D == 1 ? SGN : 0, D == 2 ? SGN : 0};
  constexpr int sQ = slot_of(qd[0], qd[1], qd[2]);
  if constexpr (E == D) {
    // vol[CP] -= g, vol[q] += g on both sides of the cell
    nrm += gv * (xv[sQ] - xv[CP]);
    diag -= gv;
  } else {
    constexpr int XI = (E < D) ? E : E - 1;
    constexpr int ee[3] = {E == 0 ? 1 : 0, E == 1 ? 1 : 0, E == 2 ? 1 : 0};
    constexpr int sC_hi = slot_of(ee[0], ee[1], ee[2]);
    constexpr int sQ_hi = slot_of(qd[0] + ee[0], qd[1] + ee[1], qd[2] + ee[2]);
    constexpr int sC_lo = slot_of(-ee[0], -ee[1], -ee[2]);
    constexpr int sQ_lo = slot_of(qd[0] - ee[0], qd[1] - ee[1], qd[2] - ee[2]);
    if constexpr (KC == X_CENTRAL) {
      // the central case's addends: +0.25 g on the upper pair, -0.25 g on the
      // lower pair, nothing on (c, q) or the diagonal.  The 0.25 is applied once
      // per face (flux_face); scaling by a power of two commutes with every rounding,
      // so the result is the same
      crs += gv * ((xv[sC_hi] + xv[sQ_hi]) - (xv[sC_lo] + xv[sQ_lo]));
    } else {
      const int cc = (int)((code >> (6 + 4 * F + 2 * XI)) & 3u);
      const T whi = cc == X_CENTRAL ? T(0.25) : (cc == X_UPPER ? T(0.5) : T(0));
      const T wlo = cc == X_CENTRAL ? T(0.25) : (cc == X_LOWER ? T(0.5) : T(0));
      const T wmd = cc == X_LOWER ? T(0.5) : (cc == X_UPPER ? T(-0.5) : T(0));
      // the cross part enters the row with the face's side sign (see face_term)
      crs += gv * (whi * (xv[sC_hi] + xv[sQ_hi]) - wlo * (xv[sC_lo] + xv[sQ_lo]) + wmd * (xv[CP] + xv[sQ]));
      diag += (SGN > 0 ? wmd : -wmd) * gv;
    }
  }
}

// a[idx] as a global-memory load through a 32-bit byte offset from a base every
// lane shares: the load takes its base from scalar registers and one offset
// register serves every array read at the same index (interior arrays hold
// fewer than 2^29 entries).  Pointers read from work-list structs would
// otherwise be loaded through the flat address space.
template <typename T>
__device__ __forceinline__ T ld_at(const T* __restrict__ a, uint32_t idx) {
  typedef const __attribute__((address_space(1))) char gchar;
  typedef const __attribute__((address_space(1))) T gT;
  return *(gT*)((gchar*)a + idx * (uint32_t)sizeof(T));
}

// The 18 face coefficients one row reads, faces in the order i+, i-, j+, j-, k+,
// k- (owners c, cw, c, cs, c, cb), three per face.  They are loaded in one group,
// before any test of the row's form, so that every load of the row -- stencil
// values and coefficients -- is in flight at once: the kernels that use them are
// bound by memory latency, and a second round of loads after a branch would
// double it.  Every owner index is a valid cell, whether or not its face is
// active.
template <typename TG>
struct RowCoef {
  TG v[6][3];
};

template <typename TG>
__device__ __forceinline__ void load_coef(const TG* __restrict__ g, size_t n, uint32_t c, uint32_t cw, uint32_t cs,
                                          uint32_t cb, RowCoef<TG>& rc) {
  const uint32_t own[6] = {c, cw, c, cs, c, cb};
#pragma unroll
  for (int f = 0; f < 6; f++)
#pragma unroll
    for (int e = 0; e < 3; e++) rc.v[f][e] = ld_at(g + (size_t)(3 * (f / 2) + e) * n, own[f]);
}

template <int D, int SGN, int KC, typename T, typename TG>
__device__ __forceinline__ void flux_face(T& ax, T& diag, const TG (&gf)[3], uint32_t code, const T (&xv)[19]) {
  constexpr int F = 2 * D + (SGN > 0 ? 0 : 1);
  const T g0 = (T)gf[0];
  const T g1 = (T)gf[1];
  const T g2 = (T)gf[2];
  T nrm = T(0), crs = T(0), dg = T(0);
  flux_term<D, SGN, 0, KC>(nrm, crs, dg, g0, code, xv);
  flux_term<D, SGN, 1, KC>(nrm, crs, dg, g1, code, xv);
  flux_term<D, SGN, 2, KC>(nrm, crs, dg, g2, code, xv);
  if constexpr (KC == X_CENTRAL) crs *= T(0.25);
  if (KC == X_CENTRAL || ((code >> F) & 1u)) {
    ax += nrm + (SGN > 0 ? crs : -crs);
    diag += dg;
  }
}

// (A x)_c and a_cc from loaded coefficients, the cross-difference cases fixed at
// compile time (KC == X_CENTRAL: an interior row) or decoded from code (KC < 0)
template <int KC, typename T, typename TG>
__device__ __forceinline__ void flux_row_v(const RowCoef<TG>& rc, uint32_t code, const T (&xv)[19], T& ax, T& diag) {
  ax = T(0);
  diag = T(0);
  flux_face<0, +1, KC>(ax, diag, rc.v[0], code, xv);
  flux_face<0, -1, KC>(ax, diag, rc.v[1], code, xv);
  flux_face<1, +1, KC>(ax, diag, rc.v[2], code, xv);
  flux_face<1, -1, KC>(ax, diag, rc.v[3], code, xv);
  flux_face<2, +1, KC>(ax, diag, rc.v[4], code, xv);
  flux_face<2, -1, KC>(ax, diag, rc.v[5], code, xv);
}

// code word of an interior cell: every face active, every cross difference
// central, a row (the Divergence bit aside)
constexpr uint32_t CODE_INTERIOR = 0x3Fu | (0x555555u << 6) | ROW_BIT;

// (A x)_c and a_cc.  A wave whose cells are all interior cells takes the
// compile-time-weight form; any other wave decodes every case.
template <typename T, typename TG>
__device__ __forceinline__ void flux_row(const TG* __restrict__ g, size_t n, uint32_t code, size_t c, size_t cw,
                                         size_t cs, size_t cb, const T (&xv)[19], T& ax, T& diag) {
  RowCoef<TG> rc;
  load_coef(g, n, (uint32_t)c, (uint32_t)cw, (uint32_t)cs, (uint32_t)cb, rc);
  if (__all((code & ~DIVOK_BIT) == CODE_INTERIOR))
    flux_row_v<X_CENTRAL>(rc, code, xv, ax, diag);
  else
    flux_row_v<-1>(rc, code, xv, ax, diag);
}

// Interior indices of the 19 stencil neighbours with periodic wrap.
struct Nbr {
  uint32_t s[19];  // systems are far below 2^32 cells
};
HDI void neighbours(const Dims& d, int i, int j, int k, Nbr& nb) {
  const uint32_t sy = (uint32_t)d.nx, sz = (uint32_t)d.nx * (uint32_t)d.ny;
  const uint32_t I = (uint32_t)i, IM = (uint32_t)wrap_m(i, d.nx, d.per[0]), IP = (uint32_t)wrap_p(i, d.nx, d.per[0]);
  const uint32_t J = (uint32_t)j * sy, JM = (uint32_t)wrap_m(j, d.ny, d.per[1]) * sy,
                 JP = (uint32_t)wrap_p(j, d.ny, d.per[1]) * sy;
  const uint32_t K = (uint32_t)k * sz, KM = (uint32_t)wrap_m(k, d.nz, d.per[2]) * sz,
                 KP = (uint32_t)wrap_p(k, d.nz, d.per[2]) * sz;
  nb.s[CP] = I + J + K;
  nb.s[EP] = IP + J + K;
  nb.s[WP] = IM + J + K;
  nb.s[NP] = I + JP + K;
  nb.s[SP] = I + JM + K;
  nb.s[TP] = I + J + KP;
  nb.s[BP] = I + J + KM;
  nb.s[NE] = IP + JP + K;
  nb.s[SE] = IP + JM + K;
  nb.s[NW] = IM + JP + K;
  nb.s[SW] = IM + JM + K;
  nb.s[TN] = I + JP + KP;
  nb.s[BN] = I + JP + KM;
  nb.s[TS] = I + JM + KP;
  nb.s[BS] = I + JM + KM;
  nb.s[TE] = IP + J + KP;
  nb.s[BE] = IP + J + KM;
  nb.s[TW] = IM + J + KP;
  nb.s[BW] = IM + J + KM;
}

template <typename T, typename TX>
__device__ __forceinline__ void load_x(const Nbr& nb, const TX* __restrict__ x, T (&xv)[19]) {
#pragma unroll
  for (int s = 0; s < 19; s++) xv[s] = (T)ld_at(x, nb.s[s]);
}

}  // namespace pp

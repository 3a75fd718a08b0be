// Host side of the pressure-projection engine: device-resident case data, the
// geometric multigrid hierarchy, the FP64 FGMRES outer loop and the C ABI of
// include/vfs_pp.h.
//
// Solve structure (one call of vfs_pp_project_device):
//   VolumeFlux on a compact list of IB faces -> RHS b (FP64) -> FGMRES(m) in FP64
//   with classical Gram-Schmidt on a lazily normalised basis (or CGS2),
//   right-preconditioned by one V-cycle of a
//   geometric multigrid in FP32 (mixed) or FP64 -> Remove_Nullspace -> scatter
//   phi to the DMDA array and P += phi -> Projection -> Contra2Cart ->
//   Divergence.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "pp_engine.hpp"

namespace pp {

// The coarsest level on the host, for its dense inverse.
struct CoarseRows {
  Dims d{};
  std::vector<double> g;       // 9*n face coefficients
  std::vector<uint32_t> code;  // n
};

static void row_of(const CoarseRows& L, int i, int j, int k, double (&vol)[19], Nbr& nb) {
  neighbours(L.d, i, j, k, nb);
  const size_t c = L.d.idx(i, j, k);
  build_row(L.g.data(), L.d.n(), L.code[c], c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
}

// The bordered coarsest operator [A E; E^T 0] (N = n + ncomp, row-major), where E
// has one indicator column per connected fluid component (the constants of each
// component span the null space of the Neumann problem); rows without a row of A
// hold 1 on the diagonal.  Its inverse's top-left n x n block is the dense solve.
static std::vector<double> coarse_bordered(const CoarseRows& L, int* ncomp_out) {
  const int n = (int)L.d.n();
  std::vector<int> comp(n, -1);
  std::vector<std::vector<std::pair<int, double>>> rows(n);
  for (int k = 0; k < L.d.nz; k++)
    for (int j = 0; j < L.d.ny; j++)
      for (int i = 0; i < L.d.nx; i++) {
        const int c = (int)L.d.idx(i, j, k);
        if (!(L.code[c] & ROW_BIT)) continue;
        double vol[19];
        Nbr nb;
        row_of(L, i, j, k, vol, nb);
        for (int s = 0; s < 19; s++)
          if (vol[s] != 0.) rows[c].push_back({(int)nb.s[s], vol[s]});
      }
  int ncomp = 0;
  for (int c = 0; c < n; c++) {
    if (!(L.code[c] & ROW_BIT) || comp[c] >= 0) continue;
    std::vector<int> st{c};
    comp[c] = ncomp;
    while (!st.empty()) {
      int u = st.back();
      st.pop_back();
      for (auto& e : rows[u])
        if (e.first != u && (L.code[e.first] & ROW_BIT) && comp[e.first] < 0) {
          comp[e.first] = ncomp;
          st.push_back(e.first);
        }
    }
    ncomp++;
  }
  const int N = n + ncomp;
  std::vector<double> B((size_t)N * N, 0.);
  for (int c = 0; c < n; c++) {
    if (!(L.code[c] & ROW_BIT)) { B[(size_t)c * N + c] = 1.; continue; }
    for (auto& e : rows[c]) B[(size_t)c * N + e.first] += e.second;
    B[(size_t)c * N + n + comp[c]] = 1.;
    B[(size_t)(n + comp[c]) * N + c] = 1.;
  }
  *ncomp_out = ncomp;
  return B;
}

// The same elimination as launch_gj_step (partial pivoting on the first largest
// |B[r][col]|, the pivot row scaled by inv = 1 / pivot, every other row with f != 0
// reduced by f times the scaled pivot row), on the host: it returns the top-left
// n x n block of B^-1, column-major.  Plain products and differences (contraction
// into fused multiply-adds is off), so it agrees bit for bit with the device kernel.
static std::vector<double> coarse_inverse_host(std::vector<double> B, int N, int n) {
#pragma clang fp contract(off)
  std::vector<double> I((size_t)N * N, 0.);
  for (int r = 0; r < N; r++) I[(size_t)r * N + r] = 1.;
  for (int col = 0; col < N; col++) {
    int piv = col;
    double best = std::fabs(B[(size_t)col * N + col]);
    for (int r = col + 1; r < N; r++) {
      const double v = std::fabs(B[(size_t)r * N + col]);
      if (v > best) { best = v; piv = r; }
    }
    if (!(best > 0.)) throw StatusError(VFS_PP_EINVAL, "coarsest operator is singular beyond its null space");
    if (piv != col)
      for (int t = 0; t < N; t++) {
        std::swap(B[(size_t)piv * N + t], B[(size_t)col * N + t]);
        std::swap(I[(size_t)piv * N + t], I[(size_t)col * N + t]);
      }
    const double inv = 1. / B[(size_t)col * N + col];
    for (int t = col + 1; t < N; t++) B[(size_t)col * N + t] *= inv;
    for (int t = 0; t < N; t++) I[(size_t)col * N + t] *= inv;
    for (int r = 0; r < N; r++) {
      if (r == col) continue;
      const double f = B[(size_t)r * N + col];
      if (f == 0.) continue;
      for (int t = col + 1; t < N; t++) B[(size_t)r * N + t] -= f * B[(size_t)col * N + t];
      for (int t = 0; t < N; t++) I[(size_t)r * N + t] -= f * I[(size_t)col * N + t];
    }
  }
  std::vector<double> out((size_t)n * n);
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) out[(size_t)c * n + r] = I[(size_t)r * N + c];
  return out;
}

// Bordered systems up to this many rows are inverted on the host: below it the
// elimination costs less there than one device launch per column.
constexpr int COARSE_HOST_MAX = 96;

// cinv (n x n, column-major, precision T) = the top-left block of B^-1, by
// Gauss-Jordan with partial pivoting on the device (launch_gj_step; on the host
// for small systems).  The working matrices come from the set-up scratch `tmp`,
// which a moving body's rebuilds keep from one call to the next.
template <typename T>
static void coarse_inverse_device(hipStream_t s, Arena& tmp, const std::vector<double>& B, int n, T* cinv) {
  const int N = (int)std::lround(std::sqrt((double)B.size()));
  if (N <= COARSE_HOST_MAX) {
    const std::vector<double> ci = coarse_inverse_host(B, N, n);
    std::vector<T> ct(ci.size());
    for (size_t i = 0; i < ci.size(); i++) ct[i] = T(ci[i]);
    h2d(s, cinv, ct.data(), ct.size());
    HIPCHECK(hipStreamSynchronize(s));
    return;
  }
  const size_t nn = (size_t)N * N;
  double* m = tmp.alloc<double>(4 * nn + 1);
  double *b[2] = {m, m + nn}, *id[2] = {m + 2 * nn, m + 3 * nn};
  int* singular = reinterpret_cast<int*>(m + 4 * nn);
  h2d(s, b[0], B.data(), nn);
  launch_gj_identity(s, N, id[0]);
  HIPCHECK(hipMemsetAsync(singular, 0, sizeof(int), s));
  for (int col = 0; col < N; col++)
    launch_gj_step(s, N, col, b[col & 1], id[col & 1], b[(col + 1) & 1], id[(col + 1) & 1], singular);
  launch_gj_out<T>(s, N, n, id[N & 1], cinv);
  int bad = 0;
  d2h(s, &bad, singular, 1);
  HIPCHECK(hipStreamSynchronize(s));
  if (bad) throw StatusError(VFS_PP_EINVAL, "coarsest operator is singular beyond its null space");
}

// --------------------------------------------------------------------- engine
double now_ms() {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Power iteration for lambda_max(M^-1 A) on one level: 12 steps from a random
// start, the norm after step `it` going to host_norms2[it] (mapped host memory);
// the caller reads them after synchronising the stream.
template <typename T>
static void queue_power_iteration(vfs_pp_handle* h, MG<T>& mg, typename MG<T>::Lev& L, double* host_norms2) {
  const size_t n = L.d.n();
  launch_random_rows<T>(h->s, n, L.code, 1234u, L.xa);
  if (std::is_same<T, float>::value) launch_fillf(h->s, n, (float*)L.f, 0.f);
  else launch_fill(h->s, n, (double*)L.f, 0.);
  int nb = 0;
  launch_normsq<T>(h->s, n, L.xa, h->partial, &nb);
  launch_finalize(h->s, h->partial, 1, nb, h->red, 0);
  launch_scale_by<T>(h->s, n, L.xa, h->red);
  for (int it = 0; it < 12; it++) {
    launch_smooth<T>(h->s, L.view(), L.lfac, SM_POWER, L.f, L.dinv, L.xa, L.xb, L.xb, L.dv, mg.scr, T(0), T(1),
                     T(0));
    launch_normsq<T>(h->s, n, L.dv, h->partial, &nb);
    launch_finalize(h->s, h->partial, 1, nb, h->red, 0, host_norms2 + it);
    launch_scale_by<T>(h->s, n, L.xa, h->red, L.dv);  // x = dv / |dv|
  }
}

// the estimate from the norms of steps 6..11
static double power_estimate(const double* norms2) {
  double best = 0.;
  for (int it = 6; it < 12; it++) best = std::max(best, std::sqrt(norms2[it]));
  return best > 0. ? best : 2.;
}

// y-line smoothing on a level: always with smoother 2; with smoother 0 when the
// wall-normal coupling dominates somewhere (|a_N| + |a_S| > 4 (|a_E| + |a_W| + |a_T|
// + |a_B|) on some row); never when j is periodic
static int line_rule(const Dims& d, const vfs_pp_opts& o, double worst) {
  if (d.per[1]) return 0;
  if (o.smoother == 2) return 1;
  if (o.smoother != 0) return 0;
  return worst > 4.0 ? 1 : 0;
}

// The multigrid hierarchy, built on the device from the finest level's arrays
// (h->g64, h->code, h->nvert).  Per level, the row statistics choose the smoother
// and whether j is kept; then one coarsening step, until the level is small enough
// for the dense solve.  Coarse FP64 coefficients stay (in h->geo) only for the FP64
// hierarchy; nvert arrays and the FP32 hierarchy's FP64 coefficients go to tmp.
static void build_hierarchy(vfs_pp_handle* h, Arena& tmp) {
  hipStream_t s = h->s;
  auto& H = h->setup;
  H.clear();
  need_partials(h, 1, level_stats_partials());
  SetupLevel L0;
  L0.d = h->d;
  L0.g = h->g64;
  L0.nv = h->nvert;
  L0.code = h->code;
  L0.cls = h->geo.alloc<uint8_t>(h->d.n());
  launch_cls_from_code(s, h->d.n(), h->code, L0.cls);
  H.push_back(L0);
  Arena& gstore = h->opts.precision == 1 ? tmp : h->geo;
  while (true) {
    SetupLevel& F = H.back();
    launch_level_stats(s, F.d, F.g, F.code, h->partial, h->hpin_dev + RED_SETUP);
    HIPCHECK(hipStreamSynchronize(s));
    const double worst = h->hpin[RED_SETUP], rows = h->hpin[RED_SETUP + 1];
    const double weak = rows > 0. ? h->hpin[RED_SETUP + 2] / rows : 0.;
    F.gmax = h->hpin[RED_SETUP + 3];
    F.line = line_rule(F.d, h->opts, worst);
    const bool can = F.d.nx >= 3 || F.d.ny >= 3 || F.d.nz >= 3;
    if ((long long)F.d.n() <= h->opts.coarse_max || !can || H.size() >= 16) break;
    // 2:1 on every axis with at least 3 cells (odd counts leave a one-child cell at
    // the high end).  j is kept on y-line levels where more than 1% of the rows
    // couple weakly along j (below 1/10 of their i-k coupling): a y-line smoother
    // barely reduces errors smooth in i-k but oscillating in j there, so the coarse
    // level must still represent them (semi-coarsening in i and k)
    int cm = (F.d.nx >= 3 ? 1 : 0) | (F.d.ny >= 3 ? 2 : 0) | (F.d.nz >= 3 ? 4 : 0);
    if (F.line && weak > 0.01 && (cm & 5)) cm &= ~2;
    F.cmask = cm;
    SetupLevel C;
    C.d = F.d;
    if (cm & 1) C.d.nx = (F.d.nx + 1) / 2;
    if (cm & 2) C.d.ny = (F.d.ny + 1) / 2;
    if (cm & 4) C.d.nz = (F.d.nz + 1) / 2;
    const size_t nc = C.d.n(), Mc = dmda_of(C.d).n();
    double* nv_tmp = tmp.alloc<double>(Mc);
    double* nvc = tmp.alloc<double>(Mc);
    double* gc = gstore.alloc<double>(9 * nc);
    C.code = h->geo.alloc<uint32_t>(nc);
    C.cls = h->geo.alloc<uint8_t>(nc);
    launch_coarsen(s, F.d, F.nv, F.g, F.code, cm, C.d, nv_tmp, nvc, gc, C.code, C.cls);
    C.nv = nvc;
    C.g = gc;
    H.push_back(C);
  }
  if ((long long)H.back().d.n() > 1024)
    throw StatusError(VFS_PP_EINVAL, "grid cannot be coarsened to a direct-solve size");
}

// VolumeFlux face list (poisson.c:2936-3010 / 2670-2700) in (k, j, i, axis) order,
// and the number of cells with nvert < 0.1
static void build_ib_faces(vfs_pp_handle* h, Arena& tmp) {
  hipStream_t s = h->s;
  const size_t nb = ibface_blocks(h->d);
  int* cnt = tmp.alloc<int>(3 * nb);
  launch_ibface_count(s, h->m, h->d, h->nvert, h->grid.immersed, cnt, cnt + nb, cnt + 2 * nb,
                      h->hpin_dev + RED_SETUP);
  HIPCHECK(hipStreamSynchronize(s));
  h->nibf = (size_t)h->hpin[RED_SETUP];
  h->n_lt = h->hpin[RED_SETUP + 1];
  h->ibf = nullptr;
  if (h->nibf) {
    h->ibf = h->geo.alloc<IbFace>(h->nibf);
    launch_ibface_write(s, h->m, h->d, h->nvert, h->met9, h->grid.immersed, cnt + 2 * nb, h->ibf);
  }
}

template <typename T>
static std::unique_ptr<MG<T>> build_mg(vfs_pp_handle* h) {
  auto mg = std::make_unique<MG<T>>();
  Arena& A = h->geo;
  mg->degree = std::max(1, h->opts.cheb_degree);
  const size_t n0 = h->d.n();
  mg->scr = A.alloc<T>(2 * n0);
  for (size_t l = 0; l < h->setup.size(); l++) {
    const SetupLevel& H = h->setup[l];
    typename MG<T>::Lev L;
    L.d = H.d;
    const size_t n = H.d.n();
    const int line = H.line;
    L.line = line;
    // coefficients: FP16 on every level of the FP32 hierarchy, scaled by a power
    // of two that puts the largest magnitude below 2^14 (FP16 keeps normal numbers
    // down to 2^-14); the FP64 hierarchy uses the set-up's FP64 coefficients (g64 on
    // the finest level)
    if constexpr (std::is_same<T, float>::value) {
      double gs = 1.;
      if (H.gmax > 0.) gs = std::ldexp(1., std::ilogb(H.gmax) + 1 - 14);
      L.gs = T(gs);
      L.gh = A.alloc<_Float16>(9 * n);
      launch_convert_scaled<double, _Float16>(h->s, 9 * n, H.g, 1. / gs, L.gh);
    } else {
      L.g = const_cast<T*>(H.g);
    }
    L.code = H.code;
    L.cls = H.cls;
    L.f = A.alloc<T>(n);
    L.xa = A.alloc<T>(n);
    L.xb = A.alloc<T>(n);
    L.dv = A.alloc<T>(n);
    L.dinv = A.alloc<T>(n);
    launch_dinv<T>(h->s, L.view(), L.dinv);
    if (line) {
      L.lfac = A.alloc<T>(3 * n);
      launch_line_factors<T>(h->s, L.view(), L.lfac);
    }
    mg->lev.push_back(L);
  }
  // the finest level's f D^-1, stored by the writers of the V-cycle input
  if (mg->lev.size() > 1 && !mg->lev[0].lfac) mg->fd = A.alloc<T>(n0);
  // Chebyshev parameters per smoothed level: the power iterations of all levels are
  // queued together and read back once
  for (size_t l = 0; l + 1 < mg->lev.size(); l++)
    queue_power_iteration<T>(h, *mg, mg->lev[l], h->hpin_dev + RED_LAMBDA + 12 * l);
  HIPCHECK(hipStreamSynchronize(h->s));
  for (size_t l = 0; l + 1 < mg->lev.size(); l++) {
    auto& L = mg->lev[l];
    L.lam = power_estimate(h->hpin + RED_LAMBDA + 12 * l);
    MG<T>::cheb(L.lam, mg->degree, h->opts.cheb_lo, h->opts.cheb_hi, L.c1, L.c2);
  }
  // coarsest dense inverse
  const SetupLevel& Hc = h->setup.back();
  CoarseRows R;
  R.d = Hc.d;
  const size_t nc = Hc.d.n();
  R.g.resize(9 * nc);
  R.code.resize(nc);
  d2h(h->s, R.g.data(), Hc.g, 9 * nc);
  d2h(h->s, R.code.data(), Hc.code, nc);
  HIPCHECK(hipStreamSynchronize(h->s));
  int ncomp = 0;
  const std::vector<double> Bc = coarse_bordered(R, &ncomp);
  mg->ncoarse = (int)nc;
  mg->ncomp = ncomp;
  mg->cinv = A.alloc<T>(nc * nc);
  coarse_inverse_device<T>(h->s, h->setup_tmp, Bc, (int)nc, mg->cinv);
  // the coarse tail's thread-to-cell order, from the codes of the level above the
  // dense one (FP32 point levels that fit one workgroup)
  std::vector<uint16_t> ord;
  if (std::is_same<T, float>::value && h->setup.size() >= 2) {
    const SetupLevel& Ht = h->setup[h->setup.size() - 2];
    if (!Ht.line && Ht.d.n() <= (size_t)TAIL_MAX_CELLS) {
      std::vector<uint32_t> tcode(Ht.d.n());
      d2h(h->s, tcode.data(), Ht.code, tcode.size());
      HIPCHECK(hipStreamSynchronize(h->s));
      ord = tail_order(tcode.size(), tcode.data());
      mg->tord = A.alloc<uint16_t>(ord.size());
      h2d(h->s, mg->tord, ord.data(), ord.size());
    }
  }
  HIPCHECK(hipStreamSynchronize(h->s));
  return mg;
}

void need_partials(vfs_pp_handle* h, size_t nvec, size_t nblk) {
  if (nvec * nblk > h->partial_cap)
    throw StatusError(VFS_PP_EINVAL, "reduction scratch too small (" + std::to_string(nvec * nblk) + " > " +
                                         std::to_string(h->partial_cap) + ")");
}

// FGMRES(m) with CGS2; right preconditioner = one V-cycle.  Solves A x = b
// (x holds the initial guess).  Returns iterations; fills rel.
template <typename TZ>
static int fgmres(vfs_pp_handle* h, double bnorm, double* rel_out, int* vcyc, bool* conv) {
  const size_t n = h->d.n();
  const int m = h->mrest;
  const double rtol = h->opts.rtol;
  TZ* Z = static_cast<TZ*>(h->Z);
  // the V-cycle input's writers also store f D^-1 for the finest level's first steps
  TZ* fd0 = (h->opts.fused & 2) ? static_cast<TZ*>(h->mg->fd0()) : nullptr;
  const TZ* dinv0 = fd0 ? static_cast<const TZ*>(h->mg->dinv0()) : nullptr;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), gv(m + 1), y(m), hcol(m + 2);
  // lazily normalised basis (orthog 2): U_l = nu_l v_l, V-cycle inputs scaled so
  // that the stored Z_l = sigma_l z_l
  std::vector<double> nu(m + 1, 1.), sigma(m + 1, 1.);
  int it = 0;
  double rel = 1.;
  *conv = false;
  double tgt = rtol * bnorm;
  // the first check of a zero-start solve needs no operator pass (r = b)
  bool x_zero = h->opts.zero_guess != 0;
  while (true) {
    int nb = 0;  // block count of the reductions below
    launch_true_residual(h, h->s, x_zero);
    HIPCHECK(hipStreamSynchronize(h->s));
    const double beta = std::sqrt(true_residual_host_norm2(h, x_zero));
    rel = bnorm > 0. ? beta / bnorm : 0.;
    if (!std::isfinite(beta)) throw StatusError(VFS_PP_ENONFINITE, "non-finite residual");
    if (h->opts.verbose)
      std::fprintf(stderr, "vfs_pp: restart at it %d  true rel %.3e  div est %.3e\n", it, rel,
                   h->div_fac * h->hpin[RED_DIVEST]);
    if (residual_verdict(h, beta, bnorm, tgt)) { *conv = true; break; }
    if (it >= h->opts.max_iters) break;
    // v_0 = r / beta and its copy as the first V-cycle input
    launch_scale_conv<TZ>(h->s, n, true_residual_src(h, x_zero), true_residual_norm2(h, x_zero), h->V,
                          static_cast<TZ*>(h->mg->f0()), dinv0, fd0);
    x_zero = false;
    if (h->opts.orthog == 2) {
      nu[0] = 1.;
      sigma[0] = 1.;
      h->hpin[RED_NUINV2] = 1.;
      h2d(h->s, h->red + RED_NUINV2, h->hpin + RED_NUINV2, 1);
    }
    std::fill(gv.begin(), gv.end(), 0.);
    gv[0] = beta;
    int j = 0;
    for (; j < m && it < h->opts.max_iters; j++) {
      double* w = h->V + (size_t)(j + 1) * n;
      // z_j = M^-1 v_j (the V-cycle reads v_j from its level-0 right-hand side)
      h->mg->set_out(Z + (size_t)j * n);
      h->mg->tail = (h->opts.fused & 1) != 0;
      h->mg->use_fd = fd0 != nullptr;
      h->mg->vcycle(h->s);
      (*vcyc)++;
      double sub = 0.;
      if (h->opts.orthog == 2) {
        // one classical Gram-Schmidt pass on the unnormalised basis, fused with the
        // operator (dots) and with the next V-cycle's input (axpy + norm).  One pass
        // leaves w' orthogonal to the basis to about eps |w| / |w'|, so a second pass
        // runs only when the first cancelled |w'| below 1e-4 |w|.
        const int nv = j + 1;
        need_partials(h, (size_t)j + 2, std::max(4 * h->stencil_blocks, (n + 255) / 256));
        launch_apply_dots<TZ>(h->s, h->d, h->g64, h->code, Z + (size_t)j * n, w, h->V, n, nv, h->partial, &nb);
        launch_finalize(h->s, h->partial, nv + 1, nb, h->red + 8, 0, h->hpin_dev + 8);
        launch_axpy_norm<TZ>(h->s, n, w, h->V, n, nv, h->red + 8, h->red + RED_NUINV2, h->red + 8 + nv,
                             static_cast<TZ*>(h->mg->f0()), h->partial, &nb, dinv0, fd0);
        launch_finalize(h->s, h->partial, 1, nb, h->red, 0, h->hpin_dev);
        HIPCHECK(hipStreamSynchronize(h->s));
        const double a2 = h->hpin[8 + nv];
        double b2 = h->hpin[0];
        for (int l = 0; l <= j; l++) H[(size_t)l * m + j] = h->hpin[8 + l] / (nu[l] * sigma[j]);
        if (b2 < 1e-8 * a2) {
          launch_multidot(h->s, n, w, h->V, n, nv, h->partial, &nb);
          launch_finalize(h->s, h->partial, nv, nb, h->red + 72, 0, h->hpin_dev + 72);
          launch_axpy_norm<TZ>(h->s, n, w, h->V, n, nv, h->red + 72, h->red + RED_NUINV2, h->red + 8 + nv,
                               static_cast<TZ*>(h->mg->f0()), h->partial, &nb, dinv0, fd0);
          launch_finalize(h->s, h->partial, 1, nb, h->red, 0, h->hpin_dev);
          HIPCHECK(hipStreamSynchronize(h->s));
          for (int l = 0; l <= j; l++) H[(size_t)l * m + j] += h->hpin[72 + l] / (nu[l] * sigma[j]);
          b2 = h->hpin[0];
        }
        nu[j + 1] = std::sqrt(b2);
        sigma[j + 1] = host_conv_scale(a2) * nu[j + 1];
        sub = nu[j + 1] / sigma[j];
        h->hpin[RED_NUINV2 + j + 1] = b2 > 0. ? 1. / b2 : 0.;
        h2d(h->s, h->red + RED_NUINV2 + j + 1, h->hpin + RED_NUINV2 + j + 1, 1);
      } else if (h->opts.orthog == 1) {
        // two-pass classical Gram-Schmidt (CGS2)
        launch_apply64<TZ>(h->s, h->d, h->g64, h->code, Z + (size_t)j * n, w, false);
        for (int pass = 0; pass < 2; pass++) {
          need_partials(h, (size_t)j + 1, (size_t)partial_blocks(n));
          launch_multidot(h->s, n, w, h->V, n, j + 1, h->partial, &nb);
          launch_finalize(h->s, h->partial, j + 1, nb, h->red + 8 + pass * 64, 0, h->hpin_dev + 8 + pass * 64);
          launch_axpy_multi(h->s, n, w, h->V, n, j + 1, h->red + 8 + pass * 64, -1.);
        }
        launch_norm2(h->s, n, w, h->partial, &nb);
        launch_finalize(h->s, h->partial, 1, nb, h->red, 0, h->hpin_dev);
        launch_scale_conv<TZ>(h->s, n, w, h->red, w, static_cast<TZ*>(h->mg->f0()), dinv0, fd0);
        HIPCHECK(hipStreamSynchronize(h->s));
        for (int l = 0; l <= j; l++) H[(size_t)l * m + j] = h->hpin[8 + l] + h->hpin[8 + 64 + l];
        sub = std::sqrt(h->hpin[0]);
      } else {
        // CGS2 in three passes over the basis: w = A z_j with h1 = V^T w fused into
        // the operator pass; w1 = w - V h1 with h2 = V^T w1 and |w1|^2 fused; then
        // v_{j+1} = (w1 - V h2) / sqrt(|w1|^2 - |h2|^2) together with the next
        // V-cycle's input.  After the first pass w1 is nearly orthogonal to V,
        // so the last norm has no cancellation.
        need_partials(h, (size_t)j + 2, std::max(4 * h->stencil_blocks, (n + 255) / 256));
        launch_apply_dots<TZ>(h->s, h->d, h->g64, h->code, Z + (size_t)j * n, w, h->V, n, j + 1, h->partial, &nb);
        launch_finalize(h->s, h->partial, j + 1, nb, h->red + 8, 0, h->hpin_dev + 8);
        launch_axpy_dots(h->s, n, w, h->V, n, j + 1, h->red + 8, h->partial, &nb);
        launch_finalize(h->s, h->partial, j + 2, nb, h->red + 72, 0, h->hpin_dev + 72);
        launch_gs_norm(h->s, h->red, 72, j + 1, h->hpin_dev);
        launch_axpy_scale<TZ>(h->s, n, w, h->V, n, j + 1, h->red + 72, h->red + 2, static_cast<TZ*>(h->mg->f0()),
                              dinv0, fd0);
        HIPCHECK(hipStreamSynchronize(h->s));
        for (int l = 0; l <= j; l++) H[(size_t)l * m + j] = h->hpin[8 + l] + h->hpin[72 + l];
        sub = h->hpin[1];
      }
      H[(size_t)(j + 1) * m + j] = sub;
      for (int l = 0; l < j; l++) {
        const double t1 = cs[l] * H[(size_t)l * m + j] + sn[l] * H[(size_t)(l + 1) * m + j];
        const double t2 = -sn[l] * H[(size_t)l * m + j] + cs[l] * H[(size_t)(l + 1) * m + j];
        H[(size_t)l * m + j] = t1;
        H[(size_t)(l + 1) * m + j] = t2;
      }
      const double a = H[(size_t)j * m + j], bb = H[(size_t)(j + 1) * m + j];
      const double rr = std::sqrt(a * a + bb * bb);
      cs[j] = rr == 0. ? 1. : a / rr;
      sn[j] = rr == 0. ? 0. : bb / rr;
      H[(size_t)j * m + j] = rr;
      H[(size_t)(j + 1) * m + j] = 0.;
      gv[j + 1] = -sn[j] * gv[j];
      gv[j] = cs[j] * gv[j];
      it++;
      if (!std::isfinite(gv[j + 1])) throw StatusError(VFS_PP_ENONFINITE, "non-finite Krylov residual");
      if (h->opts.verbose)
        std::fprintf(stderr, "vfs_pp: it %3d  est rel %.3e  h_jj %.3e  sub %.3e\n", it, std::fabs(gv[j + 1]) / bnorm,
                     H[(size_t)j * m + j], sub);
      if (std::fabs(gv[j + 1]) <= tgt) { j++; break; }
    }
    for (int l = j - 1; l >= 0; l--) {
      double s = gv[l];
      for (int q = l + 1; q < j; q++) s -= H[(size_t)l * m + q] * y[q];
      y[l] = s / H[(size_t)l * m + l];
    }
    for (int l = 0; l < j; l++) h->hpin[l] = h->opts.orthog == 2 ? y[l] / sigma[l] : y[l];
    h2d(h->s, h->red + 200, h->hpin, (size_t)j);
    launch_update_x<TZ>(h->s, n, h->x, Z, n, j, h->red + 200);
  }
  *rel_out = rel;
  return it;
}

void ensure_geom(vfs_pp_handle* h) {
  if (!h->have_geom) throw StatusError(VFS_PP_ESTATE, "vfs_pp_set_geometry has not been called");
}

// Everything that depends on nvert (device DMDA array): code words and masks, the
// VolumeFlux face list and the multigrid hierarchy.  Built on the device; the host
// only reads back a few statistics per level and inverts the coarsest level (at
// most 1024 cells).  Replaces the previous mask's hierarchy and face list, reusing
// their buffers where the sizes repeat; keep_scratch keeps the set-up's scratch
// buffers for the next call (a body that moves every step).
static void set_mask_core(vfs_pp_handle* h, const double* dnvert, bool keep_scratch) {
  hipStream_t s = h->s;
  HIPCHECK(hipStreamSynchronize(s));
  h->have_geom = false;
  h->mg.reset();
  h->ibf = nullptr;
  h->nibf = 0;
  h->geo.recycle();
  Arena& tmp = h->setup_tmp;
  tmp.recycle();
  HIPCHECK(hipMemcpyAsync(h->nvert, dnvert, h->m.n() * sizeof(double), hipMemcpyDeviceToDevice, s));
  launch_codes(s, h->m, h->d, h->nvert, h->code, h->cls);
  build_ib_faces(h, tmp);
  build_hierarchy(h, tmp);
  if (h->opts.precision == 1) h->mg = build_mg<float>(h);
  else h->mg = build_mg<double>(h);
  h->mg->tail = (h->opts.fused & 1) != 0;
  h->desc = h->mg->describe();
  h->setup.clear();
  HIPCHECK(hipStreamSynchronize(s));
  h->geo.trim();
  if (keep_scratch) {
    tmp.trim();
    tmp.recycle();
  } else {
    tmp.release();
  }
  h->have_geom = true;
}

// Set-up from the 16 metric arrays on the device (dmet: icsi..kzet, iaj, jaj, kaj,
// aj, csi, eta, zet) and nvert on the device.
static void set_geometry_core(vfs_pp_handle* h, const double* const* dmet, const double* dnvert) {
  const Dmda m = h->m;
  const Dims d = h->d;
  const size_t M = m.n(), n = d.n();
  hipStream_t s = h->s;
  HIPCHECK(hipStreamSynchronize(s));
  h->have_geom = false;
  h->have_metrics = false;
  // resident arrays
  if (!h->g64) {
    h->g64 = h->ar.alloc<double>(9 * n);
    h->code = h->ar.alloc<uint32_t>(n);
    h->cls = h->ar.alloc<uint8_t>(n);
    h->met9 = h->ar.alloc<double>(9 * n);
    h->aj = h->ar.alloc<double>(n);
    h->nvert = h->ar.alloc<double>(M);
    h->x = h->ar.alloc<double>(n);
    h->b = h->ar.alloc<double>(n);
    h->r = h->ar.alloc<double>(n);
    h->mrest = std::max(1, std::min(h->opts.restart, 60));
    // per-block partial sums: up to mrest+2 vectors per launch (fused Krylov
    // reductions over the cell grid), or 1024 blocks for grid-stride reductions
    h->stencil_blocks = (size_t)((d.nx + 63) / 64) * ((d.ny + 3) / 4) * d.nz;
    // the operator pass with the basis dots writes one partial per wave (4 per block)
    h->partial_cap =
        (size_t)(h->mrest + 4) * std::max<size_t>(std::max<size_t>(4 * h->stencil_blocks, (n + 255) / 256), 1024) +
        4096;
    h->partial = h->ar.alloc<double>(h->partial_cap);
    h->red = h->ar.alloc<double>(512);
    h->snap = h->ar.alloc<double>(3 * M);
    if (!h->external_ws) {
      h->V = h->ar.alloc<double>((size_t)(h->mrest + 1) * n);
      if (h->opts.precision == 1) h->Z = h->ar.alloc<float>((size_t)h->mrest * n);
      else h->Z = h->ar.alloc<double>((size_t)h->mrest * n);
    }
    HIPCHECK(hipHostMalloc(&h->hpin, 4096 * sizeof(double)));
    HIPCHECK(hipHostGetDevicePointer((void**)&h->hpin_dev, h->hpin, 0));
  }
  // face coefficients, and the cell metrics and aj on interior cells (Contra2Cart,
  // Divergence, VolumeFlux areas)
  launch_face_coeffs(s, m, d, dmet[0], dmet[1], dmet[2], dmet[3], dmet[4], dmet[5], dmet[6], dmet[7], dmet[8], dmet[9],
                     dmet[10], dmet[11], h->g64);
  launch_gather_cell_metrics(s, m, d, dmet[13], dmet[14], dmet[15], dmet[12], h->met9, h->aj);
  h->have_metrics = true;
  set_mask_core(h, dnvert, false);
}

// ------------------------------------------------------------ pressure path
// VolumeFlux and RHS, launches only; ||b||^2 lands in hpin[RED_B2] once the
// stream is synchronised.
void solve_prologue_host(vfs_pp_handle* h, double dt, double st, double coeff) {
  ensure_geom(h);
  // div u after projection = dt/coeff (b - A phi) holds for st = 1 (the RHS scales
  // by st, the projection by st as well); with another st the rule is off
  h->div_fac = (h->opts.div_tol > 0. && st == 1.) ? dt / coeff : 0.;
  h->div_goal = 0.5 * h->opts.div_tol;
}

void solve_prologue(vfs_pp_handle* h, double* ucont, double dt, double st, double coeff) {
  solve_prologue_host(h, dt, st, coeff);
  hipStream_t s = h->s;
  int nb = 0;
  if (h->nibf) {
    launch_ibflux_sum(s, h->ibf, h->nibf, ucont, h->grid.immersed, 1, h->partial, &nb);
    launch_finalize(s, h->partial, 2, nb, h->red + RED_IBF, 0, h->hpin_dev + RED_IBF);
    launch_ibflux_correct(s, h->ibf, h->nibf, ucont, h->grid.immersed, h->red + RED_IBF);
    launch_ibflux_sum(s, h->ibf, h->nibf, ucont, h->grid.immersed, 0, h->partial, &nb);
    launch_finalize(s, h->partial, 2, nb, h->red + RED_IBF + 2, 0, h->hpin_dev + RED_IBF + 2);
  }
  // RHS: val *= -1.0 / dt * st * coeff  (poisson_hypre.c:376)
  const double fac = -1.0 / dt * st * coeff;
  launch_rhs(s, h->m, h->d, ucont, h->cls, fac, h->b, h->partial, &nb);
  launch_finalize(s, h->partial, 1, nb, h->red + RED_B2, 0, h->hpin_dev + RED_B2);
}

// After the prologue's stream sync: ||b|| and the VolumeFlux figures.
void solve_start_stats(vfs_pp_handle* h, vfs_pp_stats* S) {
  if (h->nibf)
    for (int q = 0; q < 4; q++) S->ibm_flux[q] = h->hpin[RED_IBF + q];
  S->bnorm = std::sqrt(h->hpin[RED_B2]);
  if (!std::isfinite(S->bnorm)) throw StatusError(VFS_PP_ENONFINITE, "non-finite right-hand side");
}

// After the prologue's stream sync: ||b||, VolumeFlux figures, initial guess.
void solve_start(vfs_pp_handle* h, double* phi, vfs_pp_stats* S) {
  solve_start_stats(h, S);
  if (h->opts.zero_guess) launch_fill(h->s, h->d.n(), h->x, 0.);
  else launch_gather_phi(h->s, h->m, h->d, phi, h->cls, h->x);
}

// Remove_Nullspace (poisson_hypre.c:207-270) and Convert_Phi2_Phi
// (poisson.c:97-185), launches only; the subtracted mean is hpin[RED_MEAN] / n_lt
// after the next stream sync.
void solve_epilogue(vfs_pp_handle* h, double* phi, vfs_pp_stats* S) {
  hipStream_t s = h->s;
  int nb = 0;
  launch_masked_sum(s, h->d, h->x, h->cls, h->partial, &nb);
  launch_finalize(s, h->partial, 1, nb, h->red + RED_MEAN, 0, h->hpin_dev + RED_MEAN);
  if (h->n_lt > 0.) launch_sub_mean(s, h->d, h->x, h->cls, h->red + RED_MEAN, h->n_lt);
  launch_scatter_update(s, h->m, h->d, h->x, h->cls, phi, nullptr);
  launch_periodic_scalar(s, h->m, phi, 1);
  (void)S;
}

static double phi_mean_of(const vfs_pp_handle* h) { return h->n_lt > 0. ? h->hpin[RED_MEAN] / h->n_lt : 0.; }

static void poisson_solve(vfs_pp_handle* h, double* ucont, double* phi, double dt, double st, double coeff,
                          vfs_pp_stats* out) {
  const double t0 = now_ms();
  vfs_pp_stats S{};
  solve_prologue(h, ucont, dt, st, coeff);
  HIPCHECK(hipStreamSynchronize(h->s));
  solve_start(h, phi, &S);
  double rel = 0.;
  int vc = 0;
  bool conv = false;
  int it;
  if (S.bnorm == 0.) { it = 0; rel = 0.; conv = true; }
  else if (h->opts.precision == 1) it = fgmres<float>(h, S.bnorm, &rel, &vc, &conv);
  else it = fgmres<double>(h, S.bnorm, &rel, &vc, &conv);
  S.iterations = it;
  S.vcycles = vc;
  S.rel_residual = rel;
  S.converged = conv ? 1 : 0;
  solve_epilogue(h, phi, &S);
  HIPCHECK(hipStreamSynchronize(h->s));
  S.phi_mean = phi_mean_of(h);
  S.ms_total = now_ms() - t0;
  if (out) *out = S;
}

void update_pressure(vfs_pp_handle* h, double* p, double* phi) {
  ensure_geom(h);
  // P += Phi on fluid interior cells; periodic ghosts of P and Phi (poisson.c:2522-2614)
  launch_add_phi(h->s, h->m, h->d, h->cls, phi, p);
  launch_periodic_scalar(h->s, h->m, p, 0);
  launch_periodic_scalar(h->s, h->m, phi, 0);
}

void projection(vfs_pp_handle* h, double* ucont, const double* phi, double* ucat, double dt, double st,
                       double coeff) {
  ensure_geom(h);
  const double fac = dt * st / coeff;
  launch_projection(h->s, h->m, h->d, h->g64, h->code, h->cls, phi, ucont, fac);
  launch_ucont_ghosts(h->s, h->m, ucont);
  launch_c2c(h->s, h->m, h->d, h->met9, h->cls, h->nvert, ucont, ucat, h->snap, h->grid.bctype);
}

// Divergence (main.c:1244-1391), launches only; the maximum is hpin[RED_DIV]
// after the next stream sync.
void divergence_launch(vfs_pp_handle* h, const double* ucont) {
  ensure_geom(h);
  int nb = 0;
  launch_divergence(h->s, h->m, h->d, h->code, h->aj, ucont, h->partial, &nb);
  launch_finalize(h->s, h->partial, 1, nb, h->red + RED_DIV, 1, h->hpin_dev + RED_DIV);
}

static double divergence(vfs_pp_handle* h, const double* ucont) {
  divergence_launch(h, ucont);
  HIPCHECK(hipStreamSynchronize(h->s));
  return h->hpin[RED_DIV];
}

vfs_pp_handle* clone_for_solves(const vfs_pp_handle* src) {
  if (!src->have_geom || !src->mg) throw StatusError(VFS_PP_ESTATE, "the case has no geometry to share");
  auto h = std::make_unique<vfs_pp_handle>();
  h->grid = src->grid;
  h->opts = src->opts;
  h->m = src->m;
  h->d = src->d;
  h->immersed_ref = src->immersed_ref;
  HIPCHECK(hipStreamCreateWithFlags(&h->s, hipStreamDefault));
  try {
    // read-only resident data of the owner
    h->g64 = src->g64;
    h->code = src->code;
    h->cls = src->cls;
    h->met9 = src->met9;
    h->aj = src->aj;
    h->nvert = src->nvert;
    h->ibf = src->ibf;
    h->nibf = src->nibf;
    h->n_lt = src->n_lt;
    h->desc = src->desc;
    // per-solve state of its own, sized as the owner's
    const size_t n = h->d.n(), M = h->m.n();
    h->mrest = src->mrest;
    h->stencil_blocks = src->stencil_blocks;
    h->partial_cap = src->partial_cap;
    h->x = h->ar.alloc<double>(n);
    h->b = h->ar.alloc<double>(n);
    h->r = h->ar.alloc<double>(n);
    h->partial = h->ar.alloc<double>(h->partial_cap);
    h->red = h->ar.alloc<double>(512);
    h->snap = h->ar.alloc<double>(3 * M);
    h->external_ws = true;
    HIPCHECK(hipHostMalloc(&h->hpin, 4096 * sizeof(double)));
    HIPCHECK(hipHostGetDevicePointer((void**)&h->hpin_dev, h->hpin, 0));
    h->mg = src->mg->clone_work(h->geo);
    h->geometry_of = src;
    h->have_metrics = true;
    h->have_geom = true;
  } catch (...) {
    vfs_pp_destroy(h.release());
    throw;
  }
  return h.release();
}

}  // namespace pp

// ======================================================================= C ABI
using namespace pp;

extern "C" {

void vfs_pp_default_opts(vfs_pp_opts* o) {
  if (!o) return;
  o->precision = 1;
  o->rtol = 1e-10;
  o->div_tol = 0.;
  o->max_iters = 200;
  o->restart = 30;
  o->smoother = 0;
  o->cheb_degree = 3;
  o->coarse_max = 512;
  o->drop_small = 0;
  o->zero_guess = 1;
  o->cheb_lo = 0.1;
  o->cheb_hi = 1.1;
  o->orthog = 2;
  o->verbose = 0;
  o->fused = 3;
}

const char* vfs_pp_error_string(int code) {
  switch (code) {
    case VFS_PP_OK: return "ok";
    case VFS_PP_EINVAL: return "invalid argument";
    case VFS_PP_ENOMEM: return "device memory exhausted";
    case VFS_PP_EHIP: return "HIP runtime error";
    case VFS_PP_ENOCONV: return "solver did not converge";
    case VFS_PP_ESTATE: return "call order violated";
    case VFS_PP_ENONFINITE: return "non-finite values";
    case VFS_PP_EIO: return "grid file unreadable";
    case VFS_PP_EBUSY: return "request refused: server queue full";
    default: return "unknown";
  }
}

int vfs_pp_create(const vfs_grid* g, const vfs_pp_opts* o, vfs_pp_handle** out) {
  ABI_BEGIN
  if (!g || !out) return VFS_PP_EINVAL;
  if (g->nx < 2 || g->ny < 2 || g->nz < 2) return VFS_PP_EINVAL;
  // the stencil kernels address interior arrays with 32-bit byte offsets
  if ((long long)(g->nx + 2) * (g->ny + 2) * (g->nz + 2) >= (1LL << 29)) return VFS_PP_EINVAL;
  for (int a = 0; a < 3; a++)
    if (g->periodic[a] && (a == 0 ? g->nx : a == 1 ? g->ny : g->nz) < 3) return VFS_PP_EINVAL;
  // VolumeFlux variants (poisson_hypre.c:662-667): 1 takes the flg form; 0 and 3
  // take the four-argument form, which zeroes the corrected faces for both
  // (poisson.c:2728); 2 adds the IB flux to the outlet (Add_IBMFlux_to_Outlet),
  // an inflow/outflow boundary this engine does not model
  if (g->immersed != 0 && g->immersed != 1 && g->immersed != 3) return VFS_PP_EINVAL;
  auto h = std::make_unique<vfs_pp_handle>();
  h->grid = *g;
  h->grid.immersed = g->immersed == 1 ? 1 : 0;
  h->immersed_ref = g->immersed;
  if (o) h->opts = *o;
  else vfs_pp_default_opts(&h->opts);
  if (!(h->opts.rtol > 0.) || h->opts.max_iters < 0 || h->opts.restart < 1) return VFS_PP_EINVAL;
  if (!(h->opts.cheb_lo > 0. && h->opts.cheb_lo < h->opts.cheb_hi)) return VFS_PP_EINVAL;
  h->d = Dims{g->nx, g->ny, g->nz, {g->periodic[0] ? 1 : 0, g->periodic[1] ? 1 : 0, g->periodic[2] ? 1 : 0}};
  h->m = Dmda{g->nx + 2, g->ny + 2, g->nz + 2, {h->d.per[0], h->d.per[1], h->d.per[2]}};
  // a blocking stream: its work is ordered after work the caller queued on the
  // legacy default stream (PyTorch's default stream, for one) and before work
  // queued there afterwards, so inputs written just before a call are seen
  HIPCHECK(hipStreamCreateWithFlags(&h->s, hipStreamDefault));
  *out = h.release();
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_destroy(vfs_pp_handle* h) {
  if (!h) return VFS_PP_OK;
  if (h->batch_refs > 0) return VFS_PP_ESTATE;  // still bound to a batch
  if (h->s) (void)hipStreamSynchronize(h->s);
  h->mg.reset();
  h->ar.release();
  if (h->hpin) (void)hipHostFree(h->hpin);
  if (h->s) (void)hipStreamDestroy(h->s);
  delete h;
  return VFS_PP_OK;
}

int vfs_pp_set_opts(vfs_pp_handle* h, const vfs_pp_opts* o) {
  if (!h || !o) return VFS_PP_EINVAL;
  if (h->batch_refs > 0) return VFS_PP_ESTATE;
  // precision and restart size the Krylov basis and partial-sum buffers, which
  // exist from the first set-up on (also after a set-up that failed); the
  // hierarchy options take effect at the next set-up, so they are fixed only
  // while a hierarchy is in place
  if (h->g64 && (o->precision != h->opts.precision || o->restart != h->opts.restart)) return VFS_PP_ESTATE;
  if (h->have_geom && (o->coarse_max != h->opts.coarse_max || o->smoother != h->opts.smoother ||
                       o->cheb_degree != h->opts.cheb_degree))
    return VFS_PP_ESTATE;
  if (!(o->rtol > 0.)) return VFS_PP_EINVAL;
  h->opts = *o;
  if (h->mg) {
    h->mg->tail = (o->fused & 1) != 0;
    h->desc = h->mg->describe();
  }
  return VFS_PP_OK;
}

int vfs_pp_set_geometry(vfs_pp_handle* h, const double* icsi, const double* ieta, const double* izet,
                        const double* jcsi, const double* jeta, const double* jzet, const double* kcsi,
                        const double* keta, const double* kzet, const double* iaj, const double* jaj,
                        const double* kaj, const double* aj, const double* csi, const double* eta, const double* zet,
                        const double* nvert) {
  ABI_BEGIN
  if (!h) return VFS_PP_EINVAL;
  ensure_unbound(h);
  const double* hmet[16] = {icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, aj, csi, eta, zet};
  for (auto p : hmet)
    if (!p) return VFS_PP_EINVAL;
  if (!nvert) return VFS_PP_EINVAL;
  const size_t M = h->m.n();
  Arena tmp;
  double* dmet[16];
  for (int q = 0; q < 16; q++) {
    const size_t len = (q < 9 || q >= 13) ? 3 * M : M;
    dmet[q] = tmp.alloc<double>(len);
    h2d(h->s, dmet[q], hmet[q], len);
  }
  double* dnv = tmp.alloc<double>(M);
  h2d(h->s, dnv, nvert, M);
  set_geometry_core(h, dmet, dnv);
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_form_metrics_device(const vfs_grid* g, const double* coor, double* icsi, double* ieta, double* izet,
                               double* jcsi, double* jeta, double* jzet, double* kcsi, double* keta, double* kzet,
                               double* iaj, double* jaj, double* kaj, double* aj, double* csi, double* eta,
                               double* zet) {
  ABI_BEGIN
  if (!g || !coor) return VFS_PP_EINVAL;
  if (g->nx < 2 || g->ny < 2 || g->nz < 2) return VFS_PP_EINVAL;
  double* out[16] = {icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, aj, csi, eta, zet};
  for (auto p : out)
    if (!p) return VFS_PP_EINVAL;
  const Dmda m{g->nx + 2, g->ny + 2, g->nz + 2, {g->periodic[0] ? 1 : 0, g->periodic[1] ? 1 : 0, g->periodic[2] ? 1 : 0}};
  launch_form_metrics(nullptr, m, coor, out);
  HIPCHECK(hipDeviceSynchronize());
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_set_grid_device(vfs_pp_handle* h, const double* coor, const double* nvert) {
  ABI_BEGIN
  if (!h || !coor || !nvert) return VFS_PP_EINVAL;
  ensure_unbound(h);
  const size_t M = h->m.n();
  Arena tmp;
  double* dmet[16];
  for (int q = 0; q < 16; q++) dmet[q] = tmp.alloc<double>((q < 9 || q >= 13) ? 3 * M : M);
  launch_form_metrics(h->s, h->m, coor, dmet);
  set_geometry_core(h, dmet, nvert);
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_set_nvert_device(vfs_pp_handle* h, const double* nvert) {
  ABI_BEGIN
  if (!h || !nvert) return VFS_PP_EINVAL;
  ensure_unbound(h);
  if (!h->have_metrics) return VFS_PP_ESTATE;
  set_mask_core(h, nvert, true);
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_poisson_solve_device(vfs_pp_handle* h, double* ucont, double* phi, double dt, double st, double coeff,
                                vfs_pp_stats* stats) {
  ABI_BEGIN
  if (!h || !ucont || !phi || !(dt > 0.) || coeff == 0.) return VFS_PP_EINVAL;
  ensure_unbound(h);
  vfs_pp_stats S{};
  poisson_solve(h, ucont, phi, dt, st, coeff, &S);
  if (stats) *stats = S;
  return S.converged ? VFS_PP_OK : VFS_PP_ENOCONV;
  ABI_END
}

int vfs_pp_update_pressure_device(vfs_pp_handle* h, double* p, double* phi) {
  ABI_BEGIN
  if (!h || !p || !phi) return VFS_PP_EINVAL;
  ensure_unbound(h);
  update_pressure(h, p, phi);
  HIPCHECK(hipStreamSynchronize(h->s));
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_projection_device(vfs_pp_handle* h, double* ucont, const double* phi, double* ucat, double dt, double st,
                             double coeff) {
  ABI_BEGIN
  if (!h || !ucont || !phi || !ucat || coeff == 0.) return VFS_PP_EINVAL;
  ensure_unbound(h);
  projection(h, ucont, phi, ucat, dt, st, coeff);
  HIPCHECK(hipStreamSynchronize(h->s));
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_divergence_device(vfs_pp_handle* h, const double* ucont, double* maxdiv) {
  ABI_BEGIN
  if (!h || !ucont || !maxdiv) return VFS_PP_EINVAL;
  ensure_unbound(h);
  *maxdiv = divergence(h, ucont);
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_project_device(vfs_pp_handle* h, double* ucont, double* phi, double* p, double* ucat, double dt,
                          double st, double coeff, vfs_pp_stats* stats) {
  ABI_BEGIN
  if (!h || !ucont || !phi || !p || !ucat || !(dt > 0.) || coeff == 0.) return VFS_PP_EINVAL;
  ensure_unbound(h);
  const double t0 = now_ms();
  vfs_pp_stats S{};
  poisson_solve(h, ucont, phi, dt, st, coeff, &S);
  update_pressure(h, p, phi);
  projection(h, ucont, phi, ucat, dt, st, coeff);
  S.max_div = divergence(h, ucont);
  S.ms_total = now_ms() - t0;
  bool ok = S.converged;
  if (h->opts.div_tol > 0. && !(S.max_div <= h->opts.div_tol)) ok = false;
  S.converged = ok ? 1 : 0;
  if (stats) *stats = S;
  return ok ? VFS_PP_OK : VFS_PP_ENOCONV;
  ABI_END
}

int vfs_pp_project(vfs_pp_handle* h, double* ucont, double* phi, double* p, double* ucat, double dt, double st,
                   double coeff, vfs_pp_stats* stats) {
  ABI_BEGIN
  if (!h || !ucont || !phi || !p || !ucat) return VFS_PP_EINVAL;
  ensure_unbound(h);
  ensure_geom(h);
  const size_t M = h->m.n();
  // device copies of the four arrays, kept by the handle for its next call (the
  // grid fixes their size); the copies and the projection follow each other on
  // the handle's stream
  if (!h->hstage) h->hstage = h->ar.alloc<double>(8 * M);
  double* du = h->hstage;
  double* dphi = du + 3 * M;
  double* dp = dphi + M;
  double* dc = dp + M;
  h2d(h->s, du, ucont, 3 * M);
  h2d(h->s, dphi, phi, M);
  h2d(h->s, dp, p, M);
  h2d(h->s, dc, ucat, 3 * M);
  const int rc = vfs_pp_project_device(h, du, dphi, dp, dc, dt, st, coeff, stats);
  if (rc != VFS_PP_OK && rc != VFS_PP_ENOCONV) return rc;
  d2h(h->s, ucont, du, 3 * M);
  d2h(h->s, phi, dphi, M);
  d2h(h->s, p, dp, M);
  d2h(h->s, ucat, dc, 3 * M);
  HIPCHECK(hipStreamSynchronize(h->s));
  return rc;
  ABI_END
}

int vfs_pp_apply_device(vfs_pp_handle* h, const double* phi, double* y, int drop_small) {
  ABI_BEGIN
  if (!h || !phi || !y) return VFS_PP_EINVAL;
  ensure_unbound(h);
  ensure_geom(h);
  launch_apply_dmda(h->s, h->m, h->d, h->g64, h->code, phi, y, drop_small != 0);
  HIPCHECK(hipStreamSynchronize(h->s));
  return VFS_PP_OK;
  ABI_END
}

// Rows of a level (device g and code) as 19 coefficients and 18 case bytes per point
// of its DMDA box; zero outside the fluid rows.
static void export_rows(vfs_pp_handle* h, const Dims& d, const double* g, const uint32_t* code, double* vol19,
                        uint8_t* cases) {
  const Dmda m = dmda_of(d);
  const size_t n = d.n(), M = m.n();
  Arena tmp;
  double* dv = tmp.alloc<double>(19 * n);
  launch_export_rows(h->s, d, g, code, dv);
  std::vector<double> hv(19 * n);
  std::vector<uint32_t> hc(n);
  d2h(h->s, hv.data(), dv, 19 * n);
  d2h(h->s, hc.data(), code, n);
  HIPCHECK(hipStreamSynchronize(h->s));
  std::memset(vol19, 0, 19 * M * sizeof(double));
  std::memset(cases, 0, 18 * M);
  for (int k = 0; k < d.nz; k++)
    for (int j = 0; j < d.ny; j++)
      for (int i = 0; i < d.nx; i++) {
        const size_t c = d.idx(i, j, k), p = m.idx(i + 1, j + 1, k + 1);
        if (!(hc[c] & ROW_BIT)) continue;
        std::memcpy(vol19 + 19 * p, hv.data() + 19 * c, 19 * sizeof(double));
        for (int f = 0; f < 6; f++) {
          const bool act = (hc[c] >> f) & 1u;
          cases[18 * p + 3 * f] = act ? 1 : 0;
          if (!act) continue;
          cases[18 * p + 3 * f + 1] = (uint8_t)((hc[c] >> (6 + 4 * f)) & 3u);
          cases[18 * p + 3 * f + 2] = (uint8_t)((hc[c] >> (8 + 4 * f)) & 3u);
        }
      }
}

int vfs_pp_export_stencil(vfs_pp_handle* h, double* vol19, uint8_t* cases) {
  ABI_BEGIN
  if (!h || !vol19 || !cases) return VFS_PP_EINVAL;
  ensure_unbound(h);
  ensure_geom(h);
  export_rows(h, h->d, h->g64, h->code, vol19, cases);
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_export_level(vfs_pp_handle* h, int level, int* nlevels, int dims[3], int* line, double* vol19,
                        uint8_t* cases) {
  ABI_BEGIN
  if (!h) return VFS_PP_EINVAL;
  ensure_unbound(h);
  ensure_geom(h);
  const int nl = h->mg->nlev();
  if (nlevels) *nlevels = nl;
  if (level < 0 || level >= nl) return VFS_PP_EINVAL;
  Dims d{};
  int ln = 0;
  const uint32_t* code = nullptr;
  h->mg->level_export(h->s, level, &d, &ln, &code, nullptr);
  if (dims) {
    dims[0] = d.nx;
    dims[1] = d.ny;
    dims[2] = d.nz;
  }
  if (line) *line = ln;
  if (vol19 && cases) {
    Arena tmp;
    double* g = tmp.alloc<double>(9 * d.n());
    h->mg->level_export(h->s, level, &d, &ln, &code, g);
    export_rows(h, d, g, code, vol19, cases);
  }
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_export_coarse_inverse(vfs_pp_handle* h, size_t cap, double* cinv, int* n) {
  ABI_BEGIN
  if (!h || !n) return VFS_PP_EINVAL;
  ensure_unbound(h);
  ensure_geom(h);
  const int nc = h->mg->coarse_cells();
  *n = nc;
  if (!cinv || cap < (size_t)nc * nc) return VFS_PP_OK;
  const std::vector<double> c = h->mg->coarse_inverse_export(h->s);
  std::copy(c.begin(), c.end(), cinv);
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_export_ib_faces(vfs_pp_handle* h, size_t cap, long long* pos, double* area, float* sign, size_t* count) {
  ABI_BEGIN
  if (!h || !count) return VFS_PP_EINVAL;
  ensure_unbound(h);
  ensure_geom(h);
  *count = h->nibf;
  if (!pos || !area || !sign || cap < h->nibf) return VFS_PP_OK;
  std::vector<IbFace> f(h->nibf);
  if (h->nibf) {
    d2h(h->s, f.data(), h->ibf, h->nibf);
    HIPCHECK(hipStreamSynchronize(h->s));
  }
  for (size_t i = 0; i < h->nibf; i++) {
    pos[i] = (long long)f[i].pos;
    area[i] = f[i].area;
    sign[i] = f[i].sign;
  }
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_classify_nvert_device(const vfs_grid* g, double* nvert) {
  ABI_BEGIN
  if (!g || !nvert) return VFS_PP_EINVAL;
  const Dmda m{g->nx + 2, g->ny + 2, g->nz + 2, {g->periodic[0], g->periodic[1], g->periodic[2]}};
  launch_classify_nvert(nullptr, m, g->bctype, nvert, nullptr);
  HIPCHECK(hipDeviceSynchronize());
  return VFS_PP_OK;
  ABI_END
}

static Dmda dmda_fine(const vfs_grid* g) {
  return Dmda{g->nx + 2, g->ny + 2, g->nz + 2, {g->periodic[0], g->periodic[1], g->periodic[2]}};
}
static Dmda dmda_coarse(const vfs_grid* g) {
  // IM_c = (IM_f + 1) / 2 with IM = m - 1 (init.c:95-127)
  auto c = [](int mf) { return ((mf - 1) + 1) / 2 + 1; };
  return Dmda{c(g->nx + 2), c(g->ny + 2), c(g->nz + 2), {g->periodic[0], g->periodic[1], g->periodic[2]}};
}

int vfs_pp_mg_restrict_device(const vfs_grid* g, const double* nvf, const double* nvc, const double* xf, double* fc) {
  ABI_BEGIN
  if (!g || !nvf || !nvc || !xf || !fc || (g->nx % 2) || (g->ny % 2) || (g->nz % 2)) return VFS_PP_EINVAL;
  launch_ref_restrict(nullptr, dmda_fine(g), dmda_coarse(g), nvf, nvc, xf, fc);
  HIPCHECK(hipDeviceSynchronize());
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_mg_prolong_device(const vfs_grid* g, const double* nvf, const double* nvc, const double* xc, double* ff) {
  ABI_BEGIN
  if (!g || !nvf || !nvc || !xc || !ff || (g->nx % 2) || (g->ny % 2) || (g->nz % 2)) return VFS_PP_EINVAL;
  launch_ref_prolong(nullptr, dmda_fine(g), dmda_coarse(g), nvf, nvc, xc, ff);
  HIPCHECK(hipDeviceSynchronize());
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_mg_nvert_restrict_device(const vfs_grid* g, const double* nvf, double* nvc) {
  ABI_BEGIN
  if (!g || !nvf || !nvc || (g->nx % 2) || (g->ny % 2) || (g->nz % 2)) return VFS_PP_EINVAL;
  const Dmda c = dmda_coarse(g);
  Arena tmp;
  double* t = tmp.alloc<double>(c.n());
  launch_ref_nvert_restrict(nullptr, dmda_fine(g), c, nvf, nvc, t);
  HIPCHECK(hipDeviceSynchronize());
  return VFS_PP_OK;
  ABI_END
}

// Algorithmic bytes per cell of one finest-level Chebyshev sweep: a point sweep
// reads x, f, the coefficients and the mask and writes x (FP32 level with FP16
// coefficients 31, FP64 97); a y-line step whose lines fit the segmented layout
// reads x, f, coefficients, mask, the three Thomas factors and the direction and
// writes the direction and x (51 / 137); on longer lines the residual goes through
// memory (x, f, g, mask -> r, then r, factors, x -> x: 55 / 145).
static double sweep_bytes_per_cell(const vfs_pp_handle* h) {
  const bool f32 = h->opts.precision == 1;
  if (!h->mg->line0()) return f32 ? (4 + 4 + 18 + 1 + 4) : (8 + 8 + 72 + 1 + 8);
  const bool seg = f32 ? line_seg_fits<float>(h->d) : line_seg_fits<double>(h->d);
  if (seg) return f32 ? (4 + 4 + 18 + 1 + 12 + 8 + 4) : (8 + 8 + 72 + 1 + 24 + 16 + 8);
  return f32 ? (4 + 4 + 18 + 1 + 4) + (4 + 12 + 4 + 4) : (8 + 8 + 72 + 1 + 8) + (8 + 24 + 8 + 8);
}

int vfs_pp_time_kernel(vfs_pp_handle* h, int kernel, int reps, double* ms, double* bytes) {
  ABI_BEGIN
  if (!h || reps < 1 || !ms || !bytes) return VFS_PP_EINVAL;
  ensure_unbound(h);
  ensure_geom(h);
  const size_t n = h->d.n(), M = h->m.n();
  Arena tmp;
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  // kernel 5 copies 2^27 doubles (1 GiB), far beyond the 256 MB last-level cache
  const size_t ncopy = (size_t)1 << 27;
  double* y = tmp.alloc<double>(kernel == 5 ? ncopy : n);
  double* src = kernel == 5 ? tmp.alloc<double>(ncopy) : nullptr;
  double* u = nullptr;
  double* phi = nullptr;
  double* uc = nullptr;
  if (kernel >= 2 && kernel <= 4) {
    u = tmp.alloc<double>(3 * M);
    phi = tmp.alloc<double>(M);
    uc = tmp.alloc<double>(3 * M);
    HIPCHECK(hipMemsetAsync(u, 0, 3 * M * sizeof(double), h->s));
    HIPCHECK(hipMemsetAsync(phi, 0, M * sizeof(double), h->s));
    HIPCHECK(hipMemsetAsync(uc, 0, 3 * M * sizeof(double), h->s));
  }
  int nb = 0;
  // kernel 6: the finest sweep over rotating copies of its operands (HBM-resident);
  // kernel 8: the streaming floor of a point sweep's operands over the same copies
  if (kernel == 6 || kernel == 8) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (kernel == 8 && h->mg->line0()) return VFS_PP_EINVAL;
    const int copies = 6;
    *ms = h->mg->time_sweep_rotating(h->s, std::max(reps, copies), copies, kernel == 8);
    const double N = (double)n;
    // the floor kernel reads the previous iterate too
    *bytes = N * (kernel == 6 ? sweep_bytes_per_cell(h) : (h->opts.precision == 1 ? 35 : 105));
    return VFS_PP_OK;
  }
  // kernel 7: the Krylov operator pass w = A z with the dots of w against kDots
  // basis vectors and itself (k_apply_dots, z in the multigrid precision)
  const int kDots = 4;
  if (kernel == 7 && (!h->V || !h->Z)) return VFS_PP_EINVAL;
  double* wv = kernel == 7 ? tmp.alloc<double>(n) : nullptr;
  if (kernel == 7) need_partials(h, kDots + 1, std::max(4 * h->stencil_blocks, (n + 255) / 256));
  auto launch = [&]() {
    switch (kernel) {
      case 7:
        if (h->opts.precision == 1)
          launch_apply_dots<float>(h->s, h->d, h->g64, h->code, static_cast<const float*>(h->Z), wv, h->V, n, kDots,
                                   h->partial, &nb);
        else
          launch_apply_dots<double>(h->s, h->d, h->g64, h->code, static_cast<const double*>(h->Z), wv, h->V, n,
                                    kDots, h->partial, &nb);
        break;
      case 0: launch_apply64<double>(h->s, h->d, h->g64, h->code, h->x, y, false); break;
      case 1: h->mg->time_sweep(h->s); break;
      case 2: launch_projection(h->s, h->m, h->d, h->g64, h->code, h->cls, phi, u, 1.0); break;
      case 3: launch_c2c(h->s, h->m, h->d, h->met9, h->cls, h->nvert, u, uc, h->snap, h->grid.bctype); break;
      case 4: launch_rhs(h->s, h->m, h->d, u, h->cls, 1.0, y, h->partial, &nb); break;
      case 5: launch_copy(h->s, ncopy, src, y); break;
      default: throw StatusError(VFS_PP_EINVAL, "unknown kernel id");
    }
  };
  for (int w = 0; w < 3; w++) launch();
  HIPCHECK(hipEventRecord(e0, h->s));
  for (int r = 0; r < reps; r++) launch();
  HIPCHECK(hipEventRecord(e1, h->s));
  HIPCHECK(hipEventSynchronize(e1));
  float tms = 0.f;
  HIPCHECK(hipEventElapsedTime(&tms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *ms = (double)tms / reps;
  // Algorithmic bytes per launch: each distinct array once per cell, halos from
  // cache, a 1-byte mask per cell.  (The engine's kernels move somewhat more:
  // a 4-byte branch-code word instead of the mask, and the Chebyshev direction.)
  const double N = (double)n;
  switch (kernel) {
    case 0: *bytes = N * (8 + 72 + 1 + 8); break;  // phi, g11..g33, mask -> y           = 89
    case 1: *bytes = N * sweep_bytes_per_cell(h); break;
    case 2: *bytes = N * (8 + 72 + 1 + 24 + 24); break;  // phi, g, mask, Ucont r/w   = 129
    case 3: *bytes = N * (24 + 72 + 1 + 24); break;      // Ucont, csi/eta/zet, mask -> Ucat = 121
    case 4: *bytes = N * (24 + 1 + 8); break;            // Ucont, mask -> b        = 33
    case 5: *bytes = (double)ncopy * 16.; break;         // read + write
    case 7:  // z, g, mask, kDots basis vectors -> w
      *bytes = N * ((h->opts.precision == 1 ? 4 : 8) + 72 + 1 + 8 * kDots + 8);
      break;
  }
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_describe(vfs_pp_handle* h, char* buf, size_t len) {
  if (!h || !buf || len == 0) return VFS_PP_EINVAL;
  std::string s = h->desc;
  std::snprintf(buf, len, "%s", s.c_str());
  return VFS_PP_OK;
}

}  // extern "C"

// Kernel launchers of the pressure-projection engine.  All launches go to the
// stream passed in; nothing here synchronises.
#pragma once

#include <vector>

#include "pp_common.hpp"

namespace pp {

// Per-cell class bits (1 byte per interior cell), the scalar nvert tests used
// outside the operator rows.
constexpr uint8_t CLS_LT = 1;         // nvert < 0.1: RHS, Remove_Nullspace, UpdatePressure, Contra2Cart
constexpr uint8_t CLS_INTGT = 2;      // (int)nvert > 0.1: Convert_Phi2_Phi writes 0 (poisson.c:131)
constexpr uint8_t CLS_SOLID = 4;      // (int)(nvert+0.1) == 3: Contra2Cart zeroes Ucat (rhs.c:257)
constexpr uint8_t CLS_PROJ_I = 8;     // Projection corrects this cell's i+1/2 face (poisson.c:2221,2259)
constexpr uint8_t CLS_PROJ_J = 16;    // ... j+1/2 face
constexpr uint8_t CLS_PROJ_K = 32;    // ... k+1/2 face
constexpr uint8_t CLS_NONFLUID = 64;  // nvert > 0.1 (multigrid masks: MyInterpolation zeroes, restriction weight)
constexpr uint8_t CLS_INTERIOR = 128;  // multigrid level masks only: the cell's code word is the interior one

// DMDA description used by kernels touching caller arrays.
struct Dmda {
  int mx, my, mz;
  int per[3];
  HD size_t n() const { return (size_t)mx * my * mz; }
  HD size_t idx(int i, int j, int k) const { return ((size_t)k * my + j) * mx + i; }
};

// One face next to the immersed boundary (VolumeFlux).
struct IbFace {
  uint64_t pos;  // index into the AoS Ucont array (3*dmda_point + component)
  double area;   // |csi|, |eta| or |zet| of the cell (poisson.c:2676-2684)
  float sign;    // +1 fluid->IB, -1 IB->fluid
  float pad;
};

// ---- set-up ------------------------------------------------------------------
void launch_face_coeffs(hipStream_t s, Dmda m, Dims d, const double* icsi, const double* ieta, const double* izet,
                        const double* jcsi, const double* jeta, const double* jzet, const double* kcsi,
                        const double* keta, const double* kzet, const double* iaj, const double* jaj,
                        const double* kaj, double* g);
void launch_codes(hipStream_t s, Dmda m, Dims d, const double* nvert, uint32_t* code, uint8_t* cls);
void launch_export_rows(hipStream_t s, Dims d, const double* g, const uint32_t* code, double* vol19);
void launch_classify_nvert(hipStream_t s, Dmda m, const int bctype[6], double* nv, uint8_t* tmp);

// ---- multigrid hierarchy and VolumeFlux face list, built on the device ----------
// out4 (written by one thread, may be mapped host memory): the largest
// (|a_N| + |a_S|) / (|a_E| + |a_W| + |a_T| + |a_B|) over the rows of the level,
// the rows with a nonzero denominator, those of them below 1/10, the largest |g|.
// partial holds level_stats_partials() doubles.
void launch_level_stats(hipStream_t s, Dims d, const double* g, const uint32_t* code, double* partial, double* out4);
size_t level_stats_partials();
// One coarsening step (axes in cm) from level (df, DMDA nvert nvf, face coefficients
// gf, codes codef) to dc: coarse DMDA nvert nvc (nv_tmp: a second DMDA buffer),
// face coefficients gc, codes and class bytes (CLS_NONFLUID, CLS_INTERIOR).
void launch_coarsen(hipStream_t s, Dims df, const double* nvf, const double* gf, const uint32_t* codef, int cm,
                    Dims dc, double* nv_tmp, double* nvc, double* gc, uint32_t* codec, uint8_t* clsc);
// the finest level's multigrid class bytes (CLS_NONFLUID, CLS_INTERIOR) from its codes
void launch_cls_from_code(hipStream_t s, size_t n, const uint32_t* code, uint8_t* cls);
void launch_half_to_double(hipStream_t s, size_t n, const _Float16* a, double gs, double* out);
// v = uniform [-1, 1) on the rows (code & ROW_BIT), 0 elsewhere; a counter-based
// generator, so the vector depends on seed and n only
template <typename T>
void launch_random_rows(hipStream_t s, size_t n, const uint32_t* code, uint64_t seed, T* v);
// VolumeFlux faces in (k, j, i, axis) order: counts per block of 256 interior cells
// (cnt, lt: ibface_blocks(d) ints each), their exclusive prefix off, and
// tot2 = {faces, cells with nvert < 0.1}; then the list itself, with face areas
// from the interior cells' csi, eta, zet (met9, planar: plane 3 v + q).
size_t ibface_blocks(Dims d);
void launch_ibface_count(hipStream_t s, Dmda m, Dims d, const double* nv, int immersed, int* cnt, int* lt, int* off,
                         double* tot2);
void launch_ibface_write(hipStream_t s, Dmda m, Dims d, const double* nv, const double* met9, int immersed,
                         const int* off, IbFace* out);

// Dense inverse of the coarsest level on the device: Gauss-Jordan with partial
// pivoting on [B | I] (N x N each, row-major), one launch per column col = 0 .. N-1
// from (bi, ii) to (bo, io) (two buffer pairs in turn); launch_gj_out writes the
// column-major top-left n x n block of the final I as T.
void launch_gj_identity(hipStream_t s, int N, double* I);
void launch_gj_step(hipStream_t s, int N, int col, const double* bi, const double* ii, double* bo, double* io,
                    int* singular);
template <typename T>
void launch_gj_out(hipStream_t s, int N, int n, const double* I, T* cinv);

// ---- FP64 operator ----------------------------------------------------------------
// y = A x on interior vectors; x may be float or double.
template <typename TX>
void launch_apply64(hipStream_t s, Dims d, const double* g, const uint32_t* code, const TX* x, double* y,
                    bool drop);
// r = b - A x and per-block partial sums of r.r (nblk entries)
template <typename TX>
void launch_residual64(hipStream_t s, Dims d, const double* g, const uint32_t* code, const double* b, const TX* x,
                       double* r, double* partial, int* nblk);
// DMDA in/out variant for the public apply entry point
void launch_apply_dmda(hipStream_t s, Dmda m, Dims d, const double* g, const uint32_t* code, const double* phi,
                       double* y, bool drop);

// ---- pressure path ----------------------------------------------------------------
void launch_rhs(hipStream_t s, Dmda m, Dims d, const double* ucont, const uint8_t* cls, double fac, double* b,
                double* partial, int* nblk);
// VolumeFlux on a compact face list.  pass 0: sums (immersed==1 also zeroes |u|<=eps);
// pass 1: correction by corr (device scalar); pass 2: sums after.
void launch_ibflux_sum(hipStream_t s, const IbFace* f, size_t nf, double* ucont, int immersed, int zero_small,
                       double* partial, int* nblk);
void launch_ibflux_correct(hipStream_t s, const IbFace* f, size_t nf, double* ucont, int immersed,
                           const double* sums /* [flux, area] */);
void launch_masked_sum(hipStream_t s, Dims d, const double* x, const uint8_t* cls, double* partial, int* nblk);
void launch_sub_mean(hipStream_t s, Dims d, double* x, const uint8_t* cls, const double* sum, double inv_count);
void launch_scatter_update(hipStream_t s, Dmda m, Dims d, const double* x, const uint8_t* cls, double* phi,
                           double* p);
void launch_periodic_scalar(hipStream_t s, Dmda m, double* a, int zero_other);
void launch_gather_cell_metrics(hipStream_t s, Dmda m, Dims d, const double* csi, const double* eta,
                                const double* zet, const double* aj, double* met9, double* ajc);
void launch_form_metrics(hipStream_t s, Dmda m, const double* coor, double* const out[16]);
void launch_gather_phi(hipStream_t s, Dmda m, Dims d, const double* phi, const uint8_t* cls, double* x);
void launch_add_phi(hipStream_t s, Dmda m, Dims d, const uint8_t* cls, const double* phi, double* p);
void launch_projection(hipStream_t s, Dmda m, Dims d, const double* g, const uint32_t* code, const uint8_t* cls,
                       const double* phi, double* ucont, double fac);
void launch_ucont_ghosts(hipStream_t s, Dmda m, double* ucont);
void launch_c2c(hipStream_t s, Dmda m, Dims d, const double* met9, const uint8_t* cls, const double* nvert,
                const double* ucont, double* ucat, double* scratch, const int bctype[6]);
void launch_divest(hipStream_t s, Dims d, const uint32_t* code, const double* aj, const double* r, double* partial,
                   int* nblk);
void launch_divergence(hipStream_t s, Dmda m, Dims d, const uint32_t* code, const double* aj, const double* ucont,
                       double* partial, int* nblk);

// ---- reductions / BLAS-1 ---------------------------------------------------------
// out[v] = sum_b partial[v*nblk + b] in a fixed order (deterministic); op 0 sum, 1 max
void launch_finalize(hipStream_t s, const double* partial, int nvec, int nblk, double* out, int op,
                     double* host_mirror = nullptr);
void launch_multidot(hipStream_t s, size_t n, const double* w, const double* V, size_t ldv, int nv, double* partial,
                     int* nblk);
void launch_axpy_multi(hipStream_t s, size_t n, double* w, const double* V, size_t ldv, int nv, const double* h,
                       double sign);
void launch_norm2(hipStream_t s, size_t n, const double* a, double* partial, int* nblk);
// The writers of the V-cycle's input f (launch_axpy_norm, launch_axpy_scale,
// launch_scale_conv) also store fd = f * dinv when fd is given: the finest level's
// fused first two Chebyshev steps then read f D^-1 at the 19 neighbours from one
// array instead of forming it from two.
template <typename TF>
void launch_axpy_norm(hipStream_t s, size_t n, double* w, const double* V, size_t ldv, int nv, const double* dots,
                      const double* nuinv2, const double* a2, TF* f, double* partial, int* nblk,
                      const TF* dinv = nullptr, TF* fd = nullptr);
double host_conv_scale(double a2);
void launch_scale_to(hipStream_t s, size_t n, const double* a, const double* norm2, double* out);
template <typename TZ>
void launch_update_x(hipStream_t s, size_t n, double* x, const TZ* Z, size_t ldz, int nz, const double* y);
template <typename TA, typename TB>
void launch_convert(hipStream_t s, size_t n, const TA* a, TB* b);
template <typename TA, typename TB>
void launch_convert_scaled(hipStream_t s, size_t n, const TA* a, double scale, TB* b);
void launch_fill(hipStream_t s, size_t n, double* a, double v);
void launch_copy(hipStream_t s, size_t n, const double* a, double* b);
void launch_fillf(hipStream_t s, size_t n, float* a, float v);

int partial_blocks(size_t n);

// fused Krylov step: w = A z with w.V_l (l < nv) and w.w partials (nv+1 vectors)
template <typename TZ>
void launch_apply_dots(hipStream_t s, Dims d, const double* g, const uint32_t* code, const TZ* z, double* w,
                       const double* V, size_t ldv, int nv, double* partial, int* nblk);
// CGS2 second pass: w -= V h and w.V_l, w.w partials (nv+1 vectors, n/256 blocks)
void launch_axpy_dots(hipStream_t s, size_t n, double* w, const double* V, size_t ldv, int nv, const double* h,
                      double* partial, int* nblk);
// red[1] = sqrt(red[off+nv] - sum red[off+l]^2), red[2] = 1/red[1]
void launch_gs_norm(hipStream_t s, double* red, int off, int nv, double* host);
template <typename TF>
void launch_axpy_scale(hipStream_t s, size_t n, double* w, const double* V, size_t ldv, int nv, const double* h,
                       const double* inv, TF* f, const TF* dinv = nullptr, TF* fd = nullptr);
template <typename TF>
void launch_scale_conv(hipStream_t s, size_t n, const double* a, const double* nrm2, double* v, TF* f,
                       const TF* dinv = nullptr, TF* fd = nullptr);

// ---- multigrid ---------------------------------------------------------------------
// A multigrid level as the kernels see it.  The levels of the FP32 hierarchy keep
// their face coefficients in gh as FP16 scaled by a power of two (gh * gs are the
// coefficients); FP64 levels keep them in g.  FP16 halves the coefficient bytes,
// the largest share of a smoothing sweep's traffic.
template <typename T>
struct LevelView {
  Dims d;
  const T* g;
  const _Float16* gh;
  const uint32_t* code;
  const uint8_t* cls;  // CLS_NONFLUID (transfers) and CLS_INTERIOR (point smoother)
  T gs;
};

// Layout of the segmented y-line step (see line_solve_seg): lines per 1024-thread
// block, segments per line and rows per segment by precision; lines of up to
// line_segs * line_seg_rows rows take it, longer ones one thread per line.
template <typename T>
constexpr int line_lanes() {
  return sizeof(T) == 4 ? 64 : 32;
}
template <typename T>
constexpr int line_segs() {
  return 1024 / line_lanes<T>();
}
template <typename T>
constexpr int line_seg_rows() {
  return sizeof(T) == 4 ? 16 : 8;
}
template <typename T>
HD inline bool line_seg_fits(const Dims& d) {
  return d.ny <= line_segs<T>() * line_seg_rows<T>();
}

// Chebyshev smoothing launch modes (see smooth_point_cell); SM_POWER is one step
// with c1 = 0, c2 = 1 from f = 0 that stores D^-1 (f - A x) in dv (power iteration)
enum { SM_NEXT = 0, SM_ZERO = 1, SM_FIRST2 = 2, SM_POWER = 3 };

// Point levels keep the previous iterate in x_prev (x_new may be the same buffer:
// each cell reads its own x_prev before writing x_new); y-line levels keep the
// Chebyshev direction in dv and ignore x_prev.  fd (optional, point levels in
// SM_FIRST2): f * dinv as stored by the writer of f.
template <typename T>
void launch_smooth(hipStream_t s, LevelView<T> L, const T* lfac, int mode, const T* f, const T* dinv,
                   const T* x_old, T* x_prev, T* x_new, T* dv, T* scr, T c1, T c2, T c0, const T* fd = nullptr);
template <typename T>
void launch_line_factors(hipStream_t s, LevelView<T> L, T* lfac);
// streaming floor of a point sweep: reads every array the sweep reads once per
// cell (class byte, nine coefficient arrays, f, x, the previous iterate) and
// writes one, with no stencil
template <typename T>
void launch_stream_floor(hipStream_t s, LevelView<T> L, const T* f, const T* x, const T* x_prev, T* out);
template <typename T>
void launch_dinv(hipStream_t s, LevelView<T> L, T* dinv);
template <typename T>
void launch_resid_restrict(hipStream_t s, LevelView<T> F, const T* f, const T* x, Dims dc, const uint8_t* cls_c,
                           T* fc);
template <typename T>
void launch_prolong_add(hipStream_t s, LevelView<T> F, Dims dc, const uint8_t* cls_c, const T* xc, T* xf);
template <typename T>
void launch_dense_solve(hipStream_t s, int n, const T* cinv, const T* f, T* x);

// The bottom of a V-cycle in one 1024-thread workgroup (k_coarse_tail): a point
// level of at most TAIL_MAX_CELLS cells whose next level is the dense one --
// pre-smoothing from zero (degree Chebyshev steps, the first two fused when
// first2), residual and restriction, dense solve, prolongation, post-smoothing --
// with the level's iterates in LDS; the final x goes to xout.  Same results, bit
// for bit, as the separate launches.  FP32 levels only.
constexpr int TAIL_MAX_CELLS = 1024;
constexpr int TAIL_MAX_DEGREE = 8;
constexpr uint16_t TAIL_NO_CELL = 0xFFFF;
// Which cell each of the tail's 1024 threads smooths: cells whose code word is the
// interior one first (waves made only of them take the compile-time-weight flux),
// then the others, and the resulting 64-cell waves placed on the workgroup's waves
// (wave w runs on SIMD w % 4) so that every SIMD carries about the same
// arithmetic.  code holds the level's n code words; the result has 1024 entries.
std::vector<uint16_t> tail_order(size_t n, const uint32_t* code);
template <typename T>
struct TailArgs {
  LevelView<T> L;
  const T* f;
  const T* dinv;
  T* xout;
  Dims dc;              // the dense level
  const uint8_t* clsc;  // its class bytes
  const T* cinv;        // its inverse, column-major dc.n() x dc.n()
  const uint16_t* order;  // the cell thread t smooths (TAIL_NO_CELL: none); null: cell t
  int degree;
  int first2;
  T c1[TAIL_MAX_DEGREE], c2[TAIL_MAX_DEGREE];
};
template <typename T>
void launch_coarse_tail(hipStream_t s, const TailArgs<T>& a);

template <typename T>
void launch_normsq(hipStream_t s, size_t n, const T* a, double* partial, int* nblk);
// a = (src, or a itself when src is null) / sqrt(norm2[0]) (0 when norm2[0] is 0)
template <typename T>
void launch_scale_by(hipStream_t s, size_t n, T* a, const double* norm2, const T* src = nullptr);

// reference-layout transfer primitives (DMDA doubles), for verification
void launch_ref_restrict(hipStream_t s, Dmda f, Dmda c, const double* nvf, const double* nvc, const double* xf,
                         double* fc);
void launch_ref_prolong(hipStream_t s, Dmda f, Dmda c, const double* nvf, const double* nvc, const double* xc,
                        double* ff);
void launch_ref_nvert_restrict(hipStream_t s, Dmda f, Dmda c, const double* nvf, double* nvc, double* tmp);

}  // namespace pp

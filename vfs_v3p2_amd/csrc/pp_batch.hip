// Batched and served pressure projections: many independent cases (each with
// its own grid size, metrics, multigrid hierarchy and Krylov storage, held by
// its own vfs_pp_handle) advance together.
//
// * Every kernel launch covers all member cases: the launch's work list holds
//   one item per case (a level of its hierarchy, or its Krylov vectors) and a
//   prefix sum of the items' tile counts; persistent workers walk the combined
//   tile range, each XCD taking one contiguous eighth of it so that neighbouring
//   tiles of a case share that XCD's L2.
// * Each case's Krylov step j and an active flag live in a small device array,
//   so the launch sequence of one FGMRES iteration (the V-cycle of every case,
//   the operator with its Gram-Schmidt dots, then one pass that orthogonalises
//   the unnormalised new basis vector and writes the next V-cycle's input) never
//   changes while the membership stays the same: it is captured once as a HIP
//   graph and replayed every iteration.  Members that ask for CGS2 (orthog 0
//   or 1) run it in three passes instead.
// * The host keeps each case's Hessenberg column, Givens rotations and restart
//   logic (the same algorithm as the single-case engine) and handles cases that
//   start, restart or finish between iterations.
//
// The batch entry point solves a fixed set of cases; the server admits new
// requests between iterations (continuous batching) as slots and device memory
// free up.
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <tuple>

#include "pp_engine.hpp"
#include "pp_krylov_dev.hpp"
#include "pp_mg_dev.hpp"
#include "pp_path_dev.hpp"

namespace pp {

namespace {

// ------------------------------------------------------------------ device side
__device__ __forceinline__ int find_item(const uint32_t* __restrict__ start, int nitems, uint32_t t) {
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Fills the per-tile table of a work list laid out as [starts (nitems) | table
// (total) | slots (nitems)]: table[t] = item of tile t << 16 | that item's slot.
__global__ __launch_bounds__(256) void k_tile_table(uint32_t* __restrict__ ws, int nitems, uint32_t total) {
  const uint32_t t = blockIdx.x * 256u + threadIdx.x;
  if (t >= total) return;
  const int q = find_item(ws, nitems, t);
  ws[nitems + t] = ((uint32_t)q << 16) | ws[(size_t)nitems + total + q];
}

// Work lists: `start` holds the first tile of each of the nitems items, followed
// by one word per tile with the tile's item (high 16 bits) and that item's case
// slot (low 16 bits), so a tile finds its item and slot with one load.
//
// Persistent workers (gridDim.x a multiple of 8): worker w runs on XCD w % 8 and
// walks that XCD's contiguous eighth of the tile range [0, total).
#define FOR_TILES(total, t)                                                               \
  const uint32_t xcd_ = blockIdx.x & 7u, wper_ = gridDim.x >> 3;                         \
  const uint32_t chunk_ = ((total) + 7u) >> 3;                                            \
  for (uint32_t u_ = blockIdx.x >> 3, t = xcd_ * chunk_ + u_; u_ < chunk_ && t < (total); \
       u_ += wper_, t = xcd_ * chunk_ + u_)

// slot state: x = Krylov step j, y = active
using SlotDev = int2;

// Cells of an item are walked flat, 256 consecutive cells per tile (i fastest),
// so a level narrower than a wave still fills every lane.  Returns false past the
// last cell.
__device__ __forceinline__ bool flat_cell(const Dims& d, uint32_t lt, int& i, int& j, int& k) {
  const uint32_t e = lt * 256u + threadIdx.x + 64u * threadIdx.y;
  const uint32_t nxy = (uint32_t)d.nx * (uint32_t)d.ny;
  if (e >= nxy * (uint32_t)d.nz) return false;
  k = (int)(e / nxy);
  const uint32_t r = e - (uint32_t)k * nxy;
  j = (int)(r / (uint32_t)d.nx);
  i = (int)(r - (uint32_t)j * (uint32_t)d.nx);
  return true;
}

// Tile shapes of the cell walk, chosen per launch by the autotuner: 0 walks 256
// consecutive cells (i fastest), so levels narrower than a wave fill every lane;
// 1, 2, 3 walk 64 x 4, 32 x 8 and 16 x 16 (i x j) cell tiles of one k plane, whose
// waves hold one, two or four i-rows and share more stencil neighbours in cache.
constexpr int kTileShapes = 4;
HDI uint32_t shape_log2_tx(int shape) { return shape == 1 ? 6u : (shape == 2 ? 5u : 4u); }
HDI uint32_t shape_tiles(int shape, const Dims& d) {
  if (shape == 0) return (uint32_t)((d.n() + 255) / 256);
  const uint32_t lw = shape_log2_tx(shape), tx = 1u << lw, ty = 256u >> lw;
  return (((uint32_t)d.nx + tx - 1) >> lw) * (((uint32_t)d.ny + ty - 1) / ty) * (uint32_t)d.nz;
}
__device__ __forceinline__ bool shaped_cell(int shape, const Dims& d, uint32_t lt, int& i, int& j, int& k) {
  if (shape == 0) return flat_cell(d, lt, i, j, k);
  const uint32_t lw = shape_log2_tx(shape), tx = 1u << lw, ty = 256u >> lw;
  const uint32_t e = threadIdx.x + 64u * threadIdx.y;
  const uint32_t nti = ((uint32_t)d.nx + tx - 1) >> lw, ntj = ((uint32_t)d.ny + ty - 1) / ty;
  const uint32_t ti = lt % nti, r = lt / nti, tj = r % ntj;
  i = (int)(ti * tx + (e & (tx - 1)));
  j = (int)(tj * ty + (e >> lw));
  k = (int)(r / ntj);
  return i < d.nx && j < d.ny && k < d.nz;
}

// y-line (i, k) of an item, 256 consecutive lines per tile
__device__ __forceinline__ bool flat_line(const Dims& d, uint32_t lt, int& i, int& k) {
  const uint32_t e = lt * 256u + threadIdx.x + 64u * threadIdx.y;
  if (e >= (uint32_t)d.nx * (uint32_t)d.nz) return false;
  k = (int)(e / (uint32_t)d.nx);
  i = (int)(e - (uint32_t)k * (uint32_t)d.nx);
  return true;
}

template <typename T>
struct SmoothItem {
  LevelView<T> L;
  const T* f;
  const T* dinv;
  const T* xin;
  T* xout;
  T* xprev;  // point levels: the previous iterate (read and/or written; may be xout)
  T* dv;     // y-line levels: the Chebyshev direction
  T* out_base;  // non-null: the output is out_base + j * out_stride (the case's Z_j)
  size_t out_stride;
  const T* lfac;  // y-line levels: Thomas factors; scr holds the residual and the forward sweep
  T* scr;
  T c1, c2, c0;
  int slot;
  uint32_t tx, ty;  // point/residual: 64 x 4 cell tiles; line solve: 64 (i) x 4 (k) line tiles
};

template <typename T>
struct RestrictItem {
  LevelView<T> F;
  const T* f;
  const T* x;
  Dims dc;
  const uint8_t* clsc;
  T* fc;
  const T* r;  // general: the fine residual f - A x (written by a kb_resid launch before)
  int cmask;
  int slot;
  uint32_t tx, ty;  // blk: 64x2x2 fine tiles; general: 64x4x1 coarse tiles
};

template <typename T>
struct ProlongItem {
  LevelView<T> F;
  Dims dc;
  const uint8_t* clsc;
  const T* xc;
  T* xf;
  int cmask;
  int slot;
  uint32_t tx, ty;
};

template <typename T>
struct DenseItem {
  const T* cinv;
  const T* f;
  T* x;
  T* out_base;
  size_t out_stride;
  int n;
  int slot;
};

template <typename T>
struct ApplyItem {
  Dims d;
  const double* g;
  const uint32_t* code;
  const T* Z;
  double* V;
  size_t n;
  int slot;
  uint32_t tx, ty;
};

template <typename T>
struct VecItem {
  double* V;
  double* red;
  double* hpin;
  T* f0;
  size_t n;
  int slot;
};

struct FinItem {
  double* red;
  double* hpin;
  uint32_t t0, t1;
  int slot;
};

template <typename T, int MODE>
__global__ __launch_bounds__(256) void kb_smooth(const SmoothItem<T>* __restrict__ items,
                                                 const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                 const SlotDev* __restrict__ sd, int shape) {
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const SmoothItem<T>& I = items[q];
    const SlotDev st = sd[tw_ & 0xffffu];
    if (!st.y) continue;
    int i, j, k;
    if (!shaped_cell(shape, I.L.d, t - start[q], i, j, k)) continue;
    T* xout = I.out_base ? I.out_base + (size_t)st.x * I.out_stride : I.xout;
    smooth_point_cell<T, MODE>(I.L, I.f, I.dinv, I.xin, I.xprev, xout, I.dv, I.c1, I.c2, I.c0, i, j, k);
  }
}

// r = f - A x into the item's scratch (y-line levels)
template <typename T>
__global__ __launch_bounds__(256) void kb_resid(const SmoothItem<T>* __restrict__ items,
                                                const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                const SlotDev* __restrict__ sd, int shape) {
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const SmoothItem<T>& I = items[q];
    if (!sd[tw_ & 0xffffu].y) continue;
    int i, j, k;
    if (!shaped_cell(shape, I.L.d, t - start[q], i, j, k)) continue;
    resid_cell(I.L, I.f, I.xin, I.scr, i, j, k);
  }
}

// one Chebyshev step with the y-line preconditioner, one thread per (i, k) line
template <typename T, bool ZG>
__global__ __launch_bounds__(256) void kb_line(const SmoothItem<T>* __restrict__ items,
                                               const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                               const SlotDev* __restrict__ sd) {
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const SmoothItem<T>& I = items[q];
    const SlotDev st = sd[tw_ & 0xffffu];
    if (!st.y) continue;
    int i, k;
    if (!flat_line(I.L.d, t - start[q], i, k)) continue;
    T* xout = I.out_base ? I.out_base + (size_t)st.x * I.out_stride : I.xout;
    const size_t n = I.L.d.n();
    line_solve_col<T, ZG>(I.L.d, ZG ? I.f : I.scr, I.lfac, I.scr + n, I.xin, xout, I.dv, I.c1, I.c2, i, k);
  }
}

// the same step with each line split over the waves of a (64, 16) block,
// line_lanes<T>() consecutive lines (i fastest, then k) per tile, the residual of a step from a
// non-zero x computed in the same pass; every thread of the block takes part in
// the barriers of the segmented sweeps, lanes past the last line too
template <typename T, bool ZG>
__global__ __launch_bounds__(1024) void kb_line_seg(const SmoothItem<T>* __restrict__ items,
                                                              const uint32_t* __restrict__ start, int nitems,
                                                              uint32_t total, const SlotDev* __restrict__ sd) {
  __shared__ LineSegLds<T> sh;
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const SmoothItem<T>& I = items[q];
    const SlotDev st = sd[tw_ & 0xffffu];
    if (!st.y) continue;  // the whole block skips the tile
    const uint32_t nx = (uint32_t)I.L.d.nx;
    const uint32_t e = (t - start[q]) * (uint32_t)line_lanes<T>() + (uint32_t)line_lane<T>();
    const bool live = e < nx * (uint32_t)I.L.d.nz;
    const int k = live ? (int)(e / nx) : 0;
    const int i = live ? (int)(e - (uint32_t)k * nx) : 0;
    T* xout = I.out_base ? I.out_base + (size_t)st.x * I.out_stride : I.xout;
    if constexpr (ZG)
      line_solve_seg<T, true>(I.L.d, live, i, k, I.f, I.lfac, I.xin, xout, I.dv, I.c1, I.c2, sh);
    else
      line_step_cell<T>(I.L, live, i, k, I.f, I.lfac, I.xin, xout, I.dv, I.c1, I.c2, sh);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void kb_restrict_blk(const RestrictItem<T>* __restrict__ items,
                                                       const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                       const SlotDev* __restrict__ sd) {
  __shared__ T lds[2][2][32];
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const RestrictItem<T>& I = items[q];
    if (!sd[tw_ & 0xffffu].y) continue;
    const uint32_t lt = t - start[q];
    resid_restrict_blk(I.F, I.f, I.x, I.dc, I.clsc, I.fc, (int)(lt % I.tx), (int)((lt / I.tx) % I.ty),
                       (int)(lt / (I.tx * I.ty)), I.cmask == 5, lds);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void kb_restrict_gen(const RestrictItem<T>* __restrict__ items,
                                                       const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                       const SlotDev* __restrict__ sd, int shape) {
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const RestrictItem<T>& I = items[q];
    if (!sd[tw_ & 0xffffu].y) continue;
    int ic, jc, kc;
    if (!shaped_cell(shape, I.dc, t - start[q], ic, jc, kc)) continue;
    restrict_sum_cell(I.F, I.r, I.dc, I.clsc, I.cmask, I.fc, ic, jc, kc);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void kb_prolong(const ProlongItem<T>* __restrict__ items,
                                                  const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                  const SlotDev* __restrict__ sd, int shape) {
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const ProlongItem<T>& I = items[q];
    if (!sd[tw_ & 0xffffu].y) continue;
    int i, j, k;
    if (!shaped_cell(shape, I.F.d, t - start[q], i, j, k)) continue;
    prolong_add_cell(I.F, I.dc, I.clsc, I.cmask, I.xc, I.xf, i, j, k);
  }
}

template <typename T>
__global__ __launch_bounds__(256) void kb_dense(const DenseItem<T>* __restrict__ items,
                                                const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                const SlotDev* __restrict__ sd) {
  __shared__ T fs[1024];
  __shared__ T part[4][64];
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const DenseItem<T>& I = items[q];
    const SlotDev st = sd[tw_ & 0xffffu];
    if (!st.y) continue;
    T* x = I.out_base ? I.out_base + (size_t)st.x * I.out_stride : I.x;
    dense_rows(I.n, I.cinv, I.f, x, (int)(t - start[q]), fs, part);
  }
}

// w = A z_j (FP64 flux form) into V_{j+1}; per-wave partial sums of w.V_l (l <= j)
// and w.w at partial[l * 4 total + 4 t + wave]
template <typename T>
__global__ __launch_bounds__(256) void kb_apply_dots(const ApplyItem<T>* __restrict__ items,
                                                     const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                     const SlotDev* __restrict__ sd, double* __restrict__ partial) {
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const ApplyItem<T>& I = items[q];
    const SlotDev st = sd[tw_ & 0xffffu];
    if (!st.y) continue;
    int i = 0, j = 0, k = 0;
    const bool in = flat_cell(I.d, t - start[q], i, j, k);
    const T* z = I.Z + (size_t)st.x * I.n;
    double* w = I.V + (size_t)(st.x + 1) * I.n;
    const size_t c = in ? I.d.idx(i, j, k) : 0;
    double v8[8];
    wave_dots_prefetch(in, c, I.V, I.n, st.x + 1, v8);
    const double wv = in ? apply_cell(I.d, I.g, I.code, z, w, i, j, k) : 0.;
    wave_dots(in, c, wv, I.V, I.n, st.x + 1, v8, partial, 4 * (size_t)total, 4 * (size_t)t + threadIdx.y);
  }
}

// w1 = w - V h1 in place over V_{j+1} (h1 = red[8..]); partial sums of w1.V_l, w1.w1
template <typename T>
__global__ __launch_bounds__(256) void kb_axpy_dots(const VecItem<T>* __restrict__ items,
                                                    const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                    const SlotDev* __restrict__ sd, double* __restrict__ partial) {
  __shared__ double lds[4][64];
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const VecItem<T>& I = items[q];
    const SlotDev st = sd[tw_ & 0xffffu];
    if (!st.y) continue;
    const size_t e = (size_t)(t - start[q]) * 256 + threadIdx.x;
    const bool in = e < I.n;
    double* w = I.V + (size_t)(st.x + 1) * I.n;
    if (st.x + 1 < 16) {
      gs_sub_block_dots<16>(in, w, I.V, I.n, st.x + 1, I.red + 8, e, partial, total, t, lds);
    } else {
      const double w1 = in ? gs_sub(w, I.V, I.n, st.x + 1, I.red + 8, e) : 0.;
      block_dots(in, e, w1, I.V, I.n, st.x + 1, partial, total, t, lds);
    }
  }
}

// v_{j+1} = (w1 - V h2) / ||.|| in place; f0 = (T)v_{j+1}, the next V-cycle's input
template <typename T>
__global__ __launch_bounds__(256) void kb_axpy_scale(const VecItem<T>* __restrict__ items,
                                                     const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                     const SlotDev* __restrict__ sd) {
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const VecItem<T>& I = items[q];
    const SlotDev st = sd[tw_ & 0xffffu];
    if (!st.y) continue;
    const size_t e = (size_t)(t - start[q]) * 256 + threadIdx.x;
    if (e >= I.n) continue;
    axpy_scale_one(I.V + (size_t)(st.x + 1) * I.n, I.V, I.n, st.x + 1, I.red + 72, I.red[2], I.f0, e);
  }
}

// Fixed-order sum of the tile partials pv[t0, t1) by one 256-thread block
__device__ __forceinline__ double tile_sum(const double* __restrict__ pv, uint32_t t0, uint32_t t1, double* lds) {
  double a0 = 0., a1 = 0.;
  uint32_t b = t0 + threadIdx.x;
  for (; b + 256 < t1; b += 512) { a0 += pv[b]; a1 += pv[b + 256]; }
  for (; b < t1; b += 256) a0 += pv[b];
  return block_sum(a0 + a1, lds);
}

// The same for one 1024-thread block (the per-case finalizes, whose largest case
// sets the launch's length): wave sums, then the 16 wave results in a fixed tree
__device__ __forceinline__ double tile_sum_wide(const double* __restrict__ pv, uint32_t t0, uint32_t t1,
                                                double* lds /* >= 16 */) {
  double a0 = 0., a1 = 0.;
  uint32_t b = t0 + threadIdx.x;
  for (; b + 1024 < t1; b += 2048) { a0 += pv[b]; a1 += pv[b + 1024]; }
  for (; b < t1; b += 1024) a0 += pv[b];
  const double v = wave_sum(a0 + a1);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  double r = 0.;
  if (threadIdx.x == 0) {
    double q[4];
#pragma unroll
    for (int g = 0; g < 4; g++) q[g] = (lds[4 * g] + lds[4 * g + 1]) + (lds[4 * g + 2] + lds[4 * g + 3]);
    r = (q[0] + q[1]) + (q[2] + q[3]);
  }
  __syncthreads();
  return r;
}

// Lazily normalised Gram-Schmidt pass over U_{j+1} = w in place: w' = w - sum_l c_l U_l
// with c_l = red[8 + l] red[RED_NUINV2 + l] (w.U_l / |U_l|^2), the next V-cycle's
// input f0 = (T)(w' s) with s = conv_scale(|w|^2 = red[8 + j + 1]), and per-wave
// partial sums of |w'|^2 at partial[4 t + wave] (no LDS and no barrier per tile).
// The same per-element arithmetic as k_axpy_norm.
template <typename T>
__global__ __launch_bounds__(256) void kb_axpy_norm(const VecItem<T>* __restrict__ items,
                                                    const uint32_t* __restrict__ start, int nitems, uint32_t total,
                                                    const SlotDev* __restrict__ sd, double* __restrict__ partial) {
  FOR_TILES(total, t) {
    const uint32_t tw_ = start[nitems + t];
    const int q = (int)(tw_ >> 16);
    const VecItem<T>& I = items[q];
    const SlotDev st = sd[tw_ & 0xffffu];
    if (!st.y) continue;
    const int nv = st.x + 1;
    const size_t e = (size_t)(t - start[q]) * 256 + threadIdx.x;
    double acc[1] = {0.};
    if (e < I.n) {
      double* w = I.V + (size_t)nv * I.n;
      double v = w[e];
      for (int l = 0; l < nv; l++) v -= (I.red[8 + l] * I.red[RED_NUINV2 + l]) * I.V[(size_t)l * I.n + e];
      w[e] = v;
      I.f0[e] = (T)(v * conv_scale(I.red[8 + nv]));
      acc[0] = v * v;
    }
    const double ws = wave_sum_to63(acc[0]);
    if ((threadIdx.x & 63) == 63) partial[4 * (size_t)t + (threadIdx.x >> 6)] = ws;
  }
}

// One block per (case, vector l < j + 1 + extra): fixed-order sum of the case's
// partials (per tiles [t0, t1), `per` of them per tile) into red[off + l],
// mirrored to pinned host memory.
__global__ __launch_bounds__(1024) void kb_finalize(const FinItem* __restrict__ items, const double* __restrict__ partial,
                                                    uint32_t total, int per, int off, int extra,
                                                    const SlotDev* __restrict__ sd) {
  __shared__ double lds[16];
  const FinItem& I = items[blockIdx.x];
  const SlotDev st = sd[I.slot];
  const int l = blockIdx.y;
  if (!st.y || l >= st.x + 1 + extra) return;
  const uint32_t pt = (uint32_t)per * total;
  const double s = tile_sum_wide(partial + (size_t)l * pt, per * I.t0, per * I.t1, lds);
  if (threadIdx.x == 0) {
    I.red[off + l] = s;
    I.hpin[off + l] = s;
  }
}

// One block per case: |U_{j+1}|^2 from the partials of kb_axpy_norm into red[0]
// (mirrored) and its reciprocal into red[RED_NUINV2 + j + 1] for the next step.
__global__ __launch_bounds__(1024) void kb_finalize_norm(const FinItem* __restrict__ items,
                                                         const double* __restrict__ partial,
                                                         const SlotDev* __restrict__ sd) {
  __shared__ double lds[16];
  const FinItem& I = items[blockIdx.x];
  const SlotDev st = sd[I.slot];
  if (!st.y) return;
  const double s = tile_sum_wide(partial, 4 * I.t0, 4 * I.t1, lds);  // kb_axpy_norm's per-wave partials
  if (threadIdx.x == 0) {
    I.red[0] = s;
    I.hpin[0] = s;
    I.red[RED_NUINV2 + st.x + 1] = s > 0. ? 1. / s : 0.;
  }
}

__global__ void kb_gs_norm(const FinItem* __restrict__ items, int nitems, const SlotDev* __restrict__ sd) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nitems) return;
  const FinItem& I = items[q];
  const SlotDev st = sd[I.slot];
  if (!st.y) return;
  gs_norm_one(I.red, 72, st.x + 1, I.hpin);
}

// ------------------------------------------------------ batched per-case phases
// The per-case work around the Krylov iterations (VolumeFlux and RHS when a case
// starts, its initial guess, the solution update and true residual at each
// cycle start, and the null space, Phi, P, Projection, Contra2Cart and
// Divergence when it finishes) runs as one launch per step over every case that
// takes the step, 256 elements of one case per tile, instead of one launch per
// case and step.
enum PhaseOp {
  PH_IBSUM, PH_IBCORR, PH_RHS, PH_GATHER, PH_ZERO, PH_UPDATE_X, PH_RESID, PH_DIVEST, PH_SCALE_CONV,
  PH_MASKED_SUM, PH_SUB_MEAN, PH_SCATTER, PH_PERIODIC, PH_ADD_PHI, PH_PROJECTION, PH_UCONT_GHOSTS,
  PH_C2C_CORE, PH_VEC_GHOSTS, PH_C2C_BC, PH_C2C_SOLID, PH_DIVERGENCE
};

// One case's operands for one phase step.  Roles of the generic fields:
//   IBSUM        w0 Ucont, faces, i0 zero_small          -> [flux, area]
//   IBCORR       w0 Ucont, faces, i0 immersed, sc [flux, area]
//   RHS          r0 Ucont, w0 b, s0 factor                -> |b|^2
//   GATHER       r0 Phi, w0 x          ZERO  w0 x
//   UPDATE_X     w0 x, z Z (ld, i0 vectors), sc y
//   RESID        r0 b, r1 x, w0 r                         -> |r|^2
//   DIVEST       r0 aj, r1 r                              -> max |aj r|
//   SCALE_CONV   r0 r, w0 V_0, f f0, sc |r|^2, one (1 / |U_0|^2 of a lazy basis)
//   MASKED_SUM   r0 x                                     -> sum over nvert < 0.1
//   SUB_MEAN     w0 x, sc sum, s0 count
//   SCATTER      r0 x, w0 Phi          ADD_PHI  r0 Phi, w0 P
//   PERIODIC     w0 array, i0 zero_other
//   PROJECTION   r0 Phi, w0 Ucont, s0 factor
//   UCONT_GHOSTS w0 Ucont
//   C2C_CORE     r0 cell metrics, r1 Ucont, w0 Ucat
//   VEC_GHOSTS   w0 Ucat, w1 snapshot or null
//   C2C_BC       r0 nvert, r1 snapshot, w0 Ucat, bc
//   C2C_SOLID    w0 Ucat
//   DIVERGENCE   r0 aj, r1 Ucont                          -> max |aj div u|
struct PhaseItem {
  Dmda m;
  Dims d;
  const double* g;
  const uint32_t* code;
  const uint8_t* cls;
  const double* r0;
  const double* r1;
  double* w0;
  double* w1;
  const void* z;
  void* f;
  const IbFace* faces;
  const double* sc;
  double* one;
  double s0;
  size_t ld;
  uint32_t n;  // elements: cells, vector entries, faces or boundary points
  int i0;
  int bc[6];
};

// reduction of one case's tile partials [t0, t1) into out[v] (and the pinned host mirror)
struct PhaseFin {
  double* out;
  double* host;
  uint32_t t0, t1;
};

__device__ __forceinline__ double tile_max(const double* __restrict__ pv, uint32_t t0, uint32_t t1, double* lds) {
  double a0 = 0., a1 = 0.;
  uint32_t b = t0 + threadIdx.x;
  for (; b + 256 < t1; b += 512) { a0 = fmax(a0, pv[b]); a1 = fmax(a1, pv[b + 256]); }
  for (; b < t1; b += 256) a0 = fmax(a0, pv[b]);
  return block_max(fmax(a0, a1), lds);
}

template <int OP, typename T>
__global__ __launch_bounds__(256) void kb_phase(const PhaseItem* __restrict__ items, const uint32_t* __restrict__ start,
                                                int nitems, uint32_t total, double* __restrict__ partial) {
  __shared__ double lds[4];
  FOR_TILES(total, t) {
    const int q = find_item(start, nitems, t);
    const PhaseItem& I = items[q];
    const uint32_t lt = t - start[q];
    const uint32_t e = lt * 256u + threadIdx.x;
    const bool in = e < I.n;
    int i = 0, j = 0, k = 0;
    if constexpr (OP == PH_IBSUM) {
      double fl = 0., ar = 0.;
      if (in) ibflux_sum_face(I.faces[e], I.w0, I.i0, fl, ar);
      fl = block_sum(fl, lds);
      ar = block_sum(ar, lds);
      if (threadIdx.x == 0) { partial[t] = fl; partial[total + t] = ar; }
    } else if constexpr (OP == PH_IBCORR) {
      if (in) ibflux_correct_face(I.faces[e], I.w0, I.i0, I.sc);
    } else if constexpr (OP == PH_RHS) {
      double v = 0.;
      if (flat_cell(I.d, lt, i, j, k)) v = rhs_cell(I.m, I.d, I.r0, I.cls, I.s0, I.w0, i, j, k);
      v = block_sum(v, lds);
      if (threadIdx.x == 0) partial[t] = v;
    } else if constexpr (OP == PH_GATHER) {
      if (flat_cell(I.d, lt, i, j, k)) gather_phi_cell(I.m, I.d, I.r0, I.cls, I.w0, i, j, k);
    } else if constexpr (OP == PH_ZERO) {
      if (in) I.w0[e] = 0.;
    } else if constexpr (OP == PH_UPDATE_X) {
      if (in) update_x_elem(I.w0, static_cast<const T*>(I.z), I.ld, I.i0, I.sc, e);
    } else if constexpr (OP == PH_RESID) {
      double v = 0.;
      if (flat_cell(I.d, lt, i, j, k)) v = residual_cell<double>(I.d, I.g, I.code, I.r0, I.r1, I.w0, i, j, k);
      v = block_sum(v, lds);
      if (threadIdx.x == 0) partial[t] = v;
    } else if constexpr (OP == PH_DIVEST) {
      double v = in ? divest_elem(I.code, I.r0, I.r1, e) : 0.;
      v = block_max(v, lds);
      if (threadIdx.x == 0) partial[t] = v;
    } else if constexpr (OP == PH_SCALE_CONV) {
      if (in) scale_conv_elem(I.r0, I.sc[0], I.w0, static_cast<T*>(I.f), e);
      if (I.one && e == 0) I.one[0] = 1.;
    } else if constexpr (OP == PH_MASKED_SUM) {
      double v = (in && (I.cls[e] & CLS_LT)) ? I.r0[e] : 0.;
      v = block_sum(v, lds);
      if (threadIdx.x == 0) partial[t] = v;
    } else if constexpr (OP == PH_SUB_MEAN) {
      if (in) {
        const double val = -I.sc[0] / I.s0;  // poisson_hypre.c:256
        if (I.cls[e] & CLS_LT) I.w0[e] += val;
      }
    } else if constexpr (OP == PH_SCATTER) {
      if (flat_cell(I.d, lt, i, j, k)) scatter_cell(I.m, I.d, I.r0, I.cls, I.w0, nullptr, i, j, k);
    } else if constexpr (OP == PH_PERIODIC) {
      if (in && boundary_point(I.m, e, i, j, k)) periodic_point(I.m, I.w0, I.i0, i, j, k);
    } else if constexpr (OP == PH_ADD_PHI) {
      if (flat_cell(I.d, lt, i, j, k)) add_phi_cell(I.m, I.d, I.cls, I.r0, I.w0, i, j, k);
    } else if constexpr (OP == PH_PROJECTION) {
      if (flat_cell(I.d, lt, i, j, k)) projection_cell(I.m, I.d, I.g, I.code, I.cls, I.r0, I.w0, I.s0, i, j, k);
    } else if constexpr (OP == PH_UCONT_GHOSTS) {
      if (in && boundary_point(I.m, e, i, j, k)) ucont_ghost_point(I.m, I.w0, i, j, k);
    } else if constexpr (OP == PH_C2C_CORE) {
      if (flat_cell(I.d, lt, i, j, k)) c2c_core_cell(I.m, I.d, I.r0, I.cls, I.r1, I.w0, i, j, k);
    } else if constexpr (OP == PH_VEC_GHOSTS) {
      if (in && boundary_point(I.m, e, i, j, k)) vec_ghost_point(I.m, I.w0, I.w1, i, j, k);
    } else if constexpr (OP == PH_C2C_BC) {
      if (in && boundary_point(I.m, e, i, j, k)) c2c_bc_point(I.m, I.r0, I.r1, I.w0, I.bc, i, j, k);
    } else if constexpr (OP == PH_C2C_SOLID) {
      if (flat_cell(I.d, lt, i, j, k)) c2c_solid_cell(I.m, I.d, I.cls, I.w0, i, j, k);
    } else if constexpr (OP == PH_DIVERGENCE) {
      double v = 0.;
      if (flat_cell(I.d, lt, i, j, k)) v = divergence_cell(I.m, I.d, I.code, I.r0, I.r1, i, j, k);
      v = block_max(v, lds);
      if (threadIdx.x == 0) partial[t] = v;
    }
  }
}

// One block per (case, value v): the case's tile partials summed (or maxed) in a
// fixed order into out[v], mirrored to pinned host memory.
template <bool MAX>
__global__ __launch_bounds__(256) void kb_phase_fin(const PhaseFin* __restrict__ fins,
                                                    const double* __restrict__ partial, uint32_t total) {
  __shared__ double lds[4];
  const PhaseFin& F = fins[blockIdx.x];
  const int v = blockIdx.y;
  const double* pv = partial + (size_t)v * total;
  const double s = MAX ? tile_max(pv, F.t0, F.t1, lds) : tile_sum(pv, F.t0, F.t1, lds);
  if (threadIdx.x == 0) {
    F.out[v] = s;
    if (F.host) F.host[v] = s;
  }
}

// workers for a launch of `total` tiles: a multiple of 8, at most `cap` (0: one
// tile per worker)
inline unsigned workers(uint32_t total, unsigned cap) {
  uint32_t w = (total + 7u) & ~7u;
  if (w < 8u) w = 8u;
  if (cap && w > cap) w = cap;
  return w;
}

// Worker counts and cell-tile shapes chosen by measurement, per (launch kind and
// precision, mode, log2 of the flat tile count); shared by every batch and server
// of the process.
struct TuneCache {
  std::mutex mu;
  std::map<std::tuple<int, int, int>, std::pair<unsigned, int>> best;  // (worker cap, tile shape)
};
TuneCache& tune_cache() {
  static TuneCache c;
  return c;
}
int log2_bucket(uint32_t x) {
  int b = 0;
  while (x > 1) { x >>= 1; b++; }
  return b;
}

}  // namespace

// ------------------------------------------------------------------- host side
template <typename T>
struct BatchCore {
  enum State { FREE = 0, START, ITER, FINISH, DONE };
  struct Slot {
    vfs_pp_handle* h = nullptr;
    MG<T>* mg = nullptr;
    hipStream_t own_stream = nullptr;
    State state = FREE;
    int ticket = -1;
    double *ucont = nullptr, *phi = nullptr, *p = nullptr, *ucat = nullptr;
    double dt = 0., st = 0., coeff = 0.;
    vfs_pp_stats S{};
    double t0 = 0.;
    int j = 0, it = 0, vc = 0, m = 0;
    double tgt = 0., rel = 1.;
    bool conv = false, pending_update = false, fresh = false;
    bool x_zero = false;  // x is 0 (a zero-start case before its first cycle): r = b
    int pending_nv = 0;
    std::vector<double> H, cs, sn, gv, y;
    std::vector<double> nu, sigma;  // lazily normalised basis: |U_l| and the scale of Z_l
  };

  hipStream_t s = nullptr;
  std::vector<Slot> slots;
  Arena fixed;            // slot state, host-mirrored
  SlotDev* sd = nullptr;  // device slot state
  SlotDev* sd_host = nullptr;
  Arena plan;  // work lists of the current membership
  double* partial = nullptr;
  size_t partial_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  bool dirty = true;
  int degree = 0;
  bool use_graph = true;

  struct Launch {
    int kind, mode;
    const void* items;
    const uint32_t* start;
    int nitems;
    uint32_t total;
    unsigned cap = 2048;  // worker count limit (0: one tile per worker), set by tune()
    // cell-tiled launches: the tile shape (see shaped_cell), each item's tile count
    // per shape and its slot, from which retile() rebuilds the work list
    int shape = 0;
    bool shaped = false;
    uint32_t total0 = 0;  // tiles of the flat walk (the tuner's size bucket)
    std::vector<uint32_t> tiles[kTileShapes];
    std::vector<uint32_t> item_slot;
  };
  std::vector<Launch> seq;  // V-cycle + Krylov passes of one iteration
  Launch apply{}, axd{}, axs{}, axn{};
  bool lazy = true;  // every member orthogonalises with the lazily normalised single pass
  const FinItem* fin = nullptr;
  int nfin = 0;
  uint32_t fin_total_apply = 0, fin_total_vec = 0;
  FinItem* fin_apply = nullptr;
  FinItem* fin_vec = nullptr;

  enum { K_SMOOTH, K_RBLK, K_RGEN, K_PROLONG, K_DENSE, K_RESID, K_LINE, K_APPLY, K_AXD, K_AXS, K_AXN, K_LSEG };
  bool tuned_once = false;
  int forced_shape = -1;  // >= 0: every cell-tiled launch walks this tile shape (no shape tuning)
  // batched phase steps: pinned staging and device copy of their work lists, and
  // the tile partials of their reductions
  char* ph_host = nullptr;
  char* ph_dev = nullptr;
  size_t ph_cap = 0;
  double* ph_partial = nullptr;
  size_t ph_partial_cap = 0;
  hipEvent_t ph_ev = nullptr;  // after the last upload from ph_host
  bool ph_recorded = false;

  explicit BatchCore(int nslots) {
    HIPCHECK(hipStreamCreateWithFlags(&s, hipStreamDefault));  // blocking: ordered with the legacy default stream
    slots.resize(nslots);
    sd = fixed.alloc<SlotDev>(nslots);
    HIPCHECK(hipHostMalloc(&sd_host, sizeof(SlotDev) * nslots));
    std::memset(sd_host, 0, sizeof(SlotDev) * nslots);
    HIPCHECK(hipEventCreateWithFlags(&ph_ev, hipEventDisableTiming));
  }
  ~BatchCore() {
    for (auto& sl : slots)
      if (sl.h) release(sl);
    drop_graph();
    if (s) (void)hipStreamSynchronize(s);
    if (sd_host) (void)hipHostFree(sd_host);
    if (ph_host) (void)hipHostFree(ph_host);
    if (ph_dev) (void)hipFree(ph_dev);
    if (ph_partial) (void)hipFree(ph_partial);
    if (ph_ev) (void)hipEventDestroy(ph_ev);
    if (s) (void)hipStreamDestroy(s);
  }

  void drop_graph() {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
  }

  // a case joins slot q: its launches go to the batch stream while it is here
  void bind(int q, vfs_pp_handle* h) {
    Slot& sl = slots[q];
    ensure_geom(h);
    ensure_unbound(h);
    auto* mg = dynamic_cast<MG<T>*>(h->mg.get());
    if (!mg) throw StatusError(VFS_PP_EINVAL, "batch members must share the multigrid precision");
    if (degree == 0) degree = mg->degree;
    if (mg->degree != degree) throw StatusError(VFS_PP_EINVAL, "batch members must share cheb_degree");
    sl.h = h;
    sl.mg = mg;
    sl.own_stream = h->s;
    h->s = s;
    h->batch_refs++;
    sl.state = DONE;
    dirty = true;
  }
  void release(Slot& sl) {
    if (sl.h) {
      sl.h->s = sl.own_stream;
      sl.h->batch_refs--;
    }
    sl = Slot{};
    dirty = true;
  }

  // ---------------------------------------------------------------- the plan
  // Items, their tile starts and the per-tile (item, slot) table on the device.
  // The table (one word per tile) is written by a kernel from the starts and the
  // items' slots, so a membership change costs no host work per tile.
  template <typename Item>
  std::pair<const Item*, const uint32_t*> upload(const std::vector<Item>& items, const std::vector<uint32_t>& start,
                                                 uint32_t total) {
    const size_t ni = items.size();
    if (ni == 0 || ni > 0xffffu || slots.size() > 0xffffu) throw StatusError(VFS_PP_EINVAL, "work list too long");
    Item* di = plan.alloc<Item>(ni);
    HIPCHECK(hipMemcpyAsync(di, items.data(), ni * sizeof(Item), hipMemcpyHostToDevice, s));
    // [starts (ni) | table (total) | slots (ni)]
    std::vector<uint32_t> sl(ni);
    for (size_t q = 0; q < ni; q++) sl[q] = (uint32_t)items[q].slot;
    uint32_t* ds = plan.alloc<uint32_t>(2 * ni + total);
    HIPCHECK(hipMemcpyAsync(ds, start.data(), ni * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(ds + ni + total, sl.data(), ni * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (total) {
      k_tile_table<<<(total + 255) / 256, 256, 0, s>>>(ds, (int)ni, total);
      HIPCHECK(hipGetLastError());
    }
    return {di, ds};
  }
  template <typename Item>
  void add_launch(int kind, int mode, const std::vector<Item>& items, const std::vector<uint32_t>& ntiles) {
    if (items.empty()) return;
    std::vector<uint32_t> start(items.size());
    uint32_t tot = 0;
    for (size_t q = 0; q < items.size(); q++) { start[q] = tot; tot += ntiles[q]; }
    if (tot == 0) return;
    auto pr = upload(items, start, tot);
    seq.push_back(Launch{kind, mode, pr.first, pr.second, (int)items.size(), tot, 2048});
  }

  static uint32_t cell_tiles(const Dims& d) { return (uint32_t)((d.n() + 255) / 256); }
  static uint32_t line_tiles(const Dims& d) { return (uint32_t)(((size_t)d.nx * d.nz + 255) / 256); }
  static uint32_t seg_line_tiles(const Dims& d) {
    return (uint32_t)(((size_t)d.nx * d.nz + line_lanes<T>() - 1) / line_lanes<T>());
  }
  static bool seg_lines(const Dims& d) { return line_seg_fits<T>(d); }

  // Launches of one smoothing step k (0-based) at one depth for a group of
  // members: point levels fuse the first two steps from a zero start; y-line
  // levels run a residual launch and a line-solve launch per step (the zero-start
  // step needs no residual).  cur[] tracks each member's current x buffer.
  void smoothing_step(int d, int k, bool zero_start, bool final_level0, const std::vector<int>& grp,
                      std::vector<std::vector<T*>>& cur) {
    std::vector<SmoothItem<T>> pt, rs, lz, lnx, sz, snx;
    std::vector<uint32_t> npt, nrs, nlz, nlnx, nsz, nsnx;
    int pmode = -1;
    for (int q : grp) {
      auto& L = slots[q].mg->lev[d];
      const bool line = L.lfac != nullptr;
      int mode, last = k;
      if (zero_start && k == 0) {
        mode = (!line && degree >= 2) ? SM_FIRST2 : SM_ZERO;
        if (mode == SM_FIRST2) last = 1;
      } else if (zero_start && k == 1 && !line && degree >= 2) {
        continue;  // done by the fused first launch
      } else {
        mode = SM_NEXT;
      }
      T* xin = (zero_start && k == 0) ? L.xa : cur[q][d];
      T* xout = (xin == L.xa) ? L.xb : L.xa;
      const bool fin = final_level0 && d == 0 && last == degree - 1;
      const T c1 = (mode == SM_ZERO || last == 0) ? T(0) : L.c1[last];
      SmoothItem<T> it{};
      it.L = L.view();
      it.f = L.f;
      it.dinv = L.dinv;
      it.xin = xin;
      it.xout = xout;
      it.xprev = (mode == SM_NEXT) ? xout : xin;
      it.dv = L.dv;
      it.out_base = fin ? static_cast<T*>(slots[q].h->Z) : nullptr;
      it.out_stride = L.d.n();
      it.lfac = L.lfac;
      it.scr = slots[q].mg->scr;
      it.c1 = c1;
      it.c2 = L.c2[last];
      it.c0 = L.c2[0];
      it.slot = q;
      cur[q][d] = xout;
      if (!line) {
        it.tx = (uint32_t)(L.d.nx + 63) / 64;
        it.ty = (uint32_t)(L.d.ny + 3) / 4;
        pmode = mode;
        pt.push_back(it);
        npt.push_back(cell_tiles(L.d));
      } else {
        // long lines: a residual launch, then one thread per line; lines that fit
        // the segmented layout take their residual in the line-step pass itself
        if (mode != SM_ZERO && !seg_lines(L.d)) {
          SmoothItem<T> r = it;
          r.tx = (uint32_t)(L.d.nx + 63) / 64;
          r.ty = (uint32_t)(L.d.ny + 3) / 4;
          rs.push_back(r);
          nrs.push_back(cell_tiles(L.d));
        }
        it.tx = (uint32_t)(L.d.nx + 63) / 64;
        it.ty = 0;
        if (seg_lines(L.d)) {
          if (mode == SM_ZERO) { sz.push_back(it); nsz.push_back(seg_line_tiles(L.d)); }
          else { snx.push_back(it); nsnx.push_back(seg_line_tiles(L.d)); }
        } else if (mode == SM_ZERO) {
          lz.push_back(it);
          nlz.push_back(line_tiles(L.d));
        } else {
          lnx.push_back(it);
          nlnx.push_back(line_tiles(L.d));
        }
      }
    }
    add_launch_cells(K_SMOOTH, pmode, pt);
    add_launch_cells(K_RESID, 0, rs);
    add_launch(K_LINE, SM_ZERO, lz, nlz);
    add_launch(K_LINE, SM_NEXT, lnx, nlnx);
    add_launch(K_LSEG, SM_ZERO, sz, nsz);
    add_launch(K_LSEG, SM_NEXT, snx, nsnx);
  }

  void build_plan() {
    HIPCHECK(hipStreamSynchronize(s));
    drop_graph();
    plan.release();
    seq.clear();
    std::vector<int> mem;
    for (int q = 0; q < (int)slots.size(); q++)
      if (slots[q].h) mem.push_back(q);
    if (mem.empty()) { dirty = false; return; }
    lazy = true;
    for (int q : mem) lazy = lazy && slots[q].h->opts.orthog == 2;
    int dmax = 0;
    for (int q : mem) dmax = std::max(dmax, (int)slots[q].mg->lev.size());
    // current x buffer per (member, level)
    std::vector<std::vector<T*>> cur(slots.size());
    for (int q : mem) cur[q].assign(slots[q].mg->lev.size(), nullptr);
    auto zbase = [&](int q) { return static_cast<T*>(slots[q].h->Z); };
    // descent
    for (int d = 0; d < dmax; d++) {
      std::vector<int> sm, dn;
      for (int q : mem) {
        const int nl = (int)slots[q].mg->lev.size();
        if (d < nl - 1) sm.push_back(q);
        else if (d == nl - 1) dn.push_back(q);
      }
      if (!sm.empty()) {
        for (int k = 0; k < degree; k++) smoothing_step(d, k, true, false, sm, cur);
        // residual + restriction to level d+1
        std::vector<RestrictItem<T>> rb, rg;
        std::vector<SmoothItem<T>> rr;  // fine residuals of the general restriction
        std::vector<uint32_t> nb, ng;
        for (int q : sm) {
          auto& L = slots[q].mg->lev[d];
          auto& C = slots[q].mg->lev[d + 1];
          RestrictItem<T> it{};
          it.F = L.view();
          it.f = L.f;
          it.x = cur[q][d];
          it.dc = C.d;
          it.clsc = C.cls;
          it.fc = C.f;
          it.cmask = (C.d.nx != L.d.nx ? 1 : 0) | (C.d.ny != L.d.ny ? 2 : 0) | (C.d.nz != L.d.nz ? 4 : 0);
          it.slot = q;
          if ((it.cmask == 7 || it.cmask == 5) && L.d.nx >= 64) {  // narrower levels: flat cells fill the lanes better
            it.tx = (uint32_t)(L.d.nx + 63) / 64;
            it.ty = (uint32_t)(L.d.ny + 1) / 2;
            rb.push_back(it);
            nb.push_back(it.tx * it.ty * (uint32_t)((L.d.nz + 1) / 2));
          } else {
            // the fine residual goes to the level's direction buffer, which no step
            // reads between the end of the pre-smoothing and the first
            // post-smoothing step (a new Chebyshev sequence, c1 = 0)
            it.tx = (uint32_t)(C.d.nx + 63) / 64;
            it.ty = (uint32_t)(C.d.ny + 3) / 4;
            it.r = L.dv;
            rg.push_back(it);
            ng.push_back(cell_tiles(C.d));
            SmoothItem<T> ri{};
            ri.L = L.view();
            ri.f = L.f;
            ri.xin = cur[q][d];
            ri.scr = L.dv;
            ri.slot = q;
            ri.tx = (uint32_t)(L.d.nx + 63) / 64;
            ri.ty = (uint32_t)(L.d.ny + 3) / 4;
            rr.push_back(ri);
          }
        }
        add_launch(K_RBLK, 0, rb, nb);
        add_launch_cells(K_RESID, 0, rr);
        add_launch_cells(K_RGEN, 0, rg);
      }
      if (!dn.empty()) {
        std::vector<DenseItem<T>> items;
        std::vector<uint32_t> nt;
        for (int q : dn) {
          auto& L = slots[q].mg->lev[d];
          DenseItem<T> it{};
          it.cinv = slots[q].mg->cinv;
          it.f = L.f;
          it.n = slots[q].mg->ncoarse;
          it.slot = q;
          if (d == 0) {
            it.out_base = zbase(q);
            it.out_stride = L.d.n();
            it.x = nullptr;
          } else {
            it.x = L.xa;
          }
          cur[q][d] = L.xa;
          items.push_back(it);
          nt.push_back((uint32_t)(it.n + 63) / 64);
        }
        add_launch(K_DENSE, 0, items, nt);
      }
    }
    // ascent
    for (int d = dmax - 2; d >= 0; d--) {
      std::vector<int> up;
      for (int q : mem)
        if (d < (int)slots[q].mg->lev.size() - 1) up.push_back(q);
      if (up.empty()) continue;
      std::vector<ProlongItem<T>> pi;
      std::vector<uint32_t> pn;
      for (int q : up) {
        auto& L = slots[q].mg->lev[d];
        auto& C = slots[q].mg->lev[d + 1];
        ProlongItem<T> it{};
        it.F = L.view();
        it.dc = C.d;
        it.clsc = C.cls;
        it.xc = cur[q][d + 1];
        it.xf = cur[q][d];
        it.cmask = (C.d.nx != L.d.nx ? 1 : 0) | (C.d.ny != L.d.ny ? 2 : 0) | (C.d.nz != L.d.nz ? 4 : 0);
        it.slot = q;
        it.tx = (uint32_t)(L.d.nx + 63) / 64;
        it.ty = (uint32_t)(L.d.ny + 3) / 4;
        pi.push_back(it);
        pn.push_back(cell_tiles(L.d));
      }
      add_launch_cells(K_PROLONG, 0, pi);
      for (int k = 0; k < degree; k++) smoothing_step(d, k, false, true, up, cur);
    }
    // Krylov passes
    {
      std::vector<ApplyItem<T>> ai;
      std::vector<VecItem<T>> vi;
      std::vector<uint32_t> an, vn;
      std::vector<FinItem> fa, fv;
      uint32_t ta = 0, tv = 0;
      int mmax = 1;
      for (int q : mem) {
        vfs_pp_handle* h = slots[q].h;
        const size_t n = h->d.n();
        ApplyItem<T> a{};
        a.d = h->d;
        a.g = h->g64;
        a.code = h->code;
        a.Z = zbase(q);
        a.V = h->V;
        a.n = n;
        a.slot = q;
        a.tx = (uint32_t)(h->d.nx + 63) / 64;
        a.ty = (uint32_t)(h->d.ny + 3) / 4;
        const uint32_t na = cell_tiles(h->d);
        fa.push_back(FinItem{h->red, h->hpin_dev, ta, ta + na, q});
        ta += na;
        ai.push_back(a);
        an.push_back(na);
        VecItem<T> v{};
        v.V = h->V;
        v.red = h->red;
        v.hpin = h->hpin_dev;
        v.f0 = static_cast<T*>(slots[q].mg->f0());
        v.n = n;
        v.slot = q;
        const uint32_t nv = (uint32_t)((n + 255) / 256);
        fv.push_back(FinItem{h->red, h->hpin_dev, tv, tv + nv, q});
        tv += nv;
        vi.push_back(v);
        vn.push_back(nv);
        mmax = std::max(mmax, h->mrest);
      }
      const size_t need = (size_t)(mmax + 2) * 4 * std::max<size_t>(ta, tv);  // up to 4 partials per tile
      partial = plan.alloc<double>(need);
      partial_cap = need;
      std::vector<uint32_t> sa(ai.size()), sv(vi.size());
      uint32_t acc = 0;
      for (size_t q = 0; q < ai.size(); q++) { sa[q] = acc; acc += an[q]; }
      acc = 0;
      for (size_t q = 0; q < vi.size(); q++) { sv[q] = acc; acc += vn[q]; }
      auto pa = upload(ai, sa, ta);
      auto pv = upload(vi, sv, tv);
      apply = Launch{K_APPLY, 0, pa.first, pa.second, (int)ai.size(), ta, 2048};
      axd = Launch{K_AXD, 0, pv.first, pv.second, (int)vi.size(), tv, 2048};
      axs = Launch{K_AXS, 0, pv.first, pv.second, (int)vi.size(), tv, 2048};
      axn = Launch{K_AXN, 0, pv.first, pv.second, (int)vi.size(), tv, 2048};
      fin_apply = plan.alloc<FinItem>(fa.size());
      fin_vec = plan.alloc<FinItem>(fv.size());
      HIPCHECK(hipMemcpyAsync(fin_apply, fa.data(), fa.size() * sizeof(FinItem), hipMemcpyHostToDevice, s));
      HIPCHECK(hipMemcpyAsync(fin_vec, fv.data(), fv.size() * sizeof(FinItem), hipMemcpyHostToDevice, s));
      nfin = (int)fa.size();
      fin_total_apply = ta;
      fin_total_vec = tv;
    }
    HIPCHECK(hipStreamSynchronize(s));
    dirty = false;
  }

  // A cell-tiled launch: items over the cells of dims[q] (the level each item
  // walks); starts with the flat walk, the tuner may pick another shape.
  static Dims cells_of(const SmoothItem<T>& I) { return I.L.d; }
  static Dims cells_of(const RestrictItem<T>& I) { return I.dc; }
  static Dims cells_of(const ProlongItem<T>& I) { return I.F.d; }
  template <typename Item>
  void add_launch_cells(int kind, int mode, const std::vector<Item>& items) {
    if (items.empty()) return;
    std::vector<Dims> dims(items.size());
    for (size_t q = 0; q < items.size(); q++) dims[q] = cells_of(items[q]);
    std::vector<uint32_t> tiles[kTileShapes];
    uint32_t flat = 0;
    for (int sh = 0; sh < kTileShapes; sh++) {
      tiles[sh].resize(items.size());
      for (size_t q = 0; q < items.size(); q++) tiles[sh][q] = shape_tiles(sh, dims[q]);
    }
    for (uint32_t v : tiles[0]) flat += v;
    // the shape a forced setting or an earlier tuning chose for this kind and size
    int sh = 0;
    if (forced_shape >= 0) {
      sh = forced_shape;
    } else {
      auto& tc = tune_cache();
      std::lock_guard<std::mutex> lk(tc.mu);
      auto it = tc.best.find(std::make_tuple(kind + 100 * (int)sizeof(T), mode, log2_bucket(flat)));
      if (it != tc.best.end()) sh = it->second.second;
    }
    const size_t before = seq.size();
    add_launch(kind, mode, items, tiles[sh]);
    if (seq.size() == before) return;
    Launch& L = seq.back();
    L.shaped = true;
    L.shape = sh;
    L.total0 = flat;
    for (int q = 0; q < kTileShapes; q++) L.tiles[q] = std::move(tiles[q]);
    L.item_slot.resize(items.size());
    for (size_t q = 0; q < items.size(); q++) L.item_slot[q] = (uint32_t)items[q].slot;
  }

  // rebuild a cell-tiled launch's starts and per-tile table for tile shape sh
  void retile(Launch& L, int sh) {
    if (!L.shaped || L.shape == sh) return;
    const size_t ni = L.item_slot.size();
    std::vector<uint32_t> start(ni);
    uint32_t tot = 0;
    for (size_t q = 0; q < ni; q++) { start[q] = tot; tot += L.tiles[sh][q]; }
    uint32_t* ds = plan.alloc<uint32_t>(2 * ni + tot);
    HIPCHECK(hipMemcpyAsync(ds, start.data(), ni * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    HIPCHECK(hipMemcpyAsync(ds + ni + tot, L.item_slot.data(), ni * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    if (tot) {
      k_tile_table<<<(tot + 255) / 256, 256, 0, s>>>(ds, (int)ni, tot);
      HIPCHECK(hipGetLastError());
    }
    HIPCHECK(hipStreamSynchronize(s));
    L.start = ds;
    L.total = tot;
    L.shape = sh;
  }

  void launch_one(const Launch& L) {
    const unsigned W = workers(L.total, L.cap);
    switch (L.kind) {
      case K_SMOOTH: {
        auto* it = static_cast<const SmoothItem<T>*>(L.items);
        if (L.mode == SM_FIRST2)
          kb_smooth<T, SM_FIRST2><<<W, dim3(64, 4, 1), 0, s>>>(it, L.start, L.nitems, L.total, sd, L.shape);
        else if (L.mode == SM_ZERO)
          kb_smooth<T, SM_ZERO><<<W, dim3(64, 4, 1), 0, s>>>(it, L.start, L.nitems, L.total, sd, L.shape);
        else
          kb_smooth<T, SM_NEXT><<<W, dim3(64, 4, 1), 0, s>>>(it, L.start, L.nitems, L.total, sd, L.shape);
        break;
      }
      case K_RESID:
        kb_resid<T><<<W, dim3(64, 4, 1), 0, s>>>(static_cast<const SmoothItem<T>*>(L.items), L.start, L.nitems,
                                                 L.total, sd, L.shape);
        break;
      case K_LINE:
        if (L.mode == SM_ZERO)
          kb_line<T, true><<<W, dim3(64, 4, 1), 0, s>>>(static_cast<const SmoothItem<T>*>(L.items), L.start,
                                                        L.nitems, L.total, sd);
        else
          kb_line<T, false><<<W, dim3(64, 4, 1), 0, s>>>(static_cast<const SmoothItem<T>*>(L.items), L.start,
                                                         L.nitems, L.total, sd);
        break;
      case K_LSEG:
        if (L.mode == SM_ZERO)
          kb_line_seg<T, true><<<W, dim3(64, 16, 1), 0, s>>>(static_cast<const SmoothItem<T>*>(L.items),
                                                                    L.start, L.nitems, L.total, sd);
        else
          kb_line_seg<T, false><<<W, dim3(64, 16, 1), 0, s>>>(static_cast<const SmoothItem<T>*>(L.items),
                                                                     L.start, L.nitems, L.total, sd);
        break;
      case K_RBLK:
        kb_restrict_blk<T><<<W, dim3(64, 2, 2), 0, s>>>(static_cast<const RestrictItem<T>*>(L.items), L.start,
                                                        L.nitems, L.total, sd);
        break;
      case K_RGEN:
        kb_restrict_gen<T><<<W, dim3(64, 4, 1), 0, s>>>(static_cast<const RestrictItem<T>*>(L.items), L.start,
                                                        L.nitems, L.total, sd, L.shape);
        break;
      case K_PROLONG:
        kb_prolong<T><<<W, dim3(64, 4, 1), 0, s>>>(static_cast<const ProlongItem<T>*>(L.items), L.start, L.nitems,
                                                   L.total, sd, L.shape);
        break;
      case K_DENSE:
        kb_dense<T><<<W, 256, 0, s>>>(static_cast<const DenseItem<T>*>(L.items), L.start, L.nitems, L.total, sd);
        break;
      case K_APPLY:
        kb_apply_dots<T><<<W, dim3(64, 4, 1), 0, s>>>(static_cast<const ApplyItem<T>*>(L.items), L.start, L.nitems,
                                                      L.total, sd, partial);
        break;
      case K_AXD:
        kb_axpy_dots<T><<<W, 256, 0, s>>>(static_cast<const VecItem<T>*>(L.items), L.start, L.nitems, L.total, sd,
                                          partial);
        break;
      case K_AXS:
        kb_axpy_scale<T><<<W, 256, 0, s>>>(static_cast<const VecItem<T>*>(L.items), L.start, L.nitems, L.total, sd);
        break;
      case K_AXN:
        kb_axpy_norm<T><<<W, 256, 0, s>>>(static_cast<const VecItem<T>*>(L.items), L.start, L.nitems, L.total, sd,
                                          partial);
        break;
    }
    HIPCHECK(hipGetLastError());
  }

  void launch_seq() {
    for (const Launch& L : seq) launch_one(L);
    launch_one(apply);
    if (lazy) {
      // one Gram-Schmidt pass on the unnormalised basis: w.U_l and |w|^2 from the
      // operator pass, then w' and |w'|^2 (the host re-orthogonalises the rare
      // case whose pass cancelled most of w)
      kb_finalize<<<dim3(nfin, 64), 1024, 0, s>>>(fin_apply, partial, fin_total_apply, 4, 8, 1, sd);
      HIPCHECK(hipGetLastError());
      launch_one(axn);
      kb_finalize_norm<<<nfin, 1024, 0, s>>>(fin_vec, partial, sd);
      HIPCHECK(hipGetLastError());
      return;
    }
    // CGS2 in three passes over each case's basis
    kb_finalize<<<dim3(nfin, 64), 1024, 0, s>>>(fin_apply, partial, fin_total_apply, 4, 8, 0, sd);
    HIPCHECK(hipGetLastError());
    launch_one(axd);
    kb_finalize<<<dim3(nfin, 64), 1024, 0, s>>>(fin_vec, partial, fin_total_vec, 1, 72, 1, sd);
    HIPCHECK(hipGetLastError());
    kb_gs_norm<<<(nfin + 63) / 64, 64, 0, s>>>(fin_vec, nfin, sd);
    HIPCHECK(hipGetLastError());
    launch_one(axs);
  }

  // Autotuner: each launch's worker count and, for the cell-tiled launches, its
  // tile shape (flat 256 cells, 64x4, 32x8 or 16x16) are picked by timing the
  // candidates (HIP events, best of 3 after a warm-up) and remembered per (kind,
  // mode, size bucket).  The timed launches run on the cases' scratch state, so
  // this only happens while no case is in the middle of an iteration.
  static int bucket_of(const Launch& L) { return log2_bucket(L.shaped ? L.total0 : L.total); }
  void tune(bool allowed) {
    std::vector<Launch*> all;
    for (Launch& L : seq) all.push_back(&L);
    all.push_back(&apply);
    if (lazy) {
      all.push_back(&axn);
    } else {
      all.push_back(&axd);
      all.push_back(&axs);
    }
    auto& tc = tune_cache();
    std::lock_guard<std::mutex> lk(tc.mu);
    bool need = false;
    for (Launch* L : all) {
      auto it = tc.best.find(std::make_tuple(L->kind + 100 * (int)sizeof(T), L->mode, bucket_of(*L)));
      if (it != tc.best.end()) {
        L->cap = it->second.first;
        retile(*L, forced_shape >= 0 ? forced_shape : it->second.second);
      } else {
        need = true;
      }
    }
    if (!need || !allowed || forced_shape >= 0) return;
    // every slot active at step 0 while timing
    for (size_t q = 0; q < slots.size(); q++) sd_host[q] = SlotDev{0, slots[q].h ? 1 : 0};
    HIPCHECK(hipMemcpyAsync(sd, sd_host, sizeof(SlotDev) * slots.size(), hipMemcpyHostToDevice, s));
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    const unsigned cands[] = {0u, 16384u, 8192u, 4096u, 2048u, 1024u};
    for (Launch* L : all) {
      const auto key = std::make_tuple(L->kind + 100 * (int)sizeof(T), L->mode, bucket_of(*L));
      auto it = tc.best.find(key);
      if (it != tc.best.end()) {
        L->cap = it->second.first;
        retile(*L, it->second.second);
        continue;
      }
      float best_ms = 1e30f;
      unsigned best = 2048u;
      int best_shape = L->shape;
      const int nsh = L->shaped ? kTileShapes : 1;
      for (int shi = 0; shi < nsh; shi++) {
        const int sh = L->shaped ? shi : L->shape;
        retile(*L, sh);
        unsigned last_w = 0;
        for (unsigned cap : cands) {
          const unsigned w = workers(L->total, cap);
          if (w == last_w) continue;  // same grid as the previous candidate
          last_w = w;
          L->cap = cap;
          launch_one(*L);
          float tmin = 1e30f;
          for (int r = 0; r < 3; r++) {
            HIPCHECK(hipEventRecord(e0, s));
            launch_one(*L);
            HIPCHECK(hipEventRecord(e1, s));
            HIPCHECK(hipEventSynchronize(e1));
            float ms = 0.f;
            HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
            tmin = std::min(tmin, ms);
          }
          if (tmin < best_ms) { best_ms = tmin; best = cap; best_shape = sh; }
        }
      }
      retile(*L, best_shape);
      L->cap = best;
      tc.best[key] = {best, best_shape};
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    HIPCHECK(hipStreamSynchronize(s));
  }

  // tile shapes the current plan's cell-tiled launches use: counts per shape
  void shape_counts(int* counts) const {
    for (int sh = 0; sh < kTileShapes; sh++) counts[sh] = 0;
    for (const Launch& L : seq)
      if (L.shaped) counts[L.shape]++;
  }

  void run_iteration() {
    if (dirty) {
      build_plan();
      tune(false);
    }
    // slot state for this iteration
    for (size_t q = 0; q < slots.size(); q++) {
      const Slot& sl = slots[q];
      sd_host[q] = SlotDev{sl.j, sl.state == ITER ? 1 : 0};
    }
    HIPCHECK(hipMemcpyAsync(sd, sd_host, sizeof(SlotDev) * slots.size(), hipMemcpyHostToDevice, s));
    if (use_graph) {
      if (!gexec) {
        HIPCHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        launch_seq();
        HIPCHECK(hipStreamEndCapture(s, &graph));
        HIPCHECK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
      }
      HIPCHECK(hipGraphLaunch(gexec, s));
    } else {
      launch_seq();
    }
  }

  // ------------------------------------------------------------ per-case logic
  void begin(int q, double* ucont, double* phi, double* p, double* ucat, double dt, double st, double coeff) {
    Slot& sl = slots[q];
    sl.ucont = ucont;
    sl.phi = phi;
    sl.p = p;
    sl.ucat = ucat;
    sl.dt = dt;
    sl.st = st;
    sl.coeff = coeff;
    sl.S = vfs_pp_stats{};
    sl.t0 = now_ms();
    sl.j = sl.it = sl.vc = 0;
    sl.conv = false;
    sl.pending_update = false;
    sl.m = sl.h->mrest;
    sl.H.assign((size_t)(sl.m + 1) * sl.m, 0.);
    sl.cs.assign(sl.m, 0.);
    sl.sn.assign(sl.m, 0.);
    sl.gv.assign(sl.m + 1, 0.);
    sl.y.assign(sl.m, 0.);
    solve_prologue_host(sl.h, dt, st, coeff);  // VolumeFlux and RHS run with the next start_cycles
    sl.fresh = true;
    sl.state = START;
  }

  // ------------------------------------------------------ batched phase steps
  struct PhaseStep {
    int op = 0;
    int nval = 0;      // values reduced per case (0: none)
    bool max = false;  // reduction by maximum instead of sum
    std::vector<PhaseItem> items;
    std::vector<PhaseFin> fins;  // destination of each item's reduced values
    std::vector<int> data;       // per item: index of a Phase::data array patched into sc, or -1
    void add(const PhaseItem& it, double* out = nullptr, double* host = nullptr, int d = -1) {
      if (it.n == 0) return;
      items.push_back(it);
      fins.push_back(PhaseFin{out, host, 0, 0});
      data.push_back(d);
    }
  };
  struct Phase {
    std::vector<PhaseStep> steps;
    std::vector<std::vector<double>> data;
    // the returned step is filled before the next one is added
    PhaseStep& step(int op, int nval = 0, bool max = false) {
      steps.emplace_back();
      PhaseStep& S = steps.back();
      S.op = op;
      S.nval = nval;
      S.max = max;
      return S;
    }
  };

  static PhaseItem case_item(const vfs_pp_handle* h) {
    PhaseItem it{};
    it.m = h->m;
    it.d = h->d;
    it.g = h->g64;
    it.code = h->code;
    it.cls = h->cls;
    it.n = (uint32_t)h->d.n();
    return it;
  }
  static PhaseItem boundary_item(const vfs_pp_handle* h) {
    PhaseItem it = case_item(h);
    it.n = boundary_count(h->m);
    return it;
  }

  void ph_launch(int op, const PhaseItem* it, const uint32_t* st, int n, uint32_t total) {
    const unsigned W = workers(total, 2048);
    switch (op) {
#define PH_CASE(OP) \
  case OP: kb_phase<OP, T><<<W, 256, 0, s>>>(it, st, n, total, ph_partial); break;
      PH_CASE(PH_IBSUM) PH_CASE(PH_IBCORR) PH_CASE(PH_RHS) PH_CASE(PH_GATHER) PH_CASE(PH_ZERO)
      PH_CASE(PH_UPDATE_X) PH_CASE(PH_RESID) PH_CASE(PH_DIVEST) PH_CASE(PH_SCALE_CONV) PH_CASE(PH_MASKED_SUM)
      PH_CASE(PH_SUB_MEAN) PH_CASE(PH_SCATTER) PH_CASE(PH_PERIODIC) PH_CASE(PH_ADD_PHI) PH_CASE(PH_PROJECTION)
      PH_CASE(PH_UCONT_GHOSTS) PH_CASE(PH_C2C_CORE) PH_CASE(PH_VEC_GHOSTS) PH_CASE(PH_C2C_BC)
      PH_CASE(PH_C2C_SOLID) PH_CASE(PH_DIVERGENCE)
#undef PH_CASE
      default: throw StatusError(VFS_PP_EINVAL, "unknown phase step");
    }
    HIPCHECK(hipGetLastError());
  }

  // Upload a phase's work lists (one copy from pinned staging) and launch its
  // steps in order on the batch stream.  Launches only.
  void ph_run(Phase& P) {
    auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
    size_t bytes = 0, need = 0;
    std::vector<size_t> doff(P.data.size());
    for (size_t q = 0; q < P.data.size(); q++) {
      doff[q] = bytes;
      bytes = al(bytes + P.data[q].size() * sizeof(double));
    }
    struct Off {
      size_t it, st, fin;
      uint32_t total;
    };
    std::vector<Off> off(P.steps.size());
    for (size_t l = 0; l < P.steps.size(); l++) {
      const PhaseStep& S = P.steps[l];
      uint32_t tot = 0;
      for (const PhaseItem& it : S.items) tot += (it.n + 255u) / 256u;
      off[l].total = tot;
      off[l].it = bytes;
      bytes = al(bytes + S.items.size() * sizeof(PhaseItem));
      off[l].st = bytes;
      bytes = al(bytes + S.items.size() * sizeof(uint32_t));
      off[l].fin = bytes;
      bytes = al(bytes + S.items.size() * sizeof(PhaseFin));
      need = std::max(need, (size_t)S.nval * tot);
    }
    if (bytes == 0) return;
    if (ph_recorded) HIPCHECK(hipEventSynchronize(ph_ev));  // the staging buffer's previous upload was read
    if (bytes > ph_cap) {
      HIPCHECK(hipStreamSynchronize(s));
      if (ph_host) (void)hipHostFree(ph_host);
      if (ph_dev) (void)hipFree(ph_dev);
      ph_host = nullptr;
      ph_dev = nullptr;
      ph_cap = 0;
      const size_t cap = std::max<size_t>(2 * bytes, (size_t)1 << 16);
      HIPCHECK(hipHostMalloc(&ph_host, cap));
      HIPCHECK(hipMalloc(&ph_dev, cap));
      ph_cap = cap;
    }
    if (need > ph_partial_cap) {
      HIPCHECK(hipStreamSynchronize(s));
      if (ph_partial) (void)hipFree(ph_partial);
      ph_partial = nullptr;
      ph_partial_cap = 0;
      const size_t cap = std::max<size_t>(2 * need, 4096);
      HIPCHECK(hipMalloc(&ph_partial, cap * sizeof(double)));
      ph_partial_cap = cap;
    }
    // serialise, with the device addresses of the data arrays patched into the items
    for (size_t q = 0; q < P.data.size(); q++)
      std::memcpy(ph_host + doff[q], P.data[q].data(), P.data[q].size() * sizeof(double));
    for (size_t l = 0; l < P.steps.size(); l++) {
      const PhaseStep& S = P.steps[l];
      auto* it = reinterpret_cast<PhaseItem*>(ph_host + off[l].it);
      auto* st = reinterpret_cast<uint32_t*>(ph_host + off[l].st);
      auto* fn = reinterpret_cast<PhaseFin*>(ph_host + off[l].fin);
      uint32_t acc = 0;
      for (size_t q = 0; q < S.items.size(); q++) {
        it[q] = S.items[q];
        if (S.data[q] >= 0) it[q].sc = reinterpret_cast<const double*>(ph_dev + doff[S.data[q]]);
        const uint32_t nt = (S.items[q].n + 255u) / 256u;
        st[q] = acc;
        fn[q] = PhaseFin{S.fins[q].out, S.fins[q].host, acc, acc + nt};
        acc += nt;
      }
    }
    HIPCHECK(hipMemcpyAsync(ph_dev, ph_host, bytes, hipMemcpyHostToDevice, s));
    HIPCHECK(hipEventRecord(ph_ev, s));
    ph_recorded = true;
    for (size_t l = 0; l < P.steps.size(); l++) {
      const PhaseStep& S = P.steps[l];
      if (off[l].total == 0) continue;
      ph_launch(S.op, reinterpret_cast<const PhaseItem*>(ph_dev + off[l].it),
                reinterpret_cast<const uint32_t*>(ph_dev + off[l].st), (int)S.items.size(), off[l].total);
      if (S.nval) {
        const auto* fn = reinterpret_cast<const PhaseFin*>(ph_dev + off[l].fin);
        const dim3 gr((unsigned)S.items.size(), (unsigned)S.nval);
        if (S.max) kb_phase_fin<true><<<gr, 256, 0, s>>>(fn, ph_partial, off[l].total);
        else kb_phase_fin<false><<<gr, 256, 0, s>>>(fn, ph_partial, off[l].total);
        HIPCHECK(hipGetLastError());
      }
    }
  }

  // Cases whose cycle starts: for new cases VolumeFlux, RHS and the initial guess;
  // for the others the x update from the finished cycle; then the true residual
  // of all of them and, for those that go on, v_0 = r / ||r||.
  void start_cycles() {
    std::vector<int> qs, fr;
    for (int q = 0; q < (int)slots.size(); q++)
      if (slots[q].state == START) {
        qs.push_back(q);
        if (slots[q].fresh) fr.push_back(q);
      }
    if (qs.empty()) return;
    if (!fr.empty()) {
      // VolumeFlux (poisson.c:2640-2980) and the RHS (poisson_hypre.c:346-376)
      Phase P;
      P.steps.reserve(4);
      {
        PhaseStep& S = P.step(PH_IBSUM, 2);
        for (int q : fr) {
          vfs_pp_handle* h = slots[q].h;
          if (!h->nibf) continue;
          PhaseItem it = case_item(h);
          it.w0 = slots[q].ucont;
          it.faces = h->ibf;
          it.n = (uint32_t)h->nibf;
          it.i0 = h->grid.immersed == 1 ? 1 : 0;
          S.add(it, h->red + RED_IBF, h->hpin_dev + RED_IBF);
        }
      }
      {
        PhaseStep& S = P.step(PH_IBCORR);
        for (int q : fr) {
          vfs_pp_handle* h = slots[q].h;
          if (!h->nibf) continue;
          PhaseItem it = case_item(h);
          it.w0 = slots[q].ucont;
          it.faces = h->ibf;
          it.n = (uint32_t)h->nibf;
          it.i0 = h->grid.immersed;
          it.sc = h->red + RED_IBF;
          S.add(it);
        }
      }
      {
        PhaseStep& S = P.step(PH_IBSUM, 2);
        for (int q : fr) {
          vfs_pp_handle* h = slots[q].h;
          if (!h->nibf) continue;
          PhaseItem it = case_item(h);
          it.w0 = slots[q].ucont;
          it.faces = h->ibf;
          it.n = (uint32_t)h->nibf;
          it.i0 = 0;
          S.add(it, h->red + RED_IBF + 2, h->hpin_dev + RED_IBF + 2);
        }
      }
      {
        PhaseStep& S = P.step(PH_RHS, 1);
        for (int q : fr) {
          const Slot& sl = slots[q];
          vfs_pp_handle* h = sl.h;
          PhaseItem it = case_item(h);
          it.r0 = sl.ucont;
          it.w0 = h->b;
          it.s0 = -1.0 / sl.dt * sl.st * sl.coeff;  // poisson_hypre.c:376
          S.add(it, h->red + RED_B2, h->hpin_dev + RED_B2);
        }
      }
      ph_run(P);
      HIPCHECK(hipStreamSynchronize(s));
      for (int q : fr) {
        Slot& sl = slots[q];
        solve_start_stats(sl.h, &sl.S);
        sl.tgt = sl.h->opts.rtol * sl.S.bnorm;
        sl.fresh = false;
      }
    }
    // zero-start cases need no operator pass for their first true residual (r = b)
    for (int q : fr) slots[q].x_zero = slots[q].h->opts.zero_guess != 0;
    Phase R;
    R.steps.reserve(5);
    {
      PhaseStep& S = R.step(PH_ZERO);
      for (int q : fr) {
        vfs_pp_handle* h = slots[q].h;
        if (!h->opts.zero_guess) continue;
        PhaseItem it = case_item(h);
        it.w0 = h->x;
        S.add(it);
      }
    }
    {
      PhaseStep& S = R.step(PH_GATHER);
      for (int q : fr) {
        vfs_pp_handle* h = slots[q].h;
        if (h->opts.zero_guess) continue;
        PhaseItem it = case_item(h);
        it.r0 = slots[q].phi;
        it.w0 = h->x;
        S.add(it);
      }
    }
    {
      PhaseStep& S = R.step(PH_UPDATE_X);
      for (int q : qs) {
        Slot& sl = slots[q];
        if (!sl.pending_update) continue;
        vfs_pp_handle* h = sl.h;
        std::vector<double> y(sl.pending_nv);
        for (int l = 0; l < sl.pending_nv; l++) y[l] = lazy ? sl.y[l] / sl.sigma[l] : sl.y[l];
        R.data.push_back(std::move(y));
        PhaseItem it = case_item(h);
        it.w0 = h->x;
        it.z = h->Z;
        it.ld = h->d.n();
        it.i0 = sl.pending_nv;
        S.add(it, nullptr, nullptr, (int)R.data.size() - 1);
        sl.pending_update = false;
      }
    }
    {
      PhaseStep& S = R.step(PH_RESID, 1);
      for (int q : qs) {
        vfs_pp_handle* h = slots[q].h;
        if (slots[q].x_zero) continue;
        PhaseItem it = case_item(h);
        it.r0 = h->b;
        it.r1 = h->x;
        it.w0 = h->r;
        S.add(it, h->red, h->hpin_dev);
      }
    }
    {
      PhaseStep& S = R.step(PH_DIVEST, 1, true);
      for (int q : qs) {
        vfs_pp_handle* h = slots[q].h;
        if (!(h->div_fac > 0.)) continue;
        PhaseItem it = case_item(h);
        it.r0 = h->aj;
        it.r1 = true_residual_src(h, slots[q].x_zero);
        S.add(it, h->red + RED_DIVEST, h->hpin_dev + RED_DIVEST);
      }
    }
    ph_run(R);
    HIPCHECK(hipStreamSynchronize(s));
    Phase C;
    PhaseStep& SC = C.step(PH_SCALE_CONV);
    for (int q : qs) {
      Slot& sl = slots[q];
      vfs_pp_handle* h = sl.h;
      const bool xz = sl.x_zero;
      sl.x_zero = false;
      if (sl.S.bnorm == 0.) { sl.conv = true; sl.rel = 0.; sl.state = FINISH; continue; }
      const double beta = std::sqrt(true_residual_host_norm2(h, xz));
      if (!std::isfinite(beta)) throw StatusError(VFS_PP_ENONFINITE, "non-finite residual");
      sl.rel = beta / sl.S.bnorm;
      if (residual_verdict(h, beta, sl.S.bnorm, sl.tgt)) { sl.conv = true; sl.state = FINISH; continue; }
      if (sl.it >= h->opts.max_iters) { sl.state = FINISH; continue; }
      // v_0 = r / beta and its copy as the first V-cycle input; a lazily normalised
      // basis reads 1 / |U_0|^2 = 1
      PhaseItem it = case_item(h);
      it.r0 = true_residual_src(h, xz);
      it.w0 = h->V;
      it.f = sl.mg->f0();
      it.sc = true_residual_norm2(h, xz);
      it.one = lazy ? h->red + RED_NUINV2 : nullptr;
      SC.add(it);
      if (lazy) {
        sl.nu.assign(sl.m + 1, 1.);
        sl.sigma.assign(sl.m + 1, 1.);
      }
      std::fill(sl.gv.begin(), sl.gv.end(), 0.);
      sl.gv[0] = beta;
      sl.j = 0;
      sl.state = ITER;
    }
    ph_run(C);
  }

  // A second Gram-Schmidt pass over U_nv of one case (its first pass cancelled
  // |w'| below 1e-4 |w|): dots into red[72..], w'' and the V-cycle input with the
  // same scale, |w''|^2 into red[0] and its reciprocal into red[RED_NUINV2 + nv].
  // Returns |w''|^2.  Launches on the batch stream and waits for it.
  double reorthogonalise(Slot& sl, int nv) {
    vfs_pp_handle* h = sl.h;
    const size_t n = h->d.n();
    double* w = h->V + (size_t)nv * n;
    int nb = 0;
    need_partials(h, (size_t)nv, (size_t)partial_blocks(n));
    launch_multidot(s, n, w, h->V, n, nv, h->partial, &nb);
    launch_finalize(s, h->partial, nv, nb, h->red + 72, 0, h->hpin_dev + 72);
    launch_axpy_norm<T>(s, n, w, h->V, n, nv, h->red + 72, h->red + RED_NUINV2, h->red + 8 + nv,
                        static_cast<T*>(sl.mg->f0()), h->partial, &nb);
    launch_finalize(s, h->partial, 1, nb, h->red, 0, h->hpin_dev);
    HIPCHECK(hipStreamSynchronize(s));
    const double b2 = h->hpin[0];
    h->hpin[RED_NUINV2 + nv] = b2 > 0. ? 1. / b2 : 0.;
    h2d(s, h->red + RED_NUINV2 + nv, h->hpin + RED_NUINV2 + nv, 1);
    return b2;
  }

  // after an iteration: Arnoldi column, Givens rotation, cycle end
  void after_iteration() {
    for (int q = 0; q < (int)slots.size(); q++) {
      Slot& sl = slots[q];
      if (sl.state != ITER) continue;
      vfs_pp_handle* h = sl.h;
      const int m = sl.m, j = sl.j;
      auto H = [&](int r, int c) -> double& { return sl.H[(size_t)r * m + c]; };
      if (lazy) {
        // H(l, j) = v_l . A z_j with U_l = nu_l v_l and Z_j = sigma_j z_j
        const int nv = j + 1;
        const double a2 = h->hpin[8 + nv];
        double b2 = h->hpin[0];
        for (int l = 0; l <= j; l++) H(l, j) = h->hpin[8 + l] / (sl.nu[l] * sl.sigma[j]);
        if (b2 < 1e-8 * a2) {
          b2 = reorthogonalise(sl, nv);
          for (int l = 0; l <= j; l++) H(l, j) += h->hpin[72 + l] / (sl.nu[l] * sl.sigma[j]);
        }
        sl.nu[j + 1] = std::sqrt(b2);
        sl.sigma[j + 1] = host_conv_scale(a2) * sl.nu[j + 1];
        H(j + 1, j) = sl.nu[j + 1] / sl.sigma[j];
      } else {
        for (int l = 0; l <= j; l++) H(l, j) = h->hpin[8 + l] + h->hpin[72 + l];
        H(j + 1, j) = h->hpin[1];
      }
      for (int l = 0; l < j; l++) {
        const double t1 = sl.cs[l] * H(l, j) + sl.sn[l] * H(l + 1, j);
        const double t2 = -sl.sn[l] * H(l, j) + sl.cs[l] * H(l + 1, j);
        H(l, j) = t1;
        H(l + 1, j) = t2;
      }
      const double a = H(j, j), bb = H(j + 1, j);
      const double rr = std::sqrt(a * a + bb * bb);
      sl.cs[j] = rr == 0. ? 1. : a / rr;
      sl.sn[j] = rr == 0. ? 0. : bb / rr;
      H(j, j) = rr;
      H(j + 1, j) = 0.;
      sl.gv[j + 1] = -sl.sn[j] * sl.gv[j];
      sl.gv[j] = sl.cs[j] * sl.gv[j];
      sl.it++;
      sl.vc++;
      if (!std::isfinite(sl.gv[j + 1])) throw StatusError(VFS_PP_ENONFINITE, "non-finite Krylov residual");
      const bool est = std::fabs(sl.gv[j + 1]) <= sl.tgt;
      if (est || j + 1 == m || sl.it >= h->opts.max_iters) {
        const int nv = j + 1;
        for (int l = nv - 1; l >= 0; l--) {
          double acc = sl.gv[l];
          for (int c = l + 1; c < nv; c++) acc -= H(l, c) * sl.y[c];
          sl.y[l] = acc / H(l, l);
        }
        sl.pending_update = true;
        sl.pending_nv = nv;
        sl.state = START;
      } else {
        sl.j = j + 1;
      }
    }
  }

  // Cases that are done iterating: Remove_Nullspace and Convert_Phi2_Phi
  // (poisson_hypre.c:207-270, poisson.c:97-185), UpdatePressure (poisson.c:2522-2614),
  // Projection (poisson.c:2190-2441), Contra2Cart (rhs.c:151-260) and Divergence
  // (main.c:1244-1391), each step one launch over all of them.
  void finish_cases(std::vector<int>* done) {
    std::vector<int> qs;
    for (int q = 0; q < (int)slots.size(); q++)
      if (slots[q].state == FINISH) qs.push_back(q);
    if (qs.empty()) return;
    auto periodic = [](const vfs_pp_handle* h) { return h->m.per[0] || h->m.per[1] || h->m.per[2]; };
    Phase F;
    F.steps.reserve(16);
    {
      PhaseStep& S = F.step(PH_MASKED_SUM, 1);
      for (int q : qs) {
        vfs_pp_handle* h = slots[q].h;
        PhaseItem it = case_item(h);
        it.r0 = h->x;
        S.add(it, h->red + RED_MEAN, h->hpin_dev + RED_MEAN);
      }
    }
    {
      PhaseStep& S = F.step(PH_SUB_MEAN);
      for (int q : qs) {
        vfs_pp_handle* h = slots[q].h;
        if (!(h->n_lt > 0.)) continue;
        PhaseItem it = case_item(h);
        it.w0 = h->x;
        it.sc = h->red + RED_MEAN;
        it.s0 = h->n_lt;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_SCATTER);
      for (int q : qs) {
        PhaseItem it = case_item(slots[q].h);
        it.r0 = slots[q].h->x;
        it.w0 = slots[q].phi;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_PERIODIC);
      for (int q : qs) {
        PhaseItem it = boundary_item(slots[q].h);
        it.w0 = slots[q].phi;
        it.i0 = 1;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_ADD_PHI);
      for (int q : qs) {
        PhaseItem it = case_item(slots[q].h);
        it.r0 = slots[q].phi;
        it.w0 = slots[q].p;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_PERIODIC);
      for (int q : qs) {
        PhaseItem it = boundary_item(slots[q].h);
        it.i0 = 0;
        it.w0 = slots[q].p;
        S.add(it);
        it.w0 = slots[q].phi;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_PROJECTION);
      for (int q : qs) {
        const Slot& sl = slots[q];
        PhaseItem it = case_item(sl.h);
        it.r0 = sl.phi;
        it.w0 = sl.ucont;
        it.s0 = sl.dt * sl.st / sl.coeff;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_UCONT_GHOSTS);
      for (int q : qs) {
        if (!periodic(slots[q].h)) continue;
        PhaseItem it = boundary_item(slots[q].h);
        it.w0 = slots[q].ucont;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_C2C_CORE);
      for (int q : qs) {
        PhaseItem it = case_item(slots[q].h);
        it.r0 = slots[q].h->met9;
        it.r1 = slots[q].ucont;
        it.w0 = slots[q].ucat;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_VEC_GHOSTS);
      for (int q : qs) {
        PhaseItem it = boundary_item(slots[q].h);
        it.w0 = slots[q].ucat;
        it.w1 = slots[q].h->snap;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_C2C_BC);
      for (int q : qs) {
        vfs_pp_handle* h = slots[q].h;
        PhaseItem it = boundary_item(h);
        it.r0 = h->nvert;
        it.r1 = h->snap;
        it.w0 = slots[q].ucat;
        for (int f = 0; f < 6; f++) it.bc[f] = h->grid.bctype[f];
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_C2C_SOLID);
      for (int q : qs) {
        PhaseItem it = case_item(slots[q].h);
        it.w0 = slots[q].ucat;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_VEC_GHOSTS);
      for (int q : qs) {
        if (!periodic(slots[q].h)) continue;
        PhaseItem it = boundary_item(slots[q].h);
        it.w0 = slots[q].ucat;
        it.w1 = nullptr;
        S.add(it);
      }
    }
    {
      PhaseStep& S = F.step(PH_DIVERGENCE, 1, true);
      for (int q : qs) {
        vfs_pp_handle* h = slots[q].h;
        PhaseItem it = case_item(h);
        it.r0 = h->aj;
        it.r1 = slots[q].ucont;
        S.add(it, h->red + RED_DIV, h->hpin_dev + RED_DIV);
      }
    }
    ph_run(F);
    HIPCHECK(hipStreamSynchronize(s));
    for (int q : qs) {
      Slot& sl = slots[q];
      vfs_pp_handle* h = sl.h;
      sl.S.iterations = sl.it;
      sl.S.vcycles = sl.vc;
      sl.S.rel_residual = sl.rel;
      sl.S.phi_mean = h->n_lt > 0. ? h->hpin[RED_MEAN] / h->n_lt : 0.;
      sl.S.max_div = h->hpin[RED_DIV];
      bool ok = sl.conv;
      if (h->opts.div_tol > 0. && !(sl.S.max_div <= h->opts.div_tol)) ok = false;
      sl.S.converged = ok ? 1 : 0;
      sl.S.ms_total = now_ms() - sl.t0;
      sl.state = DONE;
      if (done) done->push_back(q);
    }
  }

  bool any(State st) const {
    for (const auto& sl : slots)
      if (sl.state == st) return true;
    return false;
  }

  // one round: start cycles, one batched iteration, finish
  void step(std::vector<int>* done) {
    if (dirty) {
      // (re)build before any cycle starts; tune only if no case is mid-iteration
      build_plan();
      tune(!any(ITER));
    }
    start_cycles();
    finish_cases(done);
    if (any(ITER)) {
      run_iteration();
      HIPCHECK(hipStreamSynchronize(s));
      after_iteration();
    }
    start_cycles();
    finish_cases(done);
  }
};

}  // namespace pp

using namespace pp;

// ======================================================================= batch
struct vfs_pp_batch {
  int precision = 1;
  std::vector<vfs_pp_handle*> members;
  std::unique_ptr<BatchCore<float>> cf;
  std::unique_ptr<BatchCore<double>> cd;
};

template <typename T>
static int batch_project(BatchCore<T>& core, vfs_pp_batch* b, double* const* ucont, double* const* phi,
                         double* const* p, double* const* ucat, const double* dt, const double* st,
                         const double* coeff, vfs_pp_stats* stats) {
  const int n = (int)b->members.size();
  for (int q = 0; q < n; q++) core.begin(q, ucont[q], phi[q], p[q], ucat[q], dt[q], st[q], coeff[q]);
  while (core.any(BatchCore<T>::ITER) || core.any(BatchCore<T>::START) || core.any(BatchCore<T>::FINISH))
    core.step(nullptr);
  bool ok = true;
  for (int q = 0; q < n; q++) {
    if (stats) stats[q] = core.slots[q].S;
    if (!core.slots[q].S.converged) ok = false;
  }
  return ok ? VFS_PP_OK : VFS_PP_ENOCONV;
}

extern "C" {

int vfs_pp_batch_create(vfs_pp_handle* const* members, int count, vfs_pp_batch** out) {
  ABI_BEGIN
  if (!members || count < 1 || !out) return VFS_PP_EINVAL;
  auto b = std::make_unique<vfs_pp_batch>();
  b->precision = members[0] ? members[0]->opts.precision : 1;
  for (int q = 0; q < count; q++) {
    if (!members[q]) return VFS_PP_EINVAL;
    for (int r = 0; r < q; r++)
      if (members[r] == members[q]) return VFS_PP_EINVAL;
    if (members[q]->opts.precision != b->precision) return VFS_PP_EINVAL;
    b->members.push_back(members[q]);
  }
  if (b->precision == 1) {
    b->cf = std::make_unique<BatchCore<float>>(count);
    for (int q = 0; q < count; q++) b->cf->bind(q, members[q]);
  } else {
    b->cd = std::make_unique<BatchCore<double>>(count);
    for (int q = 0; q < count; q++) b->cd->bind(q, members[q]);
  }
  *out = b.release();
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_batch_destroy(vfs_pp_batch* b) {
  if (!b) return VFS_PP_EINVAL;
  delete b;
  return VFS_PP_OK;
}

int vfs_pp_batch_set_graphs(vfs_pp_batch* b, int on) {
  if (!b) return VFS_PP_EINVAL;
  if (b->cf) { b->cf->use_graph = on != 0; b->cf->drop_graph(); }
  if (b->cd) { b->cd->use_graph = on != 0; b->cd->drop_graph(); }
  return VFS_PP_OK;
}

int vfs_pp_batch_set_tile_shape(vfs_pp_batch* b, int shape) {
  if (!b || shape < -1 || shape >= kTileShapes) return VFS_PP_EINVAL;
  if (b->cf) { b->cf->forced_shape = shape; b->cf->dirty = true; }
  if (b->cd) { b->cd->forced_shape = shape; b->cd->dirty = true; }
  return VFS_PP_OK;
}

int vfs_pp_batch_tile_shapes(vfs_pp_batch* b, int counts[4]) {
  if (!b || !counts) return VFS_PP_EINVAL;
  if (b->cf) b->cf->shape_counts(counts);
  else b->cd->shape_counts(counts);
  return VFS_PP_OK;
}

int vfs_pp_batch_project_device(vfs_pp_batch* b, double* const* ucont, double* const* phi, double* const* p,
                                double* const* ucat, const double* dt, const double* st, const double* coeff,
                                vfs_pp_stats* stats) {
  ABI_BEGIN
  if (!b || !ucont || !phi || !p || !ucat || !dt || !st || !coeff) return VFS_PP_EINVAL;
  const int n = (int)b->members.size();
  for (int q = 0; q < n; q++)
    if (!ucont[q] || !phi[q] || !p[q] || !ucat[q] || !(dt[q] > 0.) || coeff[q] == 0.) return VFS_PP_EINVAL;
  if (b->cf) return batch_project(*b->cf, b, ucont, phi, p, ucat, dt, st, coeff, stats);
  return batch_project(*b->cd, b, ucont, phi, p, ucat, dt, st, coeff, stats);
  ABI_END
}

}  // extern "C"

// ====================================================================== server
// Continuous batching: requests queue FIFO; between batched iterations the
// server drops requests whose admission deadline has passed, admits queued
// requests into free slots as long as a Krylov workspace fits the HBM budget
// (resident cases plus the workspace pool), and retires finished ones, whose
// workspaces go back to the pool for the next request of the same size.  A case
// runs up to max_per_case requests at once: the first on the handle that holds
// its geometry, the others on solve clones of it, made on demand and kept.
struct vfs_pp_server {
  struct Solver {
    vfs_pp_handle* h = nullptr;
    bool busy = false;
  };
  struct Case {
    std::vector<Solver> solvers;  // [0] owns the geometry, the rest are clones of it
    size_t ws = 0;
    size_t foot = 0;  // device bytes the case holds while registered (clones included)
    int busy() const {
      int b = 0;
      for (const Solver& v : solvers) b += v.busy ? 1 : 0;
      return b;
    }
  };
  struct Req {
    int ticket, cid;
    double *ucont, *phi, *p, *ucat;
    double dt, st, coeff;
    double t_submit;
  };
  vfs_pp_server_opts o{};
  std::map<int, Case> cases;
  std::deque<Req> queue;
  std::deque<int> expired;
  std::vector<int> slot_ticket, slot_case, slot_solver;
  std::vector<void*> slot_ws;
  std::multimap<size_t, void*> pool;  // free workspaces by size
  size_t resident = 0, ws_held = 0, ws_max = 0;
  int next_case = 0, next_ticket = 0;
  std::unique_ptr<BatchCore<float>> cf;
  std::unique_ptr<BatchCore<double>> cd;

  ~vfs_pp_server() {
    cf.reset();
    cd.reset();
    for (auto& kv : pool) (void)hipFree(kv.second);
    for (void* w : slot_ws)
      if (w) (void)hipFree(w);
    for (auto& kv : cases) destroy_case(kv.second);
  }

  static void destroy_case(Case& c) {
    // clones first: they read the owner's arrays
    for (size_t v = c.solvers.size(); v-- > 1;) vfs_pp_destroy(c.solvers[v].h);
    vfs_pp_destroy(c.solvers[0].h);
    c.solvers.clear();
  }

  void* take_ws(size_t bytes) {
    auto it = pool.find(bytes);
    if (it != pool.end()) {
      void* w = it->second;
      pool.erase(it);
      return w;
    }
    // make room by releasing pooled workspaces of other sizes
    while (resident + ws_held + bytes > o.hbm_budget && !pool.empty()) {
      auto big = std::prev(pool.end());
      ws_held -= big->first;
      (void)hipFree(big->second);
      pool.erase(big);
    }
    if (resident + ws_held + bytes > o.hbm_budget) return nullptr;
    void* w = nullptr;
    if (hipMalloc(&w, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    ws_held += bytes;
    return w;
  }

  // a free solver of the case, making a clone when all are busy and the case
  // may run one more request; -1 when none can be had now
  int take_solver(Case& c) {
    for (size_t v = 0; v < c.solvers.size(); v++)
      if (!c.solvers[v].busy) return (int)v;
    if ((int)c.solvers.size() >= o.max_per_case) return -1;
    vfs_pp_handle* h = nullptr;
    try {
      h = clone_for_solves(c.solvers[0].h);
    } catch (const StatusError& e) {
      if (e.code == VFS_PP_ENOMEM) return -1;
      throw;
    }
    // the clone and a workspace must fit beside everything resident: pooled
    // workspaces of other sizes go first
    const size_t foot = h->resident_bytes();
    while (!pool.empty() && resident + foot + ws_held + c.ws > o.hbm_budget) {
      auto big = std::prev(pool.end());
      ws_held -= big->first;
      (void)hipFree(big->second);
      pool.erase(big);
    }
    if (resident + foot + ws_held + c.ws > o.hbm_budget) {
      vfs_pp_destroy(h);
      return -1;
    }
    resident += foot;
    c.foot += foot;
    c.solvers.push_back(Solver{h, false});
    return (int)c.solvers.size() - 1;
  }

  void expire(double now) {
    if (!(o.deadline_ms > 0.)) return;
    for (auto it = queue.begin(); it != queue.end();) {
      if (now - it->t_submit > o.deadline_ms) {
        expired.push_back(it->ticket);
        it = queue.erase(it);
      } else {
        ++it;
      }
    }
  }

  template <typename T>
  int progress(BatchCore<T>& core, int* tickets, vfs_pp_stats* stats, int cap) {
    expire(now_ms());
    // admit: FIFO, skipping requests whose case has no solver free now
    for (auto it = queue.begin(); it != queue.end();) {
      int q = -1;
      for (int r = 0; r < (int)slot_ticket.size(); r++)
        if (slot_ticket[r] < 0) { q = r; break; }
      if (q < 0) break;
      Case& c = cases.at(it->cid);
      const int v = take_solver(c);
      if (v < 0) { ++it; continue; }
      void* w = take_ws(c.ws);
      if (!w) break;  // wait for memory to free up
      vfs_pp_handle* h = c.solvers[v].h;
      h->V = static_cast<double*>(w);
      h->Z = static_cast<char*>(w) + (size_t)(h->mrest + 1) * h->d.n() * sizeof(double);
      core.bind(q, h);
      core.begin(q, it->ucont, it->phi, it->p, it->ucat, it->dt, it->st, it->coeff);
      c.solvers[v].busy = true;
      slot_ticket[q] = it->ticket;
      slot_case[q] = it->cid;
      slot_solver[q] = v;
      slot_ws[q] = w;
      it = queue.erase(it);
    }
    std::vector<int> done;
    if (core.any(BatchCore<T>::START) || core.any(BatchCore<T>::ITER) || core.any(BatchCore<T>::FINISH))
      core.step(&done);
    int nd = 0;
    for (int q : done) {
      if (nd < cap) {
        if (tickets) tickets[nd] = slot_ticket[q];
        if (stats) stats[nd] = core.slots[q].S;
      }
      nd++;
      Case& c = cases.at(slot_case[q]);
      Solver& v = c.solvers[slot_solver[q]];
      v.busy = false;
      core.release(core.slots[q]);
      v.h->V = nullptr;
      v.h->Z = nullptr;
      pool.emplace(c.ws, slot_ws[q]);
      slot_ws[q] = nullptr;
      slot_ticket[q] = -1;
      slot_case[q] = -1;
      slot_solver[q] = -1;
    }
    if (nd > cap) throw StatusError(VFS_PP_EINVAL, "more finished requests than the output arrays hold");
    return nd;
  }
};

extern "C" {

void vfs_pp_server_default_opts(vfs_pp_server_opts* o) {
  if (!o) return;
  o->hbm_budget = 0;
  o->max_active = 64;
  vfs_pp_default_opts(&o->solver);
  o->max_per_case = 4;
  o->max_queue = 0;
  o->deadline_ms = 0.;
}

int vfs_pp_server_create(const vfs_pp_server_opts* o, vfs_pp_server** out) {
  ABI_BEGIN
  if (!out) return VFS_PP_EINVAL;
  auto sv = std::make_unique<vfs_pp_server>();
  if (o) sv->o = *o;
  else vfs_pp_server_default_opts(&sv->o);
  if (sv->o.max_active < 1 || sv->o.max_active > 4096) return VFS_PP_EINVAL;
  if (sv->o.max_per_case < 1 || sv->o.max_queue < 0 || !(sv->o.deadline_ms >= 0.)) return VFS_PP_EINVAL;
  if (sv->o.hbm_budget == 0) {
    size_t fr = 0, tot = 0;
    HIPCHECK(hipMemGetInfo(&fr, &tot));
    sv->o.hbm_budget = fr / 10 * 8;  // 80% of what is free now
  }
  const int n = sv->o.max_active;
  sv->slot_ticket.assign(n, -1);
  sv->slot_case.assign(n, -1);
  sv->slot_solver.assign(n, -1);
  sv->slot_ws.assign(n, nullptr);
  if (sv->o.solver.precision == 1) sv->cf = std::make_unique<BatchCore<float>>(n);
  else sv->cd = std::make_unique<BatchCore<double>>(n);
  *out = sv.release();
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_server_destroy(vfs_pp_server* sv) {
  if (!sv) return VFS_PP_EINVAL;
  delete sv;
  return VFS_PP_OK;
}

int vfs_pp_server_add_case(vfs_pp_server* sv, const vfs_grid* g, const double* icsi, const double* ieta,
                           const double* izet, const double* jcsi, const double* jeta, const double* jzet,
                           const double* kcsi, const double* keta, const double* kzet, const double* iaj,
                           const double* jaj, const double* kaj, const double* aj, const double* csi,
                           const double* eta, const double* zet, const double* nvert, int* case_id) {
  ABI_BEGIN
  if (!sv || !g || !case_id) return VFS_PP_EINVAL;
  vfs_pp_handle* h = nullptr;
  int rc = vfs_pp_create(g, &sv->o.solver, &h);
  if (rc != VFS_PP_OK) return rc;
  h->external_ws = true;
  rc = vfs_pp_set_geometry(h, icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, aj, csi, eta,
                           zet, nvert);
  if (rc != VFS_PP_OK) {
    vfs_pp_destroy(h);
    return rc;
  }
  const size_t foot = h->resident_bytes();
  const size_t ws = workspace_bytes(h);
  // every registered case must still be able to run alone
  if (sv->resident + foot + std::max(ws, sv->ws_max) > sv->o.hbm_budget) {
    vfs_pp_destroy(h);
    return VFS_PP_ENOMEM;
  }
  sv->resident += foot;
  sv->ws_max = std::max(sv->ws_max, ws);
  const int id = sv->next_case++;
  vfs_pp_server::Case c;
  c.solvers.push_back(vfs_pp_server::Solver{h, false});
  c.ws = ws;
  c.foot = foot;
  sv->cases[id] = std::move(c);
  *case_id = id;
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_server_remove_case(vfs_pp_server* sv, int case_id) {
  ABI_BEGIN
  if (!sv) return VFS_PP_EINVAL;
  auto it = sv->cases.find(case_id);
  if (it == sv->cases.end()) return VFS_PP_EINVAL;
  if (it->second.busy()) return VFS_PP_ESTATE;
  for (const auto& r : sv->queue)
    if (r.cid == case_id) return VFS_PP_ESTATE;
  sv->resident -= it->second.foot;
  vfs_pp_server::destroy_case(it->second);
  sv->cases.erase(it);
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_server_submit(vfs_pp_server* sv, int case_id, double* ucont, double* phi, double* p, double* ucat,
                         double dt, double st, double coeff, int* ticket) {
  ABI_BEGIN
  if (!sv || !ucont || !phi || !p || !ucat || !ticket || !(dt > 0.) || coeff == 0.) return VFS_PP_EINVAL;
  if (!sv->cases.count(case_id)) return VFS_PP_EINVAL;
  if (sv->o.max_queue > 0 && (int)sv->queue.size() >= sv->o.max_queue) return VFS_PP_EBUSY;
  const int t = sv->next_ticket++;
  sv->queue.push_back(vfs_pp_server::Req{t, case_id, ucont, phi, p, ucat, dt, st, coeff, now_ms()});
  *ticket = t;
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_server_progress(vfs_pp_server* sv, int* tickets, vfs_pp_stats* stats, int cap, int* ndone) {
  ABI_BEGIN
  if (!sv || !ndone || cap < 0) return VFS_PP_EINVAL;
  *ndone = sv->cf ? sv->progress(*sv->cf, tickets, stats, cap) : sv->progress(*sv->cd, tickets, stats, cap);
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_server_expired(vfs_pp_server* sv, int* tickets, int cap, int* n) {
  if (!sv || !n || cap < 0 || (cap > 0 && !tickets)) return VFS_PP_EINVAL;
  int k = 0;
  while (k < cap && !sv->expired.empty()) {
    tickets[k++] = sv->expired.front();
    sv->expired.pop_front();
  }
  *n = k;
  return VFS_PP_OK;
}

int vfs_pp_server_status(vfs_pp_server* sv, int* queued, int* active, size_t* resident_bytes,
                         size_t* workspace_bytes_held) {
  if (!sv) return VFS_PP_EINVAL;
  int act = 0;
  for (int t : sv->slot_ticket)
    if (t >= 0) act++;
  if (queued) *queued = (int)sv->queue.size();
  if (active) *active = act;
  if (resident_bytes) *resident_bytes = sv->resident;
  if (workspace_bytes_held) *workspace_bytes_held = sv->ws_held;
  return VFS_PP_OK;
}

}  // extern "C"

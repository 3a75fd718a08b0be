// Shared host-side definitions of the pressure-projection engine: device memory
// arena, the host multigrid hierarchy, the device multigrid (MG<T>) and the
// per-case handle.  Used by the single-case engine and by the batched/served
// engine.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <type_traits>
#include <vector>

#include "pp_kernels.hpp"
#include "pp_stencil.hpp"
#include "vfs_pp.h"

namespace pp {

// --------------------------------------------------------------------- memory
class Arena {
 public:
  ~Arena() { release(); }
  template <typename T>
  T* alloc(size_t n) {
    void* p = nullptr;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    auto it = pool_.find(bytes);
    if (it != pool_.end()) {
      p = it->second;
      pool_.erase(it);
    } else {
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        throw StatusError(VFS_PP_ENOMEM, "hipMalloc of " + std::to_string(bytes) + " bytes failed");
      }
    }
    // every buffer starts at zero so no result can depend on earlier contents of
    // the device heap; the memset goes to the null stream and is waited for here
    hipError_t e = hipMemsetAsync(p, 0, bytes, nullptr);
    if (e == hipSuccess) e = hipStreamSynchronize(nullptr);
    if (e != hipSuccess) {
      (void)hipFree(p);
      throw StatusError(VFS_PP_EHIP, "zeroing a new buffer failed");
    }
    ptrs_.push_back({p, bytes});
    total_ += bytes;
    return static_cast<T*>(p);
  }
  // hand every buffer back for reuse by later allocations of the same size (the
  // caller no longer uses them)
  void recycle() {
    for (const Buf& b : ptrs_) pool_.emplace(b.bytes, b.p);
    ptrs_.clear();
    total_ = 0;
  }
  // free the recycled buffers that no allocation has taken back
  void trim() {
    for (auto& e : pool_) (void)hipFree(e.second);
    pool_.clear();
  }
  void release() {
    trim();
    for (const Buf& b : ptrs_) (void)hipFree(b.p);
    ptrs_.clear();
    total_ = 0;
  }
  size_t total() const { return total_; }

 private:
  struct Buf {
    void* p;
    size_t bytes;
  };
  std::vector<Buf> ptrs_;
  std::multimap<size_t, void*> pool_;
  size_t total_ = 0;
};

template <typename T>
inline void h2d(hipStream_t s, T* d, const T* h, size_t n) {
  HIPCHECK(hipMemcpyAsync(d, h, n * sizeof(T), hipMemcpyHostToDevice, s));
}
template <typename T>
inline void d2h(hipStream_t s, T* h, const T* d, size_t n) {
  HIPCHECK(hipMemcpyAsync(h, d, n * sizeof(T), hipMemcpyDeviceToHost, s));
}

// ------------------------------------------------------------- host hierarchy
// One level of the multigrid hierarchy while it is being built (device arrays).
struct SetupLevel {
  Dims d{};
  const double* g = nullptr;   // 9*n FP64 face coefficients
  const double* nv = nullptr;  // DMDA nvert of this level (with ghost planes)
  uint32_t* code = nullptr;    // n
  uint8_t* cls = nullptr;      // n, CLS_NONFLUID where nvert > 0.1
  double gmax = 0.;            // largest |g|
  int cmask = 0;               // axes coarsened towards the next level
  int line = 0;                // the level is smoothed with y-lines
};

inline Dmda dmda_of(const Dims& d) { return Dmda{d.nx + 2, d.ny + 2, d.nz + 2, {d.per[0], d.per[1], d.per[2]}}; }

// ----------------------------------------------------------- device multigrid
struct MGBase {
  virtual ~MGBase() = default;
  virtual void vcycle(hipStream_t s) = 0;  // f0 -> out
  bool tail = true;    // run the bottom point level and the dense solve as one workgroup (k_coarse_tail)
  bool use_fd = true;  // the finest level's fused first steps read fd0() (its writers stored it)
  virtual bool line0() const = 0;           // the finest level uses the y-line smoother
  virtual void* f0() = 0;
  // f0 * D^-1 of the finest level, stored by the writers of f0 for the fused first
  // Chebyshev steps (nullptr when the finest level is smoothed with y-lines), and
  // that level's D^-1
  virtual void* fd0() = 0;
  virtual const void* dinv0() const = 0;
  virtual void set_out(void* out) = 0;
  virtual std::string describe() const = 0;
  virtual void time_sweep(hipStream_t s) = 0;
  // ms per finest-level sweep when `copies` copies of the level's operands are
  // swept in turn (together larger than the 256 MB last-level cache, so every
  // sweep streams its operands from HBM)
  virtual double time_sweep_rotating(hipStream_t s, int reps, int copies, bool floor = false) = 0;
  virtual int nlev() const = 0;
  // level l: dimensions, smoother, code words; its face coefficients as FP64 into
  // g64 (device, 9 n)
  virtual void level_export(hipStream_t s, int l, Dims* d, int* line, const uint32_t** code, double* g64) const = 0;
  // the coarsest level's cell count and its dense inverse (column-major) as doubles
  virtual int coarse_cells() const = 0;
  virtual std::vector<double> coarse_inverse_export(hipStream_t s) const = 0;
  // a hierarchy that shares this one's coefficients, masks, Chebyshev parameters
  // and coarsest inverse and owns fresh work vectors (allocated from A)
  virtual std::unique_ptr<MGBase> clone_work(Arena& A) const = 0;
};

template <typename T>
struct MG : MGBase {
  struct Lev {
    Dims d{};
    T* g = nullptr;            // face coefficients, or
    _Float16* gh = nullptr;    // FP16 coefficients / gs (the levels of the FP32 hierarchy)
    T gs = T(1);
    const uint32_t* code = nullptr;
    const uint8_t* cls = nullptr;
    T *f = nullptr, *xa = nullptr, *xb = nullptr, *dv = nullptr, *dinv = nullptr;
    T* lfac = nullptr;  // y-line Thomas factors [a | cp | inv] when line smoothing
    int line = 0;
    std::vector<T> c1, c2;  // Chebyshev coefficients per smoothing step
    double lam = 0.;
    LevelView<T> view() const { return LevelView<T>{d, g, gh, code, cls, gs}; }
  };
  std::vector<Lev> lev;
  T* scr = nullptr;
  T* fd = nullptr;  // finest level's f D^-1 (point smoothing only)
  T* cinv = nullptr;
  uint16_t* tord = nullptr;  // thread-to-cell order of the coarse tail's level (tail_order)
  int ncoarse = 0, ncomp = 0;
  T* out = nullptr;
  int degree = 2;

  void* f0() override { return lev[0].f; }
  void* fd0() override { return fd; }
  const void* dinv0() const override { return lev.empty() ? nullptr : lev[0].dinv; }
  void set_out(void* o) override { out = static_cast<T*>(o); }

  // Chebyshev coefficients on [lo*lam, hi*lam]
  static void cheb(double lam, int deg, double lo, double hi, std::vector<T>& c1, std::vector<T>& c2) {
    const double a = lo * lam, b = hi * lam;
    const double theta = 0.5 * (b + a), delta = 0.5 * (b - a), sigma = theta / delta;
    double rho = 1. / sigma;
    c1.assign(deg, T(0));
    c2.assign(deg, T(0));
    c2[0] = T(1. / theta);
    for (int k = 1; k < deg; k++) {
      const double rn = 1. / (2. * sigma - rho);
      c1[k] = T(rn * rho);
      c2[k] = T(2. * rn / delta);
      rho = rn;
    }
  }

  // `degree` Chebyshev steps; from a zero start the point smoother fuses the
  // first two steps into one pass
  void smooth(hipStream_t s, Lev& L, bool zero_start, T*& x, T* final_out, const T* fdl = nullptr) {
    int k = 0;
    while (k < degree) {
      int mode = SM_NEXT, adv = 1;
      if (zero_start && k == 0) {
        if (!L.line && degree >= 2) { mode = SM_FIRST2; adv = 2; }
        else mode = SM_ZERO;
      }
      const int last = k + adv - 1;  // the last step this launch performs
      T* xin = x;
      T* other = (x == L.xa) ? L.xb : L.xa;
      T* xout = other;
      if (last == degree - 1 && final_out) xout = final_out;
      // point levels: the previous iterate lives in the other buffer (which the next
      // iterate overwrites); a step from zero leaves its x_prev where x was
      T* xprev = (mode == SM_NEXT) ? other : xin;
      const T c1 = (mode == SM_ZERO || last == 0) ? T(0) : L.c1[last];
      launch_smooth<T>(s, L.view(), L.lfac, mode, L.f, L.dinv, xin, xprev, xout, L.dv, scr, c1, L.c2[last],
                       L.c2[0], mode == SM_FIRST2 ? fdl : nullptr);
      x = xout;
      k += adv;
    }
  }

  void vcycle(hipStream_t s) override { cycle(s, 0, out); }
  bool line0() const override { return !lev.empty() && lev[0].line; }

  // the level just above the dense one, when it can run as the coarse tail
  bool tail_level(size_t l) const {
    const Lev& L = lev[l];
    return std::is_same<T, float>::value && tail && l + 2 == lev.size() && !L.line && !L.lfac && L.d.n() <= (size_t)TAIL_MAX_CELLS &&
           ncoarse <= TAIL_MAX_CELLS && degree <= TAIL_MAX_DEGREE && (int)L.c1.size() >= degree;
  }

  // returns the buffer holding this level's result
  T* cycle(hipStream_t s, size_t l, T* final_out) {
    Lev& L = lev[l];
    if (l + 1 == lev.size()) {
      T* x = final_out ? final_out : L.xa;
      launch_dense_solve<T>(s, ncoarse, cinv, L.f, x);
      return x;
    }
    if (tail_level(l)) {
      TailArgs<T> a{};
      a.L = L.view();
      a.f = L.f;
      a.dinv = L.dinv;
      a.xout = final_out ? final_out : L.xa;
      a.dc = lev[l + 1].d;
      a.clsc = lev[l + 1].cls;
      a.cinv = cinv;
      a.order = tord;
      a.degree = degree;
      a.first2 = degree >= 2 ? 1 : 0;
      for (int k = 0; k < degree; k++) {
        a.c1[k] = L.c1[k];
        a.c2[k] = L.c2[k];
      }
      launch_coarse_tail<T>(s, a);
      return a.xout;
    }
    T* x = L.xa;
    smooth(s, L, true, x, nullptr, l == 0 && use_fd ? fd : nullptr);
    Lev& C = lev[l + 1];
    launch_resid_restrict<T>(s, L.view(), L.f, x, C.d, C.cls, C.f);
    T* xc = cycle(s, l + 1, nullptr);
    launch_prolong_add<T>(s, L.view(), C.d, C.cls, xc, x);
    smooth(s, L, false, x, final_out);
    return x;
  }

  std::string describe() const override {
    std::ostringstream o;
    for (size_t l = 0; l < lev.size(); l++) {
      o << "level " << l << ": " << lev[l].d.nx << "x" << lev[l].d.ny << "x" << lev[l].d.nz;
      if (l + 1 == lev.size()) o << " dense(" << ncoarse << ", components " << ncomp << ")";
      else o << (lev[l].line ? " y-line" : " point") << " chebyshev deg " << degree << " lambda " << lev[l].lam;
      if (l + 1 < lev.size() && tail_level(l)) o << " (coarse tail: with the dense solve in one workgroup)";
      o << "\n";
    }
    return o.str();
  }

  int nlev() const override { return (int)lev.size(); }
  void level_export(hipStream_t s, int l, Dims* d, int* line, const uint32_t** code, double* g64) const override {
    const Lev& L = lev[l];
    *d = L.d;
    *line = L.line;
    *code = L.code;
    if (!g64) return;
    const size_t n9 = 9 * L.d.n();
    if (L.gh) launch_half_to_double(s, n9, L.gh, (double)L.gs, g64);
    else if constexpr (std::is_same<T, double>::value)
      HIPCHECK(hipMemcpyAsync(g64, L.g, n9 * sizeof(double), hipMemcpyDeviceToDevice, s));
    else launch_convert<T, double>(s, n9, L.g, g64);
  }

  int coarse_cells() const override { return ncoarse; }
  std::vector<double> coarse_inverse_export(hipStream_t s) const override {
    std::vector<T> c((size_t)ncoarse * ncoarse);
    d2h(s, c.data(), cinv, c.size());
    HIPCHECK(hipStreamSynchronize(s));
    return std::vector<double>(c.begin(), c.end());
  }

  std::unique_ptr<MGBase> clone_work(Arena& A) const override {
    auto c = std::make_unique<MG<T>>(*this);
    c->scr = A.alloc<T>(2 * (lev.empty() ? 1 : lev[0].d.n()));
    if (fd) c->fd = A.alloc<T>(lev[0].d.n());
    for (auto& L : c->lev) {
      const size_t n = L.d.n();
      L.f = A.alloc<T>(n);
      L.xa = A.alloc<T>(n);
      L.xb = A.alloc<T>(n);
      L.dv = A.alloc<T>(n);
    }
    c->out = nullptr;
    return c;
  }

  void time_sweep(hipStream_t s) override {
    Lev& L = lev[0];
    launch_smooth<T>(s, L.view(), L.lfac, SM_NEXT, L.f, L.dinv, L.xa, L.xb, L.xb, L.dv, scr,
                     L.c1.size() > 1 ? L.c1[1] : T(0), L.c2[0], T(0));
  }

  double time_sweep_rotating(hipStream_t s, int reps, int copies, bool floor = false) override {
    const Lev& L0 = lev[0];
    const size_t n = L0.d.n();
    Arena tmp;
    auto dup = [&](auto* src, size_t len) {
      using E = std::remove_cv_t<std::remove_pointer_t<decltype(src)>>;
      if (!src) return static_cast<E*>(nullptr);
      E* d = tmp.alloc<E>(len);
      HIPCHECK(hipMemcpyAsync(d, src, len * sizeof(E), hipMemcpyDeviceToDevice, s));
      return d;
    };
    std::vector<Lev> cp(copies, L0);
    std::vector<T*> scrs(copies);
    for (auto& C : cp) {
      C.g = dup(L0.g, 9 * n);
      C.gh = dup(L0.gh, 9 * n);
      C.code = dup(L0.code, n);
      C.cls = dup(L0.cls, n);
      C.f = dup(L0.f, n);
      C.xa = dup(L0.xa, n);
      C.xb = dup(L0.xb, n);
      C.dv = dup(L0.dv, n);
      C.dinv = dup(L0.dinv, n);
      C.lfac = dup(L0.lfac, 3 * n);
    }
    for (auto& sc : scrs) sc = tmp.alloc<T>(2 * n);
    auto sweep = [&](int q) {
      Lev& C = cp[q];
      if (floor)
        launch_stream_floor<T>(s, C.view(), C.f, C.xa, C.xb, scrs[q]);
      else
        launch_smooth<T>(s, C.view(), C.lfac, SM_NEXT, C.f, C.dinv, C.xa, C.xb, C.xb, C.dv, scrs[q],
                         C.c1.size() > 1 ? C.c1[1] : T(0), C.c2[0], T(0));
    };
    for (int q = 0; q < copies; q++) sweep(q);
    hipEvent_t e0, e1;
    HIPCHECK(hipEventCreate(&e0));
    HIPCHECK(hipEventCreate(&e1));
    HIPCHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) sweep(r % copies);
    HIPCHECK(hipEventRecord(e1, s));
    HIPCHECK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return (double)ms / reps;
  }
};

}  // namespace pp

struct vfs_pp_handle {
  vfs_grid grid{};
  vfs_pp_opts opts{};
  pp::Dmda m{};
  pp::Dims d{};
  hipStream_t s = nullptr;
  pp::Arena ar;   // buffers sized by the grid alone
  pp::Arena geo;  // buffers of the current geometry (hierarchy, face list), renewed by each set-up
  pp::Arena setup_tmp;  // set-up scratch, kept between set-ups of a moving body
  // device bytes the handle holds: grid-sized buffers, the geometry's and set-up scratch
  size_t resident_bytes() const { return ar.total() + geo.total() + setup_tmp.total(); }
  double* hstage = nullptr;   // device copies of vfs_pp_project's host arrays (8 M doubles, from ar)
  int batch_refs = 0;         // batches / server slots the handle is bound to (0 or 1)
  int immersed_ref = 0;       // the caller's `immersed` as given (grid.immersed holds the VolumeFlux variant)
  bool have_geom = false;     // metrics, masks and hierarchy are in place
  bool have_metrics = false;  // face coefficients and cell metrics are in place
  // fine-level resident data
  double* g64 = nullptr;
  uint32_t* code = nullptr;
  uint8_t* cls = nullptr;
  double* met9 = nullptr;
  double* aj = nullptr;
  double* nvert = nullptr;
  pp::IbFace* ibf = nullptr;
  size_t nibf = 0;
  double n_lt = 0.;
  // vectors
  double *x = nullptr, *b = nullptr, *r = nullptr;
  double* partial = nullptr;
  size_t partial_cap = 0;  // doubles in `partial`
  size_t stencil_blocks = 0;  // blocks of one cell-grid launch
  double* red = nullptr;   // device scratch scalars (>= 256)
  double* snap = nullptr;  // DMDA vector scratch for Contra2Cart
  double* V = nullptr;
  void* Z = nullptr;
  int mrest = 0;
  bool external_ws = false;  // V and Z are attached per solve by a server
  // divergence part of the stopping rule for the current solve: max |aj div u|
  // after projection = div_fac * max |aj r|; stop when that is <= div_goal
  double div_fac = 0., div_goal = 0.;
  std::unique_ptr<pp::MGBase> mg;
  std::vector<pp::SetupLevel> setup;  // the hierarchy while set-up runs
  double* hpin = nullptr;  // pinned host scratch
  double* hpin_dev = nullptr;  // the same memory as seen by kernels
  std::string desc;
  // A solve clone (made by clone_for_solves) reads this handle's geometry, masks,
  // IB face list and hierarchy coefficients and owns only its solve vectors; the
  // owner must outlive it and keep its geometry unchanged meanwhile.
  const vfs_pp_handle* geometry_of = nullptr;
};

namespace pp {
// Slots of red[] / hpin[]: Krylov scalars at 0..2 (norm^2, norm, 1/norm), the
// two Gram-Schmidt coefficient sets at 8.. and 72.., the small solve's y at
// 200..; outside the Krylov loop:
constexpr int RED_IBF = 300;   // VolumeFlux: 4 values
constexpr int RED_MEAN = 310;  // Remove_Nullspace sum
constexpr int RED_B2 = 320;    // ||b||^2
constexpr int RED_DIV = 330;   // max |aj div u|
constexpr int RED_DIVEST = 340;  // max |aj r| at the last true-residual check
constexpr int RED_SETUP = 350;   // set-up statistics: 4 values
constexpr int RED_LAMBDA = 1024;  // set-up power iterations: 12 norms per level (hpin only)
constexpr int RED_NUINV2 = 400;  // 1 / |U_l|^2 of the lazily normalised basis (up to 64)

// engine functions shared with the batched engine (pp_engine.hip)
double now_ms();
void ensure_geom(vfs_pp_handle* h);
void need_partials(vfs_pp_handle* h, size_t nvec, size_t nblk);
void solve_prologue_host(vfs_pp_handle* h, double dt, double st, double coeff);
void solve_prologue(vfs_pp_handle* h, double* ucont, double dt, double st, double coeff);
void solve_start_stats(vfs_pp_handle* h, vfs_pp_stats* S);
void solve_start(vfs_pp_handle* h, double* phi, vfs_pp_stats* S);
void solve_epilogue(vfs_pp_handle* h, double* phi, vfs_pp_stats* S);
void update_pressure(vfs_pp_handle* h, double* p, double* phi);
void projection(vfs_pp_handle* h, double* ucont, const double* phi, double* ucat, double dt, double st,
                double coeff);
void divergence_launch(vfs_pp_handle* h, const double* ucont);
// A handle that solves on src's resident geometry with its own vectors, reduction
// scratch, pinned scalars, stream and multigrid work arrays (no Krylov basis: it
// is attached per solve, as for a server case).  Internal to the served form.
vfs_pp_handle* clone_for_solves(const vfs_pp_handle* src);
// A handle bound to a batch or a server slot runs on the batch's stream and its
// buffers are referenced by the batch's work lists and captured graph: the
// single-case entry points refuse it until the batch lets it go.
inline void ensure_unbound(const vfs_pp_handle* h) {
  if (h->batch_refs > 0) throw StatusError(VFS_PP_ESTATE, "the handle belongs to a batch; destroy the batch first");
}
}  // namespace pp

// C-ABI wrappers: exceptions become status codes
#define ABI_BEGIN try {
#define ABI_END                                     \
  }                                                 \
  catch (const pp::StatusError& e) {                \
    std::fprintf(stderr, "vfs_pp: %s\n", e.what()); \
    return e.code;                                  \
  }                                                 \
  catch (const pp::HipError& e) {                   \
    std::fprintf(stderr, "vfs_pp: %s\n", e.what()); \
    return VFS_PP_EHIP;                             \
  }                                                 \
  catch (const std::exception& e) {                 \
    std::fprintf(stderr, "vfs_pp: %s\n", e.what()); \
    return VFS_PP_EINVAL;                           \
  }

namespace pp {
// bytes of the Krylov basis (V in FP64, Z in the multigrid precision)
inline size_t workspace_bytes(const vfs_pp_handle* h) {
  const size_t n = h->d.n();
  return (size_t)(h->mrest + 1) * n * sizeof(double) +
         (size_t)h->mrest * n * (h->opts.precision == 1 ? sizeof(float) : sizeof(double));
}
}  // namespace pp

namespace pp {
// True residual r = b - A x in FP64 with ||r||^2 into red[0] / hpin[0] and, when
// the divergence rule is on, max |aj r| into hpin[RED_DIVEST]; launches only.
// With x_zero (x is known to be 0) r is b itself: no operator pass, ||b||^2 is
// the RHS's own reduction (red / hpin[RED_B2]), and the divergence estimate reads
// b.  true_residual_src / _norm2 name the vector and the device norm^2 to use.
inline void launch_true_residual(vfs_pp_handle* h, hipStream_t s, bool x_zero = false) {
  int nb = 0;
  if (!x_zero) {
    launch_residual64<double>(s, h->d, h->g64, h->code, h->b, h->x, h->r, h->partial, &nb);
    launch_finalize(s, h->partial, 1, nb, h->red, 0, h->hpin_dev);
  }
  if (h->div_fac > 0.) {
    launch_divest(s, h->d, h->code, h->aj, x_zero ? h->b : h->r, h->partial, &nb);
    launch_finalize(s, h->partial, 1, nb, h->red + RED_DIVEST, 1, h->hpin_dev + RED_DIVEST);
  }
}
inline const double* true_residual_src(const vfs_pp_handle* h, bool x_zero) { return x_zero ? h->b : h->r; }
inline const double* true_residual_norm2(const vfs_pp_handle* h, bool x_zero) {
  return x_zero ? h->red + RED_B2 : h->red;
}
// ||r||^2 on the host after the stream sync
inline double true_residual_host_norm2(const vfs_pp_handle* h, bool x_zero) {
  return x_zero ? h->hpin[RED_B2] : h->hpin[0];
}

// Stopping rule at a true-residual check (after the stream sync): ||r|| <= tgt
// and, with the divergence rule on, div_fac * max |aj r| <= div_goal.  When only
// the divergence part fails, tgt is lowered in proportion and the solve goes on;
// at the FP64 rounding floor it stops and the post-projection check reports.
inline bool residual_verdict(vfs_pp_handle* h, double beta, double bnorm, double& tgt) {
  if (!(beta <= tgt)) return false;
  if (h->div_fac <= 0.) return true;
  const double dmax = h->div_fac * h->hpin[RED_DIVEST];
  if (dmax <= h->div_goal) return true;
  const double floor_t = 1e-14 * bnorm;
  if (beta <= floor_t) return true;
  tgt = std::max(floor_t, 0.5 * beta * h->div_goal / dmax);
  return false;
}
}  // namespace pp

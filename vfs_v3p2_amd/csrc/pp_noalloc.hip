// Divergence check that allocates nothing: vfs_pp_divergence_noalloc.
//
// The handle-based check (vfs_pp_divergence_device) reduces its per-block maxima
// in the handle's partial-sum buffer, which the set-up allocates.  Callers that
// cannot allocate device memory at all (work recorded into a HIP graph, tight
// serving loops, harnesses that forbid hipMalloc) also cannot create a handle, so
// this variant takes the reference's own arrays (lUcont, lAj, lNvert, main.c:1244)
// and reduces into a fixed device array of this file, `scratch`.
//
// Host code never passes `scratch` itself to a kernel: in host code the name is a
// stand-in, not a device address.  The device address is looked up once per
// device with hipGetSymbolAddress (each device holds its own copy of the array)
// and kept in `g_scratch`.
//
// Every call on a device shares that device's array.  Calls that overlap in time
// on one device (two host threads, or streams) race on it; calls that follow one
// another do not, since each launch writes every partial it reads.
#include <mutex>

#include "pp_engine.hpp"

namespace pp {

// Block partials of the grid-stride reduction: 4096 doubles (32 KiB per device).
constexpr int kScratchPartials = 4096;

static __device__ __align__(256) double scratch[kScratchPartials];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the first
// call there.
double* scratch_ptr() {
  int dev = -1;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw StatusError(VFS_PP_EINVAL, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) throw HipError(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_scratch[dev] = static_cast<double*>(p);
  }
  return g_scratch[dev];
}

// max |aj div u| (main.c:1283-1285) over the interior cells whose 7-point nvert
// sum is not above 0.1, one partial per block; cells in flat (k, j, i) order,
// grid-stride.  Same operand order as the handle-based kernel, whose code word
// carries the same nvert rule, so the two give the same bits.
__global__ __launch_bounds__(256) void k_divergence_raw(Dmda m, Dims d, const double* __restrict__ u,
                                                        const double* __restrict__ aj,
                                                        const double* __restrict__ nv,
                                                        double* __restrict__ partial) {
#pragma clang fp contract(off)
  __shared__ double lds[4];
  const size_t n = d.n(), sy = (size_t)m.mx, sz = (size_t)m.mx * m.my;
  double mv = 0.;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)gridDim.x * 256) {
    const int i = (int)(t % d.nx);
    const size_t r = t / d.nx;
    const int j = (int)(r % d.ny), k = (int)(r / d.ny);
    const size_t p = m.idx(i + 1, j + 1, k + 1);
    const double s7 = nv[p] + nv[p + sz] + nv[p - sz] + nv[p + sy] + nv[p - sy] + nv[p + 1] + nv[p - 1];
    if (s7 > 0.1) continue;
    const double v =
        fabs((u[3 * p] - u[3 * (p - 1)] + u[3 * p + 1] - u[3 * (p - sy) + 1] + u[3 * p + 2] - u[3 * (p - sz) + 2]) *
             aj[p]);
    mv = fmax(mv, v);
  }
  mv = block_max(mv, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = mv;
}

}  // namespace
}  // namespace pp

using namespace pp;

extern "C" int vfs_pp_divergence_noalloc(const vfs_grid* g, const double* ucont, const double* aj,
                                         const double* nvert, double* maxdiv) {
  ABI_BEGIN
  if (!g || !ucont || !aj || !nvert || !maxdiv) return VFS_PP_EINVAL;
  if (g->nx < 1 || g->ny < 1 || g->nz < 1) return VFS_PP_EINVAL;
  if ((long long)(g->nx + 2) * (g->ny + 2) * (g->nz + 2) >= (1LL << 29)) {
    std::fprintf(stderr, "vfs_pp_divergence_noalloc: grid %dx%dx%d exceeds the 2^29-point DMDA limit\n", g->nx,
                 g->ny, g->nz);
    return VFS_PP_EINVAL;
  }
  const Dims d{g->nx, g->ny, g->nz, {g->periodic[0] ? 1 : 0, g->periodic[1] ? 1 : 0, g->periodic[2] ? 1 : 0}};
  const Dmda m{g->nx + 2, g->ny + 2, g->nz + 2, {d.per[0], d.per[1], d.per[2]}};
  double* part = scratch_ptr();
  const size_t want = (d.n() + 255) / 256;
  const int nblk = (int)std::min<size_t>(want, kScratchPartials);
  hipLaunchKernelGGL(k_divergence_raw, dim3(nblk), dim3(256), 0, nullptr, m, d, ucont, aj, nvert, part);
  HIPCHECK(hipGetLastError());
  launch_finalize(nullptr, part, 1, nblk, maxdiv, 1, nullptr);
  HIPCHECK(hipStreamSynchronize(nullptr));
  return VFS_PP_OK;
  ABI_END
}

// Common definitions for the pressure-projection engine (host and device).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

#define HD __host__ __device__
#define HDI __host__ __device__ __forceinline__

namespace pp {

// poisson_threshold, poisson.c:16
constexpr double TH = 0.1;

struct HipError : std::runtime_error {
  hipError_t err;
  HipError(hipError_t e, const char* what, const char* file, int line)
      : std::runtime_error(std::string(what) + " failed at " + file + ":" + std::to_string(line) + ": " +
                           hipGetErrorString(e)),
        err(e) {}
};

struct StatusError : std::runtime_error {
  int code;
  StatusError(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHECK(x)                                                  \
  do {                                                               \
    hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) throw ::pp::HipError(e_, #x, __FILE__, __LINE__); \
  } while (0)

// Interior-cell grid of one multigrid level (0-based interior indices).
struct Dims {
  int nx, ny, nz;
  int per[3];
  HD size_t n() const { return (size_t)nx * ny * nz; }
  HD size_t idx(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }
};

// Neighbour index with periodic wrap; clamps on non-periodic axes (those
// neighbours only ever carry a zero coefficient).
HD inline int wrap_m(int i, int n, int per) { return i > 0 ? i - 1 : (per ? n - 1 : 0); }
HD inline int wrap_p(int i, int n, int per) { return i < n - 1 ? i + 1 : (per ? 0 : n - 1); }

constexpr int BLOCK = 256;

// Workgroups go to the 8 XCDs round-robin by linear id (workgroup w runs on XCD
// w % 8), and each XCD has its own L2.  xcd_block() renumbers the grid so that
// XCD x runs one contiguous range of tiles in (i, j, k) order: tiles adjacent in
// j and k, which re-read each other's halo rows and lower-face coefficients, then
// run close together in time on one L2.  Any renumbering is a bijection of the
// grid, so results do not depend on it.
__device__ inline uint3 xcd_block() {
  const uint32_t nbx = gridDim.x, nby = gridDim.y, pl = nbx * nby;
  const uint32_t nb = pl * gridDim.z;
  const uint32_t w = blockIdx.x + nbx * (blockIdx.y + nby * blockIdx.z);
  const uint32_t q = nb >> 3, r = nb & 7u, x = w & 7u, l = w >> 3;
  const uint32_t t = x * q + (x < r ? x : r) + l;
  return make_uint3(t % nbx, (t % pl) / nbx, t / pl);
}
// (i, j, k) of this thread for the 64 x 4 x 1 cell tiles of a cell_grid() launch.
#define CELL_IJK(i, j, k)                                                                          \
  const uint3 xb_ = xcd_block();                                                                   \
  const int i = (int)xb_.x * 64 + (int)threadIdx.x, j = (int)xb_.y * 4 + (int)threadIdx.y,        \
            k = (int)xb_.z

// wave64 reductions
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// Block-wide sum of one double; result valid in thread 0.  blockDim.x*y*z == 256.
__device__ inline double block_sum(double v, double* lds /* >= 4 */) {
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  v = wave_sum(v);
  if ((t & 63) == 0) lds[t >> 6] = v;
  __syncthreads();
  double r = 0.;
  if (t == 0) r = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  __syncthreads();
  return r;
}
__device__ inline double block_max(double v, double* lds) {
  const int t = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
  v = wave_max(v);
  if ((t & 63) == 0) lds[t >> 6] = v;
  __syncthreads();
  double r = 0.;
  if (t == 0) r = fmax(fmax(lds[0], lds[1]), fmax(lds[2], lds[3]));
  __syncthreads();
  return r;
}

}  // namespace pp

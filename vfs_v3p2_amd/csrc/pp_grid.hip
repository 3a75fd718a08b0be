// Grid-file ingestion (init.c:251-389): the three grid-file forms the reference
// reads -- text grid.dat (block count, then per block "IM JM KM" and all x, all
// y, all z values with i fastest), the binary form selected by -binary 1 (the
// same sequence as native int32 and float64 words), and xyz.dat selected by
// -xyz 1 (one Cartesian block: IM, JM, KM lines whose first, second and third
// columns give X[i], Y[j], Z[k]).  Values become coordinates as value / cl *
// L_dim (init.c:339-366).
//
// The file is parsed on the host (text values in parallel over byte ranges,
// each with strtod, the conversion fscanf's %le performs); the node values then
// go to the device as planar arrays and one kernel lays them out as the DMDA
// coordinate array coor[KM+1][JM+1][IM+1][3].  The plane of index IM (JM, KM)
// that the reference's DMDA carries beyond the last node is never read by the
// file or by FormMetrics; it is filled by linear extrapolation, i then j then k,
// so that every cell of the array holds a finite, non-degenerate point.
#include <algorithm>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "pp_engine.hpp"

namespace pp {

namespace {

[[noreturn]] void io_error(const std::string& path, const std::string& what) {
  throw StatusError(VFS_PP_EIO, "file " + path + ": " + what);
}

inline bool is_ws(char c) { return c == ' ' || c == '\n' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

std::string slurp(const std::string& path) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) io_error(path, std::string("cannot open (") + std::strerror(errno) + ")");
  std::string s;
  if (std::fseek(f, 0, SEEK_END) == 0) {
    const long n = std::ftell(f);
    if (n > 0) {
      s.resize((size_t)n);
      std::fseek(f, 0, SEEK_SET);
      if (std::fread(&s[0], 1, (size_t)n, f) != (size_t)n) {
        std::fclose(f);
        io_error(path, "short read");
      }
    }
  }
  std::fclose(f);
  return s;  // std::string keeps a terminating NUL after size()
}

// text cursor: integers as fscanf %i reads them (base prefix honoured)
struct Text {
  const std::string& path;
  const char* p;
  const char* e;
  long next_int() {
    while (p < e && is_ws(*p)) p++;
    if (p >= e) io_error(path, "unexpected end of file in a header");
    char* q = nullptr;
    errno = 0;
    const long v = std::strtol(p, &q, 0);
    if (q == p || errno) io_error(path, "malformed integer in a header");
    p = q;
    return v;
  }
  double next_double() {
    while (p < e && is_ws(*p)) p++;
    if (p >= e) io_error(path, "unexpected end of file");
    char* q = nullptr;
    const double v = std::strtod(p, &q);
    if (q == p || (q < e && !is_ws(*q))) io_error(path, "malformed number");
    p = q;
    return v;
  }
};

int host_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  return (int)std::max(1u, std::min(h ? h : 1u, 16u));
}

// Parse the next n whitespace-separated numbers after t.p into out[0..n) and
// advance t.p past the last one.  The byte range is cut into pieces at
// whitespace; each piece counts the tokens that start in it, a prefix sum gives
// every piece its first token index, and the pieces are parsed in parallel.
void parse_numbers(Text& t, size_t n, double* out) {
  if (n == 0) return;
  const char* p = t.p;
  const char* e = t.e;
  const size_t len = (size_t)(e - p);
  const int T = len < (size_t(1) << 20) ? 1 : host_threads();
  std::vector<const char*> cut(T + 1);
  cut[0] = p;
  cut[T] = e;
  for (int q = 1; q < T; q++) {
    const char* c = p + len * q / T;
    if (c < cut[q - 1]) c = cut[q - 1];
    while (c < e && !is_ws(*c)) c++;
    cut[q] = c;
  }
  std::vector<size_t> count(T, 0), first(T + 1, 0);
  auto run = [&](auto&& body) {
    std::vector<std::thread> th;
    for (int q = 1; q < T; q++) th.emplace_back(body, q);
    body(0);
    for (auto& x : th) x.join();
  };
  run([&](int q) {
    size_t c = 0;
    bool in = false;
    for (const char* s = cut[q]; s < cut[q + 1]; s++) {
      const bool w = is_ws(*s);
      if (!w && !in) c++;
      in = !w;
    }
    count[q] = c;
  });
  for (int q = 0; q < T; q++) first[q + 1] = first[q] + count[q];
  if (first[T] < n) io_error(t.path, "fewer values than the header announces");
  std::vector<const char*> stop(T, nullptr);
  std::vector<int> bad(T, 0);
  run([&](int q) {
    size_t idx = first[q];
    const char* s = cut[q];
    const char* end = cut[q + 1];
    while (idx < n) {
      while (s < end && is_ws(*s)) s++;
      if (s >= end) break;
      char* r = nullptr;
      const double v = std::strtod(s, &r);
      if (r == s || (r < e && !is_ws(*r))) {
        bad[q] = 1;
        return;
      }
      out[idx++] = v;
      s = r;
      if (idx == n) stop[q] = s;
    }
  });
  for (int q = 0; q < T; q++)
    if (bad[q]) io_error(t.path, "malformed number");
  for (int q = 0; q < T; q++)
    if (stop[q]) t.p = stop[q];
}

struct GridBlock {
  int nblocks = 0, im = 0, jm = 0, km = 0;
  int separable = 0;        // xyz.dat: raw holds X[IM], Y[JM], Z[KM]
  std::vector<double> raw;  // else x[KM][JM][IM], y[..], z[..] (planar)
};

void check_dims(const std::string& path, long im, long jm, long km) {
  if (im < 2 || jm < 2 || km < 2 || im > (1 << 20) || jm > (1 << 20) || km > (1 << 20) ||
      (double)im * jm * km > 4e9)
    io_error(path, "implausible node counts " + std::to_string(im) + " " + std::to_string(jm) + " " +
                       std::to_string(km));
}

// Read the header (and, when want_values, the node values) of block `block`.
GridBlock read_block(const std::string& path, int format, int block, bool want_values) {
  GridBlock g;
  const std::string s = slurp(path);
  if (format == VFS_PP_GRID_BINARY) {
    size_t off = 0;
    auto rd_int = [&]() {
      int32_t v;
      if (off + 4 > s.size()) io_error(path, "unexpected end of file in a header");
      std::memcpy(&v, s.data() + off, 4);
      off += 4;
      return (long)v;
    };
    const long nb = rd_int();
    if (nb < 1) io_error(path, "block count " + std::to_string(nb));
    g.nblocks = (int)nb;
    if (block < 0 || block >= g.nblocks) throw StatusError(VFS_PP_EINVAL, "grid block index out of range");
    for (int b = 0; b <= block; b++) {
      const long im = rd_int(), jm = rd_int(), km = rd_int();
      check_dims(path, im, jm, km);
      const size_t bytes = (size_t)3 * im * jm * km * sizeof(double);
      if (off + bytes > s.size()) io_error(path, "fewer values than the header announces");
      if (b == block) {
        g.im = (int)im, g.jm = (int)jm, g.km = (int)km;
        if (want_values) {
          g.raw.resize(bytes / sizeof(double));
          std::memcpy(g.raw.data(), s.data() + off, bytes);
        }
      }
      off += bytes;
    }
    return g;
  }
  Text t{path, s.data(), s.data() + s.size()};
  if (format == VFS_PP_GRID_XYZ) {
    g.nblocks = 1;
    if (block != 0) throw StatusError(VFS_PP_EINVAL, "an xyz grid file holds one block");
    const long im = t.next_int(), jm = t.next_int(), km = t.next_int();
    check_dims(path, im, jm, km);
    g.im = (int)im, g.jm = (int)jm, g.km = (int)km;
    g.separable = 1;
    if (want_values) {
      g.raw.resize((size_t)im + jm + km);
      const long n[3] = {im, jm, km};
      size_t o = 0;
      for (int ax = 0; ax < 3; ax++)
        for (long q = 0; q < n[ax]; q++) {
          double v[3];
          for (int c = 0; c < 3; c++) v[c] = t.next_double();
          g.raw[o++] = v[ax];
        }
    }
    return g;
  }
  if (format != VFS_PP_GRID_TEXT) throw StatusError(VFS_PP_EINVAL, "unknown grid file format");
  const long nb = t.next_int();
  if (nb < 1) io_error(path, "block count " + std::to_string(nb));
  g.nblocks = (int)nb;
  if (block < 0 || block >= g.nblocks) throw StatusError(VFS_PP_EINVAL, "grid block index out of range");
  std::vector<double> skip;
  for (int b = 0; b <= block; b++) {
    const long im = t.next_int(), jm = t.next_int(), km = t.next_int();
    check_dims(path, im, jm, km);
    const size_t n = (size_t)3 * im * jm * km;
    if (b == block) {
      g.im = (int)im, g.jm = (int)jm, g.km = (int)km;
      if (want_values) {
        g.raw.resize(n);
        parse_numbers(t, n, g.raw.data());
      }
    } else {
      skip.resize(n);
      parse_numbers(t, n, skip.data());
    }
  }
  return g;
}

// node value of axis c at node (i, j, k), scaled as init.c:339-366 does
struct NodeSrc {
  const double* raw;
  int im, jm, km, separable;
  double cl, ldim;
  __device__ double node(int i, int j, int k, int c) const {
    double v;
    if (separable)
      v = c == 0 ? raw[i] : (c == 1 ? raw[im + j] : raw[im + jm + k]);
    else
      v = raw[(size_t)c * im * jm * km + ((size_t)k * jm + j) * im + i];
    return v / cl * ldim;
  }
  __device__ double xi(int i, int j, int k, int c) const {
    return i < im ? node(i, j, k, c) : 2.0 * node(im - 1, j, k, c) - node(im - 2, j, k, c);
  }
  __device__ double yj(int i, int j, int k, int c) const {
    return j < jm ? xi(i, j, k, c) : 2.0 * xi(i, jm - 1, k, c) - xi(i, jm - 2, k, c);
  }
  __device__ double zk(int i, int j, int k, int c) const {
    return k < km ? yj(i, j, k, c) : 2.0 * yj(i, j, km - 1, c) - yj(i, j, km - 2, c);
  }
};

#pragma clang fp contract(off)

__global__ __launch_bounds__(256) void k_grid_layout(NodeSrc src, double* __restrict__ coor) {
  const int mx = src.im + 1, my = src.jm + 1;
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= mx || j >= my) return;
  const size_t p = 3 * (((size_t)k * my + j) * mx + i);
  coor[p] = src.zk(i, j, k, 0);
  coor[p + 1] = src.zk(i, j, k, 1);
  coor[p + 2] = src.zk(i, j, k, 2);
}

}  // namespace

}  // namespace pp

using namespace pp;

int vfs_pp_grid_file_info(const char* path, int format, int block, int* nblocks, int* im, int* jm, int* km) {
  ABI_BEGIN
  if (!path) return VFS_PP_EINVAL;
  const GridBlock g = read_block(path, format, block, false);
  if (nblocks) *nblocks = g.nblocks;
  if (im) *im = g.im;
  if (jm) *jm = g.jm;
  if (km) *km = g.km;
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_read_grid_device(const char* path, int format, int block, double cl, double l_dim, const vfs_grid* grid,
                            double* coor) {
  ABI_BEGIN
  if (!path || !grid || !coor) return VFS_PP_EINVAL;
  GridBlock g = read_block(path, format, block, true);
  if (grid->nx != g.im - 1 || grid->ny != g.jm - 1 || grid->nz != g.km - 1)
    throw StatusError(VFS_PP_EINVAL, "grid dimensions differ from the file's node counts minus one");
  Arena tmp;
  double* raw = tmp.alloc<double>(g.raw.size());
  HIPCHECK(hipMemcpy(raw, g.raw.data(), g.raw.size() * sizeof(double), hipMemcpyHostToDevice));
  const NodeSrc src{raw, g.im, g.jm, g.km, g.separable, cl, l_dim};
  const int mx = g.im + 1, my = g.jm + 1, mz = g.km + 1;
  k_grid_layout<<<dim3((mx + 63) / 64, (my + 3) / 4, mz), dim3(64, 4, 1), 0, nullptr>>>(src, coor);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipDeviceSynchronize());
  return VFS_PP_OK;
  ABI_END
}

// PETSc binary vector files, the form of the vfield / pfield / nvfield / ufield
// checkpoints (written with VecView and read with VecLoad on a binary viewer,
// main.c:291-335, 780-797, 865-935): big-endian int32 class id 1211214,
// big-endian int32 length, then the values as big-endian float64.  A DMDA vector
// of one rank is stored in natural order, the [k][j][i][component] layout of the
// engine's DMDA arrays.
namespace {

constexpr int32_t kVecClassId = 1211214;

inline uint32_t be32(uint32_t v) { return __builtin_bswap32(v); }
inline uint64_t be64(uint64_t v) { return __builtin_bswap64(v); }

long long petsc_vec_header(const std::string& path, FILE* f) {
  uint32_t h[2];
  if (std::fread(h, 4, 2, f) != 2) pp::io_error(path, "short PETSc vector header");
  if ((int32_t)be32(h[0]) != kVecClassId) pp::io_error(path, "not a PETSc binary vector (class id)");
  const int32_t n = (int32_t)be32(h[1]);
  if (n < 0) pp::io_error(path, "negative vector length");
  return n;
}

}  // namespace

int vfs_pp_petsc_vec_info(const char* path, long long* n) {
  ABI_BEGIN
  if (!path || !n) return VFS_PP_EINVAL;
  FILE* f = std::fopen(path, "rb");
  if (!f) pp::io_error(path, "cannot open");
  long long len = 0;
  try {
    len = petsc_vec_header(path, f);
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  *n = len;
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_read_petsc_vec_device(const char* path, double* v, size_t n) {
  ABI_BEGIN
  if (!path || !v) return VFS_PP_EINVAL;
  FILE* f = std::fopen(path, "rb");
  if (!f) pp::io_error(path, "cannot open");
  std::vector<uint64_t> buf;
  try {
    const long long len = petsc_vec_header(path, f);
    if ((size_t)len != n) throw pp::StatusError(VFS_PP_EINVAL, "vector length " + std::to_string(len) +
                                                                   " differs from the array length " + std::to_string(n));
    buf.resize(n);
    if (n && std::fread(buf.data(), 8, n, f) != n) pp::io_error(path, "fewer values than the header announces");
  } catch (...) {
    std::fclose(f);
    throw;
  }
  std::fclose(f);
  for (auto& x : buf) x = be64(x);
  if (n) HIPCHECK(hipMemcpy(v, buf.data(), n * sizeof(double), hipMemcpyHostToDevice));
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_write_petsc_vec_device(const char* path, const double* v, size_t n, int block_size) {
  ABI_BEGIN
  if (!path || (!v && n) || n > 0x7fffffffu || block_size < 1) return VFS_PP_EINVAL;
  std::vector<uint64_t> buf(n);
  if (n) HIPCHECK(hipMemcpy(buf.data(), v, n * sizeof(double), hipMemcpyDeviceToHost));
  for (auto& x : buf) x = be64(x);
  FILE* f = std::fopen(path, "wb");
  if (!f) pp::io_error(path, "cannot create");
  const uint32_t h[2] = {be32((uint32_t)kVecClassId), be32((uint32_t)n)};
  const bool ok = std::fwrite(h, 4, 2, f) == 2 && (!n || std::fwrite(buf.data(), 8, n, f) == n);
  if (std::fclose(f) != 0 || !ok) pp::io_error(path, "write failed");
  if (block_size > 1) {  // PETSc's companion file for VecLoad of a blocked vector
    const std::string info = std::string(path) + ".info";
    FILE* g = std::fopen(info.c_str(), "w");
    if (!g) pp::io_error(info, "cannot create");
    std::fprintf(g, "-vecload_block_size %d\n", block_size);
    std::fclose(g);
  }
  return VFS_PP_OK;
  ABI_END
}

// The [0, m) window of a ghosted local DMDA array and the dense [mz][my][mx]
// block the engine takes (host memory, ncomp values per point, i fastest).
int vfs_pp_repack(const double* src, double* dst, const int ghost_start[3], const int ghost_dims[3],
                  const int dims[3], int ncomp, int to_dense) {
  if (!src || !dst || !ghost_start || !ghost_dims || !dims || ncomp < 1) return VFS_PP_EINVAL;
  for (int a = 0; a < 3; a++) {
    if (dims[a] < 1 || ghost_dims[a] < 1) return VFS_PP_EINVAL;
    // the window [0, dims) must lie inside [ghost_start, ghost_start + ghost_dims)
    if (ghost_start[a] > 0 || ghost_start[a] + ghost_dims[a] < dims[a]) return VFS_PP_EINVAL;
  }
  const size_t gx = (size_t)ghost_dims[0], gy = (size_t)ghost_dims[1];
  const size_t row = (size_t)dims[0] * ncomp;
  for (int k = 0; k < dims[2]; k++)
    for (int j = 0; j < dims[1]; j++) {
      const size_t gofs = ((((size_t)(k - ghost_start[2]) * gy) + (size_t)(j - ghost_start[1])) * gx +
                           (size_t)(-ghost_start[0])) * ncomp;
      const size_t dofs = (((size_t)k * dims[1]) + j) * row;
      if (to_dense) std::memcpy(dst + dofs, src + gofs, row * sizeof(double));
      else std::memcpy(dst + gofs, src + dofs, row * sizeof(double));
    }
  return VFS_PP_OK;
}

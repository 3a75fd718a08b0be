// Immersed-body classification from a triangulated surface, on the device:
// ibm_search_advanced (ibm.c:81-570) without its interpolation list, which only
// the momentum step's IB forcing reads, and without the `thin` and moving-body
// branches.
//
// Reference algorithm:
// - The surface's bounding box, widened by 0.05, is cut into 40^3 bins.
//   Each triangle goes into every bin its box touches (ibm.c:128-207).
// - Each interior cell centre inside the box casts a ray towards a point above
//   the box, inside its own bin column. The point is chosen by srand(try) /
//   rand() (randomdirection, ibm.c:48-79).
// - The ray is tested against every triangle of that column from the cell's bin
//   upwards (point_cell_advanced, ibm.c:1777-1902; Moller-Trumbore,
//   compgeom.c:17-60).
// - An odd number of hits with t > 0 means inside (4). If two hits coincide
//   (|dt| < 1e-8) on faces that are not opposed (n.n' > -0.95), the ray is
//   retried with the next direction.
// - Dilation to 2 around 4s, the remap 2 -> 1, 4 -> 3 and the wall layers then
//   follow (ibm.c:325-353, 540-568).
//
// Layout on the device:
// - Each bin column (i, j) holds the triangles whose box covers it, with the
//   exclusive upper bin index kv_max of each. A cell in bin kc tests exactly the
//   column's triangles with kv_max > kc. This is the set of distinct triangles
//   the reference's walk over bins kc..39 meets.
// - The parity and the coincidence test do not depend on the order of that set,
//   so each column is sorted by kv_max, largest first, and a thread stops at the
//   first triangle it does not need.
// - The rand() values of the tries are drawn on the host with the C library's
//   generator, seeded as the reference seeds it, and read from a small table.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pp_engine.hpp"

namespace pp {

namespace {

constexpr int kBins = 40;       // ncx = ncy = ncz, ibm.c:100
constexpr int kMaxHits = 32;    // hits kept per ray for the coincidence test
constexpr int kTries = 64;      // ray directions tried per cell

struct SearchParams {
  double x0, x1, y0, y1, z0, z1;  // widened bounding box
  double dcx, dcy, dcz;           // bin sizes
  int body;                       // 0: reset nvert < 5 first (ibm.c:247)
};

#pragma clang fp contract(off)

__device__ __forceinline__ int tri_hit(const double o[3], const double d[3], const double a[3], const double b[3],
                                       const double c[3], double& t) {
  double e1[3], e2[3], tv[3], pv[3], qv[3];
  e1[0] = b[0] - a[0]; e1[1] = b[1] - a[1]; e1[2] = b[2] - a[2];
  e2[0] = c[0] - a[0]; e2[1] = c[1] - a[1]; e2[2] = c[2] - a[2];
  pv[0] = d[1] * e2[2] - d[2] * e2[1];
  pv[1] = d[2] * e2[0] - d[0] * e2[2];
  pv[2] = d[0] * e2[1] - d[1] * e2[0];
  const double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
  if (det > -1.e-15 && det < 1.e-15) return 0;
  const double inv = 1.0 / det;
  tv[0] = o[0] - a[0]; tv[1] = o[1] - a[1]; tv[2] = o[2] - a[2];
  const double u = (tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2]) * inv;
  if (u < 0.0 || u > 1.0) return 0;
  qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
  qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
  qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
  const double v = (d[0] * qv[0] + d[1] * qv[1] + d[2] * qv[2]) * inv;
  if (v < 0.0 || u + v > 1.0) return 0;
  t = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
  return 1;
}

// One thread per DMDA point; interior cells only.  err[0] = 1: a ray met more
// than kMaxHits triangles; err[1] = 1: no direction free of coincident hits
// (plain stores of one constant, so concurrent writers agree).
__global__ __launch_bounds__(256) void k_ib_search(Dmda m, const double* __restrict__ co, SearchParams P,
                                                   const double* __restrict__ vx, const double* __restrict__ vy,
                                                   const double* __restrict__ vz, const int3* __restrict__ tri,
                                                   const double* __restrict__ nf, const int* __restrict__ col_off,
                                                   const int2* __restrict__ col, const double* __restrict__ stab,
                                                   double* __restrict__ nvert, int* __restrict__ err) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i < 1 || j < 1 || k < 1 || i >= m.mx - 1 || j >= m.my - 1 || k >= m.mz - 1) return;
  const size_t c = m.idx(i, j, k);
  double nv = nvert[c];
  if (P.body == 0 && nv < 5.) nv = 0;
  // cell centre, metrics.c:80-99
  double p[3];
  {
    const size_t n0 = 3 * m.idx(i, j, k), n1 = 3 * m.idx(i, j - 1, k), n2 = 3 * m.idx(i, j, k - 1),
                 n3 = 3 * m.idx(i, j - 1, k - 1), n4 = 3 * m.idx(i - 1, j, k), n5 = 3 * m.idx(i - 1, j - 1, k),
                 n6 = 3 * m.idx(i - 1, j, k - 1), n7 = 3 * m.idx(i - 1, j - 1, k - 1);
#pragma unroll
    for (int v = 0; v < 3; v++)
      p[v] = 0.125 * (co[n0 + v] + co[n1 + v] + co[n2 + v] + co[n3 + v] + co[n4 + v] + co[n5 + v] + co[n6 + v] +
                      co[n7 + v]);
  }
  if (p[0] > P.x0 && p[0] < P.x1 && p[1] > P.y0 && p[1] < P.y1 && p[2] > P.z0 && p[2] < P.z1) {
    const int ic = (int)floor((p[0] - P.x0) / P.dcx), jc = (int)floor((p[1] - P.y0) / P.dcy),
              kc = (int)floor((p[2] - P.z0) / P.dcz);
    const int cb = col_off[jc * kBins + ic], ce = col_off[jc * kBins + ic + 1];
    double val = -1.;
    double ht[kMaxHits];
    int he[kMaxHits];
    for (int st = 1; st <= kTries && val < 0.; st++) {
      const double s = stab[st];
      double d[3];
      d[0] = (P.dcx * (ic + 0.5) + P.x0 + s * P.dcx) - p[0];
      d[1] = (P.dcy * (jc + 0.5) + P.y0 + s * P.dcy) - p[1];
      d[2] = (P.z1 + 0.2) - p[2];
      const double len = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      d[0] /= len;
      d[1] /= len;
      d[2] /= len;
      int nh = 0;
      bool sing = false;
      for (int q = cb; q < ce && !sing; q++) {
        const int2 ent = col[q];
        if (ent.y <= kc) break;  // sorted by kv_max, largest first
        const int3 t3 = tri[ent.x];
        const double A[3] = {vx[t3.x], vy[t3.x], vz[t3.x]};
        const double B[3] = {vx[t3.y], vy[t3.y], vz[t3.y]};
        const double C[3] = {vx[t3.z], vy[t3.z], vz[t3.z]};
        double t;
        if (tri_hit(p, d, A, B, C, t) && t > 0) {
          if (nh == kMaxHits) {
            err[0] = 1;
            sing = true;
            val = 0.;
            break;
          }
          const double* n1 = &nf[3 * ent.x];
          for (int r = 0; r < nh; r++) {
            const double* n0 = &nf[3 * he[r]];
            const double nd = n0[0] * n1[0] + n0[1] * n1[1] + n0[2] * n1[2];
            if (fabs(t - ht[r]) < 1.e-8 && nd > -0.95) sing = true;
          }
          ht[nh] = t;
          he[nh] = ent.x;
          nh++;
        }
      }
      if (!sing) val = (nh % 2) ? 4. : 0.;
    }
    if (val < 0.) {
      err[1] = 1;
      val = 0.;
    }
    nv = nv < val ? val : nv;
    if (nv < 0) nv = 0;
  }
  nvert[c] = nv;
}

struct Surface {
  std::vector<double> x, y, z, nf;
  std::vector<int3> tri;
};

// element normals, ibm_io.c:249-266
void normals(Surface& S) {
  S.nf.resize(3 * S.tri.size());
  for (size_t e = 0; e < S.tri.size(); e++) {
    const int a = S.tri[e].x, b = S.tri[e].y, c = S.tri[e].z;
    const double dx12 = S.x[b] - S.x[a], dy12 = S.y[b] - S.y[a], dz12 = S.z[b] - S.z[a];
    const double dx13 = S.x[c] - S.x[a], dy13 = S.y[c] - S.y[a], dz13 = S.z[c] - S.z[a];
    const double nx = dy12 * dz13 - dz12 * dy13, ny = -dx12 * dz13 + dz12 * dx13, nz = dx12 * dy13 - dy12 * dx13;
    const double dr = std::sqrt(nx * nx + ny * ny + nz * nz);
    S.nf[3 * e] = nx / dr;
    S.nf[3 * e + 1] = ny / dr;
    S.nf[3 * e + 2] = nz / dr;
  }
}

// rand() / (RAND_MAX + 1) - 0.5 after srand(seed), for seeds 1..kTries, drawn
// from a private generator state that the C library seeds exactly as srand does
// (TYPE_3 additive feedback over a 128-byte state), so the caller's rand()
// sequence is left alone.
std::vector<double> ray_table() {
  std::vector<double> t(kTries + 1, 0.);
  for (int seed = 1; seed <= kTries; seed++) {
    char state[128];
    random_data rd;
    std::memset(&rd, 0, sizeof(rd));
    std::memset(state, 0, sizeof(state));
    if (initstate_r((unsigned)seed, state, sizeof(state), &rd) != 0)
      throw StatusError(VFS_PP_EINVAL, "random generator set-up failed");
    int32_t r = 0;
    random_r(&rd, &r);
    t[seed] = r / ((double)RAND_MAX + 1) - 0.5;
  }
  return t;
}

void search_body(const Dmda& m, const int bctype[6], const double* coor, const Surface& S, int body,
                 double* nvert) {
  if (S.x.empty() || S.tri.empty()) throw StatusError(VFS_PP_EINVAL, "empty surface");
  SearchParams P;
  P.x0 = 1.e23, P.x1 = -1.e23, P.y0 = 1.e23, P.y1 = -1.e23, P.z0 = 1.e23, P.z1 = -1.e23;
  for (size_t q = 0; q < S.x.size(); q++) {  // PetscMin / PetscMax, ibm.c:132-144
    P.x0 = P.x0 < S.x[q] ? P.x0 : S.x[q];
    P.x1 = P.x1 > S.x[q] ? P.x1 : S.x[q];
    P.y0 = P.y0 < S.y[q] ? P.y0 : S.y[q];
    P.y1 = P.y1 > S.y[q] ? P.y1 : S.y[q];
    P.z0 = P.z0 < S.z[q] ? P.z0 : S.z[q];
    P.z1 = P.z1 > S.z[q] ? P.z1 : S.z[q];
  }
  P.x0 -= 0.05; P.x1 += 0.05;
  P.y0 -= 0.05; P.y1 += 0.05;
  P.z0 -= 0.05; P.z1 += 0.05;
  P.dcx = (P.x1 - P.x0) / (kBins - 1.);
  P.dcy = (P.y1 - P.y0) / (kBins - 1.);
  P.dcz = (P.z1 - P.z0) / (kBins - 1.);
  P.body = body;
  // column lists (ibm.c:170-207): (triangle, kv_max) per covered column
  std::vector<std::vector<int2>> cols((size_t)kBins * kBins);
  for (size_t e = 0; e < S.tri.size(); e++) {
    const int a = S.tri[e].x, b = S.tri[e].y, c = S.tri[e].z;
    auto mn = [&](const std::vector<double>& v) {
      const double ab = v[a] < v[b] ? v[a] : v[b];
      return ab < v[c] ? ab : v[c];
    };
    auto mx = [&](const std::vector<double>& v) {
      const double ab = v[a] > v[b] ? v[a] : v[b];
      return ab > v[c] ? ab : v[c];
    };
    int i0 = (int)std::floor((mn(S.x) - P.x0) / P.dcx), i1 = (int)std::floor((mx(S.x) - P.x0) / P.dcx) + 1;
    int j0 = (int)std::floor((mn(S.y) - P.y0) / P.dcy), j1 = (int)std::floor((mx(S.y) - P.y0) / P.dcy) + 1;
    int k0 = (int)std::floor((mn(S.z) - P.z0) / P.dcz), k1 = (int)std::floor((mx(S.z) - P.z0) / P.dcz) + 1;
    i0 = std::max(i0, 0), i1 = std::min(i1, kBins);
    j0 = std::max(j0, 0), j1 = std::min(j1, kBins);
    k0 = std::max(k0, 0), k1 = std::min(k1, kBins);
    if (k1 <= k0) continue;
    for (int j = j0; j < j1; j++)
      for (int i = i0; i < i1; i++) cols[(size_t)j * kBins + i].push_back(int2{(int)e, k1});
  }
  std::vector<int> off(cols.size() + 1, 0);
  std::vector<int2> flat;
  for (size_t q = 0; q < cols.size(); q++) {
    auto& v = cols[q];
    std::stable_sort(v.begin(), v.end(), [](const int2& a, const int2& b) { return a.y > b.y; });
    off[q + 1] = off[q] + (int)v.size();
    flat.insert(flat.end(), v.begin(), v.end());
  }
  const std::vector<double> stab = ray_table();
  Arena tmp;
  const size_t nvx = S.x.size(), ne = S.tri.size();
  double* dx = tmp.alloc<double>(nvx);
  double* dy = tmp.alloc<double>(nvx);
  double* dz = tmp.alloc<double>(nvx);
  int3* dt = tmp.alloc<int3>(ne);
  double* dn = tmp.alloc<double>(3 * ne);
  int* doff = tmp.alloc<int>(off.size());
  int2* dcol = tmp.alloc<int2>(flat.size());
  double* ds = tmp.alloc<double>(stab.size());
  int* derr = tmp.alloc<int>(2);
  hipStream_t s = nullptr;
  h2d(s, dx, S.x.data(), nvx);
  h2d(s, dy, S.y.data(), nvx);
  h2d(s, dz, S.z.data(), nvx);
  h2d(s, dt, S.tri.data(), ne);
  h2d(s, dn, S.nf.data(), 3 * ne);
  h2d(s, doff, off.data(), off.size());
  if (!flat.empty()) h2d(s, dcol, flat.data(), flat.size());
  h2d(s, ds, stab.data(), stab.size());
  k_ib_search<<<dim3((m.mx + 63) / 64, (m.my + 3) / 4, m.mz), dim3(64, 4, 1), 0, s>>>(
      m, coor, P, dx, dy, dz, dt, dn, doff, dcol, ds, nvert, derr);
  HIPCHECK(hipGetLastError());
  launch_classify_nvert(s, m, bctype, nvert, nullptr);
  int herr[2] = {0, 0};
  HIPCHECK(hipMemcpy(herr, derr, sizeof(herr), hipMemcpyDeviceToHost));
  HIPCHECK(hipDeviceSynchronize());
  if (herr[0])
    throw StatusError(VFS_PP_EINVAL, "a search ray met more than " + std::to_string(kMaxHits) + " surface triangles");
  if (herr[1])
    throw StatusError(VFS_PP_EINVAL, "no ray direction free of coincident surface hits after " +
                                         std::to_string(kTries) + " tries");
}

// ibmdata file (ibm_io.c:106-240): three header lines, "n_v n_elmt . . .",
// n_v lines "index x y z", n_elmt lines "index index type n1 n2 n3" (1-based)
struct UcdFile {
  FILE* f = nullptr;
  ~UcdFile() {
    if (f) std::fclose(f);
  }
};

void ucd_header(UcdFile& u, const std::string& path, int& n_v, int& n_elmt) {
  u.f = std::fopen(path.c_str(), "r");
  if (!u.f) throw StatusError(VFS_PP_EIO, "surface file " + path + ": cannot open");
  char line[128];
  for (int q = 0; q < 3; q++)
    if (!std::fgets(line, 128, u.f)) throw StatusError(VFS_PP_EIO, "surface file " + path + ": short header");
  int d0, d1, d2;
  if (std::fscanf(u.f, "%i %i %i %i %i", &n_v, &n_elmt, &d0, &d1, &d2) != 5 || n_v < 3 || n_elmt < 1)
    throw StatusError(VFS_PP_EIO, "surface file " + path + ": bad counts line");
}

}  // namespace

}  // namespace pp

using namespace pp;

int vfs_pp_ucd_info(const char* path, int* n_v, int* n_elmt) {
  ABI_BEGIN
  if (!path) return VFS_PP_EINVAL;
  UcdFile u;
  int a = 0, b = 0;
  ucd_header(u, path, a, b);
  if (n_v) *n_v = a;
  if (n_elmt) *n_elmt = b;
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_read_ucd(const char* path, double cl, double cx, double cy, double cz, double* x, double* y, double* z,
                    int* nv1, int* nv2, int* nv3) {
  ABI_BEGIN
  if (!path || !x || !y || !z || !nv1 || !nv2 || !nv3) return VFS_PP_EINVAL;
  UcdFile u;
  int n_v = 0, n_elmt = 0;
  ucd_header(u, path, n_v, n_elmt);
  for (int q = 0; q < n_v; q++) {
    int id;
    if (std::fscanf(u.f, "%i %le %le %le", &id, &x[q], &y[q], &z[q]) != 4)
      throw StatusError(VFS_PP_EIO, std::string("surface file ") + path + ": bad vertex line");
    x[q] = x[q] / cl + cx;  // ibm_io.c:146-148
    y[q] = y[q] / cl + cy;
    z[q] = z[q] / cl + cz;
  }
  for (int q = 0; q < n_elmt; q++) {
    int a, b;
    char type[64];
    if (std::fscanf(u.f, "%i %i %63s %i %i %i\n", &a, &b, type, &nv1[q], &nv2[q], &nv3[q]) != 6)
      throw StatusError(VFS_PP_EIO, std::string("surface file ") + path + ": bad element line");
    nv1[q] -= 1, nv2[q] -= 1, nv3[q] -= 1;
    if (nv1[q] < 0 || nv2[q] < 0 || nv3[q] < 0 || nv1[q] >= n_v || nv2[q] >= n_v || nv3[q] >= n_v)
      throw StatusError(VFS_PP_EIO, std::string("surface file ") + path + ": vertex index out of range");
  }
  return VFS_PP_OK;
  ABI_END
}

int vfs_pp_ib_search_device(const vfs_grid* g, const double* coor, int n_v, const double* x, const double* y,
                             const double* z, int n_elmt, const int* nv1, const int* nv2, const int* nv3, int body,
                             double* nvert) {
  ABI_BEGIN
  if (!g || !coor || !nvert || !x || !y || !z || !nv1 || !nv2 || !nv3 || n_v < 3 || n_elmt < 1 || body < 0)
    return VFS_PP_EINVAL;
  if (g->nx < 1 || g->ny < 1 || g->nz < 1) return VFS_PP_EINVAL;
  Surface S;
  S.x.assign(x, x + n_v);
  S.y.assign(y, y + n_v);
  S.z.assign(z, z + n_v);
  S.tri.resize(n_elmt);
  for (int e = 0; e < n_elmt; e++) {
    if (nv1[e] < 0 || nv2[e] < 0 || nv3[e] < 0 || nv1[e] >= n_v || nv2[e] >= n_v || nv3[e] >= n_v)
      return VFS_PP_EINVAL;
    S.tri[e] = int3{nv1[e], nv2[e], nv3[e]};
  }
  normals(S);
  const Dmda m{g->nx + 2, g->ny + 2, g->nz + 2, {g->periodic[0], g->periodic[1], g->periodic[2]}};
  search_body(m, g->bctype, coor, S, body, nvert);
  return VFS_PP_OK;
  ABI_END
}

// Device set-up of the multigrid hierarchy and of the VolumeFlux face list.
//
// One coarsening step F -> C: coarse nvert by the product rule (a coarse cell is IB
// when the product of its children's nvert exceeds 0.1, the MyNvertRestriction
// rule of poisson.c:2042-2097), periodic ghosts, the isolated-fluid rule (a fluid
// cell with IB on both sides along every axis becomes IB), periodic ghosts again;
// coarse face coefficients summed over the active fine faces that make up each
// coarse upper face and divided by the coarsening factor of the difference
// direction; code words; then three rounds of "a row whose diagonal vanishes
// becomes solid".  Every value comes from the same operations in the same order as
// a sequential loop over the cells would use, so the hierarchy does not depend on
// the launch shape.  Reductions (row coupling, largest coefficient, counts) are
// maxima and sums of small integers, exact in any order.
#include <algorithm>

#include "pp_kernels.hpp"
#include "pp_stencil.hpp"

namespace pp {

namespace {

constexpr unsigned kRedBlocks = 1024;

inline unsigned lin_blocks(size_t n) { return (unsigned)std::max<size_t>(1, (n + 255) / 256); }
inline unsigned red_blocks(size_t n) { return std::min(kRedBlocks, lin_blocks(n)); }
inline dim3 cells(const Dims& d) { return dim3((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz); }

__device__ inline void cell_of(const Dims& d, size_t c, int& i, int& j, int& k) {
  const size_t pl = (size_t)d.nx * d.ny;
  k = (int)(c / pl);
  const size_t r = c - (size_t)k * pl;
  j = (int)(r / (size_t)d.nx);
  i = (int)(r - (size_t)j * d.nx);
}

// children of coarse index ci along an axis of n fine cells (coarsened or not)
__device__ inline void children(bool coarsened, int ci, int n, int& lo, int& hi) {
  if (coarsened) { lo = 2 * ci; hi = min(2 * ci + 1, n - 1); }
  else { lo = hi = ci; }
}

}  // namespace

// ------------------------------------------------------------- level statistics
// Per fluid row: yc = |a_N| + |a_S| against oc = |a_E| + |a_W| + |a_T| + |a_B|.
// part[3 b ..]: largest yc / oc, rows with oc > 0, rows with yc < oc / 10.
__global__ __launch_bounds__(256) void k_level_coupling(Dims d, const double* __restrict__ g,
                                                        const uint32_t* __restrict__ code, double* __restrict__ part) {
  __shared__ double lds[4];
  const size_t n = d.n();
  double w = 0., rows = 0., weak = 0.;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)gridDim.x * 256) {
    const uint32_t cd = code[c];
    if (!(cd & ROW_BIT)) continue;
    int i, j, k;
    cell_of(d, c, i, j, k);
    Nbr nb;
    neighbours(d, i, j, k, nb);
    double vol[19];
    build_row(g, n, cd, c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
    const double yc = fabs(vol[NP]) + fabs(vol[SP]);
    const double oc = fabs(vol[EP]) + fabs(vol[WP]) + fabs(vol[TP]) + fabs(vol[BP]);
    if (!(oc > 0.)) continue;
    const double q = yc / oc;
    w = w < q ? q : w;
    rows += 1.;
    if (yc < 0.1 * oc) weak += 1.;
  }
  w = block_max(w, lds);
  rows = block_sum(rows, lds);
  weak = block_sum(weak, lds);
  if (threadIdx.x == 0) {
    part[3 * blockIdx.x] = w;
    part[3 * blockIdx.x + 1] = rows;
    part[3 * blockIdx.x + 2] = weak;
  }
}

__global__ __launch_bounds__(256) void k_absmax(size_t n, const double* __restrict__ a, double* __restrict__ part) {
  __shared__ double lds[4];
  double m = 0.;
  for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) m = fmax(m, fabs(a[i]));
  m = block_max(m, lds);
  if (threadIdx.x == 0) part[blockIdx.x] = m;
}

// out[0..3]: largest yc / oc, rows, weak rows, largest |g|
__global__ __launch_bounds__(256) void k_level_stats_fin(const double* __restrict__ pc, int nbc,
                                                         const double* __restrict__ pm, int nbm, double* out) {
  __shared__ double lds[4];
  double w = 0., r = 0., wk = 0., gm = 0.;
  for (int b = threadIdx.x; b < nbc; b += 256) {
    w = fmax(w, pc[3 * b]);
    r += pc[3 * b + 1];
    wk += pc[3 * b + 2];
  }
  for (int b = threadIdx.x; b < nbm; b += 256) gm = fmax(gm, pm[b]);
  w = block_max(w, lds);
  r = block_sum(r, lds);
  wk = block_sum(wk, lds);
  gm = block_max(gm, lds);
  if (threadIdx.x == 0) {
    out[0] = w;
    out[1] = r;
    out[2] = wk;
    out[3] = gm;
  }
}

// ------------------------------------------------------------------ coarsening
// coarse nvert at every point of the coarse DMDA box (0 on ghost points)
__global__ __launch_bounds__(256) void k_coarse_nvert(Dmda mf, Dims df, Dmda mc, int cm, const double* __restrict__ nvf,
                                                      double* __restrict__ nvc) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= mc.n()) return;
  const size_t pl = (size_t)mc.mx * mc.my;
  const int K = (int)(p / pl), r = (int)(p - (size_t)K * pl), J = r / mc.mx, I = r - J * mc.mx;
  double v = 0.;
  if (I >= 1 && I <= mc.mx - 2 && J >= 1 && J <= mc.my - 2 && K >= 1 && K <= mc.mz - 2) {
    int i0, i1, j0, j1, k0, k1;
    children(cm & 1, I - 1, df.nx, i0, i1);
    children(cm & 2, J - 1, df.ny, j0, j1);
    children(cm & 4, K - 1, df.nz, k0, k1);
    double prod = 1.;
    for (int kk = k0; kk <= k1; kk++)
      for (int jj = j0; jj <= j1; jj++)
        for (int ii = i0; ii <= i1; ii++) prod *= nvf[mf.idx(ii + 1, jj + 1, kk + 1)];
    if (prod > 0.1) v = 1.;
  }
  nvc[p] = v;
}

// a fluid cell with IB on both sides along every axis becomes IB (reads pre, writes nv)
__global__ __launch_bounds__(256) void k_isolated_fluid(Dmda m, const double* __restrict__ pre,
                                                        double* __restrict__ nv) {
  const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (p >= m.n()) return;
  const size_t sy = m.mx, sz = (size_t)m.mx * m.my;
  const int K = (int)(p / sz), r = (int)(p - (size_t)K * sz), J = r / m.mx, I = r - J * m.mx;
  double v = pre[p];
  if (I >= 1 && I <= m.mx - 2 && J >= 1 && J <= m.my - 2 && K >= 1 && K <= m.mz - 2 && v < 0.1 &&
      pre[p + 1] + pre[p - 1] > 1.1 && pre[p + sy] + pre[p - sy] > 1.1 && pre[p + sz] + pre[p - sz] > 1.1)
    v = 1.;
  nv[p] = v;
}

// coarse face coefficients: the active fine faces on each coarse upper face,
// summed in (k, j, i) order and divided by the coarsening factor of the difference
// direction
__global__ __launch_bounds__(256) void k_coarse_faces(Dims df, Dims dc, int cm, const double* __restrict__ gf,
                                                      const uint32_t* __restrict__ codef, double* __restrict__ gc) {
  const size_t cc = (size_t)blockIdx.x * 256 + threadIdx.x;
  const size_t nc = dc.n(), nf = df.n();
  if (cc >= nc) return;
  int i, j, k;
  cell_of(dc, cc, i, j, k);
  int lo0[3], hi0[3];
  children(cm & 1, i, df.nx, lo0[0], hi0[0]);
  children(cm & 2, j, df.ny, lo0[1], hi0[1]);
  children(cm & 4, k, df.nz, lo0[2], hi0[2]);
  const double fac[3] = {(cm & 1) ? 2. : 1., (cm & 2) ? 2. : 1., (cm & 4) ? 2. : 1.};
  for (int d = 0; d < 3; d++) {
    int lo[3] = {lo0[0], lo0[1], lo0[2]};
    lo[d] = hi0[d];  // the fine faces on the coarse upper face belong to the last child
    double acc[3] = {0., 0., 0.};
    for (int kk = lo[2]; kk <= hi0[2]; kk++)
      for (int jj = lo[1]; jj <= hi0[1]; jj++)
        for (int ii = lo[0]; ii <= hi0[0]; ii++) {
          const size_t fc = df.idx(ii, jj, kk);
          if (!((codef[fc] >> (2 * d)) & 1u)) continue;
          for (int e = 0; e < 3; e++) acc[e] += gf[(size_t)(3 * d + e) * nf + fc];
        }
    for (int e = 0; e < 3; e++) gc[(size_t)(3 * d + e) * nc + cc] = acc[e] / fac[e];
  }
}

// the multigrid class byte of a cell: CLS_INTERIOR for the interior code word
__device__ inline uint8_t interior_bit(uint32_t cd) {
  return (cd & ~DIVOK_BIT) == CODE_INTERIOR ? CLS_INTERIOR : 0;
}

// code words of a coarse level; cls marks nvert > 0.1 and interior cells
__global__ __launch_bounds__(256) void k_level_codes(Dmda m, Dims d, const double* __restrict__ nv,
                                                     uint32_t* __restrict__ code, uint8_t* __restrict__ cls) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const size_t c = d.idx(i, j, k);
  const uint32_t cd = make_code(nv, m.mx, m.my, m.mz, m.per, i + 1, j + 1, k + 1);
  code[c] = cd;
  cls[c] = (nv[m.idx(i + 1, j + 1, k + 1)] > 0.1 ? CLS_NONFLUID : 0) | interior_bit(cd);
}

// rows whose diagonal vanishes (no active face left) become solid
__global__ __launch_bounds__(256) void k_dead_rows(Dmda m, Dims d, const double* __restrict__ g,
                                                   const uint32_t* __restrict__ code, double* __restrict__ nv) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const size_t c = d.idx(i, j, k);
  const uint32_t cd = code[c];
  if (!(cd & ROW_BIT)) return;
  Nbr nb;
  neighbours(d, i, j, k, nb);
  double vol[19];
  build_row(g, d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
  if (vol[CP] == 0.) nv[m.idx(i + 1, j + 1, k + 1)] = 1.;
}

// multigrid mask of the finest level from its code words
__global__ __launch_bounds__(256) void k_cls_from_code(size_t n, const uint32_t* __restrict__ code,
                                                       uint8_t* __restrict__ cls) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const uint32_t cd = code[c];
  cls[c] = ((cd & ROW_BIT) ? 0 : CLS_NONFLUID) | interior_bit(cd);
}

__global__ __launch_bounds__(256) void k_half_to_double(size_t n, const _Float16* __restrict__ a, double gs,
                                                        double* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = (double)a[i] * gs;
}

// Start vector of the power iteration for lambda_max(D^-1 A): uniform in [-1, 1) on
// the rows, 0 elsewhere, from a counter-based generator (splitmix64 of seed + index,
// Steele, Lea and Flood, OOPSLA 2014), so every cell draws independently.
template <typename T>
__global__ __launch_bounds__(256) void k_random_rows(size_t n, const uint32_t* __restrict__ code, uint64_t seed,
                                                     T* __restrict__ v) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  uint64_t z = seed + (uint64_t)(c + 1) * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  const double u = (double)(z >> 11) * 0x1.0p-53;
  v[c] = (code[c] & ROW_BIT) ? T(2. * u - 1.) : T(0);
}

// ------------------------------------------------------------ VolumeFlux faces
// The faces of DMDA point (I, J, K) that VolumeFlux corrects (poisson.c:2936-3010
// for immersed == 1, 2670-2700 otherwise): bit dd of the result; bit dd of neg
// when the face goes from IB to fluid.
__device__ inline int ib_face_mask(const Dmda& m, const double* __restrict__ nv, int immersed, int I, int J, int K,
                                   int& neg) {
  const size_t c = m.idx(I, J, K);
  const size_t st[3] = {1, (size_t)m.mx, (size_t)m.mx * m.my};
  const int pos[3] = {I, J, K}, mm[3] = {m.mx, m.my, m.mz};
  const double nc = nv[c];
  int mask = 0;
  neg = 0;
  for (int dd = 0; dd < 3; dd++) {
    int sgn = 0;
    if (immersed == 1) {
      if (!(pos[dd] < mm[dd] - 2)) continue;
      const double nq = nv[c + st[dd]];
      if (nc < 0.1 && nq > 0.1 && nq < 1.1) sgn = 1;
      else if (nc > 0.1 && nc < 1.1 && nq < 0.1) sgn = -1;
    } else {
      if (pos[dd] > mm[dd] - 2) continue;
      const double nq = nv[c + st[dd]];
      if (nc < TH && nq > TH) sgn = 1;
      else if (nc > TH && nq < TH) sgn = -1;
    }
    if (sgn) {
      mask |= 1 << dd;
      if (sgn < 0) neg |= 1 << dd;
    }
  }
  return mask;
}

// per block of 256 interior cells (interior order): faces and cells with nvert < 0.1
__global__ __launch_bounds__(256) void k_ibface_count(Dmda m, Dims d, const double* __restrict__ nv, int immersed,
                                                      int* __restrict__ cnt, int* __restrict__ lt) {
  __shared__ double lds[4];
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  double nf = 0., l = 0.;
  if (c < d.n()) {
    int i, j, k, neg;
    cell_of(d, c, i, j, k);
    nf = (double)__popc(ib_face_mask(m, nv, immersed, i + 1, j + 1, k + 1, neg));
    l = nv[m.idx(i + 1, j + 1, k + 1)] < 0.1 ? 1. : 0.;
  }
  nf = block_sum(nf, lds);
  l = block_sum(l, lds);
  if (threadIdx.x == 0) {
    cnt[blockIdx.x] = (int)nf;
    lt[blockIdx.x] = (int)l;
  }
}

// exclusive prefix of the block counts (one workgroup; each thread owns a
// contiguous run of blocks); tot[0] = faces, tot[1] = cells with nvert < 0.1
__global__ __launch_bounds__(1024) void k_scan_counts(int nb, const int* __restrict__ cnt, const int* __restrict__ lt,
                                                      int* __restrict__ off, double* tot) {
  __shared__ long long sa[1024], sb[1024];
  const int t = threadIdx.x, per = (nb + 1023) / 1024;
  const int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  long long a = 0, l = 0;
  for (int b = b0; b < b1; b++) {
    a += cnt[b];
    l += lt[b];
  }
  sa[t] = a;
  sb[t] = l;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const long long xa = t >= o ? sa[t - o] : 0, xb = t >= o ? sb[t - o] : 0;
    __syncthreads();
    sa[t] += xa;
    sb[t] += xb;
    __syncthreads();
  }
  long long run = sa[t] - a;
  for (int b = b0; b < b1; b++) {
    off[b] = (int)run;
    run += cnt[b];
  }
  if (t == 1023) {
    tot[0] = (double)sa[1023];
    tot[1] = (double)sb[1023];
  }
}

// met9: the cell metrics csi, eta, zet of the interior cells, plane 3 v + q holding
// component q of metric v
__global__ __launch_bounds__(256) void k_ibface_write(Dmda m, Dims d, const double* __restrict__ nv,
                                                      const double* __restrict__ met9, int immersed,
                                                      const int* __restrict__ off, IbFace* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ int wsum[4];
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  int mask = 0, neg = 0, I = 0, J = 0, K = 0;
  if (c < d.n()) {
    int i, j, k;
    cell_of(d, c, i, j, k);
    I = i + 1; J = j + 1; K = k + 1;
    mask = ib_face_mask(m, nv, immersed, I, J, K, neg);
  }
  const int nf = __popc(mask), lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int x = nf;
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int pos = off[blockIdx.x] + x - nf;
  for (int q = 0; q < w; q++) pos += wsum[q];
  if (!mask) return;
  const size_t p = m.idx(I, J, K), n = d.n();
  for (int dd = 0; dd < 3; dd++) {
    if (!((mask >> dd) & 1)) continue;
    const double m0 = met9[(size_t)(3 * dd) * n + c], m1 = met9[(size_t)(3 * dd + 1) * n + c],
                 m2 = met9[(size_t)(3 * dd + 2) * n + c];
    IbFace f;
    f.pos = (uint64_t)(3 * p + dd);
    f.area = sqrt(m0 * m0 + m1 * m1 + m2 * m2);
    f.sign = ((neg >> dd) & 1) ? -1.f : 1.f;
    f.pad = 0.f;
    out[pos++] = f;
  }
}

// -------------------------------------------------------------------- launchers
void launch_level_stats(hipStream_t s, Dims d, const double* g, const uint32_t* code, double* partial, double* out4) {
  const size_t n = d.n();
  const unsigned nbc = red_blocks(n), nbm = red_blocks(9 * n);
  k_level_coupling<<<nbc, 256, 0, s>>>(d, g, code, partial);
  HIPCHECK(hipGetLastError());
  k_absmax<<<nbm, 256, 0, s>>>(9 * n, g, partial + 3 * kRedBlocks);
  HIPCHECK(hipGetLastError());
  k_level_stats_fin<<<1, 256, 0, s>>>(partial, (int)nbc, partial + 3 * kRedBlocks, (int)nbm, out4);
  HIPCHECK(hipGetLastError());
}

size_t level_stats_partials() { return 4 * (size_t)kRedBlocks; }

void launch_coarsen(hipStream_t s, Dims df, const double* nvf, const double* gf, const uint32_t* codef, int cm,
                    Dims dc, double* nv_tmp, double* nvc, double* gc, uint32_t* codec, uint8_t* clsc) {
  const Dmda mf = Dmda{df.nx + 2, df.ny + 2, df.nz + 2, {df.per[0], df.per[1], df.per[2]}};
  const Dmda mc = Dmda{dc.nx + 2, dc.ny + 2, dc.nz + 2, {dc.per[0], dc.per[1], dc.per[2]}};
  k_coarse_nvert<<<lin_blocks(mc.n()), 256, 0, s>>>(mf, df, mc, cm, nvf, nv_tmp);
  HIPCHECK(hipGetLastError());
  launch_periodic_scalar(s, mc, nv_tmp, 0);
  k_isolated_fluid<<<lin_blocks(mc.n()), 256, 0, s>>>(mc, nv_tmp, nvc);
  HIPCHECK(hipGetLastError());
  launch_periodic_scalar(s, mc, nvc, 0);
  k_coarse_faces<<<lin_blocks(dc.n()), 256, 0, s>>>(df, dc, cm, gf, codef, gc);
  HIPCHECK(hipGetLastError());
  k_level_codes<<<cells(dc), dim3(64, 4), 0, s>>>(mc, dc, nvc, codec, clsc);
  HIPCHECK(hipGetLastError());
  // a round that changes nothing leaves nvert and the codes as they are, so three
  // unconditional rounds equal "at most three, stop when nothing changes"
  for (int it = 0; it < 3; it++) {
    k_dead_rows<<<cells(dc), dim3(64, 4), 0, s>>>(mc, dc, gc, codec, nvc);
    HIPCHECK(hipGetLastError());
    launch_periodic_scalar(s, mc, nvc, 0);
    k_level_codes<<<cells(dc), dim3(64, 4), 0, s>>>(mc, dc, nvc, codec, clsc);
    HIPCHECK(hipGetLastError());
  }
}

void launch_cls_from_code(hipStream_t s, size_t n, const uint32_t* code, uint8_t* cls) {
  k_cls_from_code<<<lin_blocks(n), 256, 0, s>>>(n, code, cls);
  HIPCHECK(hipGetLastError());
}

void launch_half_to_double(hipStream_t s, size_t n, const _Float16* a, double gs, double* out) {
  k_half_to_double<<<lin_blocks(n), 256, 0, s>>>(n, a, gs, out);
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_random_rows(hipStream_t s, size_t n, const uint32_t* code, uint64_t seed, T* v) {
  k_random_rows<T><<<lin_blocks(n), 256, 0, s>>>(n, code, seed, v);
  HIPCHECK(hipGetLastError());
}
template void launch_random_rows<float>(hipStream_t, size_t, const uint32_t*, uint64_t, float*);
template void launch_random_rows<double>(hipStream_t, size_t, const uint32_t*, uint64_t, double*);

size_t ibface_blocks(Dims d) { return lin_blocks(d.n()); }

void launch_ibface_count(hipStream_t s, Dmda m, Dims d, const double* nv, int immersed, int* cnt, int* lt, int* off,
                         double* tot2) {
  const unsigned nb = lin_blocks(d.n());
  k_ibface_count<<<nb, 256, 0, s>>>(m, d, nv, immersed, cnt, lt);
  HIPCHECK(hipGetLastError());
  k_scan_counts<<<1, 1024, 0, s>>>((int)nb, cnt, lt, off, tot2);
  HIPCHECK(hipGetLastError());
}

void launch_ibface_write(hipStream_t s, Dmda m, Dims d, const double* nv, const double* met9, int immersed,
                         const int* off, IbFace* out) {
  k_ibface_write<<<lin_blocks(d.n()), 256, 0, s>>>(m, d, nv, met9, immersed, off, out);
  HIPCHECK(hipGetLastError());
}

// ---- dense inverse of the coarsest level: Gauss-Jordan on [B | I], one launch per column

__global__ __launch_bounds__(256) void k_gj_identity(int N, double* __restrict__ I) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)N * N) return;
  I[e] = (e / N == e % N) ? 1. : 0.;
}

// Column col: the pivot p is the first row r >= col with the largest |B[r][col]|
// (every block finds it itself); row p, scaled by inv = 1 / B[p][col], becomes row
// col, row col's old content moves to row p, and every other row (after that swap)
// loses f = B[r][col] times the scaled pivot row, where f != 0.  Each product and
// difference is rounded on its own (contraction into fused multiply-adds is off in
// this kernel), in the order of the classic elimination, so the inverse is the same
// to the last bit as that elimination run on the host.  B's columns up to col are not written: no later step reads them.
// Blocks take rows blockIdx.x, blockIdx.x + gridDim.x, ...; a zero (or NaN) pivot
// column sets *singular = col + 1 and the step writes nothing.
__global__ __launch_bounds__(256) void k_gj_step(int N, int col, const double* __restrict__ bi,
                                                 const double* __restrict__ ii, double* __restrict__ bo,
                                                 double* __restrict__ io, int* __restrict__ singular) {
#pragma clang fp contract(off)
  __shared__ double sv[4];
  __shared__ int sr[4];
  double best = -1.;
  int piv = N;
  for (int r = col + (int)threadIdx.x; r < N; r += 256) {
    const double v = fabs(bi[(size_t)r * N + col]);
    if (v > best) {  // a thread's rows ascend: a tie keeps the first
      best = v;
      piv = r;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_xor(best, off, 64);
    const int orow = __shfl_xor(piv, off, 64);
    if (ov > best || (ov == best && orow < piv)) {
      best = ov;
      piv = orow;
    }
  }
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = best;
    sr[threadIdx.x >> 6] = piv;
  }
  __syncthreads();
  best = sv[0];
  piv = sr[0];
  for (int q = 1; q < 4; q++)
    if (sv[q] > best || (sv[q] == best && sr[q] < piv)) {
      best = sv[q];
      piv = sr[q];
    }
  if (!(best > 0.)) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && *singular == 0) *singular = col + 1;
    return;
  }
  const double inv = 1. / bi[(size_t)piv * N + col];
  const double* __restrict__ pb = bi + (size_t)piv * N;
  const double* __restrict__ pi = ii + (size_t)piv * N;
  for (int r = blockIdx.x; r < N; r += gridDim.x) {
    double* __restrict__ ob = bo + (size_t)r * N;
    double* __restrict__ oi = io + (size_t)r * N;
    if (r == col) {
      for (int t = col + 1 + (int)threadIdx.x; t < N; t += 256) ob[t] = pb[t] * inv;
      for (int t = (int)threadIdx.x; t < N; t += 256) oi[t] = pi[t] * inv;
      continue;
    }
    const int src = (r == piv) ? col : r;
    const double* __restrict__ sb = bi + (size_t)src * N;
    const double* __restrict__ si = ii + (size_t)src * N;
    const double f = sb[col];
    if (f == 0.) {
      for (int t = col + 1 + (int)threadIdx.x; t < N; t += 256) ob[t] = sb[t];
      for (int t = (int)threadIdx.x; t < N; t += 256) oi[t] = si[t];
    } else {
      for (int t = col + 1 + (int)threadIdx.x; t < N; t += 256)
        ob[t] = sb[t] - f * (pb[t] * inv);
      for (int t = (int)threadIdx.x; t < N; t += 256) oi[t] = si[t] - f * (pi[t] * inv);
    }
  }
}

// cinv[c n + r] = (T) I[r N + c] for r, c < n (column-major top-left block)
template <typename T>
__global__ __launch_bounds__(256) void k_gj_out(int N, int n, const double* __restrict__ I, T* __restrict__ cinv) {
  const size_t e = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (size_t)n * n) return;
  const size_t c = e / n, r = e % n;
  cinv[e] = (T)I[r * N + c];
}

void launch_gj_identity(hipStream_t s, int N, double* I) {
  k_gj_identity<<<lin_blocks((size_t)N * N), 256, 0, s>>>(N, I);
  HIPCHECK(hipGetLastError());
}

void launch_gj_step(hipStream_t s, int N, int col, const double* bi, const double* ii, double* bo, double* io,
                    int* singular) {
  k_gj_step<<<(unsigned)std::min(N, 256), 256, 0, s>>>(N, col, bi, ii, bo, io, singular);
  HIPCHECK(hipGetLastError());
}

template <typename T>
void launch_gj_out(hipStream_t s, int N, int n, const double* I, T* cinv) {
  k_gj_out<T><<<lin_blocks((size_t)n * n), 256, 0, s>>>(N, n, I, cinv);
  HIPCHECK(hipGetLastError());
}
template void launch_gj_out<float>(hipStream_t, int, int, const double*, float*);
template void launch_gj_out<double>(hipStream_t, int, int, const double*, double*);

}  // namespace pp

// Set-up, FP64 operator and pressure-path kernels of the projection engine.
//
// Data layout in HBM: solver vectors and per-cell coefficients are dense
// interior arrays (nx*ny*nz, i fastest) with periodic neighbours reached by
// index wrap, so no ghost planes need refreshing between sweeps.  Caller fields
// (Ucont, Phi, P, Ucat, nvert) stay in the reference's DMDA layout and are only
// touched by the kernels that read or write them in the reference.
#include <algorithm>

#include "pp_kernels.hpp"
#include "pp_stencil.hpp"
#include "pp_krylov_dev.hpp"
#include "pp_path_dev.hpp"

namespace pp {

namespace {

inline dim3 cell_grid(const Dims& d) { return dim3((d.nx + 63) / 64, (d.ny + 3) / 4, d.nz); }
inline dim3 dmda_grid(const Dmda& m) { return dim3((m.mx + 63) / 64, (m.my + 3) / 4, m.mz); }
inline unsigned bnd_blocks(const Dmda& m) { return (boundary_count(m) + 255u) / 256u; }
const dim3 kBlk(64, 4, 1);

__device__ inline int blin() { return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z); }

// ------------------------------------------------------------------------ set-up
__device__ inline double dot3_exact(const double* a, const double* b) {
#pragma clang fp contract(off)
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

__global__ __launch_bounds__(256) void k_face_coeffs(Dmda m, Dims d, const double* icsi, const double* ieta,
                                                     const double* izet, const double* jcsi, const double* jeta,
                                                     const double* jzet, const double* kcsi, const double* keta,
                                                     const double* kzet, const double* iaj, const double* jaj,
                                                     const double* kaj, double* g) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const size_t n = d.n(), c = d.idx(i, j, k), p = m.idx(i + 1, j + 1, k + 1);
  // poisson.c:1132-1140, operand order (v.x*w.x + v.y*w.y + v.z*w.z) * aj
  g[0 * n + c] = dot3_exact(&icsi[3 * p], &icsi[3 * p]) * iaj[p];
  g[1 * n + c] = dot3_exact(&ieta[3 * p], &icsi[3 * p]) * iaj[p];
  g[2 * n + c] = dot3_exact(&izet[3 * p], &icsi[3 * p]) * iaj[p];
  g[3 * n + c] = dot3_exact(&jcsi[3 * p], &jeta[3 * p]) * jaj[p];
  g[4 * n + c] = dot3_exact(&jeta[3 * p], &jeta[3 * p]) * jaj[p];
  g[5 * n + c] = dot3_exact(&jzet[3 * p], &jeta[3 * p]) * jaj[p];
  g[6 * n + c] = dot3_exact(&kcsi[3 * p], &kzet[3 * p]) * kaj[p];
  g[7 * n + c] = dot3_exact(&keta[3 * p], &kzet[3 * p]) * kaj[p];
  g[8 * n + c] = dot3_exact(&kzet[3 * p], &kzet[3 * p]) * kaj[p];
}

__global__ __launch_bounds__(256) void k_codes(Dmda m, Dims d, const double* nv, uint32_t* code, uint8_t* cls) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const int I = i + 1, J = j + 1, K = k + 1;
  const size_t c = d.idx(i, j, k), p = m.idx(I, J, K);
  code[c] = make_code(nv, m.mx, m.my, m.mz, m.per, I, J, K);
  const double v = nv[p];
  uint8_t b = 0;
  if (v < 0.1) b |= CLS_LT;
  if ((int)v > TH) b |= CLS_INTGT;
  if ((int)(v + 0.1) == 3) b |= CLS_SOLID;
  if (v > 0.1) b |= CLS_NONFLUID;
  // Projection face conditions, poisson.c:2221-2222,2259 (and rotations)
  if ((I < m.mx - 2 || (m.per[0] && I == m.mx - 2)) && (v + nv[p + 1]) < TH) b |= CLS_PROJ_I;
  if ((J < m.my - 2 || (m.per[1] && J == m.my - 2)) && (v + nv[p + m.mx]) < TH) b |= CLS_PROJ_J;
  if ((K < m.mz - 2 || (m.per[2] && K == m.mz - 2)) && (v + nv[p + (size_t)m.mx * m.my]) < TH) b |= CLS_PROJ_K;
  cls[c] = b;
}

__global__ __launch_bounds__(256) void k_export_rows(Dims d, const double* g, const uint32_t* code, double* vol19) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const size_t c = d.idx(i, j, k);
  Nbr nb;
  neighbours(d, i, j, k, nb);
  double vol[19];
  build_row(g, d.n(), code[c], c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
  for (int s = 0; s < 19; s++) vol19[19 * c + s] = vol[s];
}

// dilation of the IB mask (ibm.c:331-353); writes only 2s, so in-place is race free
__global__ __launch_bounds__(256) void k_dilate(Dmda m, double* nv) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= m.mx || j >= m.my) return;
  const size_t c = m.idx(i, j, k);
  double v = nv[c];
  if (v < 0) v = 0;
  const int ip = i < m.mx - 1 ? i + 1 : i, im = i > 0 ? i - 1 : i;
  const int jp = j < m.my - 1 ? j + 1 : j, jm = j > 0 ? j - 1 : j;
  const int kp = k < m.mz - 1 ? k + 1 : k, km = k > 0 ? k - 1 : k;
  if ((int)(v + 0.5) != 4) {
    bool hit = false;
    for (int kk = km; kk <= kp; kk++)
      for (int jj = jm; jj <= jp; jj++)
        for (int ii = im; ii <= ip; ii++)
          if ((int)(nv[m.idx(ii, jj, kk)] + 0.5) == 4) hit = true;
    if (hit) v = v > 2 ? v : 2;
  }
  nv[c] = v;
}

// remap 2->1, 4->3 and wall-model layers (ibm.c:540-568), interior points
__global__ __launch_bounds__(256) void k_remap(Dmda m, int b0, int b1, int b2, int b3, int b4, int b5, double* nv) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i < 1 || j < 1 || k < 1 || i >= m.mx - 1 || j >= m.my - 1 || k >= m.mz - 1) return;
  const size_t c = m.idx(i, j, k);
  double v = nv[c];
  if ((int)(v + 0.5) == 2) v = 1;
  if ((int)(v + 0.5) == 4) v = 3;
  const int bc[6] = {b0, b1, b2, b3, b4, b5};
  const bool at[6] = {i == 1, i == m.mx - 2, j == 1, j == m.my - 2, k == 1, k == m.mz - 2};
  for (int f = 0; f < 6; f++)
    if ((bc[f] == -1 || bc[f] == -2) && at[f]) v = 1;
  nv[c] = v;
}

// ------------------------------------------------------------------ operator
template <typename TX, bool DROP>
__global__ __launch_bounds__(256) void k_apply64(Dims d, const double* __restrict__ g,
                                                 const uint32_t* __restrict__ code, const TX* __restrict__ x,
                                                 double* __restrict__ y) {
  CELL_IJK(i, j, k);
  if (i >= d.nx || j >= d.ny) return;
  const size_t c = d.idx(i, j, k);
  const uint32_t cd = code[c];
  Nbr nb;
  neighbours(d, i, j, k, nb);
  double acc = 0.;
  if constexpr (DROP) {
    // the assembled matrix of the reference: coefficients with |v| <= 1e-10 dropped
    // (poisson.c:1608); needs the coefficients one by one
    double vol[19];
    build_row(g, d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
#pragma unroll
    for (int s = 0; s < 19; s++) {
      double v = vol[s];
      if (s != CP && !(fabs(v) > 1.e-10)) v = 0.;
      if (v != 0.) acc += v * (double)x[nb.s[s]];
    }
  } else {
    double xv[19], dg;
    load_x(nb, x, xv);
    flux_row(g, d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], xv, acc, dg);
  }
  y[c] = (cd & ROW_BIT) ? acc : 0.;
}

template <typename TX>
__global__ __launch_bounds__(256) void k_residual64(Dims d, const double* __restrict__ g,
                                                    const uint32_t* __restrict__ code, const double* __restrict__ b,
                                                    const TX* __restrict__ x, double* __restrict__ r,
                                                    double* __restrict__ partial) {
  __shared__ double lds[4];
  CELL_IJK(i, j, k);
  double rr = 0.;
  if (i < d.nx && j < d.ny) rr = residual_cell(d, g, code, b, x, r, i, j, k);
  rr = block_sum(rr, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blin()] = rr;
}

__global__ __launch_bounds__(256) void k_apply_dmda(Dmda m, Dims d, const double* __restrict__ g,
                                                    const uint32_t* __restrict__ code, const double* __restrict__ phi,
                                                    double* __restrict__ y, int drop) {
  CELL_IJK(i, j, k);
  if (i >= d.nx || j >= d.ny) return;
  const size_t c = d.idx(i, j, k), p = m.idx(i + 1, j + 1, k + 1);
  const uint32_t cd = code[c];
  if (!(cd & ROW_BIT)) { y[p] = 0.; return; }
  Nbr nb;
  neighbours(d, i, j, k, nb);
  double vol[19];
  build_row(g, d.n(), cd, c, nb.s[WP], nb.s[SP], nb.s[BP], vol);
  double acc = 0.;
#pragma unroll
  for (int s = 0; s < 19; s++) {
    double v = vol[s];
    if (drop && s != CP && !(fabs(v) > 1.e-10)) v = 0.;
    const Off3 o = slot_off(s);
    // DMDA neighbour: periodic images live in the ghost planes
    if (v != 0.) acc += v * phi[p + o.d[0] + (ptrdiff_t)o.d[1] * m.mx + (ptrdiff_t)o.d[2] * m.mx * m.my];
  }
  y[p] = acc;
}

// ------------------------------------------------------------------------- RHS
__global__ __launch_bounds__(256) void k_rhs(Dmda m, Dims d, const double* __restrict__ u,
                                             const uint8_t* __restrict__ cls, double fac, double* __restrict__ b,
                                             double* __restrict__ partial) {
  __shared__ double lds[4];
  CELL_IJK(i, j, k);
  double bb = 0.;
  if (i < d.nx && j < d.ny) bb = rhs_cell(m, d, u, cls, fac, b, i, j, k);
  bb = block_sum(bb, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blin()] = bb;
}

// ------------------------------------------------------------------ VolumeFlux
__global__ __launch_bounds__(256) void k_ibflux_sum(const IbFace* __restrict__ f, size_t nf, double* __restrict__ u,
                                                    int zero_small, double* __restrict__ partial, int nblk) {
  __shared__ double lds[4];
  double fl = 0., ar = 0.;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < nf; t += (size_t)nblk * 256)
    ibflux_sum_face(f[t], u, zero_small, fl, ar);
  fl = block_sum(fl, lds);
  ar = block_sum(ar, lds);
  if (threadIdx.x == 0) { partial[blockIdx.x] = fl; partial[nblk + blockIdx.x] = ar; }
}

__global__ __launch_bounds__(256) void k_ibflux_correct(const IbFace* __restrict__ f, size_t nf,
                                                        double* __restrict__ u, int immersed,
                                                        const double* __restrict__ sums) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= nf) return;
  ibflux_correct_face(f[t], u, immersed, sums);
}

// ------------------------------------------------------------ null space, scatter
__global__ __launch_bounds__(256) void k_masked_sum(Dims d, const double* __restrict__ x,
                                                    const uint8_t* __restrict__ cls, double* __restrict__ partial,
                                                    int nblk) {
  __shared__ double lds[4];
  const size_t n = d.n();
  double s = 0.;
  for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < n; c += (size_t)nblk * 256)
    if (cls[c] & CLS_LT) s += x[c];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_sub_mean(Dims d, double* __restrict__ x, const uint8_t* __restrict__ cls,
                                                  const double* __restrict__ sum, double cnt) {
  const size_t c = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (c >= d.n()) return;
  const double val = -sum[0] / cnt;  // poisson_hypre.c:256
  if (cls[c] & CLS_LT) x[c] += val;
}

__global__ __launch_bounds__(256) void k_scatter_update(Dmda m, Dims d, const double* __restrict__ x,
                                                        const uint8_t* __restrict__ cls, double* __restrict__ phi,
                                                        double* __restrict__ p) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  scatter_cell(m, d, x, cls, phi, p, i, j, k);
}

// cell metrics csi/eta/zet and aj of the interior cells, SoA (met9[3 v + q])
__global__ __launch_bounds__(256) void k_gather_cell_metrics(Dmda m, Dims d, const double* __restrict__ csi,
                                                             const double* __restrict__ eta,
                                                             const double* __restrict__ zet,
                                                             const double* __restrict__ aj, double* __restrict__ met9,
                                                             double* __restrict__ ajc) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  const size_t n = d.n(), c = d.idx(i, j, k), p = m.idx(i + 1, j + 1, k + 1);
  const double* src[3] = {csi, eta, zet};
#pragma unroll
  for (int v = 0; v < 3; v++)
#pragma unroll
    for (int q = 0; q < 3; q++) met9[(size_t)(3 * v + q) * n + c] = src[v][3 * p + q];
  ajc[c] = aj[p];
}

__global__ __launch_bounds__(256) void k_gather_phi(Dmda m, Dims d, const double* __restrict__ phi,
                                                    const uint8_t* __restrict__ cls, double* __restrict__ x) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  gather_phi_cell(m, d, phi, cls, x, i, j, k);
}

__global__ __launch_bounds__(256) void k_add_phi(Dmda m, Dims d, const uint8_t* __restrict__ cls,
                                                 const double* __restrict__ phi, double* __restrict__ p) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  add_phi_cell(m, d, cls, phi, p, i, j, k);
}

// Boundary-point kernels: one thread per boundary point of the DMDA box
// (boundary_point), 256-thread blocks.
__global__ __launch_bounds__(256) void k_periodic_scalar(Dmda m, double* __restrict__ a, int zero_other) {
  int i, j, k;
  if (boundary_point(m, blockIdx.x * 256u + threadIdx.x, i, j, k)) periodic_point(m, a, zero_other, i, j, k);
}

// ------------------------------------------------------------------ Projection
__global__ __launch_bounds__(256) void k_projection(Dmda m, Dims d, const double* __restrict__ g,
                                                    const uint32_t* __restrict__ code,
                                                    const uint8_t* __restrict__ cls, const double* __restrict__ phi,
                                                    double* __restrict__ u, double fac) {
  CELL_IJK(i, j, k);
  if (i >= d.nx || j >= d.ny) return;
  projection_cell(m, d, g, code, cls, phi, u, fac, i, j, k);
}

__global__ __launch_bounds__(256) void k_ucont_ghosts(Dmda m, double* __restrict__ u) {
  int i, j, k;
  if (boundary_point(m, blockIdx.x * 256u + threadIdx.x, i, j, k)) ucont_ghost_point(m, u, i, j, k);
}

// ----------------------------------------------------------------- Contra2Cart
__global__ __launch_bounds__(256) void k_c2c_core(Dmda m, Dims d, const double* __restrict__ met,
                                                  const uint8_t* __restrict__ cls, const double* __restrict__ u,
                                                  double* __restrict__ ucat) {
  CELL_IJK(i, j, k);
  if (i >= d.nx || j >= d.ny) return;
  c2c_core_cell(m, d, met, cls, u, ucat, i, j, k);
}

__global__ __launch_bounds__(256) void k_vec_ghosts(Dmda m, double* __restrict__ a, double* __restrict__ snap) {
  int i, j, k;
  if (boundary_point(m, blockIdx.x * 256u + threadIdx.x, i, j, k)) vec_ghost_point(m, a, snap, i, j, k);
}

__global__ __launch_bounds__(256) void k_c2c_bc(Dmda m, const double* __restrict__ nv, const double* __restrict__ snap,
                                                double* __restrict__ ucat, int b0, int b1, int b2, int b3, int b4,
                                                int b5) {
  int i, j, k;
  if (!boundary_point(m, blockIdx.x * 256u + threadIdx.x, i, j, k)) return;
  const int bc[6] = {b0, b1, b2, b3, b4, b5};
  c2c_bc_point(m, nv, snap, ucat, bc, i, j, k);
}

__global__ __launch_bounds__(256) void k_c2c_solid(Dims d, Dmda m, const uint8_t* __restrict__ cls,
                                                   double* __restrict__ ucat) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= d.nx || j >= d.ny) return;
  c2c_solid_cell(m, d, cls, ucat, i, j, k);
}

// ------------------------------------------------------------------- Divergence
__global__ __launch_bounds__(256) void k_divergence(Dmda m, Dims d, const uint32_t* __restrict__ code,
                                                    const double* __restrict__ aj, const double* __restrict__ u,
                                                    double* __restrict__ partial) {
  __shared__ double lds[4];
  CELL_IJK(i, j, k);
  double mv = 0.;
  if (i < d.nx && j < d.ny) mv = divergence_cell(m, d, code, aj, u, i, j, k);
  mv = block_max(mv, lds);
  if (threadIdx.x == 0 && threadIdx.y == 0) partial[blin()] = mv;
}

// max over Divergence-counted cells of |aj r|
__global__ __launch_bounds__(256) void k_divest(Dims d, const uint32_t* __restrict__ code,
                                                const double* __restrict__ aj, const double* __restrict__ r,
                                                double* __restrict__ partial) {
  __shared__ double lds[4];
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  double mv = 0.;
  if (t < d.n()) mv = divest_elem(code, aj, r, t);
  mv = block_max(mv, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = mv;
}

// ---------------------------------------------------------------- reductions
// One 1024-thread block per vector sums (op 0) or maxes (op 1) its nblk block
// partials in a fixed order: four independent accumulators per thread keep the
// loads in flight, then the 16 wave results meet in LDS.
__global__ __launch_bounds__(1024) void k_finalize(const double* __restrict__ partial, int nblk,
                                                   double* __restrict__ out, int op, double* __restrict__ host) {
  __shared__ double lds[16];
  const double* pv = partial + (size_t)blockIdx.x * nblk;
  const int t = threadIdx.x;
  double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
  int b = t;
  if (op) {
    for (; b + 3072 < nblk; b += 4096) {
      a0 = fmax(a0, pv[b]); a1 = fmax(a1, pv[b + 1024]); a2 = fmax(a2, pv[b + 2048]); a3 = fmax(a3, pv[b + 3072]);
    }
    for (; b < nblk; b += 1024) a0 = fmax(a0, pv[b]);
    a0 = wave_max(fmax(fmax(a0, a1), fmax(a2, a3)));
  } else {
    for (; b + 3072 < nblk; b += 4096) {
      a0 += pv[b]; a1 += pv[b + 1024]; a2 += pv[b + 2048]; a3 += pv[b + 3072];
    }
    for (; b < nblk; b += 1024) a0 += pv[b];
    a0 = wave_sum((a0 + a1) + (a2 + a3));
  }
  if ((t & 63) == 0) lds[t >> 6] = a0;
  __syncthreads();
  if (t == 0) {
    double s = lds[0];
    for (int w = 1; w < 16; w++) s = op ? fmax(s, lds[w]) : s + lds[w];
    out[blockIdx.x] = s;
    if (host) host[blockIdx.x] = s;  // pinned host mirror, read after the stream syncs
  }
}

template <int NV>
__global__ __launch_bounds__(256) void k_multidot(size_t n, const double* __restrict__ w, const double* __restrict__ V,
                                                  size_t ldv, double* __restrict__ partial, int nblk) {
  __shared__ double lds[4];
  double acc[NV];
#pragma unroll
  for (int v = 0; v < NV; v++) acc[v] = 0.;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)nblk * 256) {
    const double wt = w[t];
#pragma unroll
    for (int v = 0; v < NV; v++) acc[v] += wt * V[(size_t)v * ldv + t];
  }
#pragma unroll
  for (int v = 0; v < NV; v++) {
    const double s = block_sum(acc[v], lds);
    if (threadIdx.x == 0) partial[(size_t)v * nblk + blockIdx.x] = s;
  }
}

__global__ __launch_bounds__(256) void k_axpy_multi(size_t n, double* __restrict__ w, const double* __restrict__ V,
                                                    size_t ldv, int nv, const double* __restrict__ h, double sign) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  double acc = 0.;
  for (int v = 0; v < nv; v++) acc += h[v] * V[(size_t)v * ldv + t];
  w[t] += sign * acc;
}

__global__ __launch_bounds__(256) void k_norm2(size_t n, const double* __restrict__ a, double* __restrict__ partial,
                                               int nblk) {
  __shared__ double lds[4];
  double s = 0.;
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += (size_t)nblk * 256) s += a[t] * a[t];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_scale_to(size_t n, const double* __restrict__ a,
                                                  const double* __restrict__ nrm2, double* __restrict__ out) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const double s = nrm2[0] > 0. ? 1. / sqrt(nrm2[0]) : 0.;
  out[t] = a[t] * s;
}

template <typename TZ>
__global__ __launch_bounds__(256) void k_update_x(size_t n, double* __restrict__ x, const TZ* __restrict__ Z,
                                                  size_t ldz, int nz, const double* __restrict__ y) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  update_x_elem(x, Z, ldz, nz, y, t);
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_convert(size_t n, const TA* __restrict__ a, TB* __restrict__ b) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) b[t] = (TB)a[t];
}

// b = (TB)(a * scale): FP16 storage of FP32-level coefficients (scale a power of two)
template <typename TA, typename TB>
__global__ __launch_bounds__(256) void k_convert_scaled(size_t n, const TA* __restrict__ a, double scale,
                                                        TB* __restrict__ b) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) b[t] = (TB)(a[t] * scale);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(size_t n, T* a, T v) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n) a[t] = v;
}

// Plain device copy, 16 bytes per work-item: the bandwidth reference that kernel
// timings and memory-counter readings are calibrated against.
__global__ __launch_bounds__(256) void k_copy16(size_t n2, const double2* __restrict__ a, double2* __restrict__ b) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t < n2) b[t] = a[t];
}

// w = A z (flux form) and per-wave partial sums of w.V_l (l < nv) and w.w (four
// per 64x4 block); the first Krylov reduction of classical Gram-Schmidt fused
// into the operator pass.
template <typename TZ>
__global__ __launch_bounds__(256) void k_apply_dots(Dims d, const double* __restrict__ g,
                                                    const uint32_t* __restrict__ code, const TZ* __restrict__ z,
                                                    double* __restrict__ w, const double* __restrict__ V, size_t ldv,
                                                    int nv, double* __restrict__ partial) {
  CELL_IJK(i, j, k);
  const size_t nblk = (size_t)gridDim.x * gridDim.y * gridDim.z, b = blin();
  const bool in = i < d.nx && j < d.ny;
  const size_t c = in ? d.idx(i, j, k) : 0;
  double v8[8];
  wave_dots_prefetch(in, c, V, ldv, nv, v8);
  const double wv = in ? apply_cell(d, g, code, z, w, i, j, k) : 0.;
  wave_dots(in, c, wv, V, ldv, nv, v8, partial, 4 * nblk, 4 * b + threadIdx.y);
}

// CGS2 second pass, fused: w1 = w - sum_l h_l V_l (in place) and per-block
// partial sums of w1.V_l (l < nv) and w1.w1.  One element per thread; the second
// read of V_l[t] hits the cache line the first one brought in.
__global__ __launch_bounds__(256) void k_axpy_dots(size_t n, double* __restrict__ w, const double* __restrict__ V,
                                                   size_t ldv, int nv, const double* __restrict__ h,
                                                   double* __restrict__ partial) {
  __shared__ double lds[4][64];
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  const double w1 = t < n ? gs_sub(w, V, ldv, nv, h, t) : 0.;
  block_dots(t < n, t, w1, V, ldv, nv, partial, gridDim.x, blockIdx.x, lds);
}

// CGS2 second pass with the block reduction amortised over a grid-stride range
// of elements per thread: the basis values are loaded once into registers for
// both w1 = w - V h and the (nv + 1) dot products, which stay in registers
// (nv + 1 <= S) and are reduced once per block.
template <int S>
__global__ __launch_bounds__(256) void k_axpy_dots_acc(size_t n, double* __restrict__ w, const double* __restrict__ V,
                                                       size_t ldv, int nv, const double* __restrict__ h,
                                                       double* __restrict__ partial) {
  __shared__ double lds[4][64];
  double acc[S];
#pragma unroll
  for (int l = 0; l < S; l++) acc[l] = 0.;
  const size_t stride = (size_t)gridDim.x * 256;
#pragma unroll 1
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) {
    double vv[S];
#pragma unroll
    for (int l = 0; l < S; l++) vv[l] = l < nv ? V[(size_t)l * ldv + t] : 0.;
    double w1 = w[t];
#pragma unroll
    for (int l = 0; l < S; l++)
      if (l < nv) w1 -= h[l] * vv[l];
    w[t] = w1;
#pragma unroll
    for (int l = 0; l < S; l++) {
      if (l < nv) acc[l] += w1 * vv[l];
      else if (l == nv) acc[l] += w1 * w1;
    }
  }
  block_sums(acc, nv + 1, partial, gridDim.x, blockIdx.x, lds);
}

// Lazily normalised Gram-Schmidt pass: w' = w - sum_l c_l U_l in place with
// c_l = dots[l] * nuinv2[l] (the projection onto U_l / |U_l|), f = (TF)(w' s) with
// s = conv_scale(*a2) for the next V-cycle, and per-block partial sums of |w'|^2.
template <typename TF>
__global__ __launch_bounds__(256) void k_axpy_norm(size_t n, double* __restrict__ w, const double* __restrict__ V,
                                                   size_t ldv, int nv, const double* __restrict__ dots,
                                                   const double* __restrict__ nuinv2, const double* __restrict__ a2,
                                                   TF* __restrict__ f, double* __restrict__ partial,
                                                   const TF* __restrict__ dinv, TF* __restrict__ fd) {
  __shared__ double lds[4][64];
  const double sc = conv_scale(a2[0]);
  double acc[1] = {0.};
  const size_t stride = (size_t)gridDim.x * 256;
#pragma unroll 1
  for (size_t t = (size_t)blockIdx.x * 256 + threadIdx.x; t < n; t += stride) {
    double v = w[t];
    for (int l = 0; l < nv; l++) v -= (dots[l] * nuinv2[l]) * V[(size_t)l * ldv + t];
    w[t] = v;
    const TF fv = (TF)(v * sc);
    f[t] = fv;
    if (fd) fd[t] = fv * dinv[t];
    acc[0] += v * v;
  }
  block_sums(acc, 1, partial, gridDim.x, blockIdx.x, lds);
}

__global__ void k_gs_norm(double* __restrict__ red, int off, int nv, double* __restrict__ host) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  gs_norm_one(red, off, nv, host);
}

// v = (w - sum_l h_l V_l) * inv (in place over w); f = (TF)v for the next V-cycle
template <typename TF>
__global__ __launch_bounds__(256) void k_axpy_scale(size_t n, double* __restrict__ w, const double* __restrict__ V,
                                                    size_t ldv, int nv, const double* __restrict__ h,
                                                    const double* __restrict__ inv, TF* __restrict__ f,
                                                    const TF* __restrict__ dinv, TF* __restrict__ fd) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  axpy_scale_one(w, V, ldv, nv, h, inv[0], f, t);
  if (fd) fd[t] = f[t] * dinv[t];
}

// v = a / sqrt(norm2); f = (TF)v
template <typename TF>
__global__ __launch_bounds__(256) void k_scale_conv(size_t n, const double* __restrict__ a,
                                                    const double* __restrict__ nrm2, double* __restrict__ v,
                                                    TF* __restrict__ f, const TF* __restrict__ dinv,
                                                    TF* __restrict__ fd) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  scale_conv_elem(a, nrm2[0], v, f, t);
  if (fd) fd[t] = f[t] * dinv[t];
}

inline unsigned nb1(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

int partial_blocks(size_t n) {
  size_t b = (n + 255) / 256;
  return (int)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

// ------------------------------------------------------------------ launchers
void launch_face_coeffs(hipStream_t s, Dmda m, Dims d, const double* icsi, const double* ieta, const double* izet,
                        const double* jcsi, const double* jeta, const double* jzet, const double* kcsi,
                        const double* keta, const double* kzet, const double* iaj, const double* jaj,
                        const double* kaj, double* g) {
  k_face_coeffs<<<cell_grid(d), kBlk, 0, s>>>(m, d, icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj,
                                              kaj, g);
  HIPCHECK(hipGetLastError());
}

void launch_codes(hipStream_t s, Dmda m, Dims d, const double* nvert, uint32_t* code, uint8_t* cls) {
  k_codes<<<cell_grid(d), kBlk, 0, s>>>(m, d, nvert, code, cls);
  HIPCHECK(hipGetLastError());
}

void launch_export_rows(hipStream_t s, Dims d, const double* g, const uint32_t* code, double* vol19) {
  k_export_rows<<<cell_grid(d), kBlk, 0, s>>>(d, g, code, vol19);
  HIPCHECK(hipGetLastError());
}

void launch_classify_nvert(hipStream_t s, Dmda m, const int bc[6], double* nv, uint8_t*) {
  k_dilate<<<dmda_grid(m), kBlk, 0, s>>>(m, nv);
  HIPCHECK(hipGetLastError());
  k_remap<<<dmda_grid(m), kBlk, 0, s>>>(m, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5], nv);
  HIPCHECK(hipGetLastError());
}

template <typename TX>
void launch_apply64(hipStream_t s, Dims d, const double* g, const uint32_t* code, const TX* x, double* y, bool drop) {
  if (drop) k_apply64<TX, true><<<cell_grid(d), kBlk, 0, s>>>(d, g, code, x, y);
  else k_apply64<TX, false><<<cell_grid(d), kBlk, 0, s>>>(d, g, code, x, y);
  HIPCHECK(hipGetLastError());
}
template void launch_apply64<double>(hipStream_t, Dims, const double*, const uint32_t*, const double*, double*, bool);
template void launch_apply64<float>(hipStream_t, Dims, const double*, const uint32_t*, const float*, double*, bool);

template <typename TX>
void launch_residual64(hipStream_t s, Dims d, const double* g, const uint32_t* code, const double* b, const TX* x,
                       double* r, double* partial, int* nblk) {
  const dim3 gr = cell_grid(d);
  *nblk = (int)(gr.x * gr.y * gr.z);
  k_residual64<TX><<<gr, kBlk, 0, s>>>(d, g, code, b, x, r, partial);
  HIPCHECK(hipGetLastError());
}
template void launch_residual64<double>(hipStream_t, Dims, const double*, const uint32_t*, const double*,
                                        const double*, double*, double*, int*);
template void launch_residual64<float>(hipStream_t, Dims, const double*, const uint32_t*, const double*, const float*,
                                       double*, double*, int*);

void launch_apply_dmda(hipStream_t s, Dmda m, Dims d, const double* g, const uint32_t* code, const double* phi,
                       double* y, bool drop) {
  k_apply_dmda<<<cell_grid(d), kBlk, 0, s>>>(m, d, g, code, phi, y, drop ? 1 : 0);
  HIPCHECK(hipGetLastError());
}

void launch_rhs(hipStream_t s, Dmda m, Dims d, const double* ucont, const uint8_t* cls, double fac, double* b,
                double* partial, int* nblk) {
  const dim3 gr = cell_grid(d);
  *nblk = (int)(gr.x * gr.y * gr.z);
  k_rhs<<<gr, kBlk, 0, s>>>(m, d, ucont, cls, fac, b, partial);
  HIPCHECK(hipGetLastError());
}

void launch_ibflux_sum(hipStream_t s, const IbFace* f, size_t nf, double* ucont, int immersed, int zero_small,
                       double* partial, int* nblk) {
  const int nb = partial_blocks(nf);
  *nblk = nb;
  k_ibflux_sum<<<nb, 256, 0, s>>>(f, nf, ucont, (immersed == 1 && zero_small) ? 1 : 0, partial, nb);
  HIPCHECK(hipGetLastError());
}

void launch_ibflux_correct(hipStream_t s, const IbFace* f, size_t nf, double* ucont, int immersed, const double* sums) {
  if (nf == 0) return;
  k_ibflux_correct<<<nb1(nf), 256, 0, s>>>(f, nf, ucont, immersed, sums);
  HIPCHECK(hipGetLastError());
}

void launch_masked_sum(hipStream_t s, Dims d, const double* x, const uint8_t* cls, double* partial, int* nblk) {
  const int nb = partial_blocks(d.n());
  *nblk = nb;
  k_masked_sum<<<nb, 256, 0, s>>>(d, x, cls, partial, nb);
  HIPCHECK(hipGetLastError());
}

void launch_sub_mean(hipStream_t s, Dims d, double* x, const uint8_t* cls, const double* sum, double cnt) {
  k_sub_mean<<<nb1(d.n()), 256, 0, s>>>(d, x, cls, sum, cnt);
  HIPCHECK(hipGetLastError());
}

void launch_scatter_update(hipStream_t s, Dmda m, Dims d, const double* x, const uint8_t* cls, double* phi, double* p) {
  k_scatter_update<<<cell_grid(d), kBlk, 0, s>>>(m, d, x, cls, phi, p);
  HIPCHECK(hipGetLastError());
}

void launch_gather_cell_metrics(hipStream_t s, Dmda m, Dims d, const double* csi, const double* eta,
                                const double* zet, const double* aj, double* met9, double* ajc) {
  k_gather_cell_metrics<<<cell_grid(d), kBlk, 0, s>>>(m, d, csi, eta, zet, aj, met9, ajc);
  HIPCHECK(hipGetLastError());
}

void launch_gather_phi(hipStream_t s, Dmda m, Dims d, const double* phi, const uint8_t* cls, double* x) {
  k_gather_phi<<<cell_grid(d), kBlk, 0, s>>>(m, d, phi, cls, x);
  HIPCHECK(hipGetLastError());
}

void launch_add_phi(hipStream_t s, Dmda m, Dims d, const uint8_t* cls, const double* phi, double* p) {
  k_add_phi<<<cell_grid(d), kBlk, 0, s>>>(m, d, cls, phi, p);
  HIPCHECK(hipGetLastError());
}

void launch_periodic_scalar(hipStream_t s, Dmda m, double* a, int zero_other) {
  k_periodic_scalar<<<bnd_blocks(m), 256, 0, s>>>(m, a, zero_other);
  HIPCHECK(hipGetLastError());
}

void launch_projection(hipStream_t s, Dmda m, Dims d, const double* g, const uint32_t* code, const uint8_t* cls,
                       const double* phi, double* ucont, double fac) {
  k_projection<<<cell_grid(d), kBlk, 0, s>>>(m, d, g, code, cls, phi, ucont, fac);
  HIPCHECK(hipGetLastError());
}

void launch_ucont_ghosts(hipStream_t s, Dmda m, double* ucont) {
  if (!(m.per[0] || m.per[1] || m.per[2])) return;
  k_ucont_ghosts<<<bnd_blocks(m), 256, 0, s>>>(m, ucont);
  HIPCHECK(hipGetLastError());
}

void launch_c2c(hipStream_t s, Dmda m, Dims d, const double* met9, const uint8_t* cls, const double* nvert,
                const double* ucont, double* ucat, double* scratch, const int bc[6]) {
  k_c2c_core<<<cell_grid(d), kBlk, 0, s>>>(m, d, met9, cls, ucont, ucat);
  HIPCHECK(hipGetLastError());
  k_vec_ghosts<<<bnd_blocks(m), 256, 0, s>>>(m, ucat, scratch);
  HIPCHECK(hipGetLastError());
  k_c2c_bc<<<bnd_blocks(m), 256, 0, s>>>(m, nvert, scratch, ucat, bc[0], bc[1], bc[2], bc[3], bc[4], bc[5]);
  HIPCHECK(hipGetLastError());
  k_c2c_solid<<<cell_grid(d), kBlk, 0, s>>>(d, m, cls, ucat);
  HIPCHECK(hipGetLastError());
  if (m.per[0] || m.per[1] || m.per[2]) {
    k_vec_ghosts<<<bnd_blocks(m), 256, 0, s>>>(m, ucat, nullptr);
    HIPCHECK(hipGetLastError());
  }
}

void launch_divest(hipStream_t s, Dims d, const uint32_t* code, const double* aj, const double* r, double* partial,
                   int* nblk) {
  const unsigned nb = nb1(d.n());
  *nblk = (int)nb;
  k_divest<<<nb, 256, 0, s>>>(d, code, aj, r, partial);
  HIPCHECK(hipGetLastError());
}

void launch_divergence(hipStream_t s, Dmda m, Dims d, const uint32_t* code, const double* aj, const double* ucont,
                       double* partial, int* nblk) {
  const dim3 gr = cell_grid(d);
  *nblk = (int)(gr.x * gr.y * gr.z);
  k_divergence<<<gr, kBlk, 0, s>>>(m, d, code, aj, ucont, partial);
  HIPCHECK(hipGetLastError());
}

void launch_finalize(hipStream_t s, const double* partial, int nvec, int nblk, double* out, int op,
                     double* host_mirror) {
  k_finalize<<<nvec, 1024, 0, s>>>(partial, nblk, out, op, host_mirror);
  HIPCHECK(hipGetLastError());
}

void launch_multidot(hipStream_t s, size_t n, const double* w, const double* V, size_t ldv, int nv, double* partial,
                     int* nblk) {
  const int nb = partial_blocks(n);
  *nblk = nb;
  int done = 0;
  while (done < nv) {
    const int left = nv - done;
    const double* Vp = V + (size_t)done * ldv;
    double* pp = partial + (size_t)done * nb;
    if (left >= 8) { k_multidot<8><<<nb, 256, 0, s>>>(n, w, Vp, ldv, pp, nb); done += 8; }
    else if (left >= 4) { k_multidot<4><<<nb, 256, 0, s>>>(n, w, Vp, ldv, pp, nb); done += 4; }
    else if (left >= 2) { k_multidot<2><<<nb, 256, 0, s>>>(n, w, Vp, ldv, pp, nb); done += 2; }
    else { k_multidot<1><<<nb, 256, 0, s>>>(n, w, Vp, ldv, pp, nb); done += 1; }
    HIPCHECK(hipGetLastError());
  }
}

void launch_axpy_multi(hipStream_t s, size_t n, double* w, const double* V, size_t ldv, int nv, const double* h,
                       double sign) {
  k_axpy_multi<<<nb1(n), 256, 0, s>>>(n, w, V, ldv, nv, h, sign);
  HIPCHECK(hipGetLastError());
}

void launch_norm2(hipStream_t s, size_t n, const double* a, double* partial, int* nblk) {
  const int nb = partial_blocks(n);
  *nblk = nb;
  k_norm2<<<nb, 256, 0, s>>>(n, a, partial, nb);
  HIPCHECK(hipGetLastError());
}

void launch_scale_to(hipStream_t s, size_t n, const double* a, const double* norm2, double* out) {
  k_scale_to<<<nb1(n), 256, 0, s>>>(n, a, norm2, out);
  HIPCHECK(hipGetLastError());
}

template <typename TZ>
void launch_update_x(hipStream_t s, size_t n, double* x, const TZ* Z, size_t ldz, int nz, const double* y) {
  k_update_x<TZ><<<nb1(n), 256, 0, s>>>(n, x, Z, ldz, nz, y);
  HIPCHECK(hipGetLastError());
}
template void launch_update_x<float>(hipStream_t, size_t, double*, const float*, size_t, int, const double*);
template void launch_update_x<double>(hipStream_t, size_t, double*, const double*, size_t, int, const double*);

template <typename TA, typename TB>
void launch_convert(hipStream_t s, size_t n, const TA* a, TB* b) {
  k_convert<TA, TB><<<nb1(n), 256, 0, s>>>(n, a, b);
  HIPCHECK(hipGetLastError());
}
template void launch_convert<double, float>(hipStream_t, size_t, const double*, float*);
template void launch_convert<float, double>(hipStream_t, size_t, const float*, double*);
template void launch_convert<double, double>(hipStream_t, size_t, const double*, double*);
template <typename TA, typename TB>
void launch_convert_scaled(hipStream_t s, size_t n, const TA* a, double scale, TB* b) {
  k_convert_scaled<TA, TB><<<nb1(n), 256, 0, s>>>(n, a, scale, b);
  HIPCHECK(hipGetLastError());
}
template void launch_convert_scaled<double, _Float16>(hipStream_t, size_t, const double*, double, _Float16*);

void launch_fill(hipStream_t s, size_t n, double* a, double v) {
  k_fill<double><<<nb1(n), 256, 0, s>>>(n, a, v);
  HIPCHECK(hipGetLastError());
}
void launch_copy(hipStream_t s, size_t n, const double* a, double* b) {
  if (n % 2) throw StatusError(1, "copy length must be even");
  k_copy16<<<nb1(n / 2), 256, 0, s>>>(n / 2, reinterpret_cast<const double2*>(a), reinterpret_cast<double2*>(b));
  HIPCHECK(hipGetLastError());
}
void launch_fillf(hipStream_t s, size_t n, float* a, float v) {
  k_fill<float><<<nb1(n), 256, 0, s>>>(n, a, v);
  HIPCHECK(hipGetLastError());
}

template <typename TZ>
void launch_apply_dots(hipStream_t s, Dims d, const double* g, const uint32_t* code, const TZ* z, double* w,
                       const double* V, size_t ldv, int nv, double* partial, int* nblk) {
  const dim3 gr = cell_grid(d);
  *nblk = (int)(4 * gr.x * gr.y * gr.z);  // one partial per wave
  k_apply_dots<TZ><<<gr, kBlk, 0, s>>>(d, g, code, z, w, V, ldv, nv, partial);
  HIPCHECK(hipGetLastError());
}
template void launch_apply_dots<float>(hipStream_t, Dims, const double*, const uint32_t*, const float*, double*,
                                       const double*, size_t, int, double*, int*);
template void launch_apply_dots<double>(hipStream_t, Dims, const double*, const uint32_t*, const double*, double*,
                                        const double*, size_t, int, double*, int*);

void launch_axpy_dots(hipStream_t s, size_t n, double* w, const double* V, size_t ldv, int nv, const double* h,
                      double* partial, int* nblk) {
  if (nv + 1 <= 32) {
    const unsigned nb = std::min<unsigned>(nb1(n), 2048u);
    *nblk = (int)nb;
    if (nv + 1 <= 8) k_axpy_dots_acc<8><<<nb, 256, 0, s>>>(n, w, V, ldv, nv, h, partial);
    else if (nv + 1 <= 16) k_axpy_dots_acc<16><<<nb, 256, 0, s>>>(n, w, V, ldv, nv, h, partial);
    else k_axpy_dots_acc<32><<<nb, 256, 0, s>>>(n, w, V, ldv, nv, h, partial);
  } else {
    *nblk = (int)nb1(n);
    k_axpy_dots<<<nb1(n), 256, 0, s>>>(n, w, V, ldv, nv, h, partial);
  }
  HIPCHECK(hipGetLastError());
}

template <typename TF>
void launch_axpy_norm(hipStream_t s, size_t n, double* w, const double* V, size_t ldv, int nv, const double* dots,
                      const double* nuinv2, const double* a2, TF* f, double* partial, int* nblk, const TF* dinv,
                      TF* fd) {
  const unsigned nb = std::min<unsigned>(nb1(n), 2048u);
  *nblk = (int)nb;
  k_axpy_norm<TF><<<nb, 256, 0, s>>>(n, w, V, ldv, nv, dots, nuinv2, a2, f, partial, dinv, fd);
  HIPCHECK(hipGetLastError());
}
template void launch_axpy_norm<float>(hipStream_t, size_t, double*, const double*, size_t, int, const double*,
                                      const double*, const double*, float*, double*, int*, const float*, float*);
template void launch_axpy_norm<double>(hipStream_t, size_t, double*, const double*, size_t, int, const double*,
                                       const double*, const double*, double*, double*, int*, const double*,
                                       double*);

double host_conv_scale(double a2) { return conv_scale(a2); }

void launch_gs_norm(hipStream_t s, double* red, int off, int nv, double* host) {
  k_gs_norm<<<1, 64, 0, s>>>(red, off, nv, host);
  HIPCHECK(hipGetLastError());
}

template <typename TF>
void launch_axpy_scale(hipStream_t s, size_t n, double* w, const double* V, size_t ldv, int nv, const double* h,
                       const double* inv, TF* f, const TF* dinv, TF* fd) {
  k_axpy_scale<TF><<<nb1(n), 256, 0, s>>>(n, w, V, ldv, nv, h, inv, f, dinv, fd);
  HIPCHECK(hipGetLastError());
}
template void launch_axpy_scale<float>(hipStream_t, size_t, double*, const double*, size_t, int, const double*,
                                       const double*, float*, const float*, float*);
template void launch_axpy_scale<double>(hipStream_t, size_t, double*, const double*, size_t, int, const double*,
                                        const double*, double*, const double*, double*);

template <typename TF>
void launch_scale_conv(hipStream_t s, size_t n, const double* a, const double* nrm2, double* v, TF* f,
                       const TF* dinv, TF* fd) {
  k_scale_conv<TF><<<nb1(n), 256, 0, s>>>(n, a, nrm2, v, f, dinv, fd);
  HIPCHECK(hipGetLastError());
}
template void launch_scale_conv<float>(hipStream_t, size_t, const double*, const double*, double*, float*,
                                       const float*, float*);
template void launch_scale_conv<double>(hipStream_t, size_t, const double*, const double*, double*, double*,
                                        const double*, double*);

}  // namespace pp

// Pressure_Gradient (momentum.c:223-452) on the device: the contravariant
// pressure-gradient flux dP on the i, j and k faces of every interior cell from P
// and the face metrics, for the momentum step that follows the projection.
// Levelset density, the -dpdz mean-gradient forcing and the -flux inlet
// correction are off, as they are by default (main.c:73, 95).  The arithmetic
// keeps the reference's operand order and the kernel is compiled without
// contraction, so dP is bit-exact with a CPU restatement.
#include "pp_engine.hpp"

namespace pp {

namespace {

#pragma clang fp contract(off)

struct GradIn {
  const double *icsi, *ieta, *izet, *jcsi, *jeta, *jzet, *kcsi, *keta, *kzet, *iaj, *jaj, *kaj, *nvert;
};

__device__ __forceinline__ double dot3(const double* a, const double* b, size_t c) {
  return a[3 * c] * b[3 * c] + a[3 * c + 1] * b[3 * c + 1] + a[3 * c + 2] * b[3 * c + 2];
}

__global__ __launch_bounds__(256) void k_pressure_gradient(Dmda m, GradIn in, const double* __restrict__ p,
                                                           double* __restrict__ dp) {
  const int i = blockIdx.x * 64 + threadIdx.x, j = blockIdx.y * 4 + threadIdx.y, k = blockIdx.z;
  if (i >= m.mx || j >= m.my) return;
  const size_t c = m.idx(i, j, k);
  if (i == 0 || j == 0 || k == 0 || i == m.mx - 1 || j == m.my - 1 || k == m.mz - 1) {
    dp[3 * c] = 0.;  // VecSet(dP, 0.): only interior points are written below
    dp[3 * c + 1] = 0.;
    dp[3 * c + 2] = 0.;
    return;
  }
  const int mx = m.mx, my = m.my, mz = m.mz;
  const bool pi = m.per[0], pj = m.per[1], pk = m.per[2];
  auto P = [&](int kk, int jj, int ii) { return p[m.idx(ii, jj, kk)]; };
  auto N1 = [&](int kk, int jj, int ii) { return (int)(in.nvert[m.idx(ii, jj, kk)] + 0.5) == 1; };
  const double g11 = dot3(in.icsi, in.icsi, c), g12 = dot3(in.ieta, in.icsi, c), g13 = dot3(in.izet, in.icsi, c);
  const double g21 = dot3(in.jcsi, in.jeta, c), g22 = dot3(in.jeta, in.jeta, c), g23 = dot3(in.jzet, in.jeta, c);
  const double g31 = dot3(in.kcsi, in.kzet, c), g32 = dot3(in.keta, in.kzet, c), g33 = dot3(in.kzet, in.kzet, c);
  double dpdc, dpde, dpdz;
  // i face (momentum.c:352-364)
  if (i == mx - 2 && pi) dpdc = P(k, j, 1) - P(k, j, i);
  else dpdc = P(k, j, i + 1) - P(k, j, i);
  if (N1(k, j + 1, i) || N1(k, j + 1, i + 1) || (j == my - 2 && !pj))
    dpde = (P(k, j, i) - P(k, j - 1, i) + P(k, j, i + 1) - P(k, j - 1, i + 1)) * 0.5;
  else if (N1(k, j - 1, i) || N1(k, j - 1, i + 1) || (j == 1 && !pj))
    dpde = (P(k, j + 1, i) - P(k, j, i) + P(k, j + 1, i + 1) - P(k, j, i + 1)) * 0.5;
  else dpde = (P(k, j + 1, i) - P(k, j - 1, i) + P(k, j + 1, i + 1) - P(k, j - 1, i + 1)) * 0.25;
  if (N1(k + 1, j, i) || N1(k + 1, j, i + 1) || (k == mz - 2 && !pk))
    dpdz = (P(k, j, i) - P(k - 1, j, i) + P(k, j, i + 1) - P(k - 1, j, i + 1)) * 0.5;
  else if (N1(k - 1, j, i) || N1(k - 1, j, i + 1) || (k == 1 && !pk))
    dpdz = (P(k + 1, j, i) - P(k, j, i) + P(k + 1, j, i + 1) - P(k, j, i + 1)) * 0.5;
  else dpdz = (P(k + 1, j, i) - P(k - 1, j, i) + P(k + 1, j, i + 1) - P(k - 1, j, i + 1)) * 0.25;
  double dx = (dpdc * g11 + dpde * g12 + dpdz * g13) * in.iaj[c];
  // j face (momentum.c:366-380)
  if (N1(k, j, i + 1) || N1(k, j + 1, i + 1) || (i == mx - 2 && !pi))
    dpdc = (P(k, j, i) - P(k, j, i - 1) + P(k, j + 1, i) - P(k, j + 1, i - 1)) * 0.5;
  else if (N1(k, j, i - 1) || N1(k, j + 1, i - 1) || (i == 1 && !pi))
    dpdc = (P(k, j, i + 1) - P(k, j, i) + P(k, j + 1, i + 1) - P(k, j + 1, i)) * 0.5;
  else dpdc = (P(k, j, i + 1) - P(k, j, i - 1) + P(k, j + 1, i + 1) - P(k, j + 1, i - 1)) * 0.25;
  if (j == my - 2 && pj) dpde = P(k, 1, i) - P(k, j, i);
  else dpde = P(k, j + 1, i) - P(k, j, i);
  if (N1(k + 1, j, i) || N1(k + 1, j + 1, i) || (k == mz - 2 && !pk))
    dpdz = (P(k, j, i) - P(k - 1, j, i) + P(k, j + 1, i) - P(k - 1, j + 1, i)) * 0.5;
  else if (N1(k - 1, j, i) || N1(k - 1, j + 1, i) || (k == 1 && !pk))
    dpdz = (P(k + 1, j, i) - P(k, j, i) + P(k + 1, j + 1, i) - P(k, j + 1, i)) * 0.5;
  else dpdz = (P(k + 1, j, i) - P(k - 1, j, i) + P(k + 1, j + 1, i) - P(k - 1, j + 1, i)) * 0.25;
  double dy = (dpdc * g21 + dpde * g22 + dpdz * g23) * in.jaj[c];
  // k face (momentum.c:382-396)
  if (N1(k, j, i + 1) || N1(k + 1, j, i + 1) || (i == mx - 2 && !pi))
    dpdc = (P(k, j, i) - P(k, j, i - 1) + P(k + 1, j, i) - P(k + 1, j, i - 1)) * 0.5;
  else if (N1(k, j, i - 1) || N1(k + 1, j, i - 1) || (i == 1 && !pi))
    dpdc = (P(k, j, i + 1) - P(k, j, i) + P(k + 1, j, i + 1) - P(k + 1, j, i)) * 0.5;
  else dpdc = (P(k, j, i + 1) - P(k, j, i - 1) + P(k + 1, j, i + 1) - P(k + 1, j, i - 1)) * 0.25;
  if (N1(k, j + 1, i) || N1(k + 1, j + 1, i) || (j == my - 2 && !pj))
    dpde = (P(k, j, i) - P(k, j - 1, i) + P(k + 1, j, i) - P(k + 1, j - 1, i)) * 0.5;
  else if (N1(k, j - 1, i) || N1(k + 1, j - 1, i) || (j == 1 && !pj))
    dpde = (P(k, j + 1, i) - P(k, j, i) + P(k + 1, j + 1, i) - P(k + 1, j, i)) * 0.5;
  else dpde = (P(k, j + 1, i) - P(k, j - 1, i) + P(k + 1, j + 1, i) - P(k + 1, j - 1, i)) * 0.25;
  if (k == mz - 2 && pk) dpdz = P(1, j, i) - P(k, j, i);
  else dpdz = P(k + 1, j, i) - P(k, j, i);
  double dz = (dpdc * g31 + dpde * g32 + dpdz * g33) * in.kaj[c];
  // faces without a fluid pair or on a non-periodic wall carry none (momentum.c:426-436)
  const double nv = in.nvert[c];
  if (nv + in.nvert[m.idx(i + 1, j, k)] > 0.1 || (!pi && i == mx - 2)) dx = 0;
  if (nv + in.nvert[m.idx(i, j + 1, k)] > 0.1 || (!pj && j == my - 2)) dy = 0;
  if (nv + in.nvert[m.idx(i, j, k + 1)] > 0.1 || (!pk && k == mz - 2)) dz = 0;
  dp[3 * c] = dx;
  dp[3 * c + 1] = dy;
  dp[3 * c + 2] = dz;
}

}  // namespace

}  // namespace pp

using namespace pp;

int vfs_pp_pressure_gradient_device(const vfs_grid* g, const double* icsi, const double* ieta, const double* izet,
                                    const double* jcsi, const double* jeta, const double* jzet, const double* kcsi,
                                    const double* keta, const double* kzet, const double* iaj, const double* jaj,
                                    const double* kaj, const double* nvert, double* p, double* dp) {
  ABI_BEGIN
  if (!g || !nvert || !p || !dp) return VFS_PP_EINVAL;
  const double* mt[12] = {icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj};
  for (auto q : mt)
    if (!q) return VFS_PP_EINVAL;
  if (g->nx < 2 || g->ny < 2 || g->nz < 2) return VFS_PP_EINVAL;
  const Dmda m{g->nx + 2, g->ny + 2, g->nz + 2, {g->periodic[0] ? 1 : 0, g->periodic[1] ? 1 : 0, g->periodic[2] ? 1 : 0}};
  // periodic ghost copies of P (momentum.c:270-307), then the face fluxes
  launch_periodic_scalar(nullptr, m, p, 0);
  const GradIn in{icsi, ieta, izet, jcsi, jeta, jzet, kcsi, keta, kzet, iaj, jaj, kaj, nvert};
  k_pressure_gradient<<<dim3((m.mx + 63) / 64, (m.my + 3) / 4, m.mz), dim3(64, 4, 1), 0, nullptr>>>(m, in, p, dp);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipDeviceSynchronize());
  return VFS_PP_OK;
  ABI_END
}

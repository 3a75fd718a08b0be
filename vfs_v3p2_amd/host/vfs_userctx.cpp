// C++ host side of the pressure path (see vfs_userctx.hpp): device fields of one
// block and the reference's per-step calls, each forwarded to the C ABI.
#include "vfs_userctx.hpp"

#include <hip/hip_runtime.h>

namespace vfs {

namespace {

int alloc_zero(double **p, size_t n) {
  if (hipMalloc(p, n * sizeof(double)) != hipSuccess) return VFS_PP_ENOMEM;
  if (hipMemset(*p, 0, n * sizeof(double)) != hipSuccess) return VFS_PP_EHIP;
  return VFS_PP_OK;
}

int copy(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  return hipMemcpy(dst, src, bytes, kind) == hipSuccess ? VFS_PP_OK : VFS_PP_EHIP;
}

}  // namespace

int UserCtxCreate(const vfs_grid &g, double dt, double st, const vfs_pp_opts *opts, UserCtx *user) {
  if (!user || !(dt > 0.)) return VFS_PP_EINVAL;
  *user = UserCtx{};
  user->grid = g;
  user->dt = dt;
  user->st = st;
  vfs_pp_opts o;
  vfs_pp_default_opts(&o);
  if (opts) o = *opts;
  int rc = vfs_pp_create(&g, &o, &user->h);
  if (rc != VFS_PP_OK) return rc;
  const size_t n = user->points();
  double **vec[3] = {&user->Coor, &user->Ucont, &user->Ucat};
  double **sca[3] = {&user->Nvert, &user->Phi, &user->P};
  for (double **a : vec)
    if ((rc = alloc_zero(a, 3 * n)) != VFS_PP_OK) return rc;
  for (double **a : sca)
    if ((rc = alloc_zero(a, n)) != VFS_PP_OK) return rc;
  return VFS_PP_OK;
}

int UserCtxDestroy(UserCtx *user) {
  if (!user) return VFS_PP_EINVAL;
  for (double *a : {user->Coor, user->Nvert, user->Ucont, user->Ucat, user->Phi, user->P})
    if (a) (void)hipFree(a);
  int rc = user->h ? vfs_pp_destroy(user->h) : VFS_PP_OK;
  *user = UserCtx{};
  return rc;
}

int UserCtxUpload(UserCtx *user, const double *coor, const double *nvert, const double *ucont) {
  if (!user || !user->h) return VFS_PP_EINVAL;
  const size_t n = user->points();
  int rc = VFS_PP_OK;
  if (coor && (rc = copy(user->Coor, coor, 3 * n * sizeof(double), hipMemcpyHostToDevice)) != VFS_PP_OK) return rc;
  if (nvert && (rc = copy(user->Nvert, nvert, n * sizeof(double), hipMemcpyHostToDevice)) != VFS_PP_OK) return rc;
  if (ucont && (rc = copy(user->Ucont, ucont, 3 * n * sizeof(double), hipMemcpyHostToDevice)) != VFS_PP_OK)
    return rc;
  return VFS_PP_OK;
}

int UserCtxDownload(const UserCtx *user, double *ucont, double *ucat, double *phi, double *p) {
  if (!user || !user->h) return VFS_PP_EINVAL;
  const size_t n = user->points();
  int rc = VFS_PP_OK;
  if (ucont && (rc = copy(ucont, user->Ucont, 3 * n * sizeof(double), hipMemcpyDeviceToHost)) != VFS_PP_OK)
    return rc;
  if (ucat && (rc = copy(ucat, user->Ucat, 3 * n * sizeof(double), hipMemcpyDeviceToHost)) != VFS_PP_OK) return rc;
  if (phi && (rc = copy(phi, user->Phi, n * sizeof(double), hipMemcpyDeviceToHost)) != VFS_PP_OK) return rc;
  if (p && (rc = copy(p, user->P, n * sizeof(double), hipMemcpyDeviceToHost)) != VFS_PP_OK) return rc;
  return VFS_PP_OK;
}

int FormMetrics(UserCtx *user) {
  if (!user || !user->h) return VFS_PP_EINVAL;
  return vfs_pp_set_grid_device(user->h, user->Coor, user->Nvert);
}

int PoissonSolver_Hypre(UserCtx *user, double coeff) {
  if (!user || !user->h) return VFS_PP_EINVAL;
  return vfs_pp_poisson_solve_device(user->h, user->Ucont, user->Phi, user->dt, user->st, coeff, &user->stats);
}

int UpdatePressure(UserCtx *user) {
  if (!user || !user->h) return VFS_PP_EINVAL;
  return vfs_pp_update_pressure_device(user->h, user->P, user->Phi);
}

int Projection(UserCtx *user, double coeff) {
  if (!user || !user->h) return VFS_PP_EINVAL;
  return vfs_pp_projection_device(user->h, user->Ucont, user->Phi, user->Ucat, user->dt, user->st, coeff);
}

int Divergence(UserCtx *user, double *maxdiv) {
  if (!user || !user->h || !maxdiv) return VFS_PP_EINVAL;
  return vfs_pp_divergence_device(user->h, user->Ucont, maxdiv);
}

int Flow_Solver_pressure_step(UserCtx *user, double coeff, double *maxdiv) {
  // a solve that misses its tolerance is reported and the step goes on, as in
  // the reference
  int solve = PoissonSolver_Hypre(user, coeff);
  if (solve != VFS_PP_OK && solve != VFS_PP_ENOCONV) return solve;
  int rc = UpdatePressure(user);
  if (rc != VFS_PP_OK) return rc;
  if ((rc = Projection(user, coeff)) != VFS_PP_OK) return rc;
  if ((rc = Divergence(user, maxdiv)) != VFS_PP_OK) return rc;
  return solve;
}

}  // namespace vfs

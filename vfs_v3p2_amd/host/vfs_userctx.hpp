// C++ host side of the pressure path over the C ABI of libvfs_pp.so, mirroring
// the reference's interface: a UserCtx holding one block's fields (the members
// of variables.h:298-591 the path reads and writes) and the calls Flow_Solver
// makes per block and time step (solvers.c:874-895):
//
//     PoissonSolver_Hypre(user, ibm, ibm_intp)   poisson_hypre.c:580
//     UpdatePressure(user)                       poisson.c:2473
//     Projection(user)                           poisson.c:2122 (ends with Contra2Cart)
//     Divergence(user)                           main.c:1244
//
// The fields live on the device in the reference's DMDA layout ([mz][my][mx], Cmpnts
// as three consecutive doubles); the arithmetic is all in the library.  Every call
// returns the library's status (0 = success, vfs_pp.h's VFS_PP_* codes), the way the
// reference's calls return a PetscErrorCode; a solve that misses its tolerance
// returns VFS_PP_ENOCONV with its statistics filled in, where the reference prints
// nothing and goes on.
#pragma once

#include <cstddef>

#include "vfs_pp.h"

namespace vfs {

struct UserCtx {
  vfs_grid grid{};             // IM - 1, JM - 1, KM - 1 interior cells, periodicity, bctype, immersed
  vfs_pp_handle *h = nullptr;  // metrics, masks, multigrid hierarchy (the HYPRE objects' place)
  double dt = 0.;              // user->dt
  double st = 1.;              // user->st
  // device DMDA arrays owned by the context
  double *Coor = nullptr;   // node coordinates [mz][my][mx][3] (DMDAGetCoordinates)
  double *Nvert = nullptr;  // lNvert [mz][my][mx]
  double *Ucont = nullptr;  // Ucont [mz][my][mx][3]
  double *Ucat = nullptr;   // Ucat
  double *Phi = nullptr;    // Phi [mz][my][mx]
  double *P = nullptr;      // P
  vfs_pp_stats stats{};     // the last solve's statistics

  size_t points() const {
    return (size_t)(grid.nx + 2) * (size_t)(grid.ny + 2) * (size_t)(grid.nz + 2);
  }
};

// Allocates the device fields (zeroed) and the engine handle for grid g; opts may
// be null (vfs_pp_default_opts).
int UserCtxCreate(const vfs_grid &g, double dt, double st, const vfs_pp_opts *opts, UserCtx *user);
int UserCtxDestroy(UserCtx *user);

// Host arrays in, device fields out (DMDA layout, points() or 3 points() doubles).
int UserCtxUpload(UserCtx *user, const double *coor, const double *nvert, const double *ucont);
int UserCtxDownload(const UserCtx *user, double *ucont, double *ucat, double *phi, double *p);

// FormMetrics (metrics.c:4-1290) and the Poisson set-up of the first time step
// (setup_lidx2 + PoissonLHSNew + the solver objects, poisson_hypre.c:609-627), from
// the context's Coor and Nvert.
int FormMetrics(UserCtx *user);

// The pressure path; coeff is time_coeff() (poisson.c:22-30).
int PoissonSolver_Hypre(UserCtx *user, double coeff);
int UpdatePressure(UserCtx *user);
int Projection(UserCtx *user, double coeff);
int Divergence(UserCtx *user, double *maxdiv);

// The three calls of Flow_Solver's pressure block and the Divergence check.
int Flow_Solver_pressure_step(UserCtx *user, double coeff, double *maxdiv);

}  // namespace vfs

/*
 * vfs_pp -- MI355X pressure-projection engine, C ABI.
 *
 * Drop-in for the pressure path of the VFS curvilinear immersed-boundary solver
 * (trunglendsu/VFS_v3p2).  The reference calls, once per block per time step
 * (solvers.c:874-895):
 *     PoissonSolver_Hypre(user, ibm, ibm_intp);   poisson_hypre.c:580  (variables.h:785)
 *     UpdatePressure(user);                        poisson.c:2473       (variables.h:752)
 *     Projection(user);                            poisson.c:2122       (variables.h:753)
 * with every field carried in UserCtx (variables.h:298-591).  This library takes
 * the same fields as plain arrays and runs all per-cell work on one GPU.
 *
 * Array layout (all entry points): the single-rank DMDA layout the reference
 * uses -- a dense [mz][my][mx] array with i fastest, mx = nx+2, my = ny+2,
 * mz = nz+2 (planes 0 and m-1 are the boundary/ghost planes, init.c:156-157).
 * Vector fields (Cmpnts, variables.h:37-39) are AoS {x,y,z} FP64 triples.
 * "_device" entry points take device pointers (hipMalloc'd or framework
 * tensors); the others take host pointers and copy over PCIe.
 *
 * Status codes: 0 ok, otherwise one of VFS_PP_E* below.  The reference returns
 * PetscErrorCode 0 unconditionally and is silent on non-convergence; this ABI
 * reports it.
 *
 * Thread safety: one handle may be used by one host thread at a time.  Each
 * handle owns one HIP stream (default priority); calls are synchronous unless
 * stated otherwise.  Engine streams are blocking streams: their work runs after
 * work queued earlier on the legacy default stream (e.g. the kernels that wrote
 * a call's input arrays) and before work queued there afterwards; callers that
 * produce inputs on other non-blocking streams synchronise those first.
 */
#ifndef VFS_PP_H
#define VFS_PP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VFS_PP_OK 0
#define VFS_PP_EINVAL 1      /* bad argument or shape */
#define VFS_PP_ENOMEM 2      /* device allocation failed / admission refused */
#define VFS_PP_EHIP 3        /* HIP runtime error */
#define VFS_PP_ENOCONV 4     /* solver did not reach the requested tolerance */
#define VFS_PP_ESTATE 5      /* call order violated (e.g. solve before set_geometry) */
#define VFS_PP_ENONFINITE 6  /* NaN/Inf in the inputs or the iterate */
#define VFS_PP_EIO 7         /* grid file missing, truncated or malformed */
#define VFS_PP_EBUSY 8       /* served form: the request queue is full, request refused */

typedef struct vfs_pp_handle vfs_pp_handle;

/* Grid description: extents of interior cells and the reference's flags. */
typedef struct {
  int nx, ny, nz;   /* interior cells; DMDA extents are nx+2 etc. (init.c:156-157) */
  int periodic[3];  /* i_periodic || ii_periodic, j..., k... (main.c:2150-2160): both
                       wrap the same way on the dense [mz][my][mx] arrays every entry
                       point takes (ii_periodic's ghost index -2 is the point mx-2 that
                       i_periodic reads, poisson_hypre.c:358-359), so one flag serves
                       both.  Under ii_periodic the reference's ghosted local arrays are
                       3 points wider on that axis (DMDA_BOUNDARY_PERIODIC, stencil
                       width 3, init.c:147-158): pass the global vectors' arrays, dense
                       on one rank, or repack the local ones with vfs_pp_repack. */
  int bctype[6];    /* user->bctype (init.c:502); Contra2Cart honours 1 (no-slip) */
  int immersed;     /* the reference's `immersed` switch (poisson_hypre.c:662-667):
                       0 or 3: VolumeFlux(user,Ucont,..) poisson.c:2634, corrected faces
                       set to zero (poisson.c:2728); 1: flg variant poisson.c:2898;
                       2 (IB flux added to the outlet, Add_IBMFlux_to_Outlet) is not
                       supported: vfs_pp_create returns VFS_PP_EINVAL */
} vfs_grid;

/* Solver options.  Defaults (vfs_pp_default_opts) give the mixed-precision form. */
typedef struct {
  int precision;     /* 0: FP64 multigrid preconditioner (plain FP64 drop-in)
                        1: FP32 multigrid inside FP64 FGMRES (mixed precision) */
  double rtol;       /* stop when ||b - A phi||_2 <= rtol ||b||_2 (flux-form residual,
                        the norm HYPRE GMRES uses, poisson_hypre.c:90-95) */
  double div_tol;    /* additionally require max|aj div u| <= div_tol after projection
                        (main.c:1283); <= 0 disables the check */
  int max_iters;     /* outer FGMRES iterations (poisson_it, main.c:126, default 10) */
  int restart;       /* FGMRES restart length (HYPRE KDim 51, poisson_hypre.c:82) */
  int smoother;      /* 0 auto, 1 point Chebyshev-Jacobi, 2 y-line Chebyshev */
  int cheb_degree;   /* smoothing steps per pre/post smoothing */
  int coarse_max;    /* cell budget of the coarsest (direct-solve) level */
  int drop_small;    /* 1: drop |a_ij| <= 1e-10 like the assembled HYPRE matrix
                        (poisson.c:1608); 0: conservative operator (default) */
  int zero_guess;    /* 1: start from phi = 0 (default);
                        0: start from the phi passed in, as the reference does from its
                        second time step on (its HYPRE solution vector persists between
                        calls, poisson_hypre.c:565-574, 722) */
  double cheb_lo;    /* Chebyshev smoothing interval [cheb_lo, cheb_hi] x lambda_max(M^-1 A) */
  double cheb_hi;
  int orthog;        /* Gram-Schmidt of the Krylov basis (default 2): 0 classical with
                        re-orthogonalisation (CGS2) fused into three passes over the
                        basis (the first with the operator apply); 1 CGS2, one pass
                        per step; 2 one classical pass on a lazily normalised basis
                        (fused with the operator and with the next V-cycle's input),
                        repeated only when it cancels |w| below 1e-4 |w| */
  int verbose;       /* 1: print the residual history of each solve to stderr */
  int fused;         /* fused launches, each giving the same results bit for bit as the
                        separate ones (default 3):
                        bit 0: in the FP32 multigrid, the V-cycle's last point-smoothed level
                               (at most 1024 cells) and the dense solve run as one workgroup
                               with the level's iterates in LDS;
                        bit 1: the writers of the V-cycle input also store f D^-1 of the
                               finest level, which its fused first two Chebyshev steps then
                               read at the 19 neighbours instead of f and D^-1 */
} vfs_pp_opts;

typedef struct {
  int iterations;        /* outer Krylov iterations */
  int vcycles;           /* preconditioner applications */
  int converged;         /* 1 if rtol (and div_tol) were met */
  double rel_residual;   /* ||b - A phi|| / ||b|| after the solve, FP64 */
  double bnorm;          /* ||b||_2 */
  double max_div;        /* max |aj div u| after projection (Divergence, main.c:1283) */
  double ibm_flux[4];    /* VolumeFlux: flux/area before and after correction */
  double phi_mean;       /* mean subtracted by Remove_Nullspace */
  double ms_total;       /* wall time of the call (ms) */
} vfs_pp_stats;

void vfs_pp_default_opts(vfs_pp_opts *opts);

/* Create a handle for one grid.  Replaces the ti==tistart set-up branch of
 * PoissonSolver_Hypre (setup_lidx2 + Create_Hypre_* + PoissonLHSNew,
 * poisson_hypre.c:609-627).  VFS_PP_EINVAL if an axis has fewer than 2 cells (3
 * if periodic) or the DMDA box (nx+2)(ny+2)(nz+2) has 2^29 points or more. */
int vfs_pp_create(const vfs_grid *grid, const vfs_pp_opts *opts, vfs_pp_handle **out);
int vfs_pp_destroy(vfs_pp_handle *h);
/* New options.  precision and restart size the Krylov storage the first
 * vfs_pp_set_geometry allocates: changing them afterwards returns VFS_PP_ESTATE
 * (also after a set-up that failed); smoother, cheb_degree and coarse_max shape
 * the hierarchy and can change only while no hierarchy is in place. */
int vfs_pp_set_opts(vfs_pp_handle *h, const vfs_pp_opts *opts);

/* Geometry: the UserCtx fields lICsi..lKZet, lIAj..lKAj, lAj, lCsi/lEta/lZet and
 * lNvert (variables.h:333-356, 378-385), host pointers in DMDA layout.  Builds the
 * face coefficients (poisson.c:1105-1141), the per-cell mask codes of
 * PoissonLHSNew's branches (poisson.c:1146-1558), the VolumeFlux face list and the
 * multigrid hierarchy, all on the device; everything stays resident there.  A
 * later call replaces the geometry and frees the previous hierarchy: the
 * reference's per-step rebuild for moving or rotating bodies, sediment and level
 * sets (setup_lidx2 + PoissonLHSNew + a new solver, poisson_hypre.c:627-648). */
int vfs_pp_set_geometry(vfs_pp_handle *h, const double *icsi, const double *ieta,
                        const double *izet, const double *jcsi, const double *jeta,
                        const double *jzet, const double *kcsi, const double *keta,
                        const double *kzet, const double *iaj, const double *jaj,
                        const double *kaj, const double *aj, const double *csi,
                        const double *eta, const double *zet, const double *nvert);

/* Pressure_Gradient (momentum.c:223-452), the step after the projection: the
 * contravariant pressure-gradient flux dP (AoS, DMDA) on the i, j and k faces of
 * every interior cell from P, with one-sided tangential differences next to
 * nvert == 1 cells and non-periodic walls; zero on faces without a fluid pair and
 * on boundary points.  All arrays on the device in the DMDA layout; P first gets
 * its periodic ghost copies in place, as the reference does.  Levelset density
 * and the -dpdz / -flux forcing terms (off by default) are not applied.
 * Bit-exact with the reference's arithmetic. */
int vfs_pp_pressure_gradient_device(const vfs_grid *grid, const double *icsi, const double *ieta,
                                    const double *izet, const double *jcsi, const double *jeta,
                                    const double *jzet, const double *kcsi, const double *keta,
                                    const double *kzet, const double *iaj, const double *jaj,
                                    const double *kaj, const double *nvert, double *p, double *dp);

/* Formfunction_2 (momentum.c:529-2034), the momentum right-hand side of the
 * step before the projection, on device DMDA arrays: convective flux by
 * fourth-order central interpolation (one-sided upwind beside IB points when
 * grid->immersed != 0; zero IB-face convection when it is 3), laminar viscous
 * flux with nu = 1/ren, in contravariant form, times scale; zero on IB faces
 * and boundary points.  Ucont is read with whole-vector periodic ghosts and
 * not modified.  The reference's default branch: no LES/RANS, level set, skew
 * form, second-order or inviscid switch, no bctype 11 force sums.  Allocates
 * 8 vector-field scratch arrays per call; bit-exact with the reference's
 * arithmetic.  Runs on the legacy default stream and returns when done. */
int vfs_pp_formfunction2_device(const vfs_grid *grid, const double *icsi, const double *ieta,
                                const double *izet, const double *jcsi, const double *jeta,
                                const double *jzet, const double *kcsi, const double *keta,
                                const double *kzet, const double *iaj, const double *jaj,
                                const double *kaj, const double *csi, const double *eta,
                                const double *zet, const double *nvert, const double *ucont,
                                const double *ucat, double ren, double scale, double *rhs);

/* RungeKutta (rhs.c:1511-1631): one explicit four-stage step, stage s setting
 * Ucont = pUcont + alfa_s dt st (Formfunction_2(Ucont, Ucat) - dP) with
 * alfa = 1/4, 1/3, 1/2, 1 and Contra2Cart before and after, on the handle's
 * grid, nvert and cell metrics (set up as for the projection) with the face and
 * cell metrics passed here.  ucont and ucat (device, DMDA) are updated in
 * place; dp is Pressure_Gradient's output (vfs_pp_pressure_gradient_device).
 * The inflow/outflow boundary (InflowFlux, FormBCS) and the IB velocity
 * interpolation (ibm_interpolation_advanced) are the caller's: Contra2Cart
 * leaves the Ucat of IB-adjacent cells as passed.  Allocates 10 vector-field
 * scratch arrays per call. */
int vfs_pp_runge_kutta_device(vfs_pp_handle *h, const double *icsi, const double *ieta, const double *izet,
                              const double *jcsi, const double *jeta, const double *jzet,
                              const double *kcsi, const double *keta, const double *kzet,
                              const double *iaj, const double *jaj, const double *kaj, const double *csi,
                              const double *eta, const double *zet, const double *dp, double *ucont,
                              double *ucat, double ren, double dt, double st);

/* The implicit momentum step (Implicit_MatrixFree, implicitsolver.c:281-470) for
 * a body at rest.  Its residual (FormFunction_SNES, momentum.c:2285-2440) is
 *   Uz = U with the wall components zeroed (bctype 1 and -1 on every face, 2 and
 *        12 on the j faces, 10 on the i and j faces away from the boundary rows),
 *   Ucat = Contra2Cart(Uz), IB and solid cells keeping the ucat passed in (the
 *        caller's IB velocity interpolation, done before the solve as the
 *        reference does),
 *   Ul = Uz after IB_BC (momentum.c:2037-2220): with immersed == 3 the faces
 *        between cells whose nvert sum exceeds 1.1 zeroed, and the slip faces of
 *        every row, ghost rows included,
 *   F  = -(Uz - Uo)/dt + 0.5 Formfunction_2(Ul, Ucat) + 0.5 RHS_o - dP
 *        for coeff in (0.9, 1.1), otherwise
 *        -(1.5 Uz - 2 Uo + 0.5 Urm1)/dt + Formfunction_2(...) - dP,
 * summed in the reference's order (bit-exact with a CPU restatement), with the
 * handle's immersed switch in Formfunction_2.  ucont_rm1 is needed only for coeff
 * outside (0.9, 1.1), rhs_o only for coeff in it (NULL: zero).  VFS_PP_EINVAL for
 * a bctype outside {0, 1, -1, 2, 10, -10, 12} or a slip face on a periodic axis.
 * All arrays are device DMDA vector fields on the handle's grid, stream-ordered
 * on the handle's stream; the call returns when done. */
int vfs_pp_momentum_residual_device(vfs_pp_handle *h, const double *icsi, const double *ieta,
                                    const double *izet, const double *jcsi, const double *jeta,
                                    const double *jzet, const double *kcsi, const double *keta,
                                    const double *kzet, const double *iaj, const double *jaj,
                                    const double *kaj, const double *csi, const double *eta,
                                    const double *zet, const double *dp, const double *ucont_o,
                                    const double *ucont_rm1, const double *rhs_o, const double *ucont,
                                    double ren, double dt, double coeff, double *ucat, double *f);

typedef struct {
  int newton_iterations;   /* Newton steps taken */
  int krylov_iterations;   /* GMRES iterations over all steps */
  int residual_evals;      /* evaluations of F (one per Jacobian-vector product) */
  int converged;           /* 1 if ||F|| <= rtol ||F0|| or the step fell below 1e-8 ||U|| */
  double fnorm0;           /* ||F|| at the initial guess */
  double fnorm;            /* ||F|| at the solution */
  double ms_total;         /* wall time of the call (ms) */
} vfs_pp_newton_stats;

/* Solve F(U) = 0 for U, starting from ucont (updated in place): Newton steps
 * with Jacobian-free products (J v = (F(U + h v) - F(U)) / h, h = sqrt(eps)
 * sqrt(1 + |U|) / |v|), unpreconditioned restarted GMRES(30) with an
 * Eisenstat-Walker forcing term (version 3, as SNESKSPSetParametersEW(snes, 3)),
 * at most max_newton steps, relative tolerance rtol on ||F|| (the reference's
 * imp_free_tol, 1e-4 by default, main.c:86).  The reference globalises with
 * PETSc's trust region; this backtracks along the Newton step, so iterates
 * differ while the solution meets the same test.  Afterwards the solution's wall
 * faces are zeroed as Implicit_MatrixFree does (implicitsolver.c:408-415) and
 * ucat receives its Contra2Cart.  VFS_PP_ENOCONV if the tolerance was not met
 * (ucont then holds the last iterate); stats may be NULL. */
int vfs_pp_implicit_momentum_device(vfs_pp_handle *h, const double *icsi, const double *ieta,
                                    const double *izet, const double *jcsi, const double *jeta,
                                    const double *jzet, const double *kcsi, const double *keta,
                                    const double *kzet, const double *iaj, const double *jaj,
                                    const double *kaj, const double *csi, const double *eta,
                                    const double *zet, const double *dp, const double *ucont_o,
                                    const double *ucont_rm1, const double *rhs_o, double *ucont,
                                    double *ucat, double ren, double dt, double coeff, double rtol,
                                    int max_newton, vfs_pp_newton_stats *stats);

/* PETSc binary vector files, the form of the vfield / pfield / nvfield / ufield
 * checkpoints (VecView / VecLoad on a binary viewer, main.c:291-335, 780-797,
 * 865-935): big-endian int32 class id 1211214, big-endian int32 length, then
 * big-endian float64 values; a one-rank DMDA vector is in the natural
 * [k][j][i][component] order of the DMDA arrays here.  _info reads the length;
 * _read fills a device array of exactly that length; _write stores a device
 * array (and, for block_size > 1, the "-vecload_block_size" companion file
 * PETSc writes next to it). */
int vfs_pp_petsc_vec_info(const char *path, long long *n);
int vfs_pp_read_petsc_vec_device(const char *path, double *v, size_t n);
int vfs_pp_write_petsc_vec_device(const char *path, const double *v, size_t n, int block_size);

/* Host-memory repack between a ghosted local DMDA array and the dense block the
 * entry points take.  The ghosted array covers, per axis a (0 = i, fastest),
 * indices ghost_start[a] .. ghost_start[a] + ghost_dims[a] - 1 (DMDAGetGhostCorners:
 * ghost_start = -3 and ghost_dims = m + 6 on an axis with DMDA_BOUNDARY_PERIODIC
 * and stencil width 3, init.c:147-158; 0 and m otherwise), ncomp doubles per
 * point (3 for Cmpnts).  to_dense = 1 copies its [0, dims) window into the dense
 * dst; 0 copies the dense src into that window of dst, leaving its ghost layers.
 * VFS_PP_EINVAL unless the window lies inside the ghosted range.  No device work. */
int vfs_pp_repack(const double *src, double *dst, const int ghost_start[3], const int ghost_dims[3],
                  const int dims[3], int ncomp, int to_dense);

/* Immersed-body surfaces (the ibmdataNN files of ibm_io.c:78-300): vertex and
 * triangle counts from the header. */
int vfs_pp_ucd_info(const char *path, int *n_v, int *n_elmt);

/* Reads an ibmdata surface into host arrays of n_v vertices (x = value / cl + cx,
 * ibm_io.c:146-148; the reference's cl is 1/L_dim or -char_length_ibm, and cx..cz
 * are -x_c/-y_c/-z_c) and n_elmt triangles (0-based vertex indices). */
int vfs_pp_read_ucd(const char *path, double cl, double cx, double cy, double cz, double *x,
                    double *y, double *z, int *nv1, int *nv2, int *nv3);

/* ibm_search_advanced (ibm.c:81-570) for body `body` on the device: the ray-parity
 * search of every interior cell centre against the triangulated surface (host
 * arrays), then the dilation, remap and wall layers of
 * vfs_pp_classify_nvert_device.  coor is the device DMDA coordinate array and
 * nvert the device DMDA nvert array, updated in place; body 0 first resets the
 * interior values below 5 (ibm.c:247), later bodies accumulate.  The
 * interpolation list the reference also builds (ibm.c:392-525) serves the
 * momentum step and is not produced.  Returns EINVAL if a ray meets more than
 * 32 triangles or 64 directions all hit coincident points. */
int vfs_pp_ib_search_device(const vfs_grid *grid, const double *coor, int n_v, const double *x,
                            const double *y, const double *z, int n_elmt, const int *nv1,
                            const int *nv2, const int *nv3, int body, double *nvert);

/* Grid files as init.c:251-389 reads them (the reference's gridfile, -binary
 * and -xyz options): */
#define VFS_PP_GRID_TEXT 0   /* grid.dat text: block count, per block "IM JM KM" then
                                all x, all y, all z values, i fastest (fscanf) */
#define VFS_PP_GRID_BINARY 1 /* the same sequence as native int32 / float64 words */
#define VFS_PP_GRID_XYZ 2    /* xyz.dat: "IM JM KM", then IM, JM, KM lines whose 1st,
                                2nd, 3rd columns are X[i], Y[j], Z[k] (one block) */

/* Block count and node counts IM, JM, KM of block `block` (the grid then has
 * nx = IM-1, ny = JM-1, nz = KM-1 cells); reads the headers only. */
int vfs_pp_grid_file_info(const char *path, int format, int block, int *nblocks, int *im,
                          int *jm, int *km);

/* Reads block `block` into the device DMDA coordinate array
 * coor[KM+1][JM+1][IM+1][3] (grid->nx == IM-1 etc.), each value as
 * value / cl * l_dim (init.c:339-366; cl is -chact_leng, l_dim is L_dim).  The
 * plane beyond the last node on each axis, which the reference leaves unread,
 * is filled by linear extrapolation.  Returns after the data is on the device. */
int vfs_pp_read_grid_device(const char *path, int format, int block, double cl, double l_dim,
                            const vfs_grid *grid, double *coor);

/* FormMetrics (metrics.c:4-1290, the NEWMETRIC branch that metrics.c:2 selects)
 * on the device: from the DMDA node coordinates coor[mz][my][mx][3] (device
 * array; the extra node plane at index m-1 of each axis is the one the reference
 * extrapolates) writes the 16 metric arrays vfs_pp_set_geometry takes, all device
 * arrays in the DMDA layout, bit-exact with the reference's arithmetic.  Stream
 * ordered on the legacy default stream; returns after the kernel finishes. */
int vfs_pp_form_metrics_device(const vfs_grid *grid, const double *coor, double *icsi,
                               double *ieta, double *izet, double *jcsi, double *jeta,
                               double *jzet, double *kcsi, double *keta, double *kzet,
                               double *iaj, double *jaj, double *kaj, double *aj, double *csi,
                               double *eta, double *zet);

/* Grid set-up from device node coordinates and a device nvert array: FormMetrics
 * on the device followed by everything vfs_pp_set_geometry does, with no host
 * copy of the face metrics (init.c:900 FormMetrics, then the solver set-up). */
int vfs_pp_set_grid_device(vfs_pp_handle *h, const double *coor, const double *nvert);

/* A new IB mask on the same grid (device DMDA nvert): the reference's per-step
 * rebuild for moving or rotating bodies, sediment and level sets (setup_lidx2 +
 * PoissonLHSNew + a new solver, poisson_hypre.c:627-648).  Code words, VolumeFlux
 * face list and multigrid hierarchy from the new nvert, on the device; the grid's
 * metrics stay from the last vfs_pp_set_geometry or vfs_pp_set_grid_device
 * (VFS_PP_ESTATE before either). */
int vfs_pp_set_nvert_device(vfs_pp_handle *h, const double *nvert);

/* PoissonSolver_Hypre (poisson_hypre.c:580-771) on device-resident arrays:
 * VolumeFlux (modifies ucont, poisson_hypre.c:662-667) -> PoissonRHS2_hypre ->
 * Krylov+MG solve -> Remove_Nullspace -> Convert_Phi2_Phi.  phi is the DMDA
 * pressure increment (lPhi, with periodic ghost images); on entry it is the
 * initial guess unless opts.zero_guess. */
int vfs_pp_poisson_solve_device(vfs_pp_handle *h, double *ucont, double *phi, double dt,
                                double st, double coeff, vfs_pp_stats *stats);

/* UpdatePressure (poisson.c:2473-2632): P += Phi on fluid cells, periodic ghosts. */
int vfs_pp_update_pressure_device(vfs_pp_handle *h, double *p, double *phi);

/* Projection (poisson.c:2122-2470) including its Contra2Cart call (rhs.c:50-675,
 * no-slip and periodic boundaries): corrects ucont in place and writes ucat
 * (cells the reference does not touch keep their input values). */
int vfs_pp_projection_device(vfs_pp_handle *h, double *ucont, const double *phi, double *ucat,
                             double dt, double st, double coeff);

/* Divergence (main.c:1244-1391): max |aj div u| over cells whose 7-point
 * neighbourhood is fluid. */
int vfs_pp_divergence_device(vfs_pp_handle *h, const double *ucont, double *maxdiv);

/* Divergence without a handle and without any device allocation.  No
 * counterpart in the reference: it serves callers that cannot allocate device
 * memory (work recorded into a HIP graph, tight serving loops, harnesses that
 * forbid hipMalloc).  Same result as vfs_pp_divergence_device, bit for bit, from
 * the reference's own DMDA arrays: ucont (AoS), aj and nvert (main.c:1244-1391),
 * all device pointers; *maxdiv is one device double.  The block maxima go to a
 * fixed 32 KiB device array of the library (4096 partials, one copy per GPU,
 * held until the process ends), reduced with a grid-stride loop, so any grid up
 * to the 2^29-point DMDA limit fits (larger ones: VFS_PP_EINVAL).  Runs on the
 * legacy default stream and returns when *maxdiv is written.  Calls that overlap
 * in time on one GPU share that array and race; sequential calls do not. */
int vfs_pp_divergence_noalloc(const vfs_grid *grid, const double *ucont, const double *aj,
                              const double *nvert, double *maxdiv);

/* The whole per-step pressure path, as solvers.c:874-895 runs it:
 * poisson_solve -> update_pressure -> projection, plus the divergence check. */
int vfs_pp_project_device(vfs_pp_handle *h, double *ucont, double *phi, double *p, double *ucat,
                          double dt, double st, double coeff, vfs_pp_stats *stats);

/* Same as vfs_pp_project_device with host arrays (copies over PCIe).  The first
 * call allocates device copies of the four arrays (8 doubles per DMDA point) that
 * the handle keeps for later calls until vfs_pp_destroy. */
int vfs_pp_project(vfs_pp_handle *h, double *ucont, double *phi, double *p, double *ucat,
                   double dt, double st, double coeff, vfs_pp_stats *stats);

/* ---- operator-level entry points (reference: the assembled HYPRE matrix and
 *      HYPRE's ParCSR SpMV inside HYPRE_ParCSRGMRESSolve, poisson_hypre.c:722-726) */

/* y = A phi on DMDA arrays (phi with periodic ghost images; y written on fluid
 * cells, zero elsewhere).  drop_small selects the poisson.c:1608 drop rule. */
int vfs_pp_apply_device(vfs_pp_handle *h, const double *phi, double *y, int drop_small);

/* Row coefficients of PoissonLHSNew for every DMDA point, as the GPU rebuilds them
 * (19 per point, slot order poisson.c:956-974, before the drop rule), and the
 * per-face branch codes (18 bytes per point: faces i+,i-,j+,j-,k+,k- x {active,
 * cross1, cross2}).  Host outputs; for verification. */
int vfs_pp_export_stencil(vfs_pp_handle *h, double *vol19, uint8_t *cases);

/* Level `level` of the multigrid hierarchy (0 = the finest), for verification; no
 * counterpart in the reference, whose preconditioner is BoomerAMG.  *nlevels (may
 * be NULL) receives the number of levels, dims (may be NULL) the level's nx, ny,
 * nz, *line (may be NULL) 1 when the level is smoothed by y-lines.  When vol19 and
 * cases are both non-NULL they receive the level's rows as vfs_pp_export_stencil
 * gives them, on the level's DMDA box (nx+2)(ny+2)(nz+2); the FP32 hierarchy holds
 * FP16 coefficients times a power of two, so its rows carry that rounding.
 * VFS_PP_EINVAL if level is out of range (after filling *nlevels). */
int vfs_pp_export_level(vfs_pp_handle *h, int level, int *nlevels, int dims[3], int *line, double *vol19,
                        uint8_t *cases);

/* The dense inverse of the coarsest multigrid level, for verification; no
 * counterpart in the reference (BoomerAMG solves its coarsest level itself).  *n
 * receives that level's cell count; when cinv is non-NULL and cap >= n * n, cinv
 * receives the top-left n x n block of the inverse of [A E; E^T 0] (A: the level's
 * rows, identity rows where a cell has none; E: one indicator column per connected
 * fluid component), column-major (cinv[c * n + r]), in the precision the hierarchy
 * holds it (FP32 values widened to double in the mixed form). */
int vfs_pp_export_coarse_inverse(vfs_pp_handle *h, size_t cap, double *cinv, int *n);

/* The VolumeFlux face list as the engine holds it, for verification: the faces
 * between a fluid cell and an IB cell that VolumeFlux corrects (poisson.c:2936-3010
 * for immersed == 1, 2670-2700 otherwise), in (k, j, i, axis) order.  *count
 * receives their number; when pos, area and sign are non-NULL and cap >= *count,
 * pos[f] = 3 * DMDA point + axis (the face's index in the AoS Ucont array),
 * area[f] = |csi|, |eta| or |zet| of the cell, sign[f] = +1 from fluid to IB, -1
 * from IB to fluid. */
int vfs_pp_export_ib_faces(vfs_pp_handle *h, size_t cap, long long *pos, double *area, float *sign,
                           size_t *count);

/* IB classification (ibm.c:331-353, 540-568) on a device DMDA array holding 4
 * for inside cells and 0 elsewhere; rewritten in place: interior cells to
 * {0,1,3}; boundary-plane points next to an inside cell keep the dilation's 2,
 * as in the reference (its 2->1 remap covers the interior only). */
int vfs_pp_classify_nvert_device(const vfs_grid *grid, double *nvert);

/* Multigrid transfer primitives of the reference on DMDA arrays, fully coarsened
 * grids (IM_c=(IM_f+1)/2, init.c:95-127).  MyRestriction poisson.c:364-471,
 * MyInterpolation poisson.c:526-669, MyNvertRestriction poisson.c:2004-2106.
 * Device pointers; fine grid described by `fine`. */
int vfs_pp_mg_restrict_device(const vfs_grid *fine, const double *nvert_f, const double *nvert_c,
                              const double *xf, double *fc);
int vfs_pp_mg_prolong_device(const vfs_grid *fine, const double *nvert_f, const double *nvert_c,
                             const double *xc, double *ff);
int vfs_pp_mg_nvert_restrict_device(const vfs_grid *fine, const double *nvert_f, double *nvert_c);

/* ---- batched form: many independent cases solved together.
 *
 * No counterpart in the reference (it solves one system per process, state in
 * process globals, poisson_hypre.c:17-21).  Every kernel launch of a batched
 * solve covers all member cases through one work list (persistent workers,
 * XCD-contiguous tile order), and one FGMRES iteration of all members is one
 * HIP graph launch.  Members keep their own geometry, hierarchy and Krylov
 * storage (and their own point or y-line smoother per level); they must share
 * opts.precision and opts.cheb_degree.  While the batch exists its members
 * belong to it: the single-case entry points (set-up, solves, exports,
 * vfs_pp_set_opts, vfs_pp_destroy) return VFS_PP_ESTATE for them, and a handle
 * can be a member of one batch at a time (vfs_pp_batch_create returns
 * VFS_PP_ESTATE otherwise). */
typedef struct vfs_pp_batch vfs_pp_batch;
int vfs_pp_batch_create(vfs_pp_handle *const *members, int count, vfs_pp_batch **out);
int vfs_pp_batch_destroy(vfs_pp_batch *b);
/* 1 (default): replay each iteration as a captured HIP graph; 0: plain launches */
int vfs_pp_batch_set_graphs(vfs_pp_batch *b, int on);
/* Tile shape of the batched launches that walk cells (smoothing sweeps, y-line
 * residuals, restriction to coarse cells, prolongation): -1 (default) lets the
 * autotuner pick, per launch kind and size, among 0 (256 consecutive cells, i
 * fastest), 1 (64 x 4 cells in i x j), 2 (32 x 8) and 3 (16 x 16), together with
 * the worker count; 0..3 force one shape (the worker counts are still tuned).
 * Results do not depend on the shape.  Takes effect at the next project call. */
int vfs_pp_batch_set_tile_shape(vfs_pp_batch *b, int shape);
/* Number of cell-tiled launches of the current plan that walk each of the four
 * shapes (after at least one project call). */
int vfs_pp_batch_tile_shapes(vfs_pp_batch *b, int counts[4]);
/* vfs_pp_project_device for every member at once; per-member device arrays and
 * scalars; stats (may be NULL) receives one entry per member.  Returns
 * VFS_PP_ENOCONV if any member missed rtol or div_tol (see its stats). */
int vfs_pp_batch_project_device(vfs_pp_batch *b, double *const *ucont, double *const *phi,
                                double *const *p, double *const *ucat, const double *dt,
                                const double *st, const double *coeff, vfs_pp_stats *stats);

/* ---- served form: continuous batching of projection requests.
 *
 * No counterpart in the reference.  Cases (grids with their metrics, masks and
 * multigrid hierarchy) are registered once and stay resident.  Requests
 * (one projection of a registered case on caller device arrays) queue FIFO;
 * each vfs_pp_server_progress call admits queued requests into free slots
 * (at most max_active at once, and at most max_per_case of one case) when the
 * Krylov workspace fits the HBM budget, runs one batched FGMRES iteration of
 * everything in flight (as vfs_pp_batch), and reports the requests that
 * finished.  Requests of one case in flight together share the case's resident
 * geometry and hierarchy coefficients; each has its own solve vectors and
 * multigrid work arrays (a per-case solver kept for reuse once made, about
 * 80 bytes per cell in the mixed form besides the Krylov workspace).  Results
 * are the ones vfs_pp_project_device gives, to rounding (the batched kernels
 * order some sums differently), and the same bit for bit however requests are
 * interleaved or batched together.
 *
 * Back-pressure: with max_queue > 0, vfs_pp_server_submit refuses a request
 * (VFS_PP_EBUSY, no ticket) while max_queue requests wait; with deadline_ms >
 * 0, a request still waiting for admission deadline_ms after its submission is
 * dropped at the next vfs_pp_server_progress, untouched (its arrays are never
 * read or written), and its ticket is reported by vfs_pp_server_expired. */
typedef struct vfs_pp_server vfs_pp_server;
typedef struct {
  size_t hbm_budget;   /* device bytes for resident cases + workspaces; 0: 80% of free memory */
  int max_active;      /* requests in flight at once */
  vfs_pp_opts solver;  /* options of every case (precision and cheb_degree shared) */
  int max_per_case;    /* requests of one case in flight at once (default 4, >= 1) */
  int max_queue;       /* waiting requests at most; 0: unbounded (default) */
  double deadline_ms;  /* admission deadline after submission; 0: none (default) */
} vfs_pp_server_opts;
void vfs_pp_server_default_opts(vfs_pp_server_opts *opts);
int vfs_pp_server_create(const vfs_pp_server_opts *opts, vfs_pp_server **out);
int vfs_pp_server_destroy(vfs_pp_server *sv);
/* geometry arrays as vfs_pp_set_geometry; VFS_PP_ENOMEM if the case and one
 * workspace would not fit the budget */
int vfs_pp_server_add_case(vfs_pp_server *sv, const vfs_grid *grid, const double *icsi,
                           const double *ieta, const double *izet, const double *jcsi,
                           const double *jeta, const double *jzet, const double *kcsi,
                           const double *keta, const double *kzet, const double *iaj,
                           const double *jaj, const double *kaj, const double *aj,
                           const double *csi, const double *eta, const double *zet,
                           const double *nvert, int *case_id);
/* VFS_PP_ESTATE while a request of the case is queued or in flight */
int vfs_pp_server_remove_case(vfs_pp_server *sv, int case_id);
/* queue one projection; the device arrays must stay valid until it finishes or
 * expires.  VFS_PP_EBUSY (no ticket) when max_queue requests already wait. */
int vfs_pp_server_submit(vfs_pp_server *sv, int case_id, double *ucont, double *phi, double *p,
                         double *ucat, double dt, double st, double coeff, int *ticket);
/* drop expired requests, admit, iterate once, retire: *ndone finished tickets
 * (and their stats) */
int vfs_pp_server_progress(vfs_pp_server *sv, int *tickets, vfs_pp_stats *stats, int cap,
                           int *ndone);
/* tickets dropped by the admission deadline since the last call: up to cap of
 * them into tickets (oldest first), their number in *n; the rest stay listed */
int vfs_pp_server_expired(vfs_pp_server *sv, int *tickets, int cap, int *n);
int vfs_pp_server_status(vfs_pp_server *sv, int *queued, int *active, size_t *resident_bytes,
                         size_t *workspace_bytes);

/* Time `reps` launches of one engine kernel on the handle's stream with HIP
 * events; kernel ids: 0 FP64 operator apply, 1 finest-level smoother sweep (in
 * the multigrid precision), 2 projection, 3 Contra2Cart, 4 RHS, 5 a plain
 * 1 GiB device copy (the bandwidth reference), 6 the finest smoother sweep over
 * six rotating copies of its operands (together beyond the 256 MB last-level
 * cache, so each sweep streams from HBM), 7 the Krylov operator pass
 * (w = A z with the dots of w against 4 basis vectors and itself), 8 the
 * streaming floor of a point sweep: every array the sweep reads streamed once
 * per cell over the same rotating copies, with no stencil (VFS_PP_EINVAL when
 * the finest level is smoothed by y-lines).  Writes the mean duration (ms) and
 * the algorithmic bytes of one launch: every distinct array once per cell, with
 * a 1-byte mask (apply 89, smoother 31 FP32 (FP16 coefficients) / 97 FP64, or
 * for a y-line finest level its fused residual + line step 51 / 137 (lines of
 * up to 256 rows) or separate residual + line solve 55 / 145, projection 129,
 * Contra2Cart 121, RHS 33 bytes per cell, Krylov pass z + 72 + 1 + 8 + 4 x 8 =
 * 117 (FP32 z) / 121, streaming floor 35 / 105; copy 16 bytes per element). */
int vfs_pp_time_kernel(vfs_pp_handle *h, int kernel, int reps, double *ms_per_launch,
                       double *bytes_per_launch);

/* Description of the multigrid hierarchy (levels, sizes, smoother) as text. */
int vfs_pp_describe(vfs_pp_handle *h, char *buf, size_t len);

const char *vfs_pp_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif

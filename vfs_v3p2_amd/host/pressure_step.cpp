// One pressure step of one block through the C++ host side (vfs_userctx.hpp), the
// call sequence of Flow_Solver's pressure block (solvers.c:874-895):
//
//   pressure_step <dir> <nx> <ny> <nz> <per_i> <per_j> <per_k> <b0..b5> <immersed>
//                 <dt> <st> <coeff> <precision> <rtol>
//
// reads <dir>/coor.bin ([mz][my][mx][3] node coordinates), <dir>/nvert.bin and
// <dir>/ucont.bin (float64, DMDA layout, mx = nx + 2 ...), runs FormMetrics, then
// PoissonSolver_Hypre, UpdatePressure, Projection and Divergence, and writes
// <dir>/{ucont_out,ucat,phi,p}.bin and one line of statistics to stdout.  Exit
// status: the library's status of the step (0 when it converged).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "vfs_userctx.hpp"

static bool read_all(const std::string &path, std::vector<double> &v) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) return false;
  const size_t got = std::fread(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
  return got == v.size();
}

static bool write_all(const std::string &path, const std::vector<double> &v) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  const size_t put = std::fwrite(v.data(), sizeof(double), v.size(), f);
  std::fclose(f);
  return put == v.size();
}

int main(int argc, char **argv) {
  if (argc != 20) {
    std::fprintf(stderr,
                 "usage: %s dir nx ny nz per_i per_j per_k b0 b1 b2 b3 b4 b5 immersed dt st coeff precision rtol\n",
                 argv[0]);
    return 2;
  }
  const std::string dir = argv[1];
  vfs_grid g{};
  g.nx = std::atoi(argv[2]);
  g.ny = std::atoi(argv[3]);
  g.nz = std::atoi(argv[4]);
  for (int a = 0; a < 3; a++) g.periodic[a] = std::atoi(argv[5 + a]);
  for (int f = 0; f < 6; f++) g.bctype[f] = std::atoi(argv[8 + f]);
  g.immersed = std::atoi(argv[14]);
  const double dt = std::atof(argv[15]), st = std::atof(argv[16]), coeff = std::atof(argv[17]);
  vfs_pp_opts o;
  vfs_pp_default_opts(&o);
  o.precision = std::atoi(argv[18]);
  o.rtol = std::atof(argv[19]);

  vfs::UserCtx user;
  int rc = vfs::UserCtxCreate(g, dt, st, &o, &user);
  if (rc != VFS_PP_OK) {
    std::fprintf(stderr, "UserCtxCreate: %s\n", vfs_pp_error_string(rc));
    return rc;
  }
  const size_t n = user.points();
  std::vector<double> coor(3 * n), nvert(n), ucont(3 * n);
  if (!read_all(dir + "/coor.bin", coor) || !read_all(dir + "/nvert.bin", nvert) ||
      !read_all(dir + "/ucont.bin", ucont)) {
    std::fprintf(stderr, "cannot read the inputs in %s\n", dir.c_str());
    vfs::UserCtxDestroy(&user);
    return VFS_PP_EIO;
  }
  double maxdiv = -1.;
  rc = vfs::UserCtxUpload(&user, coor.data(), nvert.data(), ucont.data());
  if (rc == VFS_PP_OK) rc = vfs::FormMetrics(&user);
  if (rc == VFS_PP_OK) rc = vfs::Flow_Solver_pressure_step(&user, coeff, &maxdiv);
  if (rc != VFS_PP_OK && rc != VFS_PP_ENOCONV) {
    std::fprintf(stderr, "pressure step: %s\n", vfs_pp_error_string(rc));
    vfs::UserCtxDestroy(&user);
    return rc;
  }
  std::vector<double> ucat(3 * n), phi(n), p(n);
  int rc2 = vfs::UserCtxDownload(&user, ucont.data(), ucat.data(), phi.data(), p.data());
  if (rc2 == VFS_PP_OK &&
      !(write_all(dir + "/ucont_out.bin", ucont) && write_all(dir + "/ucat.bin", ucat) &&
        write_all(dir + "/phi.bin", phi) && write_all(dir + "/p.bin", p)))
    rc2 = VFS_PP_EIO;
  std::printf("iterations %d converged %d rel_residual %.6e max_div %.6e\n", user.stats.iterations,
              user.stats.converged, user.stats.rel_residual, maxdiv);
  vfs::UserCtxDestroy(&user);
  return rc2 != VFS_PP_OK ? rc2 : rc;
}

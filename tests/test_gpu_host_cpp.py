"""One pressure step through the C++ host side (vfs_v3p2_amd/host/pressure_step:
vfs::FormMetrics, PoissonSolver_Hypre, UpdatePressure, Projection, Divergence on a
vfs::UserCtx) equals, bit for bit, the same calls made through the Python mirror
on the same inputs, and meets the divergence bar."""
import os
import subprocess

import numpy as np
import pytest

from vfs_v3p2_amd import cases

pytestmark = pytest.mark.gpu

HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "vfs_v3p2_amd", "host")


@pytest.mark.parametrize("precision", [1, 0])
def test_cpp_pressure_step_equals_the_python_path(torch_cuda, tmp_path, precision):
    torch = torch_cuda
    from vfs_v3p2_amd import engine as E

    exe = os.path.join(HOST, "pressure_step")
    assert os.path.exists(exe), "host side not built"
    c = cases.hill_case(24, 16, 16, h=0.3, sphere=(0.6, 0.75, 0.0, 0.3), seed=4, keep_coords=True)
    coor = np.ascontiguousarray(c.coor, dtype=np.float64)
    nvert = np.ascontiguousarray(c.fields["nvert"], dtype=np.float64)
    ucont = np.ascontiguousarray(c.fields["ucont"], dtype=np.float64)
    for name, a in (("coor", coor), ("nvert", nvert), ("ucont", ucont)):
        a.tofile(str(tmp_path / f"{name}.bin"))
    rtol = 1e-12
    args = [exe, str(tmp_path), str(c.nx), str(c.ny), str(c.nz), *map(str, c.per), *map(str, c.bctype),
            str(c.immersed), repr(c.dt), repr(c.st), repr(c.coeff), str(precision), repr(rtol)]
    print("running", " ".join(args[1:]), flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    stats = dict(zip(r.stdout.split()[0::2], r.stdout.split()[1::2]))
    shp = c.shape
    got = {k: np.fromfile(str(tmp_path / f"{k}.bin")).reshape(shp + ((3,) if k in ("ucont_out", "ucat") else ()))
           for k in ("ucont_out", "ucat", "phi", "p")}

    # the same calls through the Python mirror of the C ABI
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed),
                 E.default_opts(precision=precision, rtol=rtol))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    h.set_grid_device(dev(coor), dev(nvert))
    u = dev(ucont)
    phi = torch.zeros(shp, dtype=torch.float64, device="cuda")
    p = torch.zeros_like(phi)
    ucat = torch.zeros(shp + (3,), dtype=torch.float64, device="cuda")
    st = h.poisson_solve(u, phi, c.dt, c.st, c.coeff, strict=True)
    h.update_pressure(p, phi)
    h.projection(u, phi, ucat, c.dt, c.st, c.coeff)
    md = h.divergence(u)
    h.close()
    assert int(stats["iterations"]) == st.iterations and int(stats["converged"]) == 1
    assert float(stats["max_div"]) == float(f"{md:.6e}") and md <= 1e-8
    np.testing.assert_array_equal(got["phi"], phi.cpu().numpy())
    np.testing.assert_array_equal(got["p"], p.cpu().numpy())
    np.testing.assert_array_equal(got["ucont_out"], u.cpu().numpy())
    np.testing.assert_array_equal(got["ucat"], ucat.cpu().numpy())

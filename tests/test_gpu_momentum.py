"""The momentum predictor on the device against the oracle (whose closed-form pins
are in tests/test_momentum.py): Formfunction_2 bit for bit on periodic, walled,
IB and non-periodic sheared grids, the four-stage RungeKutta step bit for bit,
and one whole time step kept on the device -- predictor, pressure projection,
Pressure_Gradient -- against the same chain on the CPU (the projection's solve
to the phi tolerance)."""
import dataclasses

import numpy as np
import pytest

from oracle import oracle as O
from vfs_v3p2_amd import cases

pytestmark = pytest.mark.gpu

METRICS = ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "iaj", "jaj", "kaj", "csi", "eta",
           "zet")


def _affine():
    M = np.array([[1.0, 0.35, -0.2], [0.15, 0.8, 0.1], [-0.25, 0.3, 1.2]])
    nx, ny, nz = 12, 10, 9
    P = np.stack(np.meshgrid(np.arange(nx + 1) / nx, np.arange(ny + 1) / ny, np.arange(nz + 1) / nz,
                             indexing="ij"), -1).transpose(2, 1, 0, 3)
    f = cases.form_metrics(cases.dmda_coords(P @ M.T), (0, 0, 0))
    f["nvert"] = np.zeros(f["aj"].shape)
    f["nvert"][4:6, 4:7, 5:8] = 1.0
    f["nvert"][5, 5, 6] = 3.0
    rng = np.random.default_rng(3)
    f["ucont"] = 0.01 * rng.standard_normal(f["aj"].shape + (3,))
    return cases.Case(name="affine_ib", nx=nx, ny=ny, nz=nz, per=(0, 0, 0), bctype=(1, 1, 1, 1, 1, 1), immersed=1,
                      fields=f)


CASES = {
    "hill_sphere": lambda: cases.hill_case(24, 16, 16, h=0.3, sphere=(0.6, 0.75, 0.0, 0.3), seed=0),
    "channel": lambda: cases.channel_case(16, 12, 14, beta=2.0, seed=1),
    "towers": lambda: cases.towers_case(40, 20, 24, ratio=1.05, rows=2, cols=2, radius_cells=3, seed=2),
    "affine_ib": _affine,
}


@pytest.fixture(scope="module")
def E(torch_cuda):
    from vfs_v3p2_amd import engine

    engine.lib()
    return engine


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def ucat_of(c, u):
    g = O.grid(c.shape, c.per)
    uc = np.zeros(c.shape + (3,))
    O.contra2cart(g, c.bctype, c.fields, np.ascontiguousarray(u), uc)
    return uc


@pytest.mark.parametrize("immersed", [None, 3])
@pytest.mark.parametrize("name", list(CASES))
def test_formfunction2_bit_exact(E, torch_cuda, name, immersed):
    torch = torch_cuda
    c = CASES[name]()
    imm = c.immersed if immersed is None else immersed
    print(f"Formfunction_2 {name} {c.nx}x{c.ny}x{c.nz} immersed {imm}", flush=True)
    g = O.grid(c.shape, c.per)
    u = np.ascontiguousarray(c.fields["ucont"])
    uc = ucat_of(c, u)
    uc[c.fields["nvert"] > 0.1] *= 0.5  # IB-adjacent cells carry the caller's interpolated values
    for ren, scale in ((100.0, 1.0), (2500.0, 0.5)):
        want = O.formfunction2(g, c.bctype, imm, c.fields, u, uc, ren, scale)
        grid = E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, imm)
        met = {k: dev(torch, c.fields[k]) for k in METRICS}
        ud = dev(torch, u)
        got = E.Formfunction_2(grid, met, dev(torch, c.fields["nvert"]), ud, dev(torch, uc), ren, scale)
        np.testing.assert_array_equal(got.cpu().numpy(), want)
        np.testing.assert_array_equal(ud.cpu().numpy(), u)  # Ucont is not modified
        assert np.abs(want).max() > 0


def _pressure(c):
    rng = np.random.default_rng(5)
    p = np.where(c.fields["nvert"] < 0.1, rng.standard_normal(c.shape), 0.0)
    cases.periodic_fill(p, c.per)
    return p


@pytest.mark.parametrize("name", list(CASES))
def test_runge_kutta_bit_exact(E, torch_cuda, name):
    torch = torch_cuda
    c = CASES[name]()
    print(f"RungeKutta {name}", flush=True)
    g = O.grid(c.shape, c.per)
    u = np.ascontiguousarray(c.fields["ucont"])
    uc = ucat_of(c, u)
    dp, _ = O.pressure_gradient(g, c.fields, _pressure(c))
    ren, dt, st = 200.0, 0.01, 1.0
    u_ref, uc_ref = O.runge_kutta(g, c.bctype, c.immersed, c.fields, dp, u, uc, ren, dt, st)
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed))
    h.set_geometry(c.fields)
    met = {k: dev(torch, c.fields[k]) for k in METRICS}
    ud, ucd = dev(torch, u), dev(torch, uc)
    E.RungeKutta(h, met, dev(torch, dp), ud, ucd, ren, dt, st)
    h.close()
    np.testing.assert_array_equal(ud.cpu().numpy(), u_ref)
    np.testing.assert_array_equal(ucd.cpu().numpy(), uc_ref)
    assert np.abs(u_ref - u).max() > 1e-6


def test_one_time_step_resident(E, torch_cuda):
    """Predictor -> projection -> Pressure_Gradient with every field on the
    device, against the oracle chain: Ucont* bit for bit, phi to 1e-8, the
    projected field divergence-free, dP of the new pressure to 1e-8."""
    torch = torch_cuda
    c0 = CASES["hill_sphere"]()
    c = dataclasses.replace(c0, dt=0.01)
    g = O.grid(c.shape, c.per)
    u = np.ascontiguousarray(c.fields["ucont"])
    uc = ucat_of(c, u)
    p0 = _pressure(c)
    dp0, p0g = O.pressure_gradient(g, c.fields, p0)
    ren = 200.0
    # CPU chain
    u_star, uc_star = O.runge_kutta(g, c.bctype, c.immersed, c.fields, dp0, u, uc, ren, c.dt, c.st)
    ref = O.reference_projection(dataclasses.replace(c, fields=dict(c.fields, ucont=u_star)), rtol=1e-13)
    p1 = p0g + ref["p"]
    dp1, _ = O.pressure_gradient(g, c.fields, p1)
    # device chain
    grid = E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed)
    h = E.Handle(grid, E.default_opts(rtol=1e-12, div_tol=1e-8))
    h.set_geometry(c.fields)
    met = {k: dev(torch, c.fields[k]) for k in METRICS}
    nv = dev(torch, c.fields["nvert"])
    pd = dev(torch, p0)
    dpd = E.Pressure_Gradient(grid, met, nv, pd)
    ud, ucd = dev(torch, u), dev(torch, uc)
    E.RungeKutta(h, met, dpd, ud, ucd, ren, c.dt, c.st)
    np.testing.assert_array_equal(ud.cpu().numpy(), u_star)
    phi = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
    st = h.project(ud, phi, pd, ucd, c.dt, c.st, c.coeff, strict=True)
    dp1d = E.Pressure_Gradient(grid, met, nv, pd)
    h.close()
    A = ref["A"]
    fl = A.fluid_cells
    rel = lambda a, b: float(np.linalg.norm((a - b)[fl]) / np.linalg.norm(b[fl]))
    assert rel(phi.cpu().numpy(), ref["phi"]) <= 1e-8
    assert st.max_div <= 1e-8
    assert rel(pd.cpu().numpy(), p1) <= 1e-8
    d1 = dp1d.cpu().numpy()
    assert float(np.linalg.norm(d1 - dp1) / np.linalg.norm(dp1)) <= 1e-8

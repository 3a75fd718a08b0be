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


# ------------------------------------------------------------ implicit form

def _case_imm0(name):
    c = CASES[name]()
    return dataclasses.replace(c, immersed=0)


def _handle0(E, c):
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, 0))
    h.set_geometry(c.fields)
    return h


@pytest.mark.parametrize("coeff", [1.0, 1.5])
@pytest.mark.parametrize("name", ["channel", "affine_ib", "hill_sphere"])
def test_momentum_residual_bit_exact(E, torch_cuda, name, coeff):
    """FormFunction_SNES's residual on the device equals the oracle's bit for bit
    (time terms, wall zeroing, Contra2Cart, accumulated Formfunction_2, RHS_o, dP)."""
    torch = torch_cuda
    c = _case_imm0(name)
    print(f"momentum residual {name} coeff {coeff}", flush=True)
    g = O.grid(c.shape, c.per)
    rng = np.random.default_rng(7)
    u = np.ascontiguousarray(c.fields["ucont"])
    uo = u * (1.0 + 0.1 * rng.standard_normal(u.shape))
    urm1 = u * (1.0 + 0.1 * rng.standard_normal(u.shape))
    rhs_o = 0.01 * rng.standard_normal(u.shape)
    dp, _ = O.pressure_gradient(g, c.fields, _pressure(c))
    ren, dt = 300.0, 0.01
    F_ref, uz_ref, ucat_ref = O.momentum_residual(g, c.bctype, c.fields, dp, uo, u, ren, dt, coeff=coeff, urm1=urm1,
                                                  rhs_o=rhs_o)
    h = _handle0(E, c)
    met = {k: dev(torch, c.fields[k]) for k in METRICS}
    F, ucat = E.momentum_residual(h, met, dev(torch, dp), dev(torch, uo), dev(torch, u), ren, dt, coeff,
                                  ucont_rm1=dev(torch, urm1), rhs_o=dev(torch, rhs_o))
    h.close()
    np.testing.assert_array_equal(F.cpu().numpy(), F_ref)
    np.testing.assert_array_equal(ucat.cpu().numpy(), ucat_ref)
    assert np.abs(F_ref).max() > 0


def _periodic_affine(nx=10, ny=9, nz=11):
    M = np.array([[1.0, 0.35, -0.2], [0.15, 0.8, 0.1], [-0.25, 0.3, 1.2]])
    P = np.stack(np.meshgrid(np.arange(nx + 1) / nx, np.arange(ny + 1) / ny, np.arange(nz + 1) / nz,
                             indexing="ij"), -1).transpose(2, 1, 0, 3)
    f = cases.form_metrics(cases.dmda_coords(P @ M.T), (1, 1, 1))
    f["nvert"] = np.zeros(f["aj"].shape)
    f["ucont"] = np.zeros(f["aj"].shape + (3,))
    return cases.Case(name="affine_periodic", nx=nx, ny=ny, nz=nz, per=(1, 1, 1), bctype=(0,) * 6, immersed=0,
                      fields=f)


def _uniform_flux(f, U):
    u = np.zeros(f["aj"].shape + (3,))
    u[..., 0] = f["icsi"] @ U
    u[..., 1] = f["jeta"] @ U
    u[..., 2] = f["kzet"] @ U
    return u


def test_implicit_step_of_a_uniform_flow(E, torch_cuda):
    """A uniform flow on a periodic sheared grid has no Formfunction_2 part, so
    F(U) = -(U - Uo)/dt - dP is linear and its root is the uniform flux of
    Uo - dt dU: the Newton-Krylov solve lands on it."""
    torch = torch_cuda
    c = _periodic_affine()
    f = c.fields
    U0, Uo, dU = np.array([0.8, -0.4, 0.3]), np.array([0.7, -0.3, 0.2]), np.array([0.2, 0.1, -0.3])
    u0, uo, dp = _uniform_flux(f, U0), _uniform_flux(f, Uo), _uniform_flux(f, dU)
    # Pressure_Gradient leaves the boundary planes at zero, where Formfunction_2
    # zeroes the residual's other terms too
    for a in range(3):
        dp[(slice(None),) * a + (0,)] = 0.0
        dp[(slice(None),) * a + (-1,)] = 0.0
    dt = 0.05
    ustar = _uniform_flux(f, Uo - dt * dU)
    h = _handle0(E, c)
    met = {k: dev(torch, f[k]) for k in METRICS}
    ud, ucd = dev(torch, u0), torch.zeros(u0.shape, dtype=torch.float64, device="cuda")
    st = E.Implicit_MatrixFree(h, met, dev(torch, dp), dev(torch, uo), ud, ucd, 100.0, dt, 1.0,
                               rhs_o=dev(torch, np.zeros(u0.shape)), rtol=1e-10)
    h.close()
    print(f"uniform flow: {st.newton_iterations} Newton, {st.krylov_iterations} GMRES", flush=True)
    assert st.converged and st.newton_iterations >= 1 and st.fnorm <= 1e-10 * st.fnorm0
    got = ud.cpu().numpy()
    inner = (slice(1, -1),) * 3  # the residual does not see the ghost planes of U
    assert np.abs(got - ustar)[inner].max() <= 1e-8 * np.abs(u0 - ustar)[inner].max()


@pytest.mark.parametrize("coeff", [1.0, 1.5])
def test_implicit_step_meets_the_reference_test(E, torch_cuda, coeff):
    """A walled channel from a perturbed start: the device's Ucont satisfies the
    reference's SNES test ||F|| <= 1e-4 ||F(U0)|| with F evaluated by the oracle,
    its ||F(U0)|| is the oracle's, and its Ucat is the oracle's Contra2Cart of it."""
    torch = torch_cuda
    c = _case_imm0("channel")
    g = O.grid(c.shape, c.per)
    rng = np.random.default_rng(11)
    uo = np.ascontiguousarray(c.fields["ucont"])
    ucat_o = ucat_of(c, uo)
    ren, dt = 400.0, 0.02
    rhs_o = O.formfunction2(g, c.bctype, 0, c.fields, uo, ucat_o, ren)
    u0 = uo * (1.0 + 0.05 * rng.standard_normal(uo.shape))
    urm1 = uo * (1.0 + 0.05 * rng.standard_normal(uo.shape))
    dp, _ = O.pressure_gradient(g, c.fields, _pressure(c))
    F0, _, _ = O.momentum_residual(g, c.bctype, c.fields, dp, uo, u0, ren, dt, coeff=coeff, urm1=urm1, rhs_o=rhs_o)
    h = _handle0(E, c)
    met = {k: dev(torch, c.fields[k]) for k in METRICS}
    ud, ucd = dev(torch, u0), torch.zeros(u0.shape, dtype=torch.float64, device="cuda")
    st = E.Implicit_MatrixFree(h, met, dev(torch, dp), dev(torch, uo), ud, ucd, ren, dt, coeff,
                               ucont_rm1=dev(torch, urm1), rhs_o=dev(torch, rhs_o), rtol=1e-4)
    h.close()
    print(f"channel coeff {coeff}: {st.newton_iterations} Newton, {st.krylov_iterations} GMRES, "
          f"{st.residual_evals} F, |F| {st.fnorm0:.3e} -> {st.fnorm:.3e}", flush=True)
    u = ud.cpu().numpy()
    F1, _, ucat_ref = O.momentum_residual(g, c.bctype, c.fields, dp, uo, u, ren, dt, coeff=coeff, urm1=urm1,
                                          rhs_o=rhs_o)
    assert st.converged
    assert abs(st.fnorm0 - np.linalg.norm(F0)) <= 1e-12 * np.linalg.norm(F0)
    assert np.linalg.norm(F1) <= 1e-4 * np.linalg.norm(F0) * (1 + 1e-6)
    assert np.linalg.norm(u - u0) > 1e-3 * np.linalg.norm(u0)
    np.testing.assert_allclose(ucd.cpu().numpy(), ucat_ref, rtol=0, atol=1e-14 * np.abs(ucat_ref).max())


def _ib_inputs(c, seed):
    """uo, urm1, rhs_o, dp and a Ucat whose IB and solid cells hold arbitrary
    values (the caller's IB interpolation), for the residual at c's Ucont."""
    g = O.grid(c.shape, c.per)
    rng = np.random.default_rng(seed)
    u = np.ascontiguousarray(c.fields["ucont"])
    uo = u * (1.0 + 0.1 * rng.standard_normal(u.shape))
    urm1 = u * (1.0 + 0.1 * rng.standard_normal(u.shape))
    rhs_o = 0.01 * rng.standard_normal(u.shape)
    dp, _ = O.pressure_gradient(g, c.fields, _pressure(c))
    ucat = 0.2 * rng.standard_normal(u.shape)
    return g, u, uo, urm1, rhs_o, dp, ucat


@pytest.mark.parametrize("coeff", [1.0, 1.5])
@pytest.mark.parametrize("immersed", [1, 3])
@pytest.mark.parametrize("name", ["affine_ib", "hill_sphere"])
def test_momentum_residual_with_a_body_bit_exact(E, torch_cuda, name, immersed, coeff):
    """The residual with an immersed body: IB-aware Formfunction_2, IB_BC (zeroed
    faces between IB and solid cells for immersed 3), the IB cells' Ucat kept as
    passed and the solid cells' zeroed, bit for bit against the oracle; immersed 3
    differs from 1 (the zeroed faces reach the residual)."""
    torch = torch_cuda
    c = dataclasses.replace(CASES[name](), immersed=immersed)
    print(f"momentum residual {name} immersed {immersed} coeff {coeff}", flush=True)
    g, u, uo, urm1, rhs_o, dp, ucat0 = _ib_inputs(c, 17)
    ren, dt = 300.0, 0.01
    F_ref, _, ucat_ref = O.momentum_residual(g, c.bctype, c.fields, dp, uo, u, ren, dt, coeff=coeff, urm1=urm1,
                                             rhs_o=rhs_o, immersed=immersed, ucat=ucat0)
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, immersed))
    h.set_geometry(c.fields)
    met = {k: dev(torch, c.fields[k]) for k in METRICS}
    F, ucat = E.momentum_residual(h, met, dev(torch, dp), dev(torch, uo), dev(torch, u), ren, dt, coeff,
                                  ucont_rm1=dev(torch, urm1), rhs_o=dev(torch, rhs_o), ucat=dev(torch, ucat0))
    h.close()
    np.testing.assert_array_equal(F.cpu().numpy(), F_ref)
    np.testing.assert_array_equal(ucat.cpu().numpy(), ucat_ref)
    nv = c.fields["nvert"]
    ibm, solid = (nv > 0.1) & (nv < 2.9), nv > 2.9
    assert ibm.any()
    np.testing.assert_array_equal(ucat_ref[ibm], ucat0[ibm])  # IB cells keep their Ucat
    assert (ucat_ref[solid] == 0).all()  # solid cells are zeroed (rhs.c:257)
    if immersed == 3:
        F1, _, _ = O.momentum_residual(g, c.bctype, c.fields, dp, uo, u, ren, dt, coeff=coeff, urm1=urm1,
                                       rhs_o=rhs_o, immersed=1, ucat=ucat0)
        assert not np.array_equal(F1, F_ref)


def test_implicit_step_with_a_body_meets_the_reference_test(E, torch_cuda):
    """The implicit step on the hill with an IB sphere (immersed 1, the IB cells'
    Ucat as the caller's interpolation left it): the device's Ucont satisfies
    ||F|| <= 1e-4 ||F(U0)|| with F evaluated by the oracle."""
    torch = torch_cuda
    c = CASES["hill_sphere"]()
    g, uo, _, _, _, dp, _ = _ib_inputs(c, 23)
    rng = np.random.default_rng(29)
    ucat0 = ucat_of(c, uo)
    ren, dt = 400.0, 0.02
    rhs_o = O.formfunction2(g, c.bctype, c.immersed, c.fields, uo, ucat0, ren)
    u0 = uo * (1.0 + 0.05 * rng.standard_normal(uo.shape))
    F0, _, _ = O.momentum_residual(g, c.bctype, c.fields, dp, uo, u0, ren, dt, rhs_o=rhs_o, immersed=c.immersed,
                                   ucat=ucat0)
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed))
    h.set_geometry(c.fields)
    met = {k: dev(torch, c.fields[k]) for k in METRICS}
    ud, ucd = dev(torch, u0), dev(torch, ucat0)
    st = E.Implicit_MatrixFree(h, met, dev(torch, dp), dev(torch, uo), ud, ucd, ren, dt, 1.0,
                               rhs_o=dev(torch, rhs_o), rtol=1e-4)
    h.close()
    print(f"hill + sphere: {st.newton_iterations} Newton, {st.krylov_iterations} GMRES, "
          f"|F| {st.fnorm0:.3e} -> {st.fnorm:.3e}", flush=True)
    u = ud.cpu().numpy()
    F1, _, _ = O.momentum_residual(g, c.bctype, c.fields, dp, uo, u, ren, dt, rhs_o=rhs_o, immersed=c.immersed,
                                   ucat=ucat0)
    assert st.converged
    assert abs(st.fnorm0 - np.linalg.norm(F0)) <= 1e-12 * np.linalg.norm(F0)
    assert np.linalg.norm(F1) <= 1e-4 * np.linalg.norm(F0) * (1 + 1e-6)


def test_implicit_form_refusals(E, torch_cuda):
    """A bctype outside the wall rules, a slip face on a periodic axis, or coeff 1.5
    without the older field: VFS_PP_EINVAL."""
    torch = torch_cuda
    c = CASES["affine_ib"]()
    met = {k: dev(torch, c.fields[k]) for k in METRICS}
    u = dev(torch, c.fields["ucont"])
    z = torch.zeros_like(u)
    cp = _periodic_affine()
    metp = {k: dev(torch, cp.fields[k]) for k in METRICS}
    up = dev(torch, cp.fields["ucont"])
    h = E.Handle(E.make_grid(cp.nx, cp.ny, cp.nz, cp.per, (10, 10, 0, 0, 0, 0), 0))
    h.set_geometry(cp.fields)
    with pytest.raises(E.VfsPPError) as ei:
        E.momentum_residual(h, metp, torch.zeros_like(up), up, up, 100.0, 0.01)
    assert ei.value.code == E.VFS_PP_EINVAL
    h.close()
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, (4, 1, 1, 1, 1, 1), 0))
    h.set_geometry(c.fields)
    with pytest.raises(E.VfsPPError) as ei:
        E.Implicit_MatrixFree(h, met, z, u, u.clone(), z.clone(), 100.0, 0.01)
    assert ei.value.code == E.VFS_PP_EINVAL
    h.close()
    h = _handle0(E, dataclasses.replace(c, immersed=0))
    with pytest.raises(E.VfsPPError) as ei:
        E.momentum_residual(h, met, z, u, u, 100.0, 0.01, coeff=1.5)
    assert ei.value.code == E.VFS_PP_EINVAL
    h.close()

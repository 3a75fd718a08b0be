"""Batched form: independent cases of uneven sizes solved together through
vfs_pp_batch_project_device must give each case the result it gets alone (and
the oracle's), call after call, with and without HIP graph replay."""
import numpy as np
import pytest

from oracle import oracle as O
from vfs_v3p2_amd import cases

pytestmark = pytest.mark.gpu

SPECS = [((20, 16, 24), 1.5, 0.2, False, 11), ((32, 24, 16), 2.5, 0.35, True, 12),
         ((16, 16, 16), 1.0, 0.0, False, 13), ((24, 20, 28), 3.0, 0.1, True, 14)]
_cases = {}


def get_cases():
    if not _cases:
        for q, sp in enumerate(SPECS):
            c = cases.ensemble_case(*sp)
            _cases[q] = (c, O.Assembled(c, drop=False))
    return [_cases[q] for q in range(len(SPECS))]


def relL2(a, b, mask):
    return float(np.linalg.norm((a - b)[mask]) / np.linalg.norm(b[mask]))


@pytest.fixture(scope="module")
def E(torch_cuda):
    from vfs_v3p2_amd import engine

    engine.lib()
    return engine


def handles(E, cs, **kw):
    out = []
    for c, _ in cs:
        h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed), E.default_opts(**kw))
        h.set_geometry(c.fields)
        out.append(h)
    return out


def fields(torch, cs):
    u = [torch.from_numpy(np.ascontiguousarray(c.fields["ucont"])).cuda() for c, _ in cs]
    phi = [torch.zeros(c.shape, dtype=torch.float64, device="cuda") for c, _ in cs]
    p = [torch.zeros(c.shape, dtype=torch.float64, device="cuda") for c, _ in cs]
    ucat = [torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda") for c, _ in cs]
    return u, phi, p, ucat


@pytest.mark.parametrize("precision,smoother,orthog", [(1, 1, 2), (0, 1, 2), (1, 2, 2), (0, 0, 2), (1, 0, 0)])
def test_batch_equals_alone_and_oracle(E, torch_cuda, precision, smoother, orthog):
    """smoother 1: point Chebyshev on every level; 2: y-line on every level; 0:
    chosen per level (y-line where the wall-normal coupling dominates).  orthog
    2: one lazily normalised Gram-Schmidt pass per step; 0: CGS2."""
    torch = torch_cuda
    cs = get_cases()
    kw = dict(precision=precision, rtol=1e-12, smoother=smoother, orthog=orthog)
    print(f"batch of {len(cs)} cases, {kw}", flush=True)
    # alone
    alone = []
    for h, (c, A) in zip(handles(E, cs, **kw), cs):
        u, phi, p, ucat = fields(torch, [(c, A)])
        st = h.project(u[0], phi[0], p[0], ucat[0], c.dt, c.st, c.coeff, strict=True)
        alone.append((st, phi[0].cpu().numpy(), u[0].cpu().numpy()))
        h.close()
    hs = handles(E, cs, **kw)
    B = E.Batch(hs)
    runs = []
    for call in range(2):
        u, phi, p, ucat = fields(torch, cs)
        sts = B.project(u, phi, p, ucat, [c.dt for c, _ in cs], [c.st for c, _ in cs], [c.coeff for c, _ in cs],
                        strict=True)
        runs.append([(s, ph.cpu().numpy(), uu.cpu().numpy(), uc.cpu().numpy()) for s, ph, uu, uc in
                     zip(sts, phi, u, ucat)])
    B.close()
    # plain launches instead of graph replay
    hs2 = handles(E, cs, **kw)
    B2 = E.Batch(hs2, graphs=False)
    u, phi, p, ucat = fields(torch, cs)
    sts2 = B2.project(u, phi, p, ucat, [c.dt for c, _ in cs], [c.st for c, _ in cs], [c.coeff for c, _ in cs],
                      strict=True)
    nograph = [ph.cpu().numpy() for ph in phi]
    B2.close()
    for q, (c, A) in enumerate(cs):
        s, ph, uu, uc = runs[0][q]
        assert s.converged and s.max_div <= 1e-8
        np.testing.assert_array_equal(ph, runs[1][q][1])  # graph replays agree bit for bit
        np.testing.assert_array_equal(ph, nograph[q])     # graph == plain launches
        assert relL2(ph, alone[q][1], A.fluid_cells) <= 1e-9
        assert abs(s.iterations - alone[q][0].iterations) <= 1
        ref = O.reference_projection(c, rtol=1e-13)
        assert relL2(ph, ref["phi"], A.fluid_cells) <= 1e-8
    for h in hs + hs2:
        h.close()


def test_batch_rejects_mixed_precision(E):
    cs = get_cases()[:2]
    h0 = handles(E, cs[:1], precision=1)[0]
    h1 = handles(E, cs[1:], precision=0)[0]
    with pytest.raises(E.VfsPPError):
        E.Batch([h0, h1])
    h0.close()
    h1.close()


def test_batch_divergence_rule(E, torch_cuda):
    torch = torch_cuda
    cs = get_cases()
    hs = handles(E, cs, precision=1, rtol=1e-6, div_tol=1e-8)
    B = E.Batch(hs)
    u, phi, p, ucat = fields(torch, cs)
    sts = B.project(u, phi, p, ucat, [c.dt for c, _ in cs], [c.st for c, _ in cs], [c.coeff for c, _ in cs],
                    strict=True)
    B.close()
    for s in sts:
        assert s.converged and s.max_div <= 1e-8
    for h in hs:
        h.close()


def _topology_cases():
    """The single-case parity cases (no periodic axis, immersed 0 and 1, towers)
    plus an all-walls box whose Contra2Cart uses the no-slip rule on every face."""
    from test_gpu_parity import CASES, affine_case, get_case

    out = [get_case(n) for n in CASES]
    c = affine_case(12, 10, 9, seed=3)
    c.bctype = (1,) * 6
    c.name = "affine-walls"
    out.append((c, O.Assembled(c, drop=False), O.reference_projection(c, rtol=1e-13)))
    return out


def test_batched_steps_on_every_topology(E, torch_cuda):
    """The batched per-case steps (VolumeFlux, RHS, initial guess from Phi, true
    residual, null space, Phi, P, Projection, Contra2Cart, Divergence) against the
    oracle: Phi and P to 1e-8, and Ucont, Ucat, the VolumeFlux figures and the
    divergence equal to the oracle's processing of the batch's own Phi."""
    torch = torch_cuda
    cs = _topology_cases()
    print("batch of " + ", ".join(c.name for c, _, _ in cs), flush=True)
    hs = []
    for c, _, _ in cs:
        h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed),
                     E.default_opts(rtol=1e-12, zero_guess=0))
        h.set_geometry(c.fields)
        hs.append(h)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    u = [dev(c.fields["ucont"]) for c, _, _ in cs]
    phi = [dev(0.5 * ref["phi"]) for _, _, ref in cs]  # a warm start read by the batched gather step
    p = [torch.zeros(c.shape, dtype=torch.float64, device="cuda") for c, _, _ in cs]
    ucat = [torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda") for c, _, _ in cs]
    B = E.Batch(hs)
    sts = B.project(u, phi, p, ucat, [c.dt for c, _, _ in cs], [c.st for c, _, _ in cs],
                    [c.coeff for c, _, _ in cs], strict=True)
    B.close()
    for h in hs:
        h.close()
    for q, (c, A, ref) in enumerate(cs):
        st = sts[q]
        fl = A.fluid_cells
        ph = phi[q].cpu().numpy()
        assert st.converged and st.max_div <= 1e-8, c.name
        assert relL2(ph, ref["phi"], fl) <= 1e-8, c.name
        assert relL2(p[q].cpu().numpy(), ref["p"], fl) <= 1e-8, c.name
        uo = np.ascontiguousarray(c.fields["ucont"].copy())
        vf = O.volume_flux(A.g, c.immersed, uo, c.fields)
        np.testing.assert_allclose(st.ibm_flux, vf[:4], rtol=1e-12, atol=1e-15)
        O.projection(A.g, c.fields, np.ascontiguousarray(ph), uo, c.dt, c.st, c.coeff)
        uc = np.zeros_like(uo)
        O.contra2cart(A.g, c.bctype, c.fields, uo, uc)
        ug = u[q].cpu().numpy()
        np.testing.assert_allclose(ug, uo, rtol=0, atol=1e-13 * np.abs(uo).max())
        np.testing.assert_allclose(ucat[q].cpu().numpy(), uc, rtol=0, atol=1e-13 * np.abs(uc).max())
        md = O.divergence(A.g, ug, np.ascontiguousarray(c.fields["aj"]), A.nvert)
        assert abs(md - st.max_div) <= 1e-12 * max(md, 1e-300) + 1e-20, c.name


def test_batch_members_are_refused_while_bound(E, torch_cuda):
    """While a batch holds a handle, its work lists and captured graph refer to the
    handle's hierarchy and buffers: the single-case entry points, a second batch
    and vfs_pp_destroy return VFS_PP_ESTATE, and after the batch is destroyed the
    handle works alone again."""
    torch = torch_cuda
    cs = get_cases()[:2]
    hs = handles(E, cs, rtol=1e-12)
    B = E.Batch(hs)
    c, _ = cs[0]
    u, phi, p, ucat = fields(torch, cs[:1])
    nv = torch.from_numpy(np.ascontiguousarray(c.fields["nvert"])).cuda()
    calls = [lambda: hs[0].set_geometry(c.fields), lambda: hs[0].set_nvert_device(nv),
             lambda: hs[0].project(u[0], phi[0], p[0], ucat[0], c.dt, c.st, c.coeff),
             lambda: hs[0].poisson_solve(u[0], phi[0], c.dt, c.st, c.coeff), lambda: hs[0].set_opts(rtol=1e-11),
             lambda: hs[0].export_stencil(), lambda: hs[0].close(), lambda: E.Batch(hs[:1])]
    for q, call in enumerate(calls):
        with pytest.raises(E.VfsPPError) as ei:
            call()
        assert ei.value.code == E.VFS_PP_ESTATE, q
    B.close()
    st = hs[0].project(u[0], phi[0], p[0], ucat[0], c.dt, c.st, c.coeff, strict=True)
    assert st.converged
    for h in hs:
        h.close()


def test_tile_shapes_give_identical_results(E, torch_cuda):
    """The cell-tiled batched launches walk 256 consecutive cells or 64x4, 32x8,
    16x16 (i x j) tiles, chosen by the autotuner; every shape, forced, gives the
    same bits, and the tuned plan reports the shapes it uses."""
    torch = torch_cuda
    cs = get_cases()
    hs = handles(E, cs, rtol=1e-12)
    B = E.Batch(hs)
    outs = {}
    for shape in (-1, 0, 1, 2, 3):
        B.set_tile_shape(shape)
        u, phi, p, ucat = fields(torch, cs)
        sts = B.project(u, phi, p, ucat, [c.dt for c, _ in cs], [c.st for c, _ in cs], [c.coeff for c, _ in cs],
                        strict=True)
        counts = B.tile_shapes()
        print(f"tile shape {shape}: launches per shape {counts}", flush=True)
        if shape >= 0:
            assert counts[shape] == sum(counts) > 0
        outs[shape] = [(ph.cpu().numpy(), uu.cpu().numpy(), s.iterations) for ph, uu, s in zip(phi, u, sts)]
    B.close()
    for h in hs:
        h.close()
    for shape in (0, 1, 2, 3):
        for (a, b, ia), (c, d, ic) in zip(outs[shape], outs[-1]):
            np.testing.assert_array_equal(a, c)
            np.testing.assert_array_equal(b, d)
            assert ia == ic

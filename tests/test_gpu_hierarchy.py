"""The set-up the engine builds on the device, against the oracle.

- The multigrid hierarchy, level by level: every level's rows and case bytes
  equal the oracle's rows of its own coarsening of the previous level (bit for bit
  in FP64; within FP16 rounding in the FP32 hierarchy, whose structure is
  bit-exact), the smoother choice follows the oracle's row coupling, and j is kept
  exactly where the oracle's rule keeps it.
- The VolumeFlux face list equals a restatement of poisson.c:2936-3010 (immersed
  == 1) and 2670-2700 (immersed == 0) in (k, j, i, axis) order, areas bit for bit.
- A second set-up on the same handle (a moved body, through vfs_pp_set_geometry
  or vfs_pp_set_nvert_device) gives the same solve as a new handle with that
  geometry, and repeated set-ups do not grow device memory.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vfs_v3p2_amd import cases

pytestmark = pytest.mark.gpu


def _sphere_immersed0():
    c = cases.hill_case(24, 16, 16, h=0.3, sphere=(0.6, 0.75, 0.0, 0.3), seed=1)
    c.immersed = 0
    return c


CASES = {
    "channel": lambda: cases.channel_case(16, 16, 12, beta=2.0, seed=0),
    "hill_sphere": lambda: cases.hill_case(24, 16, 16, h=0.3, sphere=(0.6, 0.75, 0.0, 0.3), seed=0),
    "hill_sphere_immersed0": _sphere_immersed0,
    "towers": lambda: cases.towers_case(48, 24, 24, ratio=1.05, rows=2, cols=2, radius_cells=3, seed=2),
    "stretched_towers": lambda: cases.towers_case(48, 32, 24, ratio=1.1, rows=2, cols=2, radius_cells=3, seed=2),
}
_cache = {}


def get_case(name):
    if name not in _cache:
        c = CASES[name]()
        _cache[name] = (c, O.Assembled(c, drop=False))
    return _cache[name]


@pytest.fixture(scope="module")
def E(torch_cuda):
    from vfs_v3p2_amd import engine

    engine.lib()
    return engine


def handle(E, c, **kw):
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed), E.default_opts(**kw))
    h.set_geometry(c.fields)
    return h


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name", list(CASES))
def test_hierarchy_matches_oracle_level_by_level(E, name, precision):
    c, A = get_case(name)
    print(f"hierarchy {name} {c.nx}x{c.ny}x{c.nz} precision {precision}", flush=True)
    h = handle(E, c, precision=precision)
    levels = h.export_levels()
    h.close()
    g, g9, nv = A.g, A.g9, A.nvert
    kept_j = 0
    for lv, L in enumerate(levels):
        assert L["dims"] == (g.mx - 2, g.my - 2, g.mz - 2), (lv, L["dims"])
        vol, _, cs = O.stencil(g, g9, nv)
        np.testing.assert_array_equal(L["cases"], cs, err_msg=f"level {lv} cases")
        if precision == 0:
            np.testing.assert_array_equal(L["vol"], vol, err_msg=f"level {lv} rows")
        else:
            # each row entry sums at most 12 terms of weight <= 1, each coefficient
            # rounded to FP16 (relative 2^-11, or 2^-25 of the level's scale below
            # the normal range)
            gmax = np.abs(g9[:, 1:-1, 1:-1, 1:-1]).max()
            np.testing.assert_allclose(L["vol"], vol, rtol=0, atol=12 * 2.0 ** -11 * gmax,
                                       err_msg=f"level {lv} rows")
        worst, weak = O.level_coupling(g, g9, nv)
        assert L["line"] == (not c.per[1] and worst > 4.0), (lv, worst)
        if lv + 1 < len(levels):
            keep_y = L["line"] and weak > 0.01
            kept_j += keep_y
            g, g9, nv, cm = O.level_coarsen(g, g9, nv, keep_y=keep_y)
    assert len(levels) >= 2 and (levels[-1]["dims"][0] * levels[-1]["dims"][1] * levels[-1]["dims"][2]) <= 1024
    if name == "stretched_towers":
        assert kept_j >= 1


def ib_faces_ref(c):
    """(pos, area, sign) of the VolumeFlux faces, poisson.c:2936-3010 / 2670-2700."""
    nv = c.fields["nvert"]
    mz, my, mx = nv.shape
    out = []
    for dd, key in enumerate(("csi", "eta", "zet")):
        K, J, I = np.meshgrid(np.arange(1, mz - 1), np.arange(1, my - 1), np.arange(1, mx - 1), indexing="ij")
        pos = (K, J, I)[2 - dd]
        m = (mx, my, mz)[dd]
        sh = [0, 0, 0]
        sh[2 - dd] = 1
        nc = nv[K, J, I]
        nq = nv[K + sh[0], J + sh[1], I + sh[2]]
        if c.immersed == 1:
            ok = pos < m - 2
            plus = ok & (nc < 0.1) & (nq > 0.1) & (nq < 1.1)
            minus = ok & ~plus & (nc > 0.1) & (nc < 1.1) & (nq < 0.1)
        else:
            ok = pos <= m - 2
            plus = ok & (nc < 0.1) & (nq > 0.1)
            minus = ok & ~plus & (nc > 0.1) & (nq < 0.1)
        sel = plus | minus
        p = (K * my + J) * mx + I
        mv = c.fields[key][K, J, I]
        area = np.sqrt(mv[..., 0] * mv[..., 0] + mv[..., 1] * mv[..., 1] + mv[..., 2] * mv[..., 2])
        out.append((3 * p[sel] + dd, area[sel], np.where(plus[sel], 1.0, -1.0)))
    pos = np.concatenate([o[0] for o in out])
    order = np.argsort(pos, kind="stable")
    return pos[order], np.concatenate([o[1] for o in out])[order], np.concatenate([o[2] for o in out])[order]


@pytest.mark.parametrize("name", ["hill_sphere", "hill_sphere_immersed0", "towers", "channel"])
def test_ib_face_list_bit_exact(E, name):
    c, _ = get_case(name)
    h = handle(E, c)
    pos, area, sign = h.export_ib_faces()
    h.close()
    rp, ra, rs = ib_faces_ref(c)
    if name == "channel":
        assert len(rp) == 0
    else:
        assert len(rp) > 20
    np.testing.assert_array_equal(pos, rp)
    np.testing.assert_array_equal(area, ra)
    np.testing.assert_array_equal(sign, rs.astype(np.float32))


def _moved(c, shift):
    """The same grid with the IB mask rolled by `shift` cells along i (a body that
    moved downstream)."""
    d = cases.Case(name=c.name + "-moved", nx=c.nx, ny=c.ny, nz=c.nz, per=c.per, bctype=c.bctype,
                   immersed=c.immersed, fields=dict(c.fields))
    nv = c.fields["nvert"].copy()
    nv[1:-1, 1:-1, 1:-1] = np.roll(nv[1:-1, 1:-1, 1:-1], shift, axis=2)
    cases.periodic_fill(nv, c.per)
    d.fields["nvert"] = nv
    return d


def _solve(E, torch, h, c):
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    u = dev(c.fields["ucont"])
    phi = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
    p = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
    ucat = torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda")
    st = h.project(u, phi, p, ucat, c.dt, c.st, c.coeff, strict=True)
    return st, phi.cpu().numpy(), u.cpu().numpy()


@pytest.mark.parametrize("precision", [1, 0])
def test_new_geometry_on_a_live_handle(E, torch_cuda, precision):
    torch = torch_cuda
    c, _ = get_case("hill_sphere")
    d = _moved(c, 3)
    print(f"moved body, precision {precision}", flush=True)
    fresh = handle(E, d, precision=precision)
    st_ref, phi_ref, u_ref = _solve(E, torch, fresh, d)
    fresh.close()
    h = handle(E, c, precision=precision)
    _solve(E, torch, h, c)
    h.set_geometry(d.fields)
    st, phi, u = _solve(E, torch, h, d)
    assert st.iterations == st_ref.iterations
    np.testing.assert_array_equal(phi, phi_ref)
    np.testing.assert_array_equal(u, u_ref)
    assert st.max_div == st_ref.max_div
    # and back: the first geometry again gives its own solve
    h.set_geometry(c.fields)
    again = handle(E, c, precision=precision)
    _, phi_a, _ = _solve(E, torch, again, c)
    again.close()
    _, phi_b, _ = _solve(E, torch, h, c)
    h.close()
    np.testing.assert_array_equal(phi_b, phi_a)


@pytest.mark.parametrize("precision", [1, 0])
def test_new_mask_on_the_same_grid(E, torch_cuda, precision):
    """vfs_pp_set_nvert_device: the moving-body step keeps the metrics and rebuilds
    masks, face list and hierarchy; the solve equals a new handle's."""
    torch = torch_cuda
    c, _ = get_case("hill_sphere")
    d = _moved(c, 3)
    fresh = handle(E, d, precision=precision)
    st_ref, phi_ref, u_ref = _solve(E, torch, fresh, d)
    fresh.close()
    h = handle(E, c, precision=precision)
    _solve(E, torch, h, c)
    h.set_nvert_device(torch.from_numpy(d.fields["nvert"]).cuda())
    st, phi, u = _solve(E, torch, h, d)
    h.close()
    assert st.iterations == st_ref.iterations and st.max_div == st_ref.max_div
    np.testing.assert_array_equal(phi, phi_ref)
    np.testing.assert_array_equal(u, u_ref)
    # before any geometry there are no metrics to keep
    h2 = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed))
    with pytest.raises(E.VfsPPError) as ei:
        h2.set_nvert_device(torch.from_numpy(d.fields["nvert"]).cuda())
    assert ei.value.code == E.VFS_PP_ESTATE
    h2.close()


def test_repeated_set_up_keeps_device_memory(E, torch_cuda):
    torch = torch_cuda
    c = cases.hill_case(96, 48, 64, h=0.3, sphere=(0.6, 0.75, 0.0, 0.3), seed=0)
    d = _moved(c, 5)
    h = handle(E, c)
    h.set_geometry(d.fields)
    torch.cuda.synchronize()
    free0 = torch.cuda.mem_get_info()[0]
    for f in (c.fields, d.fields, c.fields, d.fields):
        h.set_geometry(f)
    nvs = [torch.from_numpy(f["nvert"]).cuda() for f in (c.fields, d.fields)]
    for q in range(4):
        h.set_nvert_device(nvs[q % 2])
    del nvs
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    free1 = torch.cuda.mem_get_info()[0]
    h.close()
    # one hierarchy of this grid holds about 20 MB; eight leaked ones would show
    assert free0 - free1 < 8 << 20, (free0, free1)


def test_new_mask_after_device_grid_set_up(E, torch_cuda):
    """Metrics formed on the device (vfs_pp_set_grid_device), then a moved body
    through vfs_pp_set_nvert_device: the same solve as a handle set up from host
    metrics with the moved mask."""
    torch = torch_cuda
    c = cases.hill_case(24, 16, 16, h=0.3, sphere=(0.6, 0.75, 0.0, 0.3), seed=0, keep_coords=True)
    d = _moved(c, 2)
    fresh = handle(E, d)
    st_ref, phi_ref, u_ref = _solve(E, torch, fresh, d)
    fresh.close()
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed), E.default_opts())
    h.set_grid_device(torch.from_numpy(c.coor).cuda(), torch.from_numpy(c.fields["nvert"]).cuda())
    _solve(E, torch, h, c)
    h.set_nvert_device(torch.from_numpy(d.fields["nvert"]).cuda())
    st, phi, u = _solve(E, torch, h, d)
    h.close()
    assert st.iterations == st_ref.iterations
    np.testing.assert_array_equal(phi, phi_ref)
    np.testing.assert_array_equal(u, u_ref)


def _oracle_coarsest(c, A, levels):
    """The oracle's coarsest level (g, g9, nvert), coarsened along the engine's
    level sequence (j kept where the engine keeps it)."""
    g, g9, nv = A.g, A.g9, A.nvert
    for L in levels[:-1]:
        _, weak = O.level_coupling(g, g9, nv)
        g, g9, nv, _ = O.level_coarsen(g, g9, nv, keep_y=L["line"] and weak > 0.01)
    return g, g9, nv


# stencil slot offsets (di, dj, dk) in the slot order of poisson.c:956-974
SLOT_OFFSETS = [(0, 0, 0), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (1, -1, 0),
                (-1, 1, 0), (-1, -1, 0), (0, 1, 1), (0, 1, -1), (0, -1, 1), (0, -1, -1), (1, 0, 1), (1, 0, -1),
                (-1, 0, 1), (-1, 0, -1)]


def _bordered_inverse(g, g9, nv, per):
    """Top-left block of [A E; E^T 0]^-1 from the oracle's 19-slot rows of the level:
    A on the interior cells in (k, j, i) order, each slot added at its neighbour
    (periodic axes wrap, so on an axis of two cells both images add up, as in the
    matrix-free operator), identity rows for cells without a row; E one indicator
    column per fluid component (reached along each row's nonzero entries, cells
    taken in index order).  Inverted by LAPACK."""
    vol, rowflag, _ = O.stencil(g, g9, nv)
    nx, ny, nz = g.mx - 2, g.my - 2, g.mz - 2
    dims = (nx, ny, nz)
    n = nx * ny * nz
    has_row = rowflag[1:-1, 1:-1, 1:-1].ravel() != 0
    Afull = np.zeros((n, n))
    nbrs = [[] for _ in range(n)]
    for c in np.nonzero(has_row)[0]:
        k, j, i = c // (nx * ny), (c // nx) % ny, c % nx
        for s, off in enumerate(SLOT_OFFSETS):
            v = vol[k + 1, j + 1, i + 1, s]
            if v == 0.0:
                continue
            q = [i + off[0], j + off[1], k + off[2]]
            for a in range(3):
                if per[a]:
                    q[a] %= dims[a]
                assert 0 <= q[a] < dims[a], (c, s)  # a wall face carries no coefficient
            m = (q[2] * ny + q[1]) * nx + q[0]
            Afull[c, m] += v
            nbrs[c].append(m)
    comp = np.full(n, -1)
    ncomp = 0
    for c0 in range(n):
        if not has_row[c0] or comp[c0] >= 0:
            continue
        comp[c0] = ncomp
        stack = [c0]
        while stack:
            u = stack.pop()
            for v in nbrs[u]:
                if v != u and has_row[v] and comp[v] < 0:
                    comp[v] = ncomp
                    stack.append(v)
        ncomp += 1
    N = n + ncomp
    M = np.zeros((N, N))
    M[:n, :n] = Afull
    for c in range(n):
        if not has_row[c]:
            M[c, c] = 1.0
        else:
            M[c, n + comp[c]] = 1.0
            M[n + comp[c], c] = 1.0
    return np.linalg.inv(M)[:n, :n], N


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,coarse_max", [("hill_sphere", 512), ("hill_sphere", 64), ("towers", 512),
                                             ("stretched_towers", 512)])
def test_coarse_inverse_matches_lapack(E, name, coarse_max, precision):
    """The dense inverse of the coarsest level -- Gauss-Jordan on the host for
    bordered systems of up to 96 rows, one launch per column on the device above --
    against LAPACK's inverse of the same bordered system assembled from the oracle's
    rows of that level: to 1e-9 of its largest entry in FP64, to FP32 rounding in the
    mixed form.  coarse_max 64 takes the hill case down to a 3x2x2 level, periodic
    on an axis of two cells."""
    c, A = get_case(name)
    h = handle(E, c, precision=precision, coarse_max=coarse_max)
    levels = h.export_levels(rows=False)
    ce = h.export_coarse_inverse()
    h.close()
    ref, N = _bordered_inverse(*_oracle_coarsest(c, A, levels), c.per)
    print(f"coarse inverse {name} coarse_max {coarse_max} precision {precision}: n {ce.shape[0]} "
          f"bordered {N} ({'host' if N <= 96 else 'device'})", flush=True)
    assert ce.shape == ref.shape
    scale = np.abs(ref).max()
    tol = 1e-9 if precision == 0 else 2.0 ** -23
    assert np.abs(ce - ref).max() <= tol * scale
    # the system is not trivial: off-diagonal coupling is present
    assert np.abs(ref - np.diag(np.diag(ref))).max() > 1e-3 * scale

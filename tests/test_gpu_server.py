"""Served form: requests for resident cases, admitted into a continuous batch
as slots and HBM free up, must each get the result they get alone."""
import numpy as np
import pytest

from vfs_v3p2_amd import cases

pytestmark = pytest.mark.gpu

SPECS = [((20, 16, 24), 1.5, 0.2, False, 21), ((32, 24, 16), 2.5, 0.35, True, 22),
         ((16, 16, 16), 1.0, 0.0, False, 23), ((24, 20, 28), 3.0, 0.1, True, 24), ((28, 18, 20), 2.0, 0.3, True, 25)]
_cases = []


def get_cases():
    if not _cases:
        _cases.extend(cases.ensemble_case(*sp) for sp in SPECS)
    return _cases


def relL2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def E(torch_cuda):
    from vfs_v3p2_amd import engine

    engine.lib()
    return engine


def requests(torch, cs, nreq, seed):
    rng = np.random.default_rng(seed)
    out = []
    for t in range(nreq):
        q = int(rng.integers(len(cs)))
        c = cs[q]
        u = np.ascontiguousarray(c.fields["ucont"]) * (1.0 + 0.05 * rng.standard_normal(c.fields["ucont"].shape))
        cases.set_periodic_flux_ghosts(u, c.per)
        out.append((q, u))
    return out


def solve_alone(E, torch, c, u, opts):
    h = E.Handle(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed), E.default_opts(**opts))
    h.set_geometry(c.fields)
    ud = torch.from_numpy(u.copy()).cuda()
    phi = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
    p = torch.zeros_like(phi)
    ucat = torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda")
    st = h.project(ud, phi, p, ucat, c.dt, c.st, c.coeff, strict=True)
    h.close()
    return st, phi.cpu().numpy(), ud.cpu().numpy()


def run_server(E, torch, cs, reqs, opts, max_active, budget=0, per_round=2):
    sv = E.Server(max_active=max_active, hbm_budget=budget, **opts)
    ids = [sv.add_case(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed), c.fields) for c in cs]
    res, arrays, tick2req = {}, {}, {}
    pending = list(enumerate(reqs))
    max_queued = 0
    rounds = 0
    while pending or len(res) < len(reqs):
        # requests keep arriving while earlier ones are on the GPU
        for _ in range(per_round):
            if not pending:
                break
            r, (q, u) = pending.pop(0)
            c = cs[q]
            ud = torch.from_numpy(u.copy()).cuda()
            phi = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
            p = torch.zeros_like(phi)
            ucat = torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda")
            t = sv.submit(ids[q], ud, phi, p, ucat, c.dt, c.st, c.coeff)
            tick2req[t] = r
            arrays[r] = (ud, phi)
        max_queued = max(max_queued, sv.status()["queued"])
        for t, st in sv.progress():
            res[tick2req[t]] = st
        rounds += 1
        assert rounds < 10000
    st = sv.status()
    sv.close()
    return res, {r: (a[1].cpu().numpy(), a[0].cpu().numpy()) for r, a in arrays.items()}, max_queued, st


@pytest.mark.parametrize("precision", [1, 0])
def test_served_requests_equal_alone(E, torch_cuda, precision):
    torch = torch_cuda
    cs = get_cases()
    opts = dict(precision=precision, rtol=1e-12)
    reqs = requests(torch, cs, 12, seed=precision)
    print(f"12 requests over {len(cs)} resident cases, 3 slots, precision {precision}", flush=True)
    res, out, max_queued, st = run_server(E, torch, cs, reqs, opts, max_active=3)
    assert max_queued > 0  # requests did wait for slots
    assert st["active"] == 0 and st["queued"] == 0
    for r, (q, u) in enumerate(reqs):
        s = res[r]
        assert s.converged and s.max_div <= 1e-8
        st_a, phi_a, u_a = solve_alone(E, torch, cs[q], u, opts)
        phi, uo = out[r]
        assert relL2(phi, phi_a) <= 1e-9, (r, q)
        assert relL2(uo, u_a) <= 1e-12, (r, q)
        assert abs(s.iterations - st_a.iterations) <= 1


def test_server_memory_admission(E, torch_cuda):
    torch = torch_cuda
    cs = get_cases()[:3]
    opts = dict(precision=1, rtol=1e-10)
    # a budget that holds the resident cases and about one workspace: requests run one at a time
    probe = E.Server(max_active=4, **opts)
    for c in cs:
        probe.add_case(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed), c.fields)
    resident = probe.status()["resident_bytes"]
    probe.close()
    n = max(c.nx * c.ny * c.nz for c in cs)
    ws = (30 + 1) * n * 8 + 30 * n * 4
    reqs = requests(torch, cs, 6, seed=7)
    res, out, max_queued, st = run_server(E, torch, cs, reqs, opts, max_active=4, budget=resident + int(1.5 * ws),
                                          per_round=6)
    assert len(res) == 6 and all(s.converged for s in res.values())
    assert st["workspace_bytes"] <= int(1.5 * ws)
    assert max_queued >= 4
    with pytest.raises(E.VfsPPError) as ei:
        small = E.Server(max_active=2, hbm_budget=1 << 20, **opts)
        c = cs[0]
        small.add_case(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed), c.fields)
    assert ei.value.code == E.VFS_PP_ENOMEM


def test_inputs_written_just_before_submit_are_seen(E, torch_cuda):
    """Inputs produced by device kernels immediately before submit (on the
    default stream) are read after those kernels finish."""
    torch = torch_cuda
    c = get_cases()[1]
    opts = dict(precision=1, rtol=1e-12)
    base = torch.from_numpy(np.ascontiguousarray(c.fields["ucont"])).cuda()
    st_a, phi_a, _ = solve_alone(E, torch, c, (base * 0.75).cpu().numpy(), opts)
    sv = E.Server(max_active=2, **opts)
    cid = sv.add_case(E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed), c.fields)
    outs = []
    for rep in range(3):
        big = torch.ones(1 << 24, dtype=torch.float64, device="cuda").cumsum(0)  # keep the default stream busy
        u = base * 0.5 + base * 0.25 + big[:1] * 0.0
        phi = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
        p = torch.zeros_like(phi)
        ucat = torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda")
        t = sv.submit(cid, u, phi, p, ucat, c.dt, c.st, c.coeff)
        got = []
        while not got:
            got = sv.progress()
        assert got[0][0] == t and got[0][1].converged
        outs.append(phi.cpu().numpy())
    sv.close()
    for ph in outs:
        assert relL2(ph, phi_a) <= 1e-9


def _arrays(torch, c, u):
    ud = torch.from_numpy(u.copy()).cuda()
    phi = torch.full(c.shape, -7.0, dtype=torch.float64, device="cuda")
    p = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
    ucat = torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda")
    return ud, phi, p, ucat


def _drain(sv):
    done = {}
    rounds = 0
    while sv.status()["queued"] or sv.status()["active"]:
        for t, s in sv.progress():
            done[t] = s
        rounds += 1
        assert rounds < 10000
    return done


@pytest.mark.parametrize("precision", [1, 0])
def test_two_tickets_of_one_case_in_flight_together(E, torch_cuda, precision):
    """Three requests of one case run at once on the case's resident hierarchy
    (the first on the case's own handle, the others on solve clones of it) and
    give, bit for bit, what each gives when the server runs it by itself; a
    second round reuses the clones."""
    torch = torch_cuda
    c = get_cases()[1]
    opts = dict(precision=precision, rtol=1e-12)
    reqs = [u for _, u in requests(torch, [c], 3, seed=40 + precision)]
    print(f"3 requests of one {c.nx}x{c.ny}x{c.nz} case, precision {precision}", flush=True)
    grid = E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed)
    # one at a time
    sv1 = E.Server(max_active=4, max_per_case=1, **opts)
    cid = sv1.add_case(grid, c.fields)
    alone = []
    for u in reqs:
        a = _arrays(torch, c, u)
        t = sv1.submit(cid, *a, c.dt, c.st, c.coeff)
        st = _drain(sv1)[t]
        assert st.converged
        alone.append([x.cpu().numpy() for x in a])
    res1 = sv1.status()["resident_bytes"]
    sv1.close()
    # together
    sv = E.Server(max_active=4, max_per_case=3, **opts)
    cid = sv.add_case(grid, c.fields)
    res0 = sv.status()["resident_bytes"]
    assert res0 == res1
    for rnd in range(2):
        arrs, tick = [], []
        for u in reqs:
            a = _arrays(torch, c, u)
            arrs.append(a)
            tick.append(sv.submit(cid, *a, c.dt, c.st, c.coeff))
        first = sv.progress()
        assert sv.status()["active"] + len(first) == 3  # all three admitted at once
        done = dict(first)
        done.update(_drain(sv))
        for q in range(3):
            assert done[tick[q]].converged
            for got, want in zip(arrs[q], alone[q]):
                np.testing.assert_array_equal(got.cpu().numpy(), want)
        if rnd == 0:
            grown = sv.status()["resident_bytes"]
            assert grown > res0  # two solve clones
        else:
            assert sv.status()["resident_bytes"] == grown  # kept and reused
    clone = (grown - res0) / 2 / (c.nx * c.ny * c.nz)
    print(f"solve clone: {clone:.1f} bytes per cell", flush=True)
    assert clone < 200
    sv.close()


def test_queue_cap_and_admission_deadline(E, torch_cuda):
    """max_queue refuses submissions beyond the cap (no ticket); a request not
    admitted within deadline_ms is dropped untouched and listed by expired();
    the admitted one finishes as usual."""
    import time

    torch = torch_cuda
    c = get_cases()[0]
    opts = dict(precision=1, rtol=1e-12)
    grid = E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed)
    u = np.ascontiguousarray(c.fields["ucont"])
    sv = E.Server(max_active=1, max_per_case=1, max_queue=2, **opts)
    cid = sv.add_case(grid, c.fields)
    arrs = [_arrays(torch, c, u) for _ in range(3)]
    t0 = sv.submit(cid, *arrs[0], c.dt, c.st, c.coeff)
    t1 = sv.submit(cid, *arrs[1], c.dt, c.st, c.coeff)
    assert t0 is not None and t1 is not None
    assert sv.submit(cid, *arrs[2], c.dt, c.st, c.coeff) is None  # queue full
    done = _drain(sv)
    assert set(done) == {t0, t1} and all(s.converged for s in done.values())
    sv.close()

    sv = E.Server(max_active=1, max_per_case=1, deadline_ms=250.0, **opts)
    cid = sv.add_case(grid, c.fields)
    a0, a1 = _arrays(torch, c, u), _arrays(torch, c, u)
    t0 = sv.submit(cid, *a0, c.dt, c.st, c.coeff)
    got = sv.progress()  # admits t0 (well inside its deadline)
    t1 = sv.submit(cid, *a1, c.dt, c.st, c.coeff)  # waits: the only slot is taken
    time.sleep(0.5)
    done = dict(got)
    done.update(_drain(sv))
    assert set(done) == {t0} and done[t0].converged
    assert sv.expired() == [t1]
    assert sv.expired() == []
    assert bool((a1[1] == -7.0).all())  # the dropped request's phi was never written
    np.testing.assert_array_equal(a1[0].cpu().numpy(), u)
    sv.close()

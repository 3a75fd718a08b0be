"""Grid-file ingestion (init.c:251-389): text grid.dat, the -binary form and
xyz.dat.  Files are written here with every value in shortest round-trip form,
so reading them back must reproduce the written doubles exactly; the device
coordinates must equal, bit for bit, value / cl * L_dim laid out as the DMDA
coordinate array (with the plane beyond the last node extrapolated)."""
import os
import struct

import numpy as np
import pytest

from vfs_v3p2_amd import cases

from test_metrics import nodes_hill, nodes_warped


# ------------------------------------------------------------------ writers

def _values_text(a, per_line=5, sep=" ", eol="\n", fmt=repr):
    v = [fmt(float(x)) for x in np.ravel(a)]
    return eol.join(sep.join(v[q:q + per_line]) for q in range(0, len(v), per_line)) + eol


def write_text(path, blocks, eol="\n", fmt=repr):
    """grid.dat text: block count, then per block 'IM JM KM' and x, y, z planes."""
    with open(path, "w", newline="") as f:
        f.write(f"{len(blocks)}{eol}")
        for N in blocks:
            km, jm, im, _ = N.shape
            f.write(f"{im} {jm} {km}{eol}")
            for c in range(3):
                f.write(_values_text(N[..., c], eol=eol, fmt=fmt))


def write_binary(path, blocks):
    with open(path, "wb") as f:
        f.write(struct.pack("<i", len(blocks)))
        for N in blocks:
            km, jm, im, _ = N.shape
            f.write(struct.pack("<3i", im, jm, km))
            for c in range(3):
                f.write(np.ascontiguousarray(N[..., c], dtype="<f8").tobytes())


def write_xyz(path, X, Y, Z):
    with open(path, "w") as f:
        f.write(f"{len(X)} {len(Y)} {len(Z)}\n")
        for x in X:
            f.write(f"{float(x)!r} 0.0 0.0\n")
        for y in Y:
            f.write(f"0.0 {float(y)!r} 0.0\n")
        for z in Z:
            f.write(f"0.0 0.0 {float(z)!r}\n")


def expected_coor(N, cl=1.0, l_dim=1.0):
    return cases.dmda_coords(N / cl * l_dim)


@pytest.fixture(scope="module")
def E():
    from vfs_v3p2_amd import engine

    if not os.path.exists(engine.LIB_PATH):
        from vfs_v3p2_amd import build as B

        B.build(verbose=False)
    engine.lib()
    return engine


# ------------------------------------------------------ headers (no GPU)

def test_headers_of_every_format(E, tmp_path):
    a, b = nodes_hill(9, 7, 5), nodes_warped(6, 5, 4)
    write_text(tmp_path / "g.dat", [a, b])
    write_binary(tmp_path / "g.bin", [a, b])
    write_xyz(tmp_path / "xyz.dat", np.arange(12.0), np.arange(8.0), np.arange(3.0))
    for fmt, p in (("text", "g.dat"), ("binary", "g.bin")):
        assert E.grid_file_info(tmp_path / p, fmt, 0) == (2, 10, 8, 6)
        assert E.grid_file_info(tmp_path / p, fmt, 1) == (2, 7, 6, 5)
    assert E.grid_file_info(tmp_path / "xyz.dat", "xyz") == (1, 12, 8, 3)


def test_header_integers_read_as_percent_i(E, tmp_path):
    """fscanf %i honours base prefixes: '0x0a' is 10, '012' is octal 10."""
    N = nodes_warped(9, 9, 9)
    p = tmp_path / "g.dat"
    write_text(p, [N])
    txt = p.read_text().split("\n", 2)
    p.write_text("0x1\n0x0a 012 10\n" + txt[2])
    assert E.grid_file_info(p, "text") == (1, 10, 10, 10)


def test_bad_files_are_reported(E, tmp_path):
    N = nodes_warped(6, 5, 4)
    with pytest.raises(E.VfsPPError) as e:
        E.grid_file_info(tmp_path / "missing.dat", "text")
    assert e.value.code == 7
    write_binary(tmp_path / "g.bin", [N])
    raw = (tmp_path / "g.bin").read_bytes()
    (tmp_path / "short.bin").write_bytes(raw[:-8])
    with pytest.raises(E.VfsPPError) as e:
        E.grid_file_info(tmp_path / "short.bin", "binary")
    assert e.value.code == 7
    with pytest.raises(E.VfsPPError) as e:
        E.grid_file_info(tmp_path / "g.bin", "binary", 1)
    assert e.value.code == 1
    (tmp_path / "dims.dat").write_text("1\n1 5 5\n")
    with pytest.raises(E.VfsPPError) as e:
        E.grid_file_info(tmp_path / "dims.dat", "text")
    assert e.value.code == 7


# ------------------------------------------------------------ values (GPU)

def _read(E, path, fmt, block=0, cl=1.0, l_dim=1.0):
    g, coor = E.read_grid_device(path, fmt, block, cl, l_dim)
    return coor.cpu().numpy()


@pytest.mark.gpu
def test_text_grid_bitwise(E, torch_cuda, tmp_path):
    N = nodes_hill(30, 21, 17)
    write_text(tmp_path / "grid.dat", [N])
    np.testing.assert_array_equal(_read(E, tmp_path / "grid.dat", "text"), expected_coor(N))


@pytest.mark.gpu
def test_large_text_grid_parsed_in_pieces(E, torch_cuda, tmp_path):
    """A file of tens of MB is parsed by several threads over byte ranges; CRLF
    line ends, tabs, explicit '+' signs and upper-case exponents included."""
    N = nodes_warped(120, 70, 50)
    p = tmp_path / "grid.dat"
    write_text(p, [nodes_warped(5, 4, 3), N], eol="\r\n", fmt=lambda x: f"{x:+.17E}")
    assert os.path.getsize(p) > 8 << 20
    np.testing.assert_array_equal(_read(E, p, "text", block=1), expected_coor(N))


@pytest.mark.gpu
def test_binary_second_block_and_scaling(E, torch_cuda, tmp_path):
    a, b = nodes_warped(8, 6, 5), nodes_hill(20, 12, 10)
    write_binary(tmp_path / "grid.bin", [a, b])
    cl, l_dim = 2.5, 1.0 / 25.4
    np.testing.assert_array_equal(_read(E, tmp_path / "grid.bin", "binary", 1, cl, l_dim),
                                  expected_coor(b, cl, l_dim))
    np.testing.assert_array_equal(_read(E, tmp_path / "grid.bin", "binary", 0), expected_coor(a))


@pytest.mark.gpu
def test_xyz_cartesian(E, torch_cuda, tmp_path):
    rng = np.random.default_rng(3)
    X, Y, Z = (np.concatenate([[0.0], np.cumsum(rng.uniform(0.5, 1.5, n))]) for n in (11, 7, 5))
    write_xyz(tmp_path / "xyz.dat", X, Y, Z)
    N = np.stack(np.meshgrid(X, Y, Z, indexing="ij"), -1).transpose(2, 1, 0, 3)
    np.testing.assert_array_equal(_read(E, tmp_path / "xyz.dat", "xyz"), expected_coor(N))


@pytest.mark.gpu
def test_wrong_grid_dims_rejected(E, torch_cuda, tmp_path):
    from vfs_v3p2_amd import engine

    torch = torch_cuda
    N = nodes_warped(6, 5, 4)
    write_binary(tmp_path / "g.bin", [N])
    g = engine.make_grid(7, 5, 4)
    coor = torch.zeros(engine.dmda_shape(g, vector=True), dtype=torch.float64, device="cuda")
    with pytest.raises(engine.VfsPPError):
        engine.lib()
        engine._check(engine.lib().vfs_pp_read_grid_device(os.fsencode(tmp_path / "g.bin"), 1, 0, 1.0, 1.0,
                                                             engine.C.byref(g), engine._tptr(coor)), "read")


@pytest.mark.gpu
def test_file_to_projection(E, torch_cuda, tmp_path):
    """grid file -> device coordinates -> device FormMetrics and set-up ->
    projection, equal bit for bit to the projection set up from host metrics."""
    torch = torch_cuda
    c = cases.hill_case(24, 16, 16, h=0.3, sphere=(0.6, 0.75, 0.0, 0.3), seed=0, keep_coords=True)
    N = c.coor[:-1, :-1, :-1]
    write_text(tmp_path / "grid.dat", [N])
    g, coor = E.read_grid_device(tmp_path / "grid.dat", "text", periodic=c.per, bctype=c.bctype,
                                 immersed=c.immersed)
    np.testing.assert_array_equal(coor.cpu().numpy(), c.coor)
    outs = []
    for how in ("host", "file"):
        h = E.Handle(g, E.default_opts(rtol=1e-12))
        if how == "host":
            h.set_geometry(c.fields)
        else:
            h.set_grid_device(coor, torch.from_numpy(c.fields["nvert"]).cuda())
        u = torch.from_numpy(np.ascontiguousarray(c.fields["ucont"])).cuda()
        phi = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
        p = torch.zeros_like(phi)
        ucat = torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda")
        h.project(u, phi, p, ucat, c.dt, c.st, c.coeff, strict=True)
        outs.append((phi.cpu().numpy(), u.cpu().numpy(), ucat.cpu().numpy()))
        h.close()
    for x, y in zip(*outs):
        np.testing.assert_array_equal(x, y)

"""PETSc binary vector files (the vfield / pfield / nvfield / ufield checkpoints,
main.c:291-335, 780-797, 865-935).  The format is PETSc's: big-endian int32
class id 1211214 (bytes 00 12 7b 4e), big-endian int32 length, big-endian
float64 values.  Files are written here byte by byte with numpy and read back
through the library, and the library's files are checked byte by byte."""
import os
import struct

import numpy as np
import pytest

from vfs_v3p2_amd import cases


def write_vec(path, a):
    a = np.ascontiguousarray(a, dtype=np.float64).ravel()
    with open(path, "wb") as f:
        f.write(struct.pack(">ii", 1211214, a.size))
        f.write(a.astype(">f8").tobytes())


@pytest.fixture(scope="module")
def E():
    from vfs_v3p2_amd import engine

    if not os.path.exists(engine.LIB_PATH):
        from vfs_v3p2_amd import build as B

        B.build(verbose=False)
    engine.lib()
    return engine


def test_header_and_errors(E, tmp_path):
    write_vec(tmp_path / "pfield000010_0.dat", np.arange(7.0))
    assert (tmp_path / "pfield000010_0.dat").read_bytes()[:8] == bytes([0, 0x12, 0x7B, 0x4E, 0, 0, 0, 7])
    assert E.petsc_vec_length(tmp_path / "pfield000010_0.dat") == 7
    (tmp_path / "bad.dat").write_bytes(struct.pack(">ii", 1211216, 4) + bytes(32))  # a matrix class id
    with pytest.raises(E.VfsPPError) as e:
        E.petsc_vec_length(tmp_path / "bad.dat")
    assert e.value.code == 7
    with pytest.raises(E.VfsPPError) as e:
        E.petsc_vec_length(tmp_path / "missing.dat")
    assert e.value.code == 7


@pytest.mark.gpu
def test_checkpoint_round_trip(E, torch_cuda, tmp_path):
    torch = torch_cuda
    c = cases.hill_case(20, 12, 10, seed=4)
    ucont = c.fields["ucont"]
    print(f"petsc vec {c.name}", flush=True)
    # a file written with numpy reads into the DMDA layout
    write_vec(tmp_path / "vfield000001_0.dat", ucont)
    u = E.read_petsc_vec(tmp_path / "vfield000001_0.dat", ucont.shape)
    np.testing.assert_array_equal(u.cpu().numpy(), ucont)
    # the library writes the same bytes, and the block-size companion file
    E.write_petsc_vec(tmp_path / "out.dat", u)
    assert (tmp_path / "out.dat").read_bytes() == (tmp_path / "vfield000001_0.dat").read_bytes()
    assert (tmp_path / "out.dat.info").read_text() == "-vecload_block_size 3\n"
    nv = torch.from_numpy(c.fields["nvert"]).cuda()
    E.write_petsc_vec(tmp_path / "nv.dat", nv)
    assert not os.path.exists(tmp_path / "nv.dat.info")
    back = np.frombuffer((tmp_path / "nv.dat").read_bytes()[8:], dtype=">f8").reshape(nv.shape)
    np.testing.assert_array_equal(back, c.fields["nvert"])
    with pytest.raises(E.VfsPPError):
        E.read_petsc_vec(tmp_path / "nv.dat", ucont.shape)  # length mismatch

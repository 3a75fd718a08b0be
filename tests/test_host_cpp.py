"""The C++ host side (vfs_v3p2_amd/host: vfs::UserCtx and the reference's per-step
calls over the C ABI) builds, links against libvfs_pp.so and exports every call;
its driver refuses a wrong command line before touching the device."""
import ctypes as C
import os
import subprocess

import pytest

HOST = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "vfs_v3p2_amd", "host")

CALLS = ("UserCtxCreate", "UserCtxDestroy", "UserCtxUpload", "UserCtxDownload", "FormMetrics",
         "PoissonSolver_Hypre", "UpdatePressure", "Projection", "Divergence", "Flow_Solver_pressure_step")


def _built():
    lib = os.path.join(HOST, "libvfs_host.so")
    exe = os.path.join(HOST, "pressure_step")
    if not (os.path.exists(lib) and os.path.exists(exe)):
        pytest.skip("host side not built (python -m vfs_v3p2_amd.build)")
    return lib, exe


def test_host_library_exports_the_reference_calls():
    lib, _ = _built()
    L = C.CDLL(lib)
    out = subprocess.run(["nm", "-DC", "--defined-only", lib], capture_output=True, text=True, check=True).stdout
    for name in CALLS:
        assert f"vfs::{name}(" in out, name
    assert L is not None


def test_driver_refuses_a_wrong_command_line():
    _, exe = _built()
    r = subprocess.run([exe, "/nonexistent"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr

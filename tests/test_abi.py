"""The C ABI boundary, checked without a GPU: the library loads, exports every
function include/vfs_pp.h declares, the Python mirror binds all of them with
struct layouts identical to the C ones, and host-only calls behave."""
import ctypes as C
import os
import re
import subprocess
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vfs_pp.h")


def _ensure_built():
    from vfs_v3p2_amd import engine as E

    if not os.path.exists(E.LIB_PATH):
        from vfs_v3p2_amd import build as B

        B.build(verbose=False)
    return E


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vfs_pp_\w+)\s*\(", text)))


def test_header_declares_the_path():
    names = declared_functions()
    for must in ("vfs_pp_create", "vfs_pp_set_geometry", "vfs_pp_poisson_solve_device",
                 "vfs_pp_update_pressure_device", "vfs_pp_projection_device", "vfs_pp_project_device",
                 "vfs_pp_divergence_device", "vfs_pp_project"):
        assert must in names


def test_library_exports_every_declared_symbol():
    E = _ensure_built()
    L = C.CDLL(E.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", E.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\b(vfs_pp_\w+)\b", out))
    assert set(declared_functions()) <= exported
    extra = exported - set(declared_functions())
    assert not extra, f"exported but undeclared: {sorted(extra)}"


def test_python_mirror_binds_every_symbol():
    E = _ensure_built()
    L = E.lib()
    for n in declared_functions():
        assert getattr(L, n).argtypes is not None or n == "vfs_pp_error_string", n


def declared_arities():
    """Number of parameters of every function the header declares."""
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    out = {}
    for name, params in re.findall(r"\b(vfs_pp_\w+)\s*\(([^)]*)\)\s*;", text):
        params = params.strip()
        out[name] = 0 if params in ("", "void") else params.count(",") + 1
    return out


def test_python_mirror_arities_match_header():
    E = _ensure_built()
    L = E.lib()
    ar = declared_arities()
    assert set(ar) == set(declared_functions())
    for n, k in ar.items():
        assert len(getattr(L, n).argtypes) == k, n


PROBE = r"""
#include <stdio.h>
#include <stddef.h>
#include "vfs_pp.h"
#define F(T, m) printf(#T "." #m " %zu\n", offsetof(T, m));
int main(void) {
  printf("vfs_grid %zu\nvfs_pp_opts %zu\nvfs_pp_stats %zu\nvfs_pp_server_opts %zu\n", sizeof(vfs_grid),
         sizeof(vfs_pp_opts), sizeof(vfs_pp_stats), sizeof(vfs_pp_server_opts));
  F(vfs_pp_server_opts, max_active) F(vfs_pp_server_opts, solver)
  F(vfs_grid, periodic) F(vfs_grid, bctype) F(vfs_grid, immersed)
  F(vfs_pp_opts, rtol) F(vfs_pp_opts, div_tol) F(vfs_pp_opts, max_iters) F(vfs_pp_opts, restart)
  F(vfs_pp_opts, smoother) F(vfs_pp_opts, cheb_degree) F(vfs_pp_opts, coarse_max) F(vfs_pp_opts, drop_small)
  F(vfs_pp_opts, zero_guess) F(vfs_pp_opts, cheb_lo) F(vfs_pp_opts, cheb_hi) F(vfs_pp_opts, orthog)
  F(vfs_pp_opts, verbose) F(vfs_pp_opts, fused)
  F(vfs_pp_stats, iterations) F(vfs_pp_stats, vcycles) F(vfs_pp_stats, converged) F(vfs_pp_stats, rel_residual)
  F(vfs_pp_stats, bnorm) F(vfs_pp_stats, max_div) F(vfs_pp_stats, ibm_flux) F(vfs_pp_stats, phi_mean)
  F(vfs_pp_stats, ms_total)
  return 0;
}
"""


def test_struct_layouts_match_header(tmp_path):
    from vfs_v3p2_amd import engine as E

    src = tmp_path / "probe.c"
    src.write_text(PROBE)
    exe = tmp_path / "probe"
    subprocess.run(["gcc", "-std=c11", f"-I{os.path.dirname(HEADER)}", str(src), "-o", str(exe)], check=True)
    lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    got = dict(l.rsplit(" ", 1) for l in lines if l)
    cls = {"vfs_grid": E.Grid, "vfs_pp_opts": E.Opts, "vfs_pp_stats": E.Stats, "vfs_pp_server_opts": E.ServerOpts}
    for k, v in got.items():
        if "." in k:
            s, m = k.split(".")
            assert getattr(cls[s], m).offset == int(v), k
        else:
            assert C.sizeof(cls[k]) == int(v), k


def test_default_opts_and_error_strings():
    E = _ensure_built()
    o = E.default_opts()
    assert o.precision == 1 and o.zero_guess == 1 and o.drop_small == 0
    assert 0 < o.rtol < 1e-6 and o.restart >= 1 and o.max_iters >= 1
    assert 0 < o.cheb_lo < o.cheb_hi
    seen = set()
    for code in range(8):
        s = E.error_string(code)
        assert s and s not in seen
        seen.add(s)
    assert E.default_opts(precision=0, rtol=1e-9).precision == 0


def test_python_mirror_rejects_bad_arrays():
    E = _ensure_built()
    g = E.make_grid(8, 6, 4)
    assert E.dmda_shape(g) == (6, 8, 10) and E.dmda_shape(g, True) == (6, 8, 10, 3)
    with pytest.raises(ValueError):
        E._tptr(np.zeros((6, 8, 10)), E.dmda_shape(g), "phi")  # host array where a device tensor is needed
    with pytest.raises(ValueError):
        E._tptr(None, None, "phi")
    with pytest.raises(ValueError):
        E.coarse_dmda_shape(E.make_grid(9, 6, 4))
    assert E.coarse_dmda_shape(g) == (4, 5, 6)


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_create_refuses_unsupported_immersed_values():
    """immersed = 2 (VolumeFlux plus Add_IBMFlux_to_Outlet, poisson_hypre.c:662-665)
    and values the reference never defines are refused before any device work, so
    this runs without a GPU; 0, 1 and 3 pass the argument checks."""
    E = _ensure_built()
    for bad in (2, 4, -1):
        h = C.c_void_p()
        rc = E.lib().vfs_pp_create(C.byref(E.make_grid(8, 6, 4, immersed=bad)), None, C.byref(h))
        assert rc == E.VFS_PP_EINVAL and not h.value, bad


def test_repack_of_a_periodic_ghosted_array():
    """Under ii_periodic (or jj/kk) the reference's local arrays are ghosted by 3
    points on that axis (DMDA_BOUNDARY_PERIODIC, stencil width 3, init.c:147-158):
    index -2 holds the image of m-2.  vfs_pp_repack takes such an array (here
    built by periodic padding of a dense block) back to the dense block, and the
    dense block back into the window of a ghosted one; misaligned windows are
    refused."""
    E = _ensure_built()
    rng = np.random.default_rng(0)
    mz, my, mx = 5, 6, 7
    dense = rng.standard_normal((mz, my, mx, 3))
    # ii_periodic and kk_periodic: 3 ghost points on i and k, none on j
    ghosted = np.pad(dense, ((3, 3), (0, 0), (3, 3), (0, 0)), mode="wrap")
    assert np.array_equal(ghosted[3:-3, :, 3 - 2], dense[:, :, mx - 2])  # index -2 is the image of mx-2
    got = E.repack_to_dense(ghosted, (-3, 0, -3), (mx, my, mz))
    np.testing.assert_array_equal(got, dense)
    scalar = rng.standard_normal((mz, my, mx))
    np.testing.assert_array_equal(
        E.repack_to_dense(np.pad(scalar, ((0, 0), (3, 3), (0, 0)), mode="wrap"), (0, -3, 0), (mx, my, mz)), scalar)
    # dense -> ghosted window, ghost layers untouched
    back = np.full(ghosted.shape, 7.0)
    gd = (C.c_int * 3)(ghosted.shape[2], ghosted.shape[1], ghosted.shape[0])
    rc = E.lib().vfs_pp_repack(dense.ctypes.data_as(C.c_void_p), back.ctypes.data_as(C.c_void_p),
                               (C.c_int * 3)(-3, 0, -3), gd, (C.c_int * 3)(mx, my, mz), 3, 0)
    assert rc == 0
    np.testing.assert_array_equal(back[3:-3, :, 3:-3], dense)
    assert np.all(back[:3] == 7.0) and np.all(back[:, :, :3] == 7.0)
    for start, dims in (((1, 0, -3), (mx, my, mz)), ((-3, 0, -3), (mx + 4, my, mz))):
        with pytest.raises(E.VfsPPError):  # the window does not lie inside the ghosted range
            E.repack_to_dense(ghosted, start, dims)

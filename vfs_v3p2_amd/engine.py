"""Host-side mirror of the reference's pressure-path interface over the C ABI.

The reference calls, per block and time step (solvers.c:874-895):

    PoissonSolver_Hypre(user, ibm, ibm_intp)   # poisson_hypre.c:580
    UpdatePressure(user)                       # poisson.c:2473
    Projection(user)                           # poisson.c:2122 (ends with Contra2Cart)
    Divergence(user)                           # main.c:1244

with every field carried by UserCtx (variables.h:298-591).  `UserCtx` below
carries the same fields as device tensors in the reference's DMDA layout and
the functions of the same names run the HIP engine (libvfs_pp.so).  PyTorch is
only used to own device memory; all arithmetic is in the library.

Error behaviour: the reference returns 0 and is silent on non-convergence; here
every call raises `VfsPPError` on a non-zero status, except that a
non-converged solve is reported through `stats.converged` unless
`strict=True`.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvfs_pp.so")

(VFS_PP_OK, VFS_PP_EINVAL, VFS_PP_ENOMEM, VFS_PP_EHIP, VFS_PP_ENOCONV, VFS_PP_ESTATE, VFS_PP_ENONFINITE, VFS_PP_EIO,
 VFS_PP_EBUSY) = range(9)


class VfsPPError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what}: {error_string(code)} ({code})")
        self.code = code


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("periodic", C.c_int * 3),
                ("bctype", C.c_int * 6), ("immersed", C.c_int)]


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("rtol", C.c_double), ("div_tol", C.c_double), ("max_iters", C.c_int),
                ("restart", C.c_int), ("smoother", C.c_int), ("cheb_degree", C.c_int), ("coarse_max", C.c_int),
                ("drop_small", C.c_int), ("zero_guess", C.c_int), ("cheb_lo", C.c_double), ("cheb_hi", C.c_double),
                ("orthog", C.c_int), ("verbose", C.c_int), ("fused", C.c_int)]


class ServerOpts(C.Structure):
    _fields_ = [("hbm_budget", C.c_size_t), ("max_active", C.c_int), ("solver", Opts), ("max_per_case", C.c_int),
                ("max_queue", C.c_int), ("deadline_ms", C.c_double)]


class NewtonStats(C.Structure):
    _fields_ = [("newton_iterations", C.c_int), ("krylov_iterations", C.c_int), ("residual_evals", C.c_int),
                ("converged", C.c_int), ("fnorm0", C.c_double), ("fnorm", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Stats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("vcycles", C.c_int), ("converged", C.c_int), ("rel_residual", C.c_double),
                ("bnorm", C.c_double), ("max_div", C.c_double), ("ibm_flux", C.c_double * 4),
                ("phi_mean", C.c_double), ("ms_total", C.c_double)]

    def as_dict(self):
        return {k: (list(getattr(self, k)) if k == "ibm_flux" else getattr(self, k)) for k, _ in self._fields_}


_lib = None


def lib():
    """Load libvfs_pp.so; raises if it has not been built (no CPU fallback exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m vfs_v3p2_amd.build` "
                           "(the engine has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, dp, u8 = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_uint8)
    H = C.c_void_p
    sig = {
        "vfs_pp_default_opts": (None, [C.POINTER(Opts)]),
        "vfs_pp_create": (C.c_int, [C.POINTER(Grid), C.POINTER(Opts), C.POINTER(H)]),
        "vfs_pp_destroy": (C.c_int, [H]),
        "vfs_pp_set_opts": (C.c_int, [H, C.POINTER(Opts)]),
        "vfs_pp_set_geometry": (C.c_int, [H] + [dp] * 17),
        "vfs_pp_form_metrics_device": (C.c_int, [C.POINTER(Grid)] + [vp] * 17),
        "vfs_pp_set_grid_device": (C.c_int, [H, vp, vp]),
        "vfs_pp_pressure_gradient_device": (C.c_int, [C.POINTER(Grid)] + [vp] * 15),
        "vfs_pp_formfunction2_device": (C.c_int, [C.POINTER(Grid)] + [vp] * 18 + [C.c_double, C.c_double, vp]),
        "vfs_pp_runge_kutta_device": (C.c_int, [H] + [vp] * 18 + [C.c_double, C.c_double, C.c_double]),
        "vfs_pp_momentum_residual_device": (C.c_int, [H] + [vp] * 20 + [C.c_double] * 3 + [vp, vp]),
        "vfs_pp_implicit_momentum_device": (C.c_int, [H] + [vp] * 21 + [C.c_double] * 4 +
                                            [C.c_int, C.POINTER(NewtonStats)]),
        "vfs_pp_petsc_vec_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_longlong)]),
        "vfs_pp_read_petsc_vec_device": (C.c_int, [C.c_char_p, vp, C.c_size_t]),
        "vfs_pp_write_petsc_vec_device": (C.c_int, [C.c_char_p, vp, C.c_size_t, C.c_int]),
        "vfs_pp_ucd_info": (C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "vfs_pp_read_ucd": (C.c_int, [C.c_char_p] + [C.c_double] * 4 + [dp] * 3 + [C.POINTER(C.c_int)] * 3),
        "vfs_pp_ib_search_device": (C.c_int, [C.POINTER(Grid), vp, C.c_int, dp, dp, dp, C.c_int] +
                                    [C.POINTER(C.c_int)] * 3 + [C.c_int, vp]),
        "vfs_pp_grid_file_info": (C.c_int, [C.c_char_p, C.c_int, C.c_int] + [C.POINTER(C.c_int)] * 4),
        "vfs_pp_read_grid_device": (C.c_int, [C.c_char_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                              C.POINTER(Grid), vp]),
        "vfs_pp_poisson_solve_device": (C.c_int, [H, vp, vp, C.c_double, C.c_double, C.c_double, C.POINTER(Stats)]),
        "vfs_pp_update_pressure_device": (C.c_int, [H, vp, vp]),
        "vfs_pp_projection_device": (C.c_int, [H, vp, vp, vp, C.c_double, C.c_double, C.c_double]),
        "vfs_pp_divergence_device": (C.c_int, [H, vp, dp]),
        "vfs_pp_divergence_noalloc": (C.c_int, [C.POINTER(Grid), vp, vp, vp, vp]),
        "vfs_pp_batch_set_tile_shape": (C.c_int, [H, C.c_int]),
        "vfs_pp_batch_tile_shapes": (C.c_int, [H, C.POINTER(C.c_int)]),
        "vfs_pp_repack": (C.c_int, [vp, vp, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int), C.c_int,
                                    C.c_int]),
        "vfs_pp_project_device": (C.c_int, [H, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                            C.POINTER(Stats)]),
        "vfs_pp_project": (C.c_int, [H, dp, dp, dp, dp, C.c_double, C.c_double, C.c_double, C.POINTER(Stats)]),
        "vfs_pp_apply_device": (C.c_int, [H, vp, vp, C.c_int]),
        "vfs_pp_export_stencil": (C.c_int, [H, dp, u8]),
        "vfs_pp_export_level": (C.c_int, [H, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                          dp, u8]),
        "vfs_pp_set_nvert_device": (C.c_int, [H, vp]),
        "vfs_pp_export_coarse_inverse": (C.c_int, [H, C.c_size_t, dp, C.POINTER(C.c_int)]),
        "vfs_pp_export_ib_faces": (C.c_int, [H, C.c_size_t, C.POINTER(C.c_longlong), dp, C.POINTER(C.c_float),
                                             C.POINTER(C.c_size_t)]),
        "vfs_pp_classify_nvert_device": (C.c_int, [C.POINTER(Grid), vp]),
        "vfs_pp_mg_restrict_device": (C.c_int, [C.POINTER(Grid), vp, vp, vp, vp]),
        "vfs_pp_mg_prolong_device": (C.c_int, [C.POINTER(Grid), vp, vp, vp, vp]),
        "vfs_pp_mg_nvert_restrict_device": (C.c_int, [C.POINTER(Grid), vp, vp]),
        "vfs_pp_time_kernel": (C.c_int, [H, C.c_int, C.c_int, dp, dp]),
        "vfs_pp_describe": (C.c_int, [H, C.c_char_p, C.c_size_t]),
        "vfs_pp_error_string": (C.c_char_p, [C.c_int]),
        "vfs_pp_batch_create": (C.c_int, [C.POINTER(H), C.c_int, C.POINTER(H)]),
        "vfs_pp_batch_destroy": (C.c_int, [H]),
        "vfs_pp_batch_set_graphs": (C.c_int, [H, C.c_int]),
        "vfs_pp_batch_project_device": (C.c_int, [H] + [C.POINTER(C.c_void_p)] * 4 + [dp] * 3 +
                                        [C.POINTER(Stats)]),
        "vfs_pp_server_default_opts": (None, [C.POINTER(ServerOpts)]),
        "vfs_pp_server_create": (C.c_int, [C.POINTER(ServerOpts), C.POINTER(H)]),
        "vfs_pp_server_destroy": (C.c_int, [H]),
        "vfs_pp_server_add_case": (C.c_int, [H, C.POINTER(Grid)] + [dp] * 17 + [C.POINTER(C.c_int)]),
        "vfs_pp_server_remove_case": (C.c_int, [H, C.c_int]),
        "vfs_pp_server_submit": (C.c_int, [H, C.c_int, vp, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                           C.POINTER(C.c_int)]),
        "vfs_pp_server_progress": (C.c_int, [H, C.POINTER(C.c_int), C.POINTER(Stats), C.c_int,
                                             C.POINTER(C.c_int)]),
        "vfs_pp_server_expired": (C.c_int, [H, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int)]),
        "vfs_pp_server_status": (C.c_int, [H, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_size_t),
                                           C.POINTER(C.c_size_t)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def error_string(code):
    return lib().vfs_pp_error_string(code).decode()


def _check(rc, what, allow=()):
    if rc != VFS_PP_OK and rc not in allow:
        raise VfsPPError(rc, what)
    return rc


def default_opts(**kw) -> Opts:
    o = Opts()
    lib().vfs_pp_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def make_grid(nx, ny, nz, periodic=(1, 0, 1), bctype=(0, 0, 1, 1, 0, 0), immersed=0) -> Grid:
    g = Grid()
    g.nx, g.ny, g.nz = nx, ny, nz
    for a in range(3):
        g.periodic[a] = int(bool(periodic[a]))
    for f in range(6):
        g.bctype[f] = int(bctype[f])
    g.immersed = int(immersed)
    return g


def _dptr(a: np.ndarray):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def _tptr(t, shape=None, what="array"):
    """Device pointer of a contiguous FP64 CUDA tensor, checked against `shape`."""
    if t is None:
        raise ValueError(f"{what}: a device tensor is required")
    if not getattr(t, "is_cuda", False):
        raise ValueError(f"{what}: expected a CUDA tensor, got {type(t).__name__} on {getattr(t, 'device', 'host')}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: tensor must be contiguous")
    if str(t.dtype) != "torch.float64":
        raise ValueError(f"{what}: expected float64, got {t.dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{what}: expected shape {tuple(shape)}, got {tuple(t.shape)}")
    return C.c_void_p(t.data_ptr())


def dmda_shape(grid: Grid, vector=False):
    """DMDA array shape (mz, my, mx[, 3]) of a grid: interior cells plus one ghost
    plane on each side (init.c:156-157)."""
    sh = (grid.nz + 2, grid.ny + 2, grid.nx + 2)
    return sh + (3,) if vector else sh


class Handle:
    """One device-resident pressure system (vfs_pp_handle)."""

    def __init__(self, grid: Grid, opts: Optional[Opts] = None):
        self.grid = grid
        self.opts = opts if opts is not None else default_opts()
        self._h = C.c_void_p()
        _check(lib().vfs_pp_create(C.byref(grid), C.byref(self.opts), C.byref(self._h)), "vfs_pp_create")

    def close(self):
        if self._h:
            # VFS_PP_ESTATE while a batch still holds the handle: it stays open
            _check(lib().vfs_pp_destroy(self._h), "vfs_pp_destroy")
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_opts(self, **kw):
        for k, v in kw.items():
            setattr(self.opts, k, v)
        _check(lib().vfs_pp_set_opts(self._h, C.byref(self.opts)), "vfs_pp_set_opts")

    def set_geometry(self, f: dict):
        keep = []
        args = []
        for k in ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "iaj", "jaj", "kaj", "aj",
                  "csi", "eta", "zet", "nvert"):
            want = dmda_shape(self.grid, vector=k not in ("iaj", "jaj", "kaj", "aj", "nvert"))
            if tuple(np.shape(f[k])) != want:
                raise ValueError(f"set_geometry: {k} has shape {np.shape(f[k])}, expected {want}")
            a, p = _dptr(f[k])
            keep.append(a)
            args.append(p)
        _check(lib().vfs_pp_set_geometry(self._h, *args), "vfs_pp_set_geometry")

    def set_grid_device(self, coor, nvert):
        """FormMetrics on the device from node coordinates (mz, my, mx, 3) and the
        set-up of vfs_pp_set_geometry, from device tensors."""
        _check(lib().vfs_pp_set_grid_device(self._h, self._v(coor, "coor"), self._s(nvert, "nvert")),
               "vfs_pp_set_grid_device")

    def set_nvert_device(self, nvert):
        """A new IB mask (device tensor, DMDA shape) on the same grid: codes, face
        list and multigrid hierarchy are rebuilt, the metrics kept."""
        _check(lib().vfs_pp_set_nvert_device(self._h, self._s(nvert, "nvert")), "vfs_pp_set_nvert_device")

    def _s(self, t, what):
        return _tptr(t, dmda_shape(self.grid), what)

    def _v(self, t, what):
        return _tptr(t, dmda_shape(self.grid, vector=True), what)

    def describe(self) -> str:
        buf = C.create_string_buffer(4096)
        lib().vfs_pp_describe(self._h, buf, len(buf))
        return buf.value.decode()

    def poisson_solve(self, ucont, phi, dt, st, coeff, strict=False):
        s = Stats()
        rc = lib().vfs_pp_poisson_solve_device(self._h, self._v(ucont, "ucont"), self._s(phi, "phi"), dt, st, coeff,
                                               C.byref(s))
        _check(rc, "poisson solve", allow=() if strict else (VFS_PP_ENOCONV,))
        return s

    def update_pressure(self, p, phi):
        _check(lib().vfs_pp_update_pressure_device(self._h, self._s(p, "p"), self._s(phi, "phi")), "update pressure")

    def projection(self, ucont, phi, ucat, dt, st, coeff):
        _check(lib().vfs_pp_projection_device(self._h, self._v(ucont, "ucont"), self._s(phi, "phi"), self._v(ucat, "ucat"),
                                             dt, st, coeff),
               "projection")

    def divergence(self, ucont) -> float:
        out = C.c_double()
        _check(lib().vfs_pp_divergence_device(self._h, self._v(ucont, "ucont"), C.byref(out)), "divergence")
        return out.value

    def project(self, ucont, phi, p, ucat, dt, st, coeff, strict=False):
        s = Stats()
        rc = lib().vfs_pp_project_device(self._h, self._v(ucont, "ucont"), self._s(phi, "phi"), self._s(p, "p"),
                                         self._v(ucat, "ucat"), dt, st, coeff, C.byref(s))
        _check(rc, "project", allow=() if strict else (VFS_PP_ENOCONV,))
        return s

    def project_host(self, ucont, phi, p, ucat, dt, st, coeff):
        """Host-array variant (PCIe copies in and out); arrays are updated in place."""
        s = Stats()
        for nm, a, vec in (("ucont", ucont, True), ("phi", phi, False), ("p", p, False), ("ucat", ucat, True)):
            if not (isinstance(a, np.ndarray) and a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
                raise ValueError(f"{nm}: expected a C-contiguous float64 numpy array")
            if a.shape != dmda_shape(self.grid, vec):
                raise ValueError(f"{nm}: expected shape {dmda_shape(self.grid, vec)}, got {a.shape}")
        ptrs = [a.ctypes.data_as(C.POINTER(C.c_double)) for a in (ucont, phi, p, ucat)]
        rc = lib().vfs_pp_project(self._h, *ptrs, dt, st, coeff, C.byref(s))
        _check(rc, "project (host arrays)", allow=(VFS_PP_ENOCONV,))
        return s

    def apply(self, phi, y, drop_small=False):
        _check(lib().vfs_pp_apply_device(self._h, self._s(phi, "phi"), self._s(y, "y"), int(drop_small)), "apply")

    def export_stencil(self):
        M = (self.grid.nz + 2) * (self.grid.ny + 2) * (self.grid.nx + 2)
        vol = np.zeros(19 * M)
        cases = np.zeros(18 * M, np.uint8)
        _check(lib().vfs_pp_export_stencil(self._h, vol.ctypes.data_as(C.POINTER(C.c_double)),
                                           cases.ctypes.data_as(C.POINTER(C.c_uint8))), "export stencil")
        shape = (self.grid.nz + 2, self.grid.ny + 2, self.grid.nx + 2)
        return vol.reshape(shape + (19,)), cases.reshape(shape + (18,))

    def export_levels(self, rows=True):
        """The multigrid hierarchy: one dict per level with dims (nx, ny, nz), line
        (y-line smoothing) and, with rows, the level's row coefficients and case
        bytes on its DMDA box as export_stencil gives them."""
        nl = C.c_int()
        rc = lib().vfs_pp_export_level(self._h, 0, C.byref(nl), None, None, None, None)
        _check(rc, "export level")
        out = []
        for lv in range(nl.value):
            dims = (C.c_int * 3)()
            line = C.c_int()
            _check(lib().vfs_pp_export_level(self._h, lv, None, dims, C.byref(line), None, None), "export level")
            nx, ny, nz = dims[0], dims[1], dims[2]
            item = {"dims": (nx, ny, nz), "line": bool(line.value)}
            if rows:
                shape = (nz + 2, ny + 2, nx + 2)
                M = shape[0] * shape[1] * shape[2]
                vol = np.zeros(19 * M)
                cases = np.zeros(18 * M, np.uint8)
                _check(lib().vfs_pp_export_level(self._h, lv, None, None, None,
                                                 vol.ctypes.data_as(C.POINTER(C.c_double)),
                                                 cases.ctypes.data_as(C.POINTER(C.c_uint8))), "export level")
                item["vol"] = vol.reshape(shape + (19,))
                item["cases"] = cases.reshape(shape + (18,))
            out.append(item)
        return out

    def export_coarse_inverse(self):
        """The coarsest level's dense inverse as an (n, n) array C[r, c] (the top-left
        block of the bordered inverse; FP32 values widened in the mixed form)."""
        n = C.c_int()
        _check(lib().vfs_pp_export_coarse_inverse(self._h, 0, None, C.byref(n)), "export coarse inverse")
        m = n.value
        out = np.zeros(m * m)
        _check(lib().vfs_pp_export_coarse_inverse(self._h, out.size, out.ctypes.data_as(C.POINTER(C.c_double)),
                                                  C.byref(n)), "export coarse inverse")
        return out.reshape(m, m).T.copy()

    def export_ib_faces(self):
        """VolumeFlux face list: (pos = 3 * DMDA point + axis, area, sign) arrays."""
        n = C.c_size_t()
        _check(lib().vfs_pp_export_ib_faces(self._h, 0, None, None, None, C.byref(n)), "export IB faces")
        pos = np.zeros(n.value, np.int64)
        area = np.zeros(n.value)
        sign = np.zeros(n.value, np.float32)
        _check(lib().vfs_pp_export_ib_faces(self._h, n.value, pos.ctypes.data_as(C.POINTER(C.c_longlong)),
                                            area.ctypes.data_as(C.POINTER(C.c_double)),
                                            sign.ctypes.data_as(C.POINTER(C.c_float)), C.byref(n)),
               "export IB faces")
        return pos, area, sign

    def time_kernel(self, kernel: int, reps: int = 20):
        ms, by = C.c_double(), C.c_double()
        _check(lib().vfs_pp_time_kernel(self._h, kernel, reps, C.byref(ms), C.byref(by)), "time kernel")
        return ms.value, by.value


class Batch:
    """Independent cases solved together (vfs_pp_batch): one work list per kernel
    launch covers every member, one FGMRES iteration of all members is one HIP
    graph launch.  Members are `Handle`s with equal precision and Chebyshev
    degree; they must not be used alone while in a batch."""

    def __init__(self, handles, graphs=True):
        self.handles = list(handles)
        arr = (C.c_void_p * len(self.handles))(*[h._h.value for h in self.handles])
        self._b = C.c_void_p()
        _check(lib().vfs_pp_batch_create(arr, len(self.handles), C.byref(self._b)), "vfs_pp_batch_create")
        if not graphs:
            _check(lib().vfs_pp_batch_set_graphs(self._b, 0), "vfs_pp_batch_set_graphs")

    def close(self):
        if self._b:
            lib().vfs_pp_batch_destroy(self._b)
            self._b = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tile_shape(self, shape: int):
        """-1: tuned (default); 0 flat 256 cells, 1 64x4, 2 32x8, 3 16x16 forced."""
        _check(lib().vfs_pp_batch_set_tile_shape(self._b, int(shape)), "vfs_pp_batch_set_tile_shape")

    def tile_shapes(self):
        """Cell-tiled launches of the current plan per tile shape (flat, 64x4, 32x8, 16x16)."""
        c = (C.c_int * 4)()
        _check(lib().vfs_pp_batch_tile_shapes(self._b, c), "vfs_pp_batch_tile_shapes")
        return list(c)

    def project(self, ucont, phi, p, ucat, dt, st, coeff, strict=False):
        """Per-member lists of device tensors and scalars; returns a list of Stats."""
        n = len(self.handles)
        ptrs = []
        for name, arrs, vec in (("ucont", ucont, True), ("phi", phi, False), ("p", p, False), ("ucat", ucat, True)):
            if len(arrs) != n:
                raise ValueError(f"{name}: expected {n} arrays")
            ptrs.append((C.c_void_p * n)(*[_tptr(a, dmda_shape(h.grid, vec), f"{name}[{q}]").value
                                           for q, (a, h) in enumerate(zip(arrs, self.handles))]))
        scal = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (n,))) for v in (dt, st, coeff)]
        stats = (Stats * n)()
        rc = lib().vfs_pp_batch_project_device(self._b, *ptrs, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in scal],
                                               stats)
        _check(rc, "batch project", allow=() if strict else (VFS_PP_ENOCONV,))
        return list(stats)


GEOMETRY_FIELDS = ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "iaj", "jaj", "kaj",
                   "aj", "csi", "eta", "zet", "nvert")


class Server:
    """Continuous batching of projection requests (vfs_pp_server): cases stay
    resident; `submit` queues a projection of a case on device tensors and
    `progress` admits queued requests into free slots as the HBM budget allows,
    runs one batched iteration of everything in flight, and returns the
    requests that finished as (ticket, Stats) pairs.  Up to `max_per_case`
    requests of one case run at once; `max_queue` (> 0) refuses submissions
    while that many wait, and `deadline_ms` (> 0) drops requests not admitted in
    time (their tickets come back from `expired`)."""

    def __init__(self, max_active=64, hbm_budget=0, max_per_case=None, max_queue=None, deadline_ms=None, **solver):
        o = ServerOpts()
        lib().vfs_pp_server_default_opts(C.byref(o))
        o.max_active = int(max_active)
        o.hbm_budget = int(hbm_budget)
        if max_per_case is not None:
            o.max_per_case = int(max_per_case)
        if max_queue is not None:
            o.max_queue = int(max_queue)
        if deadline_ms is not None:
            o.deadline_ms = float(deadline_ms)
        for k, v in solver.items():
            setattr(o.solver, k, v)
        self.max_active = o.max_active
        self._s = C.c_void_p()
        _check(lib().vfs_pp_server_create(C.byref(o), C.byref(self._s)), "vfs_pp_server_create")
        self._keep = {}
        self.grids = {}

    def close(self):
        if self._s:
            lib().vfs_pp_server_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_case(self, grid: Grid, fields: dict) -> int:
        arrs, ptrs = [], []
        for k in GEOMETRY_FIELDS:
            want = dmda_shape(grid, vector=k not in ("iaj", "jaj", "kaj", "aj", "nvert"))
            if tuple(np.shape(fields[k])) != want:
                raise ValueError(f"add_case: {k} has shape {np.shape(fields[k])}, expected {want}")
            a, p = _dptr(fields[k])
            arrs.append(a)
            ptrs.append(p)
        cid = C.c_int()
        _check(lib().vfs_pp_server_add_case(self._s, C.byref(grid), *ptrs, C.byref(cid)), "vfs_pp_server_add_case")
        self.grids[cid.value] = grid
        return cid.value

    def remove_case(self, cid: int):
        _check(lib().vfs_pp_server_remove_case(self._s, cid), "vfs_pp_server_remove_case")
        self.grids.pop(cid, None)

    def submit(self, cid, ucont, phi, p, ucat, dt, st, coeff) -> Optional[int]:
        """The request's ticket, or None when the full queue refused it."""
        g = self.grids[cid]
        t = C.c_int()
        rc = _check(lib().vfs_pp_server_submit(self._s, cid, _tptr(ucont, dmda_shape(g, True), "ucont"),
                                               _tptr(phi, dmda_shape(g), "phi"), _tptr(p, dmda_shape(g), "p"),
                                               _tptr(ucat, dmda_shape(g, True), "ucat"), dt, st, coeff, C.byref(t)),
                    "vfs_pp_server_submit", allow=(VFS_PP_EBUSY,))
        if rc == VFS_PP_EBUSY:
            return None
        self._keep[t.value] = (ucont, phi, p, ucat)  # the arrays stay alive until the request finishes or expires
        return t.value

    def expired(self):
        """Tickets dropped by the admission deadline since the last call."""
        out = []
        buf = (C.c_int * 256)()
        n = C.c_int()
        while True:
            _check(lib().vfs_pp_server_expired(self._s, buf, 256, C.byref(n)), "vfs_pp_server_expired")
            for q in range(n.value):
                self._keep.pop(buf[q], None)
                out.append(buf[q])
            if n.value < 256:
                return out

    def progress(self):
        cap = self.max_active
        tickets = (C.c_int * cap)()
        stats = (Stats * cap)()
        nd = C.c_int()
        _check(lib().vfs_pp_server_progress(self._s, tickets, stats, cap, C.byref(nd)), "vfs_pp_server_progress")
        out = []
        for q in range(nd.value):
            self._keep.pop(tickets[q], None)
            out.append((tickets[q], stats[q]))
        return out

    def status(self):
        q, a = C.c_int(), C.c_int()
        r, w = C.c_size_t(), C.c_size_t()
        lib().vfs_pp_server_status(self._s, C.byref(q), C.byref(a), C.byref(r), C.byref(w))
        return dict(queued=q.value, active=a.value, resident_bytes=r.value, workspace_bytes=w.value)


# ----------------------------------------------------------- reference-named API

class UserCtx:
    """Device-resident counterpart of the UserCtx fields the pressure path uses
    (variables.h:305-309, 333-336, 352-356, 378-385, 398-400): Ucont, Ucat, P,
    Phi as torch CUDA tensors in DMDA layout, plus the engine handle that holds
    the metrics, nvert masks and multigrid hierarchy."""

    def __init__(self, case, opts: Optional[Opts] = None, device="cuda"):
        import torch

        self.case = case
        self.grid = make_grid(case.nx, case.ny, case.nz, case.per, case.bctype, case.immersed)
        self.handle = Handle(self.grid, opts)
        self.handle.set_geometry(case.fields)
        self.dt, self.st, self.coeff = case.dt, case.st, case.coeff
        sh = case.shape
        f = case.fields
        self.Ucont = torch.from_numpy(np.ascontiguousarray(f["ucont"])).to(device)
        self.Phi = torch.zeros(sh, dtype=torch.float64, device=device)
        self.P = torch.zeros(sh, dtype=torch.float64, device=device)
        self.Ucat = torch.zeros(sh + (3,), dtype=torch.float64, device=device)
        self.stats = None


def PoissonSolver(user: UserCtx, strict=False):
    """PoissonSolver_Hypre (poisson_hypre.c:580): VolumeFlux, RHS, solve,
    Remove_Nullspace, Convert_Phi2_Phi.  Returns the solve statistics."""
    user.stats = user.handle.poisson_solve(user.Ucont, user.Phi, user.dt, user.st, user.coeff, strict=strict)
    return user.stats


def UpdatePressure(user: UserCtx):
    """UpdatePressure (poisson.c:2473): P += Phi on fluid cells."""
    user.handle.update_pressure(user.P, user.Phi)


def Projection(user: UserCtx):
    """Projection (poisson.c:2122) followed by Contra2Cart (rhs.c:50)."""
    user.handle.projection(user.Ucont, user.Phi, user.Ucat, user.dt, user.st, user.coeff)


def Divergence(user: UserCtx) -> float:
    """Divergence (main.c:1244): max |aj div u| over cells not next to the IB."""
    return user.handle.divergence(user.Ucont)


def repack_to_dense(ghosted: np.ndarray, ghost_start, dims) -> np.ndarray:
    """The [0, m) window of a ghosted local DMDA array (numpy, shape (gz, gy, gx[, c])
    with ghost corners ghost_start = (gxs, gys, gzs)) as the dense block
    (mz, my, mx[, c]) the engine takes (vfs_pp_repack)."""
    g = np.ascontiguousarray(ghosted, dtype=np.float64)
    nc = g.shape[3] if g.ndim == 4 else 1
    gd = (C.c_int * 3)(g.shape[2], g.shape[1], g.shape[0])
    gs = (C.c_int * 3)(*ghost_start)
    dm = (C.c_int * 3)(*dims)
    out = np.zeros((dims[2], dims[1], dims[0]) + ((nc,) if g.ndim == 4 else ()))
    _check(lib().vfs_pp_repack(g.ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), gs, gd, dm, nc, 1),
           "vfs_pp_repack")
    return out


def divergence_noalloc(grid: Grid, ucont, aj, nvert, out) -> float:
    """Divergence (main.c:1244) from the DMDA arrays themselves, without a handle
    and without any device allocation: max |aj div u| is written to the one-element
    float64 device tensor `out` (and returned)."""
    S = dmda_shape(grid)
    _check(lib().vfs_pp_divergence_noalloc(C.byref(grid), _tptr(ucont, dmda_shape(grid, True), "ucont"),
                                           _tptr(aj, S, "aj"), _tptr(nvert, S, "nvert"), _tptr(out, (1,), "out")),
           "vfs_pp_divergence_noalloc")
    return float(out.item())


def Flow_Solver_pressure_step(user: UserCtx, strict=False):
    """The pressure block of Flow_Solver (solvers.c:874-895) as one call."""
    user.stats = user.handle.project(user.Ucont, user.Phi, user.P, user.Ucat, user.dt, user.st, user.coeff,
                                     strict=strict)
    return user.stats


# ----------------------------------------------------------- reference primitives

def coarse_dmda_shape(grid: Grid):
    """DMDA shape of the reference's next coarser level, IM_c = (IM_f + 1) / 2 with
    IM = m - 1 (init.c:95-127); the fine interior extents must be even."""
    if grid.nx % 2 or grid.ny % 2 or grid.nz % 2:
        raise ValueError("the reference multigrid primitives need even interior extents")
    return tuple((m - 1 + 1) // 2 + 1 for m in dmda_shape(grid))


METRIC_FIELDS = GEOMETRY_FIELDS[:16]
GRID_FORMATS = {"text": 0, "binary": 1, "xyz": 2}


def grid_file_info(path, fmt="text", block=0):
    """(nblocks, IM, JM, KM) of a grid file block (init.c:251-314 headers)."""
    v = [C.c_int() for _ in range(4)]
    _check(lib().vfs_pp_grid_file_info(os.fsencode(path), GRID_FORMATS[fmt], block, *[C.byref(x) for x in v]),
           "vfs_pp_grid_file_info")
    return tuple(x.value for x in v)


def read_grid_device(path, fmt="text", block=0, cl=1.0, l_dim=1.0, device="cuda", **grid_kw):
    """Read a grid-file block into device DMDA coordinates: returns (Grid, coor)
    with coor of shape (KM+1, JM+1, IM+1, 3); grid_kw go to make_grid."""
    import torch

    _, im, jm, km = grid_file_info(path, fmt, block)
    g = make_grid(im - 1, jm - 1, km - 1, **grid_kw)
    coor = torch.empty(dmda_shape(g, vector=True), dtype=torch.float64, device=device)
    _check(lib().vfs_pp_read_grid_device(os.fsencode(path), GRID_FORMATS[fmt], block, cl, l_dim, C.byref(g),
                                         _tptr(coor, None, "coor")), "vfs_pp_read_grid_device")
    return g, coor


def form_metrics_device(grid: Grid, coor):
    """FormMetrics on the device: dict of the 16 metric tensors (float64, on the
    device of `coor`) from DMDA node coordinates (mz, my, mx, 3)."""
    import torch

    V, S = dmda_shape(grid, vector=True), dmda_shape(grid)
    cp = _tptr(coor, V, "coor")
    out = {k: torch.empty(V if k in ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "csi",
                                     "eta", "zet") else S, dtype=torch.float64, device=coor.device)
           for k in METRIC_FIELDS}
    _check(lib().vfs_pp_form_metrics_device(C.byref(grid), cp, *[C.c_void_p(out[k].data_ptr()) for k in METRIC_FIELDS]),
           "vfs_pp_form_metrics_device")
    return out


def device_metrics_fn(device="cuda"):
    """A metrics function for cases.build_case that runs FormMetrics on the device
    (bit-identical to cases.form_metrics) and returns host arrays: generating a
    large synthetic case then costs one PCIe round trip instead of the numpy pass."""
    import torch

    def fn(co, per):
        mz, my, mx, _ = co.shape
        g = make_grid(mx - 2, my - 2, mz - 2, per)
        coor = torch.from_numpy(np.ascontiguousarray(co, dtype=np.float64)).to(device)
        out = form_metrics_device(g, coor)
        del coor
        host = {}
        for k in METRIC_FIELDS:
            host[k] = out.pop(k).cpu().numpy()
        torch.cuda.empty_cache()
        return host

    return fn


def Pressure_Gradient(grid: Grid, metrics: dict, nvert, p):
    """momentum.c:223 on the device: dP (mz, my, mx, 3) from P (periodic ghosts of P
    are filled in place) and the face metrics (device tensors keyed icsi..kaj, as
    form_metrics_device returns them)."""
    import torch

    V, S = dmda_shape(grid, vector=True), dmda_shape(grid)
    names = ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "iaj", "jaj", "kaj")
    ptrs = [_tptr(metrics[k], V if k[1:] in ("csi", "eta", "zet") else S, k) for k in names]
    dp = torch.empty(V, dtype=torch.float64, device=p.device)
    _check(lib().vfs_pp_pressure_gradient_device(C.byref(grid), *ptrs, _tptr(nvert, S, "nvert"), _tptr(p, S, "p"),
                                                 _tptr(dp, V, "dp")), "vfs_pp_pressure_gradient_device")
    return dp


MOMENTUM_METRICS = ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "iaj", "jaj", "kaj",
                    "csi", "eta", "zet")


def _momentum_ptrs(grid: Grid, metrics: dict):
    V, S = dmda_shape(grid, vector=True), dmda_shape(grid)
    return [_tptr(metrics[k], S if k.endswith("aj") else V, k) for k in MOMENTUM_METRICS]


def Formfunction_2(grid: Grid, metrics: dict, nvert, ucont, ucat, ren: float, scale: float = 1.0):
    """momentum.c:529 on the device: the momentum right-hand side (mz, my, mx, 3)
    from Ucont and Ucat and the 15 metric tensors (icsi..kzet, iaj, jaj, kaj, csi,
    eta, zet); grid.immersed is the reference's `immersed` switch."""
    import torch

    V, S = dmda_shape(grid, vector=True), dmda_shape(grid)
    rhs = torch.empty(V, dtype=torch.float64, device=ucont.device)
    _check(lib().vfs_pp_formfunction2_device(C.byref(grid), *_momentum_ptrs(grid, metrics), _tptr(nvert, S, "nvert"),
                                             _tptr(ucont, V, "ucont"), _tptr(ucat, V, "ucat"), float(ren),
                                             float(scale), _tptr(rhs, V, "rhs")), "vfs_pp_formfunction2_device")
    return rhs


def RungeKutta(handle: "Handle", metrics: dict, dp, ucont, ucat, ren: float, dt: float, st: float):
    """rhs.c:1511 on the device: one explicit four-stage step of Ucont (and Ucat)
    in place, with the handle's grid, masks and cell metrics."""
    g = handle.grid
    V = dmda_shape(g, vector=True)
    _check(lib().vfs_pp_runge_kutta_device(handle._h, *_momentum_ptrs(g, metrics), _tptr(dp, V, "dp"),
                                           _tptr(ucont, V, "ucont"), _tptr(ucat, V, "ucat"), float(ren), float(dt),
                                           float(st)), "vfs_pp_runge_kutta_device")


def _opt_ptr(t, shape, what):
    return None if t is None else _tptr(t, shape, what)


def momentum_residual(handle: "Handle", metrics: dict, dp, ucont_o, ucont, ren: float, dt: float, coeff: float = 1.0,
                      ucont_rm1=None, rhs_o=None, ucat=None):
    """FormFunction_SNES (momentum.c:2285) on the device: (F, Ucat) of the implicit
    momentum equation at Ucont (new device tensors).  ucat (default zero; copied)
    holds the Ucat that IB and solid cells keep."""
    import torch

    g = handle.grid
    V = dmda_shape(g, vector=True)
    F = torch.zeros(V, dtype=torch.float64, device=ucont.device)
    ucat = (torch.zeros(V, dtype=torch.float64, device=ucont.device) if ucat is None
            else ucat.detach().clone().contiguous())
    _check(lib().vfs_pp_momentum_residual_device(
        handle._h, *_momentum_ptrs(g, metrics), _tptr(dp, V, "dp"), _tptr(ucont_o, V, "ucont_o"),
        _opt_ptr(ucont_rm1, V, "ucont_rm1"), _opt_ptr(rhs_o, V, "rhs_o"), _tptr(ucont, V, "ucont"), float(ren),
        float(dt), float(coeff), _tptr(ucat, V, "ucat"), _tptr(F, V, "F")), "vfs_pp_momentum_residual_device")
    return F, ucat


def Implicit_MatrixFree(handle: "Handle", metrics: dict, dp, ucont_o, ucont, ucat, ren: float, dt: float,
                        coeff: float = 1.0, ucont_rm1=None, rhs_o=None, rtol: float = 1e-4, max_newton: int = 50,
                        strict: bool = True) -> NewtonStats:
    """implicitsolver.c:281 on the device: Ucont (in place) solving the implicit
    momentum equation by Jacobian-free Newton-Krylov, Ucat its Contra2Cart.
    strict=False returns the stats of a solve that missed rtol instead of raising."""
    g = handle.grid
    V = dmda_shape(g, vector=True)
    st = NewtonStats()
    rc = lib().vfs_pp_implicit_momentum_device(
        handle._h, *_momentum_ptrs(g, metrics), _tptr(dp, V, "dp"), _tptr(ucont_o, V, "ucont_o"),
        _opt_ptr(ucont_rm1, V, "ucont_rm1"), _opt_ptr(rhs_o, V, "rhs_o"), _tptr(ucont, V, "ucont"),
        _tptr(ucat, V, "ucat"), float(ren), float(dt), float(coeff), float(rtol), int(max_newton), C.byref(st))
    _check(rc, "vfs_pp_implicit_momentum_device", allow=() if strict else (VFS_PP_ENOCONV,))
    return st


def petsc_vec_length(path) -> int:
    n = C.c_longlong()
    _check(lib().vfs_pp_petsc_vec_info(os.fsencode(path), C.byref(n)), "vfs_pp_petsc_vec_info")
    return n.value


def read_petsc_vec(path, shape, device="cuda"):
    """A PETSc binary vector file (vfield/pfield/nvfield/ufield) into a new device
    tensor of the given DMDA shape (mz, my, mx[, 3])."""
    import torch

    t = torch.empty(tuple(shape), dtype=torch.float64, device=device)
    _check(lib().vfs_pp_read_petsc_vec_device(os.fsencode(path), _tptr(t, None, "v"), t.numel()),
           "vfs_pp_read_petsc_vec_device")
    return t


def write_petsc_vec(path, t, block_size=None):
    """Write a device tensor as a PETSc binary vector (block size 3 for AoS vector
    fields by default)."""
    bs = block_size if block_size is not None else (3 if t.dim() == 4 and t.shape[-1] == 3 else 1)
    _check(lib().vfs_pp_write_petsc_vec_device(os.fsencode(path), _tptr(t, None, "v"), t.numel(), int(bs)),
           "vfs_pp_write_petsc_vec_device")


def read_ucd(path, cl=1.0, centre=(0.0, 0.0, 0.0)):
    """ibmdata surface file -> (vertices (n_v, 3), triangles (n_elmt, 3), 0-based)."""
    nv, ne = C.c_int(), C.c_int()
    _check(lib().vfs_pp_ucd_info(os.fsencode(path), C.byref(nv), C.byref(ne)), "vfs_pp_ucd_info")
    xyz = [np.zeros(nv.value) for _ in range(3)]
    tri = [np.zeros(ne.value, dtype=np.int32) for _ in range(3)]
    _check(lib().vfs_pp_read_ucd(os.fsencode(path), cl, *centre, *[a.ctypes.data_as(C.POINTER(C.c_double)) for a in xyz],
                                 *[t.ctypes.data_as(C.POINTER(C.c_int)) for t in tri]), "vfs_pp_read_ucd")
    return np.stack(xyz, -1), np.stack(tri, -1)


def ib_search_device(grid: Grid, coor, verts, tris, nvert, body=0):
    """ibm_search_advanced for one triangulated body on the device: updates the
    device nvert tensor in place (ray parity, dilation, remap, wall layers)."""
    V = np.ascontiguousarray(verts, dtype=np.float64)
    T = np.ascontiguousarray(tris, dtype=np.int32)
    if V.ndim != 2 or V.shape[1] != 3 or T.ndim != 2 or T.shape[1] != 3:
        raise ValueError("ib_search_device: verts and tris must be (n, 3)")
    x, y, z = (np.ascontiguousarray(V[:, q]) for q in range(3))
    a, b, c = (np.ascontiguousarray(T[:, q]) for q in range(3))
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int)
    _check(lib().vfs_pp_ib_search_device(C.byref(grid), _tptr(coor, dmda_shape(grid, vector=True), "coor"), len(x),
                                         x.ctypes.data_as(dp), y.ctypes.data_as(dp), z.ctypes.data_as(dp), len(a),
                                         a.ctypes.data_as(ip), b.ctypes.data_as(ip), c.ctypes.data_as(ip), int(body),
                                         _tptr(nvert, dmda_shape(grid), "nvert")), "vfs_pp_ib_search_device")


def classify_nvert_device(grid: Grid, nvert_dev):
    _check(lib().vfs_pp_classify_nvert_device(C.byref(grid), _tptr(nvert_dev, dmda_shape(grid), "nvert")),
           "classify nvert")


def mg_restrict_device(grid: Grid, nvf, nvc, xf, fc):
    F, Cs = dmda_shape(grid), coarse_dmda_shape(grid)
    _check(lib().vfs_pp_mg_restrict_device(C.byref(grid), _tptr(nvf, F, "nvert_f"), _tptr(nvc, Cs, "nvert_c"),
                                           _tptr(xf, F, "xf"), _tptr(fc, Cs, "fc")), "restrict")


def mg_prolong_device(grid: Grid, nvf, nvc, xc, ff):
    F, Cs = dmda_shape(grid), coarse_dmda_shape(grid)
    _check(lib().vfs_pp_mg_prolong_device(C.byref(grid), _tptr(nvf, F, "nvert_f"), _tptr(nvc, Cs, "nvert_c"),
                                          _tptr(xc, Cs, "xc"), _tptr(ff, F, "ff")), "prolong")


def mg_nvert_restrict_device(grid: Grid, nvf, nvc):
    F, Cs = dmda_shape(grid), coarse_dmda_shape(grid)
    _check(lib().vfs_pp_mg_nvert_restrict_device(C.byref(grid), _tptr(nvf, F, "nvert_f"), _tptr(nvc, Cs, "nvert_c")),
           "nvert restrict")

"""ctypes wrapper of the CPU oracle (liborvfs.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may
import this module.  All arrays are numpy float64 in the reference DMDA layout
([mz][my][mx] with i fastest; vector fields carry a trailing axis of 3).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liborvfs.so")
_lib = None


class Grid(C.Structure):
    _fields_ = [("mx", C.c_int), ("my", C.c_int), ("mz", C.c_int), ("per", C.c_int * 3)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE, "liborvfs.so"])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        d = C.POINTER(C.c_double)
        u8 = C.POINTER(C.c_uint8)
        ip = C.POINTER(C.c_int)
        g = C.POINTER(Grid)
        _lib.or_setup_gid.restype = C.c_int
        _lib.or_setup_gid.argtypes = [g, d, d]
        _lib.or_face_coeffs.argtypes = [g] + [d] * 13
        _lib.or_stencil.argtypes = [g, d, d, d, u8, u8]
        _lib.or_assemble_csr.restype = C.c_int
        _lib.or_assemble_csr.argtypes = [g, d, u8, d, C.c_int, ip, ip, d]
        _lib.or_csr_matvec.argtypes = [C.c_int, ip, ip, d, d, d]
        _lib.or_stencil_apply.argtypes = [g, d, u8, d, d]
        _lib.or_rhs.argtypes = [g, d, d, C.c_double, C.c_double, C.c_double, d]
        _lib.or_volume_flux.argtypes = [g, C.c_int, d, d, d, d, d, d]
        _lib.or_remove_nullspace.argtypes = [g, d, d, d]
        _lib.or_interior_to_phi.argtypes = [g, d, d, d]
        _lib.or_convert_phi.argtypes = [g, d, d, d]
        _lib.or_update_pressure.argtypes = [g, d, d, d]
        _lib.or_projection.argtypes = [g] + [d] * 12 + [d, d, d, C.c_double, C.c_double, C.c_double]
        _lib.or_contra2cart.argtypes = [g, ip, d, d, d, d, d, d]
        _lib.or_divergence.restype = C.c_double
        _lib.or_divergence.argtypes = [g, d, d, d, d]
        _lib.or_nvert_classify.argtypes = [g, ip, d]
        _lib.or_mg_restrict.argtypes = [g, g, d, d, d, d]
        _lib.or_mg_prolong.argtypes = [g, g, d, d, d, d]
        _lib.or_mg_nvert_restrict.argtypes = [g, g, d, d]
        _lib.or_gmres.restype = C.c_int
        _lib.or_gmres.argtypes = [C.c_int, ip, ip, d, d, d, C.c_int, C.c_double, C.c_int, C.c_int, d]
        _lib.or_num_threads.restype = C.c_int
        _lib.or_form_metrics.argtypes = [g, d, C.POINTER(d)]
        _lib.or_ib_search.restype = C.c_int
        _lib.or_ib_search.argtypes = [g, ip, d, C.c_int, d, d, d, C.c_int, ip, ip, ip, C.c_int, C.c_int, d]
        _lib.or_pressure_gradient.argtypes = [g] + [d] * 15
        _lib.or_formfunction2.argtypes = [g, ip, C.c_int] + [d] * 18 + [C.c_double, C.c_double, d]
        _lib.or_runge_kutta.argtypes = [g, ip, C.c_int] + [d] * 19 + [C.c_double, C.c_double, C.c_double]
        _lib.or_momentum_bc_zero.argtypes = [g, ip, d]
        _lib.or_momentum_ib_bc.argtypes = [g, ip, C.c_int, d, d]
        _lib.or_momentum_residual.argtypes = [g, ip, C.c_int] + [d] * 16 + [d] * 5 + [C.c_double] * 3 + [d] * 3
        _lib.or_ray_s.restype = C.c_double
        _lib.or_ray_s.argtypes = [C.c_int]
        _lib.or_level_coupling.argtypes = [g, d, d, d, d]
        _lib.or_level_coarsen.restype = C.c_int
        _lib.or_level_coarsen.argtypes = [g, d, d, C.c_int, g, d, d]
        _lib.or_mg_create.restype = C.c_void_p
        _lib.or_mg_create.argtypes = [g, d, d, C.c_int, C.c_double, C.c_double, C.c_int]
        _lib.or_mg_destroy.argtypes = [C.c_void_p]
        _lib.or_mg_vcycle.argtypes = [C.c_void_p, d, d]
        _lib.or_mg_levels.restype = C.c_int
        _lib.or_mg_levels.argtypes = [C.c_void_p]
        _lib.or_mg_fgmres.restype = C.c_int
        _lib.or_mg_fgmres.argtypes = [C.c_void_p, d, d, C.c_int, C.c_double, C.c_int, d, ip]
    return _lib


def _p(a, t=C.c_double):
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(t))


def grid(shape, per):
    mz, my, mx = shape[:3]
    g = Grid(mx, my, mz)
    g.per[0], g.per[1], g.per[2] = [int(bool(x)) for x in per]
    return g


def setup_gid(g, nvert):
    gid = np.empty_like(nvert)
    n = lib().or_setup_gid(C.byref(g), _p(nvert), _p(gid))
    return gid, n


def face_coeffs(g, f):
    shape = (g.mz, g.my, g.mx)
    g9 = np.zeros((9,) + shape)
    args = [np.ascontiguousarray(f[k]) for k in ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta",
                                                 "kzet", "iaj", "jaj", "kaj")]
    lib().or_face_coeffs(C.byref(g), *[_p(a) for a in args], _p(g9))
    return g9


def stencil(g, g9, nvert):
    shape = (g.mz, g.my, g.mx)
    vol = np.zeros(shape + (19,))
    rowflag = np.zeros(shape, np.uint8)
    cases = np.zeros(shape + (18,), np.uint8)
    lib().or_stencil(C.byref(g), _p(g9), _p(nvert), _p(vol), _p(rowflag, C.c_uint8), _p(cases, C.c_uint8))
    return vol, rowflag, cases


def assemble_csr(g, vol, rowflag, gid, nrow, drop=True):
    rowptr = np.zeros(nrow + 1, np.int32)
    col = np.zeros(19 * max(nrow, 1), np.int32)
    val = np.zeros(19 * max(nrow, 1))
    nnz = lib().or_assemble_csr(C.byref(g), _p(vol), _p(rowflag, C.c_uint8), _p(gid), int(drop), _p(rowptr, C.c_int),
                                _p(col, C.c_int), _p(val))
    if nnz < 0:
        raise RuntimeError("row numbering mismatch")
    return rowptr, col[:nnz].copy(), val[:nnz].copy()


def csr_matvec(rowptr, col, val, x):
    y = np.zeros(len(rowptr) - 1)
    lib().or_csr_matvec(len(rowptr) - 1, _p(rowptr, C.c_int), _p(col, C.c_int), _p(val), _p(x), _p(y))
    return y


def stencil_apply(g, vol, rowflag, phi):
    y = np.zeros_like(phi)
    lib().or_stencil_apply(C.byref(g), _p(vol), _p(rowflag, C.c_uint8), _p(phi), _p(y))
    return y


def rhs(g, ucont, nvert, dt, st, coeff):
    b = np.zeros(ucont.shape[:3])
    lib().or_rhs(C.byref(g), _p(ucont), _p(nvert), dt, st, coeff, _p(b))
    return b


def volume_flux(g, immersed, ucont, f):
    out = np.zeros(4)
    lib().or_volume_flux(C.byref(g), int(immersed), _p(ucont), _p(f["csi"]), _p(f["eta"]), _p(f["zet"]),
                         _p(f["nvert"]), _p(out))
    return out


def remove_nullspace(g, nvert, gid, x):
    lib().or_remove_nullspace(C.byref(g), _p(nvert), _p(gid), _p(x))


def interior_to_phi(g, nvert, x):
    """or_interior_to_phi: removes the fluid mean from the interior array x (in
    place) and returns phi on the DMDA with periodic ghost images."""
    phi = np.zeros((g.mz, g.my, g.mx))
    lib().or_interior_to_phi(C.byref(g), _p(nvert), _p(x), _p(phi))
    return phi


def convert_phi(g, nvert, phi2):
    phi = np.zeros(nvert.shape)
    lib().or_convert_phi(C.byref(g), _p(nvert), _p(phi2), _p(phi))
    return phi


def update_pressure(g, nvert, p, phi):
    lib().or_update_pressure(C.byref(g), _p(nvert), _p(p), _p(phi))


def projection(g, f, phi, ucont, dt, st, coeff):
    args = [f[k] for k in ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "iaj", "jaj",
                           "kaj")]
    lib().or_projection(C.byref(g), *[_p(a) for a in args], _p(f["nvert"]), _p(phi), _p(ucont), dt, st, coeff)


def contra2cart(g, bctype, f, ucont, ucat):
    bc = np.asarray(bctype, np.int32)
    lib().or_contra2cart(C.byref(g), _p(bc, C.c_int), _p(f["csi"]), _p(f["eta"]), _p(f["zet"]), _p(f["nvert"]),
                         _p(ucont), _p(ucat))


def divergence(g, ucont, aj, nvert, want_field=False):
    div = np.zeros(nvert.shape) if want_field else None
    m = lib().or_divergence(C.byref(g), _p(ucont), _p(aj), _p(nvert), _p(div) if want_field else None)
    return (m, div) if want_field else m


def nvert_classify(g, bctype, nvert):
    bc = np.asarray(bctype, np.int32)
    lib().or_nvert_classify(C.byref(g), _p(bc, C.c_int), _p(nvert))


def mg_restrict(gf, gc, nvf, nvc, xf):
    fc = np.zeros(nvc.shape)
    lib().or_mg_restrict(C.byref(gf), C.byref(gc), _p(nvf), _p(nvc), _p(xf), _p(fc))
    return fc


def mg_prolong(gf, gc, nvf, nvc, xc):
    ff = np.zeros(nvf.shape)
    lib().or_mg_prolong(C.byref(gf), C.byref(gc), _p(nvf), _p(nvc), _p(xc), _p(ff))
    return ff


def mg_nvert_restrict(gf, gc, nvf):
    nvc = np.zeros((gc.mz, gc.my, gc.mx))
    lib().or_mg_nvert_restrict(C.byref(gf), C.byref(gc), _p(nvf), _p(nvc))
    return nvc


def level_coupling(g, g9, nvert):
    """(largest |a_N|+|a_S| over |a_E|+|a_W|+|a_T|+|a_B|, fraction of rows below 1/10)
    of a multigrid level given by its DMDA face coefficients and nvert."""
    w, f = C.c_double(), C.c_double()
    lib().or_level_coupling(C.byref(g), _p(g9), _p(nvert), C.byref(w), C.byref(f))
    return w.value, f.value


def level_coarsen(g, g9, nvert, keep_y=False):
    """One coarsening step of the CPU multigrid: (coarse grid, DMDA face
    coefficients, nvert, coarsened-axes mask)."""
    gc = Grid()
    lib().or_level_coarsen(C.byref(g), _p(g9), _p(nvert), int(keep_y), C.byref(gc), None, None)
    g9c = np.zeros((9, gc.mz, gc.my, gc.mx))
    nvc = np.zeros((gc.mz, gc.my, gc.mx))
    cm = lib().or_level_coarsen(C.byref(g), _p(g9), _p(nvert), int(keep_y), C.byref(gc), _p(g9c), _p(nvc))
    return gc, g9c, nvc, cm


def gmres(rowptr, col, val, b, x0=None, m=51, rtol=1e-13, maxit=100000, precond=1):
    n = len(b)
    x = np.zeros(n) if x0 is None else x0.copy()
    rel = np.zeros(1)
    it = lib().or_gmres(n, _p(rowptr, C.c_int), _p(col, C.c_int), _p(val), _p(np.ascontiguousarray(b)), _p(x), m,
                        rtol, maxit, precond, _p(rel))
    return x, it, float(rel[0])


def num_threads():
    return lib().or_num_threads()


# ----------------------------------------------------------------- composite steps

class Assembled:
    """The reference's live system for one case: Gid numbering, face coefficients,
    row coefficients before/after the drop rule, CSR."""

    def __init__(self, case, drop=True):
        f = case.fields
        self.case = case
        self.g = grid(case.shape, case.per)
        self.nvert = np.ascontiguousarray(f["nvert"])
        self.gid, self.n = setup_gid(self.g, self.nvert)
        self.g9 = face_coeffs(self.g, f)
        self.vol, self.rowflag, self.cases = stencil(self.g, self.g9, self.nvert)
        self.rowptr, self.col, self.val = assemble_csr(self.g, self.vol, self.rowflag, self.gid, self.n, drop)
        fluid = self.gid >= 0
        inner = np.zeros_like(fluid)
        inner[1:-1, 1:-1, 1:-1] = True
        self.fluid_cells = fluid & inner  # DMDA points holding unknowns (gid order = C order)

    def to_compact(self, a):
        return np.ascontiguousarray(a[self.fluid_cells])


def reference_projection(case, rtol=1e-13, maxit=200000, precond=1, ucont=None, drop=False):
    """The reference's full pressure projection in FP64 on the CPU:
    VolumeFlux -> RHS -> GMRES solve -> Remove_Nullspace -> Convert_Phi2_Phi ->
    UpdatePressure -> Projection -> Contra2Cart.  Returns a dict of outputs.

    drop=False solves the conservative (undropped) operator, which the reference's
    projection is consistent with; drop=True solves the assembled HYPRE matrix
    (|v|<=1e-10 dropped), whose solution differs by ~1e-10 relative after mean
    removal and which is nearly singular, so tight solves on it stall."""
    A = Assembled(case, drop=drop)
    f = dict(case.fields)
    g = A.g
    u = np.ascontiguousarray((case.fields["ucont"] if ucont is None else ucont).copy())
    vf = volume_flux(g, case.immersed, u, f)
    b = rhs(g, u, A.nvert, case.dt, case.st, case.coeff)
    bc = A.to_compact(b)
    x, it, rel = gmres(A.rowptr, A.col, A.val, bc, m=51, rtol=rtol, maxit=maxit, precond=precond)
    remove_nullspace(g, A.nvert, A.gid, x)
    phi = convert_phi(g, A.nvert, x)
    p = np.zeros_like(phi)
    update_pressure(g, A.nvert, p, phi)
    u_flux = u.copy()
    projection(g, f, phi, u, case.dt, case.st, case.coeff)
    ucat = np.zeros_like(u)
    contra2cart(g, case.bctype, f, u, ucat)
    maxdiv = divergence(g, u, np.ascontiguousarray(f["aj"]), A.nvert)
    return dict(A=A, b=b, phi=phi, p=p, ucont=u, ucont_in=u_flux, ucat=ucat, maxdiv=maxdiv, iters=it, rel=rel,
                vflux=vf)


# ------------------------------------------------------------ CPU solver port

class CpuMG:
    """CPU port of the engine's multigrid-preconditioned FGMRES (FP64)."""

    def __init__(self, A: "Assembled", degree=3, lo=0.1, hi=1.1, coarse_max=512):
        self.A = A
        self.ptr = lib().or_mg_create(C.byref(A.g), _p(A.g9), _p(A.nvert), degree, lo, hi, coarse_max)
        self.levels = lib().or_mg_levels(self.ptr)
        c = A.case
        self.shape = (c.nz, c.ny, c.nx)

    def close(self):
        if self.ptr:
            lib().or_mg_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        self.close()

    def vcycle(self, r):
        z = np.zeros(self.shape)
        r = np.ascontiguousarray(r, dtype=np.float64)
        lib().or_mg_vcycle(self.ptr, _p(r), _p(z))
        return z

    def solve(self, b_int, x0=None, rtol=1e-10, maxit=500, restart=30):
        x = np.zeros(self.shape) if x0 is None else np.ascontiguousarray(x0.copy())
        rel = np.zeros(1)
        vc = np.zeros(1, np.int32)
        b = np.ascontiguousarray(b_int, dtype=np.float64)
        it = lib().or_mg_fgmres(self.ptr, _p(b), _p(x), restart, rtol, maxit, _p(rel), _p(vc, C.c_int))
        return x, it, float(rel[0])


def cpu_projection(case, A: "Assembled", mg: CpuMG, rtol=1e-10, ucont=None):
    """The full pressure projection on the CPU with the engine's solver port."""
    f = case.fields
    g = A.g
    u = np.ascontiguousarray((f["ucont"] if ucont is None else ucont).copy())
    vf = volume_flux(g, case.immersed, u, f)
    b = rhs(g, u, A.nvert, case.dt, case.st, case.coeff)
    x, it, rel = mg.solve(b[1:-1, 1:-1, 1:-1], rtol=rtol)
    phi = interior_to_phi(g, A.nvert, x)  # Remove_Nullspace + Convert_Phi2_Phi + ghost images
    p = np.zeros_like(phi)
    update_pressure(g, A.nvert, p, phi)
    projection(g, f, phi, u, case.dt, case.st, case.coeff)
    ucat = np.zeros_like(u)
    contra2cart(g, case.bctype, f, u, ucat)
    maxdiv = divergence(g, u, np.ascontiguousarray(f["aj"]), A.nvert)
    return dict(phi=phi, p=p, ucont=u, ucat=ucat, maxdiv=maxdiv, iters=it, rel=rel, vflux=vf)


METRIC_NAMES = ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "iaj", "jaj", "kaj", "aj",
                "csi", "eta", "zet")


def form_metrics(co, per):
    """or_form_metrics on DMDA node coordinates co (mz, my, mx, 3): dict of the 16
    metric arrays in the UserCtx layout."""
    co = np.ascontiguousarray(co, dtype=np.float64)
    mz, my, mx, _ = co.shape
    g = Grid(mx, my, mz, (C.c_int * 3)(*[int(bool(v)) for v in per]))
    out = {nm: np.zeros((mz, my, mx, 3) if (q < 9 or q >= 13) else (mz, my, mx)) for q, nm in
           enumerate(METRIC_NAMES)}
    ptrs = (C.POINTER(C.c_double) * 16)(*[_p(out[nm]) for nm in METRIC_NAMES])
    lib().or_form_metrics(C.byref(g), _p(co), ptrs)
    return out


def ib_search(co, per, bctype, verts, tris, nvert, body=0, classify=True):
    """or_ib_search: ibm_search_advanced for one triangulated body on DMDA node
    coordinates co; updates and returns nvert (a copy)."""
    co = np.ascontiguousarray(co, dtype=np.float64)
    g = grid(co.shape[:3], per)
    nv = np.ascontiguousarray(nvert, dtype=np.float64).copy()
    V = np.ascontiguousarray(verts, dtype=np.float64)
    T = np.ascontiguousarray(tris, dtype=np.int32)
    x, y, z = (np.ascontiguousarray(V[:, q]) for q in range(3))
    a, b, c = (np.ascontiguousarray(T[:, q]) for q in range(3))
    bc = np.ascontiguousarray(bctype, dtype=np.int32)
    rc = lib().or_ib_search(C.byref(g), _p(bc, C.c_int), _p(co), len(x), _p(x), _p(y), _p(z), len(a),
                            _p(a, C.c_int), _p(b, C.c_int), _p(c, C.c_int), int(body), int(bool(classify)), _p(nv))
    if rc:
        raise RuntimeError(f"or_ib_search status {rc}")
    return nv


def ray_s(seed):
    return lib().or_ray_s(int(seed))


def pressure_gradient(g, f, p):
    """or_pressure_gradient: (dP (mz,my,mx,3), p with periodic ghosts) from P and the
    face metrics in f."""
    pp = np.ascontiguousarray(p, dtype=np.float64).copy()
    dp = np.zeros(pp.shape + (3,))
    a = [np.ascontiguousarray(f[k]) for k in ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet",
                                              "iaj", "jaj", "kaj", "nvert")]
    lib().or_pressure_gradient(C.byref(g), *[_p(x) for x in a], _p(pp), _p(dp))
    return dp, pp


MOMENTUM_METRICS = ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet", "iaj", "jaj", "kaj",
                    "csi", "eta", "zet", "nvert")


def formfunction2(g, bctype, immersed, f, ucont, ucat, ren, scale=1.0):
    """or_formfunction2: the momentum right-hand side (mz, my, mx, 3) from Ucont,
    Ucat and the metrics in f (Formfunction_2, momentum.c:529-2034)."""
    a = [np.ascontiguousarray(f[k], dtype=np.float64) for k in MOMENTUM_METRICS]
    u = np.ascontiguousarray(ucont, dtype=np.float64)
    uc = np.ascontiguousarray(ucat, dtype=np.float64)
    rhs = np.zeros(u.shape)
    bc = np.ascontiguousarray(bctype, dtype=np.int32)
    lib().or_formfunction2(C.byref(g), _p(bc, C.c_int), int(immersed), *[_p(x) for x in a], _p(u), _p(uc),
                           float(ren), float(scale), _p(rhs))
    return rhs


def runge_kutta(g, bctype, immersed, f, dp, ucont, ucat, ren, dt, st):
    """or_runge_kutta: one explicit four-stage step (rhs.c:1511-1631); returns
    (Ucont, Ucat) after it, inputs untouched."""
    a = [np.ascontiguousarray(f[k], dtype=np.float64) for k in MOMENTUM_METRICS]
    u = np.ascontiguousarray(ucont, dtype=np.float64).copy()
    uc = np.ascontiguousarray(ucat, dtype=np.float64).copy()
    d = np.ascontiguousarray(dp, dtype=np.float64)
    bc = np.ascontiguousarray(bctype, dtype=np.int32)
    lib().or_runge_kutta(C.byref(g), _p(bc, C.c_int), int(immersed), *[_p(x) for x in a], _p(d), _p(u), _p(uc),
                         float(ren), float(dt), float(st))
    return u, uc


def momentum_bc_zero(g, bctype, u):
    """or_momentum_bc_zero: U with the wall components FormFunction_SNES zeroes
    (momentum.c:2323-2356) set to zero (a copy)."""
    uz = np.ascontiguousarray(u, dtype=np.float64).copy()
    bc = np.ascontiguousarray(bctype, dtype=np.int32)
    lib().or_momentum_bc_zero(C.byref(g), _p(bc, C.c_int), _p(uz))
    return uz


def momentum_ib_bc(g, bctype, immersed, nvert, u):
    """or_momentum_ib_bc: IB_BC (momentum.c:2037-2220) applied to a copy of the
    Ucont field u, for a body at rest."""
    ul = np.ascontiguousarray(u, dtype=np.float64).copy()
    bc = np.ascontiguousarray(bctype, dtype=np.int32)
    nv = np.ascontiguousarray(nvert, dtype=np.float64)
    lib().or_momentum_ib_bc(C.byref(g), _p(bc, C.c_int), int(immersed), _p(nv), _p(ul))
    return ul


def momentum_residual(g, bctype, f, dp, uo, u, ren, dt, coeff=1.0, urm1=None, rhs_o=None, immersed=0,
                      ucat=None):
    """or_momentum_residual: (F, Uz, Ucat) of the implicit momentum equation
    (FormFunction_SNES, momentum.c:2285-2440) at U; ucat (default zero) holds the
    Ucat that IB and solid cells keep."""
    a = [np.ascontiguousarray(f[k], dtype=np.float64) for k in MOMENTUM_METRICS]
    u = np.ascontiguousarray(u, dtype=np.float64)
    z = np.zeros(u.shape)
    uo = np.ascontiguousarray(uo, dtype=np.float64)
    urm1 = z if urm1 is None else np.ascontiguousarray(urm1, dtype=np.float64)
    rhs_o = z if rhs_o is None else np.ascontiguousarray(rhs_o, dtype=np.float64)
    dp = np.ascontiguousarray(dp, dtype=np.float64)
    uz, out = np.zeros(u.shape), np.zeros(u.shape)
    ucat = np.zeros(u.shape) if ucat is None else np.ascontiguousarray(ucat, dtype=np.float64).copy()
    bc = np.ascontiguousarray(bctype, dtype=np.int32)
    lib().or_momentum_residual(C.byref(g), _p(bc, C.c_int), int(immersed), *[_p(x) for x in a], _p(dp), _p(uo),
                               _p(urm1), _p(rhs_o), _p(u), float(ren), float(dt), float(coeff), _p(uz), _p(ucat),
                               _p(out))
    return out, uz, ucat

"""Workload generators: curvilinear grids, metrics, IB masks and divergent fluxes.

These produce the inputs the reference's pressure path receives from its own
setup code, in the reference's single-rank DMDA layout ([mz][my][mx], i fastest,
vector fields as AoS {x,y,z}):

* node coordinates for the channel / hill / tower-farm grids named by the
  benchmark configurations,
* cell metrics csi/eta/zet and aj, face metrics i/j/k{csi,eta,zet} and i/j/kaj,
  restating FormMetrics with NEWMETRIC (metrics.c:80-99, 183-371, 374-509,
  613-1004): area vectors from face-centre differences, one-sided at the
  non-periodic ends and the two-sided average at a periodic seam,
* the IB mask nvert in {0,1,3} from an analytic inside test on cell centres
  followed by the reference's dilation and remap rules (ibm.c:331-353, 540-547),
* a seeded contravariant flux field Ucont: the discrete curl of a random edge
  vector potential (exactly divergence-free, zero wall flux) plus seeded white
  noise on interior faces.

This is host-side set-up tooling (the caller's side of the boundary); the
engine itself receives these arrays through the C ABI.
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional, Tuple

import numpy as np


@dataclasses.dataclass
class Case:
    """One pressure-projection system in reference layout."""

    name: str
    nx: int
    ny: int
    nz: int
    per: Tuple[int, int, int]
    bctype: Tuple[int, int, int, int, int, int]
    immersed: int
    fields: dict  # icsi..kzet (mz,my,mx,3), iaj,jaj,kaj,aj (mz,my,mx), csi,eta,zet, nvert, ucont
    dt: float = 0.04  # init.c:58
    st: float = 1.0  # init.c:76
    coeff: float = 1.0  # time_coeff() with rans=0, poisson.c:22-30
    coor: Optional[np.ndarray] = None  # DMDA node coordinates (mz,my,mx,3), kept on request

    @property
    def shape(self):
        return (self.nz + 2, self.ny + 2, self.nx + 2)

    @property
    def ncell(self):
        return self.nx * self.ny * self.nz


# ----------------------------------------------------------------------------- grids

def tanh_wall_nodes(n: int, beta: float) -> np.ndarray:
    """y_j = -tanh(beta (1 - 2j/n)) / tanh(beta) on [-1, 1]; uniform when beta == 0."""
    s = np.arange(n + 1, dtype=np.float64) / n
    if beta == 0.0:
        return 2.0 * s - 1.0
    return -np.tanh(beta * (1.0 - 2.0 * s)) / math.tanh(beta)


def geometric_nodes(n: int, ratio: float, length: float) -> np.ndarray:
    """Nodes 0..length with spacing growing by `ratio` per cell."""
    if ratio == 1.0:
        return np.linspace(0.0, length, n + 1)
    d0 = length * (ratio - 1.0) / (ratio ** n - 1.0)
    d = d0 * ratio ** np.arange(n)
    y = np.concatenate([[0.0], np.cumsum(d)])
    y[-1] = length
    return y


def make_nodes(nx, ny, nz, x, eta, z, hill_h=0.0, hill_s=0.5, top=2.0):
    """Node array N[k][j][i] (nz+1, ny+1, nx+1, 3).

    x, z: 1-D node positions; eta: 1-D wall-normal parameter in [0,1].
    y = y0(x,z) + eta (top - y0(x,z)), y0 = h exp(-(x^2+z^2)/(2 s^2)).
    """
    X = np.broadcast_to(x[None, None, :], (nz + 1, ny + 1, nx + 1))
    Z = np.broadcast_to(z[:, None, None], (nz + 1, ny + 1, nx + 1))
    y0 = hill_h * np.exp(-(x[None, :] ** 2 + z[:, None] ** 2) / (2.0 * hill_s ** 2))  # (nz+1, nx+1)
    Y = y0[:, None, :] + eta[None, :, None] * (top - y0[:, None, :])
    N = np.empty((nz + 1, ny + 1, nx + 1, 3))
    N[..., 0] = X
    N[..., 1] = Y
    N[..., 2] = Z
    return N


def dmda_coords(N: np.ndarray) -> np.ndarray:
    """coor[k][j][i] = node (i,j,k) for i<=mx-2; the extra plane is extrapolated."""
    nz1, ny1, nx1, _ = N.shape
    co = np.empty((nz1 + 1, ny1 + 1, nx1 + 1, 3))
    co[:nz1, :ny1, :nx1] = N
    co[:nz1, :ny1, nx1] = 2 * N[:, :, -1] - N[:, :, -2]
    co[:nz1, ny1, :] = 2 * co[:nz1, ny1 - 1, :] - co[:nz1, ny1 - 2, :]
    co[nz1, :, :] = 2 * co[nz1 - 1] - co[nz1 - 2]
    return co


# --------------------------------------------------------------------------- metrics

def _cross_metrics(dc, de, dz):
    """(csi, eta, zet, 1/J) from the three coordinate derivative vectors."""
    dxdc, dydc, dzdc = dc[..., 0], dc[..., 1], dc[..., 2]
    dxde, dyde, dzde = de[..., 0], de[..., 1], de[..., 2]
    dxdz, dydz, dzdz = dz[..., 0], dz[..., 1], dz[..., 2]
    csi = np.stack([dyde * dzdz - dzde * dydz, -dxde * dzdz + dzde * dxdz, dxde * dydz - dyde * dxdz], -1)
    eta = np.stack([dydz * dzdc - dzdz * dydc, -dxdz * dzdc + dzdz * dxdc, dxdz * dydc - dydz * dxdc], -1)
    zet = np.stack([dydc * dzde - dzdc * dyde, -dxdc * dzde + dzdc * dxde, dxdc * dyde - dydc * dxde], -1)
    det = dxdc * (dyde * dzdz - dzde * dydz) - dydc * (dxde * dzdz - dzde * dxdz) + dzdc * (dxde * dydz - dyde * dxdz)
    with np.errstate(divide="ignore"):  # the extrapolated outer plane is not used
        return csi, eta, zet, 1.0 / det


def _diff(c, axis, first, last, periodic_seam):
    """Derivative of face-centre array c along `axis` at indices first..last (inclusive):
    forward at `first`, backward at `last` (two-sided seam average when periodic_seam),
    central inside (metrics.c:652-700)."""
    c = np.moveaxis(c, axis, 0)
    out = np.zeros_like(c)
    out[first + 1:last] = 0.5 * (c[first + 2:last + 1] - c[first:last - 1])
    out[first] = c[first + 1] - c[first]
    if periodic_seam:
        out[last] = 0.5 * (c[1] - c[0] + c[last] - c[last - 1])
    else:
        out[last] = c[last] - c[last - 1]
    return np.moveaxis(out, 0, axis)


def form_metrics(co: np.ndarray, per) -> dict:
    """Face and cell metrics on the DMDA grid from ghosted node coordinates co."""
    mz, my, mx, _ = co.shape
    f = {}
    # ---- face metrics -----------------------------------------------------------
    # i faces: face centre of node plane i for cell column (j,k), metrics.c:613-724
    cx = np.zeros_like(co)
    cx[1:, 1:, :] = 0.25 * (co[1:, 1:, :] + co[:-1, 1:, :] + co[1:, :-1, :] + co[:-1, :-1, :])
    dc = _diff(cx, 2, 0, mx - 2, per[0])
    de = _diff(cx, 1, 1, my - 2, False)
    dz = _diff(cx, 0, 1, mz - 2, False)
    csi, eta, zet, aj = _cross_metrics(dc, de, dz)
    m = np.zeros((mz, my, mx), bool)
    m[1:mz - 1, 1:my - 1, 0:mx - 1] = True
    for nm, v in (("icsi", csi), ("ieta", eta), ("izet", zet)):
        f[nm] = np.where(m[..., None], v, 0.0)
    f["iaj"] = np.where(m, aj, 0.0)
    # j faces, metrics.c:739-846
    cy = np.zeros_like(co)
    cy[1:, :, 1:] = 0.25 * (co[1:, :, 1:] + co[:-1, :, 1:] + co[1:, :, :-1] + co[:-1, :, :-1])
    dc = _diff(cy, 2, 1, mx - 2, False)
    de = _diff(cy, 1, 0, my - 2, per[1])
    dz = _diff(cy, 0, 1, mz - 2, False)
    csi, eta, zet, aj = _cross_metrics(dc, de, dz)
    m[:] = False
    m[1:mz - 1, 0:my - 1, 1:mx - 1] = True
    for nm, v in (("jcsi", csi), ("jeta", eta), ("jzet", zet)):
        f[nm] = np.where(m[..., None], v, 0.0)
    f["jaj"] = np.where(m, aj, 0.0)
    # k faces, metrics.c:859-1004
    cz = np.zeros_like(co)
    cz[:, 1:, 1:] = 0.25 * (co[:, 1:, 1:] + co[:, :-1, 1:] + co[:, 1:, :-1] + co[:, :-1, :-1])
    dc = _diff(cz, 2, 1, mx - 2, False)
    de = _diff(cz, 1, 1, my - 2, False)
    dz = _diff(cz, 0, 0, mz - 2, per[2])
    csi, eta, zet, aj = _cross_metrics(dc, de, dz)
    m[:] = False
    m[0:mz - 1, 1:my - 1, 1:mx - 1] = True
    for nm, v in (("kcsi", csi), ("keta", eta), ("kzet", zet)):
        f[nm] = np.where(m[..., None], v, 0.0)
    f["kaj"] = np.where(m, aj, 0.0)
    del cx, cy, cz, dc, de, dz, csi, eta, zet, aj, m

    # ---- cell metrics (NEWMETRIC): average of node-plane area vectors, metrics.c:185-296.
    # Each derivative is 0.5*(a + b - c - d) and each average 0.5*x + 0.5*y (AxByC,
    # momentum.c:35-40), in the reference's operand order.
    def d4(a, b, c, d):
        return 0.5 * (a + b - c - d)

    def cross(u, v):
        return np.stack([u[..., 1] * v[..., 2] - u[..., 2] * v[..., 1],
                         -u[..., 0] * v[..., 2] + u[..., 2] * v[..., 0],
                         u[..., 0] * v[..., 1] - u[..., 1] * v[..., 0]], -1)

    # csitmp of node plane i for cell column (j,k): index [k-1, j-1, i]
    dxde = d4(co[1:, 1:, :], co[:-1, 1:, :], co[1:, :-1, :], co[:-1, :-1, :])
    dxdz = d4(co[1:, :-1, :], co[1:, 1:, :], co[:-1, :-1, :], co[:-1, 1:, :])
    csit = cross(dxde, dxdz)  # (mz-1,my-1,mx,3)
    # etatmp of node plane j for cell row (i,k): index [k-1, j, i-1]
    dxdc = d4(co[1:, :, 1:], co[:-1, :, 1:], co[1:, :, :-1], co[:-1, :, :-1])
    dxdz = d4(co[1:, :, 1:], co[1:, :, :-1], co[:-1, :, 1:], co[:-1, :, :-1])
    etat = cross(dxdz, dxdc)  # (mz-1,my,mx-1,3)
    # zettmp of node plane k for cell column (i,j): index [k, j-1, i-1]
    dxdc = d4(co[:, 1:, 1:], co[:, :-1, 1:], co[:, 1:, :-1], co[:, :-1, :-1])
    dxde = d4(co[:, 1:, 1:], co[:, 1:, :-1], co[:, :-1, 1:], co[:, :-1, :-1])
    zett = cross(dxdc, dxde)  # (mz,my-1,mx-1,3)
    I = slice(1, mx - 1)
    J = slice(1, my - 1)
    K = slice(1, mz - 1)
    csi = np.zeros((mz, my, mx, 3))
    eta = np.zeros_like(csi)
    zet = np.zeros_like(csi)
    csi[K, J, I] = 0.5 * csit[0:mz - 2, 0:my - 2, 1:mx - 1] + 0.5 * csit[0:mz - 2, 0:my - 2, 0:mx - 2]
    eta[K, J, I] = 0.5 * etat[0:mz - 2, 1:my - 1, 0:mx - 2] + 0.5 * etat[0:mz - 2, 0:my - 2, 0:mx - 2]
    zet[K, J, I] = 0.5 * zett[1:mz - 1, 0:my - 2, 0:mx - 2] + 0.5 * zett[0:mz - 2, 0:my - 2, 0:mx - 2]
    del csit, etat, zett, dxdc, dxde, dxdz
    # cell Jacobian from the 8 corner nodes, metrics.c:305-371
    c8 = [co[1:, 1:, 1:], co[1:, :-1, 1:], co[:-1, 1:, 1:], co[:-1, :-1, 1:],
          co[1:, 1:, :-1], co[1:, :-1, :-1], co[:-1, 1:, :-1], co[:-1, :-1, :-1]]
    dC = 0.25 * (c8[0] + c8[1] + c8[2] + c8[3] - c8[4] - c8[5] - c8[6] - c8[7])
    dE = 0.25 * (c8[0] + c8[4] + c8[2] + c8[6] - c8[1] - c8[5] - c8[3] - c8[7])
    dZ = 0.25 * (c8[0] + c8[1] + c8[4] + c8[5] - c8[2] - c8[3] - c8[6] - c8[7])
    _, _, _, ajc = _cross_metrics(dC, dE, dZ)
    aj = np.zeros((mz, my, mx))
    aj[1:, 1:, 1:] = ajc
    aj[:, :, 0] = 0
    del c8, dC, dE, dZ, ajc
    # mirror to boundary planes (metrics.c:374-454), then periodic aliasing (477-509)
    for arr in (csi, eta, zet, aj):
        arr[:, :, 0] = arr[:, :, 1]
        arr[:, :, mx - 1] = arr[:, :, mx - 2]
        arr[:, 0, :] = arr[:, 1, :]
        arr[:, my - 1, :] = arr[:, my - 2, :]
        arr[0, :, :] = arr[1, :, :]
        arr[mz - 1, :, :] = arr[mz - 2, :, :]
        periodic_fill(arr, per)
    f["csi"], f["eta"], f["zet"], f["aj"] = csi, eta, zet, aj
    return f


def periodic_fill(a: np.ndarray, per) -> None:
    """In-place ghost-plane copies: plane 0 <- m-2, plane m-1 <- 1 (poisson.c:153-174)."""
    mz, my, mx = a.shape[:3]
    if per[0]:
        a[:, :, 0] = a[:, :, mx - 2]
        a[:, :, mx - 1] = a[:, :, 1]
    if per[1]:
        a[:, 0, :] = a[:, my - 2, :]
        a[:, my - 1, :] = a[:, 1, :]
    if per[2]:
        a[0, :, :] = a[mz - 2, :, :]
        a[mz - 1, :, :] = a[1, :, :]


def cell_centres(co: np.ndarray) -> np.ndarray:
    """cent[k][j][i] for interior cells (metrics.c:80-99); zeros elsewhere."""
    mz, my, mx, _ = co.shape
    c = np.zeros((mz, my, mx, 3))
    c[1:, 1:, 1:] = 0.125 * (co[1:, 1:, 1:] + co[1:, :-1, 1:] + co[:-1, 1:, 1:] + co[:-1, :-1, 1:] +
                             co[1:, 1:, :-1] + co[1:, :-1, :-1] + co[:-1, 1:, :-1] + co[:-1, :-1, :-1])
    c[:, :, mx - 1] = 0
    c[:, my - 1, :] = 0
    c[mz - 1, :, :] = 0
    return c


# ------------------------------------------------------------------------------ IB

def classify_nvert(inside: np.ndarray, bctype=(0, 0, 0, 0, 0, 0)) -> np.ndarray:
    """nvert in {0,1,3} from a boolean inside mask over the DMDA array.

    Dilation over the clamped 27-neighbourhood marks near-boundary cells 2
    (ibm.c:331-353); the interior remap turns 2 into 1 and 4 into 3
    (ibm.c:540-547); wall-model face layers become 1 (ibm.c:550-568).
    """
    nv = np.where(inside, 4.0, 0.0)
    four = inside.copy()
    near = np.zeros_like(four)
    mz, my, mx = nv.shape
    pad = np.pad(four, 1, mode="edge")  # clamped neighbourhood
    for dk in range(3):
        for dj in range(3):
            for di in range(3):
                near |= pad[dk:dk + mz, dj:dj + my, di:di + mx]
    nv = np.where(~four & near, np.maximum(2.0, nv), nv)
    inner = (slice(1, mz - 1), slice(1, my - 1), slice(1, mx - 1))
    sub = nv[inner]
    sub[sub == 2.0] = 1.0
    sub[sub == 4.0] = 3.0
    nv[inner] = sub
    for f, (ax, idx) in enumerate(((2, 1), (2, mx - 2), (1, 1), (1, my - 2), (0, 1), (0, mz - 2))):
        if bctype[f] in (-1, -2):
            sl = [slice(1, mz - 1), slice(1, my - 1), slice(1, mx - 1)]
            sl[ax] = idx
            nv[tuple(sl)] = 1.0
    return nv


def inside_sphere(cent, centre, r):
    d = cent - np.asarray(centre)[None, None, None, :]
    return (d ** 2).sum(-1) < r * r


def inside_towers(cent, xs, zs, radius, height):
    m = np.zeros(cent.shape[:3], bool)
    for xc in xs:
        for zc in zs:
            m |= ((cent[..., 0] - xc) ** 2 + (cent[..., 2] - zc) ** 2 < radius ** 2) & (cent[..., 1] < height)
    return m


# ------------------------------------------------------- triangulated surfaces

def icosphere(centre, r, level=3):
    """Closed triangulated sphere (icosahedron subdivided `level` times), outward
    orientation: (vertices (n, 3), triangles (m, 3) 0-based)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    V = [np.array(v, float) / np.linalg.norm(v) for v in V]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        mid = {}

        def m(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                v = V[a] + V[b]
                V.append(v / np.linalg.norm(v))
                mid[key] = len(V) - 1
            return mid[key]

        F = [f for a, b, c in F for f in ((a, m(a, b), m(c, a)), (b, m(b, c), m(a, b)), (c, m(c, a), m(b, c)),
                                           (m(a, b), m(b, c), m(c, a)))]
    return np.asarray(centre, float)[None, :] + r * np.array(V), np.array(F, dtype=np.int32)


def cylinder_surface(xc, zc, r, y0, y1, nseg=24, nlayers=4):
    """Closed cylinder about the y axis through (xc, zc), from y0 to y1, outward
    orientation."""
    a = 2 * np.pi * np.arange(nseg) / nseg
    ys = np.linspace(y0, y1, nlayers + 1)
    V = [(xc + r * np.cos(q), y, zc + r * np.sin(q)) for y in ys for q in a]
    V += [(xc, y0, zc), (xc, y1, zc)]
    bot, top = len(V) - 2, len(V) - 1
    F = []
    for L in range(nlayers):
        for q in range(nseg):
            p0, p1 = L * nseg + q, L * nseg + (q + 1) % nseg
            F += [(p0, p0 + nseg, p1), (p1, p0 + nseg, p1 + nseg)]
    for q in range(nseg):
        F.append((bot, q, (q + 1) % nseg))
        F.append((top, nlayers * nseg + (q + 1) % nseg, nlayers * nseg + q))
    return np.array(V, float), np.array(F, dtype=np.int32)


def write_ucd(path, verts, tris, title="vfs_pp surface"):
    """ibmdata file as ibm_io.c:106-240 reads it (1-based triangle indices)."""
    with open(path, "w") as f:
        f.write(f"# {title}\n#\n#\n{len(verts)} {len(tris)} 0 0 0\n")
        for q, v in enumerate(verts):
            f.write(f"{q + 1} {float(v[0])!r} {float(v[1])!r} {float(v[2])!r}\n")
        for q, t in enumerate(tris):
            f.write(f"{q + 1} 0 tri {int(t[0]) + 1} {int(t[1]) + 1} {int(t[2]) + 1}\n")


def interior_only(mask):
    m = np.zeros_like(mask)
    m[1:-1, 1:-1, 1:-1] = mask[1:-1, 1:-1, 1:-1]
    return m


# ------------------------------------------------------------------------- fluxes

def _periodic_field(rng, nz, ny1, nx, kmax, ybasis):
    """Random smooth field on (nz, ny1, nx): Fourier in i and k (|k|<=kmax),
    the given basis functions along j."""
    nky = ybasis.shape[0]
    out = np.zeros((nz, ny1, nx))
    kx = np.fft.fftfreq(nx, 1.0 / nx)
    kz = np.fft.fftfreq(nz, 1.0 / nz)
    for ky in range(nky):
        spec = np.zeros((nz, nx), complex)
        KZ, KX = np.meshgrid(kz, kx, indexing="ij")
        keep = (KX ** 2 + KZ ** 2 + (ky + 1) ** 2) <= kmax * kmax
        spec[keep] = rng.standard_normal(keep.sum()) + 1j * rng.standard_normal(keep.sum())
        plane = np.fft.ifft2(spec).real * nx * nz
        out += plane[:, None, :] * ybasis[ky][None, :, None]
    return out


def solenoidal_flux(nx, ny, nz, rng, kmax=8):
    """Contravariant fluxes (3 arrays in DMDA layout) = discrete curl of a random edge
    potential; periodic in i and k, zero flux through the j walls.  Exactly
    divergence-free in index space."""
    j = np.arange(ny + 1)
    nky = kmax
    sinb = np.stack([np.sin((m + 1) * np.pi * j / ny) for m in range(nky)])  # zero at both walls
    cosb = np.stack([np.cos((m + 1) * np.pi * (j - 0.5) / ny) for m in range(nky)])
    px = _periodic_field(rng, nz, ny + 1, nx, kmax, sinb)
    py = _periodic_field(rng, nz, ny + 1, nx, kmax, cosb)
    pz = _periodic_field(rng, nz, ny + 1, nx, kmax, sinb)
    mz, my, mx = nz + 2, ny + 2, nx + 2
    U = np.zeros((mz, my, mx, 3))
    ii = np.arange(mx - 1) % nx  # node planes 0..nx
    kk = np.arange(mz - 1) % nz
    # U.x at node plane i, column (j,k): psi_z(j) - psi_z(j-1) - psi_y(k) + psi_y(k-1)
    K = np.arange(1, mz - 1)
    Jj = np.arange(1, my - 1)
    PZ = pz[np.ix_(kk[K], np.arange(ny + 1), ii)]
    PY_k = py[np.ix_(kk[K], np.arange(ny + 1), ii)]
    PY_km = py[np.ix_(kk[K - 1], np.arange(ny + 1), ii)]
    U[1:mz - 1, 1:my - 1, 0:mx - 1, 0] = (PZ[:, Jj] - PZ[:, Jj - 1]) - (PY_k[:, Jj] - PY_km[:, Jj])
    # U.y at node plane j, column (i,k): psi_x(k) - psi_x(k-1) - psi_z(i) + psi_z(i-1)
    I = np.arange(1, mx - 1)
    PX_k = px[np.ix_(kk[K], np.arange(ny + 1), ii[I])]
    PX_km = px[np.ix_(kk[K - 1], np.arange(ny + 1), ii[I])]
    PZ_i = pz[np.ix_(kk[K], np.arange(ny + 1), ii[I])]
    PZ_im = pz[np.ix_(kk[K], np.arange(ny + 1), ii[I - 1])]
    U[1:mz - 1, 0:my - 1, 1:mx - 1, 1] = (PX_k - PX_km) - (PZ_i - PZ_im)
    # U.z at node plane k, column (i,j): psi_y(i) - psi_y(i-1) - psi_x(j) + psi_x(j-1)
    Kn = np.arange(0, mz - 1)
    PY_i = py[np.ix_(kk[Kn], Jj, ii[I])]
    PY_im = py[np.ix_(kk[Kn], Jj, ii[I - 1])]
    PX_j = px[np.ix_(kk[Kn], Jj, ii[I])]
    PX_jm = px[np.ix_(kk[Kn], Jj - 1, ii[I])]
    U[0:mz - 1, 1:my - 1, 1:mx - 1, 2] = (PY_i - PY_im) - (PX_j - PX_jm)
    return U


def contra2cart_np(U, csi, eta, zet):
    """Cartesian velocity of interior cells by the Cramer solve of rhs.c:144-189."""
    q0 = 0.5 * (U[1:-1, 1:-1, :-2, 0] + U[1:-1, 1:-1, 1:-1, 0])
    q1 = 0.5 * (U[1:-1, :-2, 1:-1, 1] + U[1:-1, 1:-1, 1:-1, 1])
    q2 = 0.5 * (U[:-2, 1:-1, 1:-1, 2] + U[1:-1, 1:-1, 1:-1, 2])
    M = np.stack([csi[1:-1, 1:-1, 1:-1], eta[1:-1, 1:-1, 1:-1], zet[1:-1, 1:-1, 1:-1]], -2)
    q = np.stack([q0, q1, q2], -1)
    return np.linalg.solve(M, q[..., None])[..., 0]


def set_periodic_flux_ghosts(U, per):
    """Per-component periodic ghost faces as Projection writes them (poisson.c:2401-2441)."""
    mz, my, mx = U.shape[:3]
    if per[0]:
        U[:, :, 0, 0] = U[:, :, mx - 2, 0]
        U[:, :, mx - 1, 0] = U[:, :, 1, 0]
    if per[1]:
        U[:, 0, :, 1] = U[:, my - 2, :, 1]
        U[:, my - 1, :, 1] = U[:, 1, :, 1]
    if per[2]:
        U[0, :, :, 2] = U[mz - 2, :, :, 2]
        U[mz - 1, :, :, 2] = U[1, :, :, 2]


def divergent_ucont(nx, ny, nz, f, per, seed, noise=0.05, mean_profile=None, fluct=1.0):
    """Seeded solenoidal field scaled to unit RMS Cartesian speed per component, plus
    `noise` x RMS face flux of Gaussian white noise on interior faces."""
    rng = np.random.default_rng(seed)
    U = solenoidal_flux(nx, ny, nz, rng)
    uc = contra2cart_np(U, f["csi"], f["eta"], f["zet"])
    rms = math.sqrt(float(np.mean(uc ** 2)))
    U *= fluct / rms
    if mean_profile is not None:
        U[..., 0] += mean_profile * f["icsi"][..., 0]
    mz, my, mx = nz + 2, ny + 2, nx + 2
    rng2 = np.random.default_rng(seed + 1)
    sig = noise * math.sqrt(float(np.mean(U[1:-1, 1:-1, 1:-1] ** 2)))
    nzn = rng2.standard_normal(U.shape) * sig
    # interior faces only: i faces at node planes 1..nx (seam included when periodic)
    msk = np.zeros(U.shape, bool)
    msk[1:mz - 1, 1:my - 1, 1:mx - 1, 0] = True
    if not per[0]:
        msk[:, :, mx - 2, 0] = False
    msk[1:mz - 1, 1:my - 2, 1:mx - 1, 1] = True
    if per[1]:
        msk[1:mz - 1, my - 2, 1:mx - 1, 1] = True
    msk[1:mz - 1, 1:my - 1, 1:mx - 1, 2] = True
    if not per[2]:
        msk[mz - 2, :, :, 2] = False
    U = U + np.where(msk, nzn, 0.0)
    set_periodic_flux_ghosts(U, per)
    return U


# ------------------------------------------------------------------------- configs

def build_case(name, nx, ny, nz, x, eta, z, *, hill_h=0.0, hill_s=0.5, top=2.0,
               ib=None, seed=0, noise=0.05, mean_profile_fn=None, fluct=1.0, keep_coords=False, metrics_fn=None):
    """metrics_fn(co, per) -> the 16 metric arrays; form_metrics by default (any
    function bit-identical to it, such as FormMetrics on the device, gives the same
    case)."""
    per = (1, 0, 1)
    bct = (0, 0, 1, 1, 0, 0)  # no-slip walls in j; periodic i and k
    N = make_nodes(nx, ny, nz, x, eta, z, hill_h=hill_h, hill_s=hill_s, top=top)
    co = dmda_coords(N)
    del N
    f = (metrics_fn or form_metrics)(co, per)
    cent = cell_centres(co)
    if ib is not None:
        inside = interior_only(ib(cent))
        nv = classify_nvert(inside, bct)
        periodic_fill(nv, per)
        immersed = 1
    else:
        nv = np.zeros((nz + 2, ny + 2, nx + 2))
        immersed = 0
    f["nvert"] = nv
    mp = None
    if mean_profile_fn is not None:
        # the mean streamwise flux is taken from the first interior cell column so
        # that it is uniform along i (divergence-free on a streamwise-uniform grid)
        mp = np.broadcast_to(mean_profile_fn(cent[:, :, 1:2, 1]), cent.shape[:3])
    f["ucont"] = divergent_ucont(nx, ny, nz, f, per, seed, noise=noise, mean_profile=mp, fluct=fluct)
    del cent
    return Case(name=name, nx=nx, ny=ny, nz=nz, per=per, bctype=bct, immersed=immersed, fields=f,
                coor=co if keep_coords else None)


def channel_case(nx=64, ny=64, nz=64, beta=2.0, seed=0, noise=0.05, keep_coords=False, metrics_fn=None):
    """Periodic channel 2pi x 2 x pi, tanh wall stretching in y, no IB."""
    x = np.arange(nx + 1) * (2 * np.pi / nx)
    z = np.arange(nz + 1) * (np.pi / nz)
    eta = 0.5 * (tanh_wall_nodes(ny, beta) + 1.0)
    return build_case(f"channel{nx}x{ny}x{nz}", nx, ny, nz, x, eta, z, top=2.0, seed=seed, noise=noise,
                      keep_coords=keep_coords, metrics_fn=metrics_fn)


def hill_case(nx=192, ny=96, nz=128, h=0.3, s=0.5, beta=0.0, sphere=(0.6, 0.75, 0.0, 0.15), seed=0,
              noise=0.05, keep_coords=False, metrics_fn=None):
    """Gaussian hill under algebraic grid-line blending, optional sphere IB."""
    x = -3.0 + np.arange(nx + 1) * (6.0 / nx)
    z = -2.0 + np.arange(nz + 1) * (4.0 / nz)
    eta = 0.5 * (tanh_wall_nodes(ny, beta) + 1.0)
    ib = None
    if sphere is not None:
        cx, cy, cz, r = sphere
        ib = lambda cent: inside_sphere(cent, (cx, cy, cz), r)
    return build_case(f"hill{nx}x{ny}x{nz}", nx, ny, nz, x, eta, z, hill_h=h, hill_s=s, top=2.0, ib=ib,
                      seed=seed, noise=noise, keep_coords=keep_coords, metrics_fn=metrics_fn)


def towers_case(nx=512, ny=256, nz=256, ratio=1.02, rows=3, cols=4, radius_cells=4, height=0.6, seed=0,
                noise=0.05, keep_coords=False, metrics_fn=None):
    """Wind-farm-like box: uniform in x and z, geometric stretching in y, an array of
    cylindrical towers, log-law mean flow with 10% fluctuations."""
    Lz = 1.0
    dx = Lz / nz
    Lx = dx * nx
    x = np.arange(nx + 1) * dx
    z = np.arange(nz + 1) * dx
    eta = geometric_nodes(ny, ratio, 1.0)
    r = radius_cells * dx
    xs = [Lx * (c + 0.5) / cols for c in range(cols)]
    zs = [Lz * (q + 0.5) / rows for q in range(rows)]
    ib = lambda cent: inside_towers(cent, xs, zs, r, height)
    y0 = 1e-3
    norm = math.log((1.0 + y0) / y0)
    prof = lambda yc: np.log((np.maximum(yc, 0.0) + y0) / y0) / norm
    return build_case(f"towers{nx}x{ny}x{nz}", nx, ny, nz, x, eta, z, top=1.0, ib=ib, seed=seed,
                      noise=noise, mean_profile_fn=prof, fluct=0.1, keep_coords=keep_coords, metrics_fn=metrics_fn)


def batch_sizes(count=64, lo=32, hi=128, jitter=0.25, seed=3):
    """Seeded log-uniform base sizes with per-axis jitter (rounded, >= 8)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        n0 = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        dims = [max(8, int(round(n0 * (1.0 + rng.uniform(-jitter, jitter))))) for _ in range(3)]
        out.append((tuple(dims), float(rng.uniform(1.0, 3.0)), float(rng.uniform(0.0, 0.4)), bool(rng.integers(2)),
                    int(rng.integers(1 << 30))))
    return out


def ensemble_case(dims, beta, h, with_sphere, seed, metrics_fn=None):
    nx, ny, nz = dims
    x = -3.0 + np.arange(nx + 1) * (6.0 / nx)
    z = -2.0 + np.arange(nz + 1) * (4.0 / nz)
    eta = 0.5 * (tanh_wall_nodes(ny, beta) + 1.0)
    ib = None
    if with_sphere:
        rng = np.random.default_rng(seed + 7)
        c = (rng.uniform(-2.0, 2.0), rng.uniform(0.8, 1.4), rng.uniform(-1.2, 1.2))
        ib = lambda cent: inside_sphere(cent, c, 0.15)
    return build_case(f"ens{nx}x{ny}x{nz}", nx, ny, nz, x, eta, z, hill_h=h, top=2.0, ib=ib, seed=seed,
                      metrics_fn=metrics_fn)


FIELD_ORDER_VEC = ("icsi", "ieta", "izet", "jcsi", "jeta", "jzet", "kcsi", "keta", "kzet")
FIELD_ORDER_SCA = ("iaj", "jaj", "kaj", "aj")


def served_case(kind, dims, seed, metrics_fn=None):
    """One case of the served stream: 'channel' (tanh-stretched periodic channel,
    no IB), 'hill' (terrain + IB sphere) or 'towers' (stretched box with an IB
    tower array), at the given interior extents."""
    nx, ny, nz = dims
    rng = np.random.default_rng(seed)
    if kind == "channel":
        c = channel_case(nx, ny, nz, beta=float(rng.uniform(1.0, 3.0)), seed=seed, metrics_fn=metrics_fn)
    elif kind == "hill":
        c = hill_case(nx, ny, nz, h=float(rng.uniform(0.1, 0.4)), s=0.5, beta=float(rng.uniform(0.0, 2.0)),
                      sphere=(float(rng.uniform(-1.5, 1.5)), float(rng.uniform(0.7, 1.3)), float(rng.uniform(-1.0, 1.0)),
                              0.25), seed=seed, metrics_fn=metrics_fn)
    elif kind == "towers":
        c = towers_case(nx, ny, nz, ratio=1.0 + 4.0 / ny, rows=2, cols=3, radius_cells=max(2, nx // 48), height=0.6,
                        seed=seed, metrics_fn=metrics_fn)
    else:
        raise ValueError(kind)
    c.name = f"{kind}{nx}x{ny}x{nz}"
    return c


def served_specs(count=16, lo=24, hi=256, seed=5):
    """Seeded (kind, dims, seed) specs: sizes log-uniform in [lo^3, hi^3] with
    +-15% per axis, kinds 60% channel / 30% hill / 10% towers."""
    rng = np.random.default_rng(seed)
    out = []
    for q in range(count):
        n0 = math.exp(rng.uniform(math.log(lo), math.log(hi)))
        dims = tuple(max(12, int(round(n0 * (1.0 + rng.uniform(-0.15, 0.15))))) for _ in range(3))
        u = rng.uniform()
        kind = "channel" if u < 0.6 else ("hill" if u < 0.9 else "towers")
        out.append((kind, dims, int(rng.integers(1 << 30))))
    return out

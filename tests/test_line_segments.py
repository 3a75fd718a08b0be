"""The segmented y-line sweeps (line_seg_sweeps in vfs_v3p2_amd/csrc/pp_mg_dev.hpp):
both Thomas sweeps are affine recurrences, so a line split into S segments, each
swept from a zero carry with the product of its multipliers kept, chained through
the segments' carries and re-swept exactly from its carry, solves the same
tridiagonal system as the sequential sweep.  Checked here in FP64 on diagonally
dominant systems of 1 to 256 rows for the two layouts the kernels use (16
segments of up to 16 rows, 32 of up to 8), including lines shorter than the
segment count (empty segments carry the identity)."""
import numpy as np
import pytest


def thomas_factors(a, b, c):
    ny = len(b)
    cp, inv = np.zeros(ny), np.zeros(ny)
    prev = 0.0
    for j in range(ny):
        den = b[j] - a[j] * prev
        inv[j] = 1.0 / den
        prev = c[j] * inv[j]
        cp[j] = prev
    return cp, inv


def sequential(a, cp, inv, r):
    ny = len(r)
    d = np.zeros(ny)
    dp = 0.0
    for j in range(ny):
        dp = (r[j] - a[j] * dp) * inv[j]
        d[j] = dp
    z = np.zeros(ny)
    zn = 0.0
    for j in range(ny - 1, -1, -1):
        zn = d[j] - cp[j] * zn
        z[j] = zn
    return z


def segmented(a, cp, inv, r, S):
    ny = len(r)
    L = (ny + S - 1) // S
    rows = [(s * L, max(0, min(L, ny - s * L))) for s in range(S)]
    part = np.zeros((2, S))
    for s, (j0, nr) in enumerate(rows):
        dl, pr = 0.0, 1.0
        for j in range(j0, j0 + nr):
            dl = (r[j] - a[j] * dl) * inv[j]
            pr = pr * -(a[j] * inv[j])
        part[:, s] = dl, pr
    carry, c = np.zeros(S), 0.0
    for t in range(S):
        carry[t] = c
        c = part[0, t] + part[1, t] * c
    d = np.zeros(ny)
    for s, (j0, nr) in enumerate(rows):
        dp = carry[s]
        for j in range(j0, j0 + nr):
            dp = (r[j] - a[j] * dp) * inv[j]
            d[j] = dp
    for s, (j0, nr) in enumerate(rows):
        zl, pr = 0.0, 1.0
        for j in range(j0 + nr - 1, j0 - 1, -1):
            zl = d[j] - cp[j] * zl
            pr = pr * -cp[j]
        part[:, s] = zl, pr
    c = 0.0
    for t in range(S - 1, -1, -1):
        carry[t] = c
        c = part[0, t] + part[1, t] * c
    z = np.zeros(ny)
    for s, (j0, nr) in enumerate(rows):
        zn = carry[s]
        for j in range(j0 + nr - 1, j0 - 1, -1):
            zn = d[j] - cp[j] * zn
            z[j] = zn
    return z


@pytest.mark.parametrize("segs", [16, 32])
@pytest.mark.parametrize("ny", [1, 5, 16, 17, 31, 100, 128, 255, 256])
def test_segmented_sweeps_solve_the_line(ny, segs):
    rng = np.random.default_rng(ny + 1000 * segs)
    a = rng.uniform(-1.0, 0.0, ny)
    c = rng.uniform(-1.0, 0.0, ny)
    b = 2.5 + rng.uniform(0.0, 1.0, ny)
    a[0] = 0.0
    c[-1] = 0.0
    if ny > 8:
        # a cell without an unknown: identity row, as the factors store it
        a[3], b[3], c[3] = 0.0, 1.0, 0.0
    cp, inv = thomas_factors(a, b, c)
    r = rng.standard_normal(ny)
    z_seq = sequential(a, cp, inv, r)
    z_seg = segmented(a, cp, inv, r, segs)
    T = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    assert np.abs(T @ z_seg - r).max() <= 1e-14 * max(1.0, np.abs(r).max())
    assert np.abs(z_seg - z_seq).max() <= 1e-14 * np.abs(z_seq).max()

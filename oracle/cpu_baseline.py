"""CPU baseline for bench.py -- uses the oracle (test infrastructure) only as the
timed CPU reference, never as part of the GPU path.

The reference itself (PETSc 3.2 + HYPRE) cannot be built here, so both legs run
the oracle's FP64 restatement of the reference's pressure path (VolumeFlux,
PoissonRHS2_hypre, Remove_Nullspace, Convert_Phi2_Phi, UpdatePressure,
Projection, Contra2Cart, Divergence), every per-cell step parallelised with
OpenMP over the host cores, and differ in the linear solve:

- ``port``: a CPU port of this engine's solver (geometric multigrid inside
  FGMRES with CGS2, FP64) to the engine's flux-form tolerance;
- ``reference_method``: the reference's Krylov method, restarted GMRES with
  KDim 51 (poisson_hypre.c:91) on the restated HYPRE matrix (|v| <= 1e-10
  dropped, poisson.c:1608) to the reference's default relative tolerance
  poisson_tol = 5e-9 (main.c:87).  Its preconditioner is the matrix diagonal
  (right Jacobi): BoomerAMG (poisson_hypre.c:56-67) is part of HYPRE, which is
  absent, so this leg is slower than the reference's own AMG-preconditioned
  solve would be on the same cores.
"""
from __future__ import annotations

import platform
import time

REFERENCE_TOL = 5e-9  # poisson_tol, main.c:87
REFERENCE_KDIM = 51  # HYPRE_ParCSRGMRESSetKDim, poisson_hypre.c:91


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def reference_method_projection(case, A, rtol=REFERENCE_TOL, cap_s=150.0, chunk_cycles=4):
    """The full pressure projection with the reference's Krylov method on the
    CPU: GMRES(51), diagonal preconditioner, on A's CSR (A = Assembled(case,
    drop=True)).  The solve runs in chunks of whole restart cycles (a restart
    from the current iterate is what GMRES(m) does at every cycle, so chunking
    changes nothing) and gives up after cap_s seconds of solving."""
    import numpy as np

    from . import oracle as O

    f = case.fields
    g = A.g
    u = np.ascontiguousarray(f["ucont"].copy())
    O.volume_flux(g, case.immersed, u, f)
    b = O.rhs(g, u, A.nvert, case.dt, case.st, case.coeff)
    bc = A.to_compact(b)
    x = np.zeros(A.n)
    it, rel = 0, 1.0
    t0 = time.perf_counter()
    while True:
        x, k, rel = O.gmres(A.rowptr, A.col, A.val, bc, x0=x, m=REFERENCE_KDIM, rtol=rtol,
                            maxit=REFERENCE_KDIM * chunk_cycles, precond=1)
        it += k
        if rel <= rtol or time.perf_counter() - t0 > cap_s:
            break
    O.remove_nullspace(g, A.nvert, A.gid, x)
    phi = O.convert_phi(g, A.nvert, x)
    p = np.zeros_like(phi)
    O.update_pressure(g, A.nvert, p, phi)
    O.projection(g, f, phi, u, case.dt, case.st, case.coeff)
    ucat = np.zeros_like(u)
    O.contra2cart(g, case.bctype, f, u, ucat)
    maxdiv = O.divergence(g, u, np.ascontiguousarray(f["aj"]), A.nvert)
    return dict(phi=phi, ucont=u, maxdiv=maxdiv, iters=it, rel=rel, converged=rel <= rtol)


def measure(case, budget_s=25.0, rtol=2e-11, reference_leg=True, reference_cap_s=150.0):
    """Both legs on the host cores; the port leg is the headline value (the
    faster, so the stronger, baseline), the reference-method leg rides along."""
    from . import oracle as O

    cores = O.num_threads()
    model = _cpu_model()
    t0 = time.perf_counter()
    A = O.Assembled(case, drop=False)
    mg = O.CpuMG(A)
    setup = time.perf_counter() - t0
    n = 0
    t1 = time.perf_counter()
    last = None
    while True:
        last = O.cpu_projection(case, A, mg, rtol=rtol)
        n += 1
        if time.perf_counter() - t1 >= budget_s * 0.5 or n >= 50:
            break
    el = time.perf_counter() - t1
    mg.close()
    del A, mg
    port = {
        "value": round(n / el, 5),
        "unit": "solves/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"{n} full pressure projection(s) of the same {case.nx}x{case.ny}x{case.nz} case on the CPU "
                   f"(oracle restatement + CPU port of the engine's FP64 multigrid/FGMRES to rtol {rtol:g}; "
                   f"{cores} OpenMP threads on {model}); set-up {setup:.1f} s excluded; "
                   f"last solve {last['iters']} iterations, max|div| {last['maxdiv']:.2e}"),
        "seconds": round(el, 3),
    }
    out = dict(port)
    out["cpu_model"] = model
    out["legs"] = {"port": port}
    if reference_leg:
        t0 = time.perf_counter()
        Ad = O.Assembled(case, drop=True)
        setup_r = time.perf_counter() - t0
        t1 = time.perf_counter()
        r = reference_method_projection(case, Ad, cap_s=reference_cap_s)
        el_r = time.perf_counter() - t1
        out["legs"]["reference_method"] = {
            "value": round(1.0 / el_r, 5) if r["converged"] else None,
            "unit": "solves/s",
            "cores": cores,
            "kind": "reference_method",
            "method": (f"GMRES(KDim {REFERENCE_KDIM}) to relative residual {REFERENCE_TOL:g}, diagonal (Jacobi) "
                       "right preconditioner in place of BoomerAMG (HYPRE absent), on the restated HYPRE matrix"),
            "sample": (f"1 full pressure projection of the same {case.nx}x{case.ny}x{case.nz} case; "
                       f"{r['iters']} GMRES iterations to relative residual {r['rel']:.2e}"
                       f"{'' if r['converged'] else ' (stopped at the time cap, not converged)'}; "
                       f"max|div| {r['maxdiv']:.2e}; {cores} OpenMP threads on {model}; "
                       f"matrix assembly {setup_r:.1f} s excluded"),
            "seconds": round(el_r, 3),
            "iterations": int(r["iters"]),
        }
    return out

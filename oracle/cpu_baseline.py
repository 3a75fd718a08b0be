"""CPU baseline for bench.py -- uses the oracle (test infrastructure) only as the
timed CPU reference, never as part of the GPU path.

The reference itself (PETSc 3.2 + HYPRE) cannot be built here, so the baseline
is the oracle's FP64 restatement of the reference's pressure path (VolumeFlux,
PoissonRHS2_hypre, Remove_Nullspace, Convert_Phi2_Phi, UpdatePressure,
Projection, Contra2Cart, Divergence) with the linear solve done by a CPU port of
this engine's solver (geometric multigrid + FGMRES/CGS2 in FP64), parallelised
with OpenMP over the host cores.
"""
from __future__ import annotations

import os
import platform
import time


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def measure(case, budget_s=25.0, rtol=2e-11):
    from . import oracle as O

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
    return {
        "value": round(n / el, 5),
        "unit": "solves/s",
        "cores": O.num_threads(),
        "kind": "port",
        "sample": (f"{n} full pressure projection(s) of the same {case.nx}x{case.ny}x{case.nz} case on the CPU "
                   f"(oracle restatement + CPU port of the engine's FP64 multigrid/FGMRES to rtol {rtol:g}; "
                   f"{O.num_threads()} OpenMP threads on {_cpu_model()}); set-up {setup:.1f} s excluded; "
                   f"last solve {last['iters']} iterations, max|div| {last['maxdiv']:.2e}"),
        "seconds": round(el, 3),
    }

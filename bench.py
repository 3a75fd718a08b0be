"""Benchmark: pressure-projection solves/s on one MI355X.

Workload (one step = one full pressure projection, the reference's
PoissonSolver_Hypre + UpdatePressure + Projection + Divergence, solvers.c:874-930):
the 192x96x128 terrain grid with a Gaussian hill (h=0.3, s=0.5), algebraic
grid-line blending (non-zero cross metrics, non-symmetric operator), a voxelised
IB sphere (r=0.15) and a seeded divergent Ucont.  Inputs are resident in HBM
before timing starts; every step projects its own copy of the field.

Forms: mixed (FP32 multigrid preconditioner inside FP64 FGMRES; the headline
`value`) and plain FP64 (FP64 multigrid), both to max|aj div u| <= 1e-8.

Prints one JSON line: solves/s of the mixed form as `value`, p50/p99 latency,
the FP64 form beside it, the per-kernel HBM figures (`roofline`) and a CPU
baseline timed on this host's cores (`cpu_baseline`: the headline leg is this
engine's solver ported to the CPU; `legs.reference_method` is the reference's
GMRES(51) with a diagonal preconditioner standing in for BoomerAMG).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E peak
DIV_TOL = 1e-8
RTOL = 2e-11  # flux-form relative residual; leaves margin under DIV_TOL


# vfs_pp_time_kernel ids
KERNELS = {0: "apply_fp64", 1: "smoother_fp32", 2: "projection", 3: "contra2cart", 4: "rhs", 5: "copy_1GiB",
           6: "smoother_fp32_hbm", 7: "krylov_apply_dots", 8: "smoother_fp32_stream_floor"}
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "smoother_fp32_hbm_hill192x96x128_hbm_traffic.txt")


def traffic_from_profile():
    """HBM bytes per launch of the finest FP32 smoother sweep, from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (copy-kernel calibrated); None if absent."""
    try:
        for line in open(TRAFFIC_FILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=25.0,
                    help="seconds of CPU work for the port leg of the baseline (half is spent timing)")
    ap.add_argument("--cpu-ref-cap", type=float, default=150.0,
                    help="cap in seconds on the reference-method leg's one GMRES(51) solve; 0 skips that leg")
    ap.add_argument("--form", choices=("single", "batch", "serve", "towers"), default="single",
                    help="single: the 192x96x128 hill projection (default); batch: 64 uneven cases solved "
                         "together; serve: a Poisson stream of requests over resident cases; towers: the "
                         "512x256x256 wind-farm box on one GPU")
    ap.add_argument("--requests", type=int, default=150, help="serve form: requests per arrival rate")
    ap.add_argument("--rates", type=str, default="10,50,100,200,500", help="serve form: arrival rates (1/s)")
    ap.add_argument("--per-case", type=int, default=4, help="serve form: requests of one case in flight at once")
    ap.add_argument("--queue-cap", type=int, default=0,
                    help="serve form: refuse submissions while this many wait (0: unbounded)")
    ap.add_argument("--cases", type=int, default=64, help="batch form: number of cases")
    ap.add_argument("--workers", type=int, default=16, help="host processes generating the batched and served cases")
    ap.add_argument("--time-kernel", type=str, default=None,
                    help="profiling aid: set up the 192x96x128 case and only time the engine kernels with "
                         "these comma-separated ids (vfs_pp_time_kernel), printing their ms and GB/s")
    ap.add_argument("--single-only", action="store_true",
                    help="default run: only the 192x96x128 projection (no towers, batch, serve or host-array blocks)")
    ap.add_argument("--block-steps", type=int, default=8,
                    help="default run: timed steps of the towers and batch blocks")
    ap.add_argument("--block-requests", type=int, default=60,
                    help="default run: requests per arrival rate of the serve block")
    ap.add_argument("--block-rates", type=str, default="10,100,500",
                    help="default run: arrival rates (1/s) of the serve block")
    return ap.parse_args()


def _gen_case(spec):
    from vfs_v3p2_amd import cases

    return cases.ensemble_case(*spec)


def towers_block(E, torch, steps, warmup):
    """512x256x256 box (33.5M cells): uniform in x and z, geometric stretching 1.02 in
    y, a 3x4 array of IB towers (radius 4 cells, 60% of the height), log-law mean
    flow with 10% fluctuations; full projections in the mixed and FP64 forms.  The
    case is generated with FormMetrics on the device (bit-identical to the numpy
    generator)."""
    from vfs_v3p2_amd import cases

    t0 = time.perf_counter()
    case = cases.towers_case(512, 256, 256, metrics_fn=E.device_metrics_fn())
    gen_s = time.perf_counter() - t0
    forms = {}
    for precision, tag in ((1, "mixed"), (0, "fp64")):
        h, r = run_form(E, torch, case, precision, steps, warmup)
        ms_sm, by_sm = h.time_kernel(1, 20)
        desc = h.describe()
        h.close()
        torch.cuda.empty_cache()
        lat = np.array(r["lat"]) * 1e3
        forms[tag] = dict(solves_per_s=round(steps / r["total"], 3), p50_ms=round(float(np.percentile(lat, 50)), 2),
                          p99_ms=round(float(np.percentile(lat, 99)), 2),
                          iterations=int(np.median([s_[0] for s_ in r["stats"]])),
                          max_div=float(max(s_[2] for s_ in r["stats"])), setup_s=round(r["setup_s"], 2),
                          set_nvert_ms=round(r["set_nvert_ms"], 1),
                          finest_step_gbs=round(by_sm / (ms_sm * 1e-3) / 1e9, 1),
                          finest_step_ms=round(ms_sm, 4), levels=desc.strip().split("\n"))
    del case
    mx = forms["mixed"]
    return dict(config={"workload": "wind-farm box 512x256x256 on one GPU: uniform x/z, geometric y stretching "
                                    "1.02, 3x4 IB towers (radius 4 cells, 60% height), log-law mean + 10% "
                                    "fluctuations; full projection", "grid": [512, 256, 256]},
                steps=steps, warmup=warmup, solves_per_s=mx["solves_per_s"], p50_ms=mx["p50_ms"],
                p99_ms=mx["p99_ms"], mixed_over_fp64=round(mx["solves_per_s"] / forms["fp64"]["solves_per_s"], 3),
                forms=forms, generate_s=round(gen_s, 1))


def bench_towers(args):
    import torch

    from vfs_v3p2_amd import engine as E

    E.lib()
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    blk = towers_block(E, torch, args.steps, args.warmup)
    forms = blk["forms"]
    mx = forms["mixed"]
    gen_s = blk["generate_s"]
    out = {
        "metric": "Poisson solves/s & p99 latency at div<=1e-8, mixed vs FP64; stencil HBM GB/s",
        "value": mx["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 / mx["solves_per_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP32 multigrid preconditioner)",
        "data": "synthetic (seeded generator, vfs_v3p2_amd/cases.py)",
        "config": {
            "workload": "wind-farm box 512x256x256 on one GPU: uniform x/z, geometric y stretching 1.02, 3x4 IB "
                        "towers (radius 4 cells, 60% height), log-law mean + 10% fluctuations; full projection",
            "grid": [512, 256, 256],
            "rtol": RTOL,
            "div_tol": DIV_TOL,
        },
        "p50_ms": mx["p50_ms"],
        "p99_ms": mx["p99_ms"],
        "forms": forms,
        "mixed_over_fp64": round(mx["solves_per_s"] / forms["fp64"]["solves_per_s"], 3),
        "generate_s": round(gen_s, 1),
    }
    print(json.dumps(out), flush=True)


def _gen_served(spec):
    from vfs_v3p2_amd import cases

    return cases.served_case(*spec)


def serve_block(E, torch, generated, rates, nreq, max_per_case=4, max_queue=0):
    """Poisson stream of projection requests over 16 resident cases (sizes seeded
    log-uniform 24^3-256^3; 60% channel, 30% hill + sphere, 10% towers), served by
    continuous batching with up to `max_per_case` requests of one case in flight;
    p50/p99 latency (arrival to completion) of the served requests, solves/s and
    (with a queue cap) refused requests per arrival rate.  `generated` yields the
    cases (served_case)."""
    t0 = time.perf_counter()
    sv = E.Server(max_active=64, max_per_case=max_per_case, max_queue=max_queue, precision=1, rtol=RTOL,
                  div_tol=DIV_TOL)
    res = []
    for c in generated:
        g = E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed)
        cid = sv.add_case(g, c.fields)
        u1 = torch.from_numpy(np.ascontiguousarray(c.fields["ucont"])).cuda()
        # a second input field of the same case: its divergent part from another seed
        rng = np.random.default_rng(cid + 99)
        u2 = u1 * torch.from_numpy(1.0 + 0.2 * rng.standard_normal(1)).cuda()
        res.append((cid, c.shape, c.dt, c.st, c.coeff, u1, u2, c.nx * c.ny * c.nz, c.name))
    setup_s = time.perf_counter() - t0
    ncell = sum(r[7] for r in res)
    rng = np.random.default_rng(2024)
    per_rate = []
    # warm-up: every case once (plans, tuning, graphs)
    for r in res:
        cid, sh = r[0], r[1]
        sv.submit(cid, r[5].clone(), torch.zeros(sh, dtype=torch.float64, device="cuda"),
                  torch.zeros(sh, dtype=torch.float64, device="cuda"),
                  torch.zeros(sh + (3,), dtype=torch.float64, device="cuda"), r[2], r[3], r[4])
    while sv.status()["queued"] or sv.status()["active"]:
        sv.progress()
    for lam in rates:
        arrivals = np.cumsum(rng.exponential(1.0 / lam, nreq))
        picks = rng.integers(len(res), size=nreq)
        mix = rng.uniform(-1.0, 1.0, size=nreq)
        lat = np.full(nreq, np.nan)
        divs, its, failed = [], [], []
        refused = 0
        t2r = {}
        i = done = 0
        torch.cuda.synchronize()
        start = time.perf_counter()
        while done < nreq:
            now = time.perf_counter() - start
            while i < nreq and arrivals[i] <= now:
                r = res[picks[i]]
                cid, sh = r[0], r[1]
                u = r[5] + mix[i] * r[6]
                t = sv.submit(cid, u, torch.zeros(sh, dtype=torch.float64, device="cuda"),
                              torch.zeros(sh, dtype=torch.float64, device="cuda"),
                              torch.zeros(sh + (3,), dtype=torch.float64, device="cuda"), r[2], r[3], r[4])
                if t is None:  # the full queue refused it
                    refused += 1
                    done += 1
                else:
                    t2r[t] = i
                i += 1
            st = sv.status()
            if st["queued"] == 0 and st["active"] == 0:
                if i < nreq:
                    time.sleep(max(0.0, min(0.002, arrivals[i] - (time.perf_counter() - start))))
                continue
            for t, s in sv.progress():
                q = t2r.pop(t)
                lat[q] = (time.perf_counter() - start) - arrivals[q]
                divs.append(s.max_div)
                its.append(s.iterations)
                if not s.converged:
                    failed.append(q)
                    r = res[picks[q]]
                    print(f"request {q} (case {r[8]}) missed the tolerance: it {s.iterations} rel {s.rel_residual:.3e} "
                          f"max_div {s.max_div:.3e} bnorm {s.bnorm:.3e}", file=sys.stderr, flush=True)
                done += 1
        span = time.perf_counter() - start
        served = lat[~np.isnan(lat)]
        per_rate.append(dict(rate=lam, requests=nreq, solves_per_s=round(len(served) / span, 2),
                             p50_ms=round(1e3 * float(np.percentile(served, 50)), 2),
                             p99_ms=round(1e3 * float(np.percentile(served, 99)), 2),
                             max_div=float(max(divs)), iterations_max=int(max(its)), failed=len(failed),
                             refused=refused))
    st = sv.status()
    sv.close()
    del res
    torch.cuda.empty_cache()
    top = per_rate[-1]
    return dict(config={"workload": "served stream: Poisson arrivals over 16 resident cases (sizes seeded "
                                    "log-uniform 24^3-256^3; 60% stretched channel, 30% hill + IB sphere, 10% IB "
                                    "towers), continuous batching, one step = one request",
                        "cells_resident": ncell, "max_active": 64, "max_per_case": max_per_case,
                        "max_queue": max_queue},
                solves_per_s=top["solves_per_s"], p50_ms=top["p50_ms"], p99_ms=top["p99_ms"], rates=per_rate,
                resident_gb=round(st["resident_bytes"] / 1e9, 2), workspace_gb=round(st["workspace_bytes"] / 1e9, 2),
                setup_s=round(setup_s, 1), failed=sum(p["failed"] for p in per_rate))


def bench_serve(args):
    import multiprocessing as mp

    from vfs_v3p2_amd import cases

    specs = cases.served_specs()
    pool = mp.get_context("fork").Pool(max(1, min(args.workers, len(specs))))
    pending = pool.imap(_gen_served, specs)
    import torch

    from vfs_v3p2_amd import engine as E

    E.lib()
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    rates = [float(x) for x in args.rates.split(",") if x]
    blk = serve_block(E, torch, pending, rates, args.requests, args.per_case, args.queue_cap)
    pool.close()
    per_rate = blk["rates"]
    top = per_rate[-1]
    out = {
        "metric": "Poisson solves/s & p99 latency at div<=1e-8, mixed vs FP64; stencil HBM GB/s",
        "value": top["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": sum(p["requests"] for p in per_rate),
        "warmup": len(specs),
        "ms_per_step": round(1e3 / top["solves_per_s"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP32 multigrid preconditioner)",
        "data": "synthetic (seeded generator, vfs_v3p2_amd/cases.py)",
        "config": dict(blk["config"], workload=blk["config"]["workload"] +
                       "; value = solves/s at the highest arrival rate"),
        "p50_ms": top["p50_ms"],
        "p99_ms": top["p99_ms"],
        "rates": per_rate,
        "resident_gb": blk["resident_gb"],
        "workspace_gb": blk["workspace_gb"],
        "setup_s": blk["setup_s"],
    }
    print(json.dumps(out), flush=True)


def batch_block(E, torch, generated, steps):
    """64 independent cases, sizes seeded log-uniform in [32^3, 128^3] with +-25% per
    axis, own stretching and hill height, half with an IB sphere: solved one at a
    time (plain FP64 drop-in, and mixed), then together as one batch.  `generated`
    yields the cases (ensemble_case)."""
    t0 = time.perf_counter()
    cs = []
    for c in generated:
        u = torch.from_numpy(np.ascontiguousarray(c.fields["ucont"])).cuda()
        geo = {k: v for k, v in c.fields.items() if k != "ucont"}
        cs.append((c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed, c.dt, c.st, c.coeff, geo, u))
    gen_s = time.perf_counter() - t0
    ncell = sum(nx * ny * nz for nx, ny, nz, *_ in cs)

    def make(precision):
        hs = []
        for nx, ny, nz, per, bct, imm, *_rest in cs:
            geo = _rest[3]
            h = E.Handle(E.make_grid(nx, ny, nz, per, bct, imm),
                         E.default_opts(precision=precision, rtol=RTOL, div_tol=DIV_TOL))
            h.set_geometry(geo)
            hs.append(h)
        return hs

    def fields():
        u = [c[10].clone() for c in cs]
        phi = [torch.zeros((c[2] + 2, c[1] + 2, c[0] + 2), dtype=torch.float64, device="cuda") for c in cs]
        p = [torch.zeros_like(f) for f in phi]
        ucat = [torch.zeros(f.shape + (3,), dtype=torch.float64, device="cuda") for f in phi]
        return u, phi, p, ucat

    dts, sts, cos = [c[6] for c in cs], [c[7] for c in cs], [c[8] for c in cs]
    out = {}
    for precision, tag in ((0, "fp64"), (1, "mixed")):
        t = time.perf_counter()
        hs = make(precision)
        setup_s = time.perf_counter() - t
        # one at a time (the drop-in path): every case through its own handle
        u, phi, p, ucat = fields()
        for q, h in enumerate(hs):  # warm-up
            h.project(u[q], phi[q], p[q], ucat[q], dts[q], sts[q], cos[q], strict=True)
        seq = []
        for rep in range(steps):
            u, phi, p, ucat = fields()
            torch.cuda.synchronize()
            t = time.perf_counter()
            its = []
            for q, h in enumerate(hs):
                st = h.project(u[q], phi[q], p[q], ucat[q], dts[q], sts[q], cos[q], strict=True)
                its.append(st.iterations)
            torch.cuda.synchronize()
            seq.append(time.perf_counter() - t)
        # all together
        B = E.Batch(hs)
        u, phi, p, ucat = fields()
        B.project(u, phi, p, ucat, dts, sts, cos, strict=True)  # warm-up, tuning, graph capture
        shapes = B.tile_shapes()
        bat = []
        for rep in range(steps):
            u, phi, p, ucat = fields()
            torch.cuda.synchronize()
            t = time.perf_counter()
            sts_b = B.project(u, phi, p, ucat, dts, sts, cos, strict=True)
            torch.cuda.synchronize()
            bat.append(time.perf_counter() - t)
        B.close()
        for h in hs:
            h.close()
        del hs
        torch.cuda.empty_cache()
        n = len(cs)
        out[tag] = dict(
            sequential_solves_per_s=round(n / float(np.median(seq)), 2),
            batched_solves_per_s=round(n / float(np.median(bat)), 2),
            sequential_ms=round(1e3 * float(np.median(seq)), 2),
            batched_ms=round(1e3 * float(np.median(bat)), 2),
            iterations_max=int(max(s.iterations for s in sts_b)),
            max_div=float(max(s.max_div for s in sts_b)),
            setup_s=round(setup_s, 2),
            tile_shapes={"flat256": shapes[0], "64x4": shapes[1], "32x8": shapes[2], "16x16": shapes[3]},
        )
    del cs
    torch.cuda.empty_cache()
    value = out["mixed"]["batched_solves_per_s"]
    base = out["fp64"]["sequential_solves_per_s"]
    return dict(config={"workload": f"batched ensemble: {len(dts)} independent projections, sizes seeded "
                                    "log-uniform 32^3-128^3 (+-25% per axis), stretching beta in [1,3], hill h in "
                                    "[0,0.4], half with an IB sphere; one step = all cases projected once",
                        "cells": ncell,
                        "form": "mixed batched (one work list per launch, HIP graph per FGMRES iteration)"},
                steps=steps, solves_per_s=value, ms_per_step=out["mixed"]["batched_ms"],
                mixed_batched_over_fp64_sequential=round(value / base, 3), forms=out, upload_s=round(gen_s, 1))


def bench_batch(args):
    import multiprocessing as mp

    from vfs_v3p2_amd import cases

    specs = cases.batch_sizes(count=args.cases)
    # cases are generated on host processes forked before anything touches the GPU
    pool = mp.get_context("fork").Pool(max(1, min(args.workers, len(specs))))
    pending = pool.imap(_gen_case, specs)
    import torch

    from vfs_v3p2_amd import engine as E

    E.lib()
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    blk = batch_block(E, torch, pending, args.steps)
    pool.close()
    res = {
        "metric": "Poisson solves/s & p99 latency at div<=1e-8, mixed vs FP64; stencil HBM GB/s",
        "value": blk["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": blk["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP32 multigrid preconditioner)",
        "data": "synthetic (seeded generator, vfs_v3p2_amd/cases.py)",
        "config": blk["config"],
        "mixed_batched_over_fp64_sequential": blk["mixed_batched_over_fp64_sequential"],
        "forms": blk["forms"],
    }
    print(json.dumps(res), flush=True)


def host_array_block(E, torch, case, steps):
    """The drop-in as a host-array caller sees it (vfs_pp_project): Ucont, Phi, P and
    Ucat cross PCIe in and out on every step, around the same projection."""
    g = E.make_grid(case.nx, case.ny, case.nz, case.per, case.bctype, case.immersed)
    h = E.Handle(g, E.default_opts(precision=1, rtol=RTOL, div_tol=DIV_TOL))
    h.set_geometry(case.fields)
    sh = case.shape
    base = np.ascontiguousarray(case.fields["ucont"])
    # the caller's arrays persist from step to step, as the reference's Vecs do
    u = np.empty_like(base)
    phi, p, ucat = np.zeros(sh), np.zeros(sh), np.zeros(sh + (3,))
    lat = []
    for k in range(steps + 1):
        u[...] = base
        for a_ in (phi, p, ucat):
            a_.fill(0.0)
        a = time.perf_counter()
        st = h.project_host(u, phi, p, ucat, case.dt, case.st, case.coeff)
        if k:  # the first call is the warm-up
            lat.append(time.perf_counter() - a)
        if not st.converged:
            raise RuntimeError("host-array projection missed the tolerance")
    h.close()
    moved = 2 * (base.nbytes + 2 * np.zeros(sh).nbytes + np.zeros(sh + (3,)).nbytes)
    ms = 1e3 * float(np.median(lat))
    return dict(solves_per_s=round(1e3 / ms, 2), ms_per_step=round(ms, 3), steps=steps,
                pcie_bytes_per_step=moved,
                note="vfs_pp_project on host arrays: Ucont, Phi, P, Ucat copied in and out over PCIe each step")


def time_step_block(E, torch, case, steps):
    """Whole time steps kept on the device (Pressure_Gradient of the previous
    pressure, the momentum predictor, the pressure projection), timed per step:
    the explicit RungeKutta predictor on the projection's own case, and the
    Jacobian-free Newton-Krylov implicit predictor (Implicit_MatrixFree,
    immersed = 0) on a walled stretched channel."""
    from vfs_v3p2_amd import cases

    def setup(c):
        g = E.make_grid(c.nx, c.ny, c.nz, c.per, c.bctype, c.immersed)
        h = E.Handle(g, E.default_opts(precision=1, rtol=RTOL, div_tol=DIV_TOL))
        h.set_geometry(c.fields)
        met = {k: torch.from_numpy(np.ascontiguousarray(c.fields[k])).cuda() for k in E.MOMENTUM_METRICS}
        nv = torch.from_numpy(np.ascontiguousarray(c.fields["nvert"])).cuda()
        u = torch.from_numpy(np.ascontiguousarray(c.fields["ucont"])).cuda()
        phi = torch.zeros(c.shape, dtype=torch.float64, device="cuda")
        p = torch.zeros_like(phi)
        ucat = torch.zeros(c.shape + (3,), dtype=torch.float64, device="cuda")
        h.project(u, phi, p, ucat, c.dt, c.st, c.coeff, strict=True)  # a divergence-free start
        return g, h, met, nv, u, phi, p, ucat

    out = {}
    ren, dt = 1000.0, 0.005
    g, h, met, nv, u, phi, p, ucat = setup(case)
    lat, its = [], []
    for k in range(steps + 1):
        torch.cuda.synchronize()
        a = time.perf_counter()
        dp = E.Pressure_Gradient(g, met, nv, p)
        E.RungeKutta(h, met, dp, u, ucat, ren, dt, 1.0)
        st = h.project(u, phi, p, ucat, dt, 1.0, 1.0, strict=True)
        torch.cuda.synchronize()
        if k:
            lat.append(time.perf_counter() - a)
            its.append(st.iterations)
    h.close()
    out["explicit"] = dict(case=f"hill {case.nx}x{case.ny}x{case.nz} + IB sphere, ren {ren:g}, dt {dt:g}",
                           ms_per_step=round(1e3 * float(np.median(lat)), 3), steps=steps,
                           projection_iterations=int(max(its)), max_div=float(st.max_div))
    c2 = cases.channel_case(128, 96, 96, beta=2.0, seed=5)
    g, h, met, nv, u, phi, p, ucat = setup(c2)
    lat, newton, gm = [], [], []
    for k in range(steps + 1):
        torch.cuda.synchronize()
        a = time.perf_counter()
        dp = E.Pressure_Gradient(g, met, nv, p)
        rhs_o = E.Formfunction_2(g, met, nv, u, ucat, ren)
        un = u.clone()
        ns = E.Implicit_MatrixFree(h, met, dp, u, un, ucat, ren, dt, 1.0, rhs_o=rhs_o, rtol=1e-4)
        st = h.project(un, phi, p, ucat, dt, 1.0, 1.0, strict=True)
        u = un
        torch.cuda.synchronize()
        if k:
            lat.append(time.perf_counter() - a)
            newton.append(ns.newton_iterations)
            gm.append(ns.krylov_iterations)
    h.close()
    torch.cuda.empty_cache()
    out["implicit"] = dict(case=f"channel {c2.nx}x{c2.ny}x{c2.nz}, walls in y, ren {ren:g}, dt {dt:g}, "
                                f"rtol 1e-4 on ||F||",
                           ms_per_step=round(1e3 * float(np.median(lat)), 3), steps=steps,
                           newton_iterations_max=int(max(newton)), gmres_iterations_max=int(max(gm)),
                           max_div=float(st.max_div))
    return out


def run_form(E, torch, case, precision, steps, warmup):
    g = E.make_grid(case.nx, case.ny, case.nz, case.per, case.bctype, case.immersed)
    opts = E.default_opts(precision=precision, rtol=RTOL, div_tol=DIV_TOL)
    h = E.Handle(g, opts)
    t0 = time.perf_counter()
    h.set_geometry(case.fields)
    setup_s = time.perf_counter() - t0
    sh = case.shape
    base = torch.from_numpy(np.ascontiguousarray(case.fields["ucont"])).cuda()
    n = steps + warmup
    us = [base.clone() for _ in range(n)]
    phi = torch.zeros(sh, dtype=torch.float64, device="cuda")
    p = torch.zeros_like(phi)
    ucat = torch.zeros(sh + (3,), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    for w in range(warmup):
        st = h.project(us[w], phi, p, ucat, case.dt, case.st, case.coeff, strict=True)
    torch.cuda.synchronize()
    lat = []
    stats = []
    t_start = time.perf_counter()
    for k in range(steps):
        a = time.perf_counter()
        st = h.project(us[warmup + k], phi, p, ucat, case.dt, case.st, case.coeff, strict=True)
        lat.append(time.perf_counter() - a)
        stats.append((st.iterations, st.rel_residual, st.max_div))
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start
    # after the timed solves: a moving body's per-step rebuild on the same grid (new
    # mask, VolumeFlux face list and multigrid hierarchy; the metrics stay)
    nv = torch.from_numpy(np.ascontiguousarray(case.fields["nvert"])).cuda()
    remask = []
    for _ in range(5):
        a = time.perf_counter()
        h.set_nvert_device(nv)
        remask.append(time.perf_counter() - a)
    return h, dict(total=total, lat=lat, stats=stats, setup_s=setup_s, set_nvert_ms=1e3 * float(np.median(remask)),
                   desc=h.describe())


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build runs every system on a single MI355X "
              f"(--gpus {args.gpus} is not supported; systems would shard one per GPU without communication)",
              file=sys.stderr)
        sys.exit(2)
    if args.form == "batch":
        bench_batch(args)
        return
    if args.form == "serve":
        bench_serve(args)
        return
    if args.form == "towers":
        bench_towers(args)
        return
    # The batched and served forms' synthetic cases are generated first, on host
    # processes forked before anything touches the GPU, and collected before any
    # GPU work starts (no process but this one ever holds the GPU, and nothing
    # competes with the timed regions for host cores).
    gen_batch, gen_serve, pre_s = [], [], 0.0
    if not args.single_only:
        import multiprocessing as mp

        from vfs_v3p2_amd import cases as _cases

        t0 = time.perf_counter()
        sspecs = _cases.served_specs()
        bspecs = _cases.batch_sizes(count=args.cases)
        with mp.get_context("fork").Pool(max(1, args.workers)) as pool:
            served_async = pool.map_async(_gen_served, sspecs, chunksize=1)
            gen_batch = pool.map(_gen_case, bspecs, chunksize=1)
            gen_serve = served_async.get()
        pre_s = time.perf_counter() - t0
    import torch

    from vfs_v3p2_amd import cases
    from vfs_v3p2_amd import engine as E

    E.lib()  # fail loudly without the HIP library
    if not torch.cuda.is_available():
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    case = cases.hill_case(192, 96, 128, h=0.3, s=0.5, sphere=(0.6, 0.75, 0.0, 0.15), seed=0)
    if args.time_kernel is not None:
        h = E.Handle(E.make_grid(case.nx, case.ny, case.nz, case.per, case.bctype, case.immersed),
                     E.default_opts(precision=1, rtol=RTOL, div_tol=DIV_TOL))
        h.set_geometry(case.fields)
        for kid in [int(x) for x in args.time_kernel.split(",") if x]:
            ms, by = h.time_kernel(kid, 100 if kid != 5 else 20)
            print(json.dumps({"kernel": KERNELS.get(kid, kid), "ms_per_launch": ms, "algorithmic_bytes": by,
                              "gbs": round(by / (ms * 1e-3) / 1e9, 1)}), flush=True)
        h.close()
        return

    h_mx, mx = run_form(E, torch, case, 1, args.steps, args.warmup)
    # every engine kernel timed in isolation with HIP events on the handle's stream
    kern = {}
    for kid, name in KERNELS.items():
        ms, by = h_mx.time_kernel(kid, 100 if kid != 5 else 20)
        kern[name] = dict(ms_per_launch=round(ms, 5), algorithmic_bytes=by,
                          gbs=round(by / (ms * 1e-3) / 1e9, 1), frac=round(by / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, 4))
    h_mx.close()
    h64, f64 = run_form(E, torch, case, 0, args.steps, args.warmup)
    ms_sm64, by_sm64 = h64.time_kernel(1, 100)
    h64.close()
    kern["smoother_fp64"] = dict(ms_per_launch=round(ms_sm64, 5), algorithmic_bytes=by_sm64,
                                 gbs=round(by_sm64 / (ms_sm64 * 1e-3) / 1e9, 1),
                                 frac=round(by_sm64 / (ms_sm64 * 1e-3) / 1e9 / PEAK_HBM_GBS, 4))

    lat = np.array(mx["lat"]) * 1e3
    value = args.steps / mx["total"]
    fp64_value = args.steps / f64["total"]
    its = [s[0] for s in mx["stats"]]
    divs = [s[2] for s in mx["stats"]]
    divs64 = [s[2] for s in f64["stats"]]
    sm = kern["smoother_fp32_hbm"]
    out = {
        "metric": "Poisson solves/s & p99 latency at div<=1e-8, mixed vs FP64; stencil HBM GB/s",
        "value": round(value, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mx["total"] / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64 (FP32 multigrid preconditioner)",
        "data": "synthetic (seeded generator, vfs_v3p2_amd/cases.py)",
        "config": {
            "workload": "terrain/hill 192x96x128 + IB sphere r=0.15, full pressure projection "
                        "(VolumeFlux, RHS, FGMRES+MG solve to max|div|<=1e-8, null space, P+=phi, "
                        "Projection, Contra2Cart, Divergence)",
            "grid": [case.nx, case.ny, case.nz],
            "form": "mixed: FP32 geometric multigrid inside FP64 FGMRES",
            "rtol": RTOL,
            "div_tol": DIV_TOL,
        },
        "p50_ms": round(float(np.percentile(lat, 50)), 4),
        "p99_ms": round(float(np.percentile(lat, 99)), 4),
        "iterations": int(np.median(its)),
        "max_div": float(max(divs)),
        "fp64": {
            "value": round(fp64_value, 3),
            "unit": "solves/s",
            "p99_ms": round(float(np.percentile(np.array(f64["lat"]) * 1e3, 99)), 4),
            "iterations": int(np.median([s[0] for s in f64["stats"]])),
            "max_div": float(max(divs64)),
        },
        "mixed_over_fp64": round(value / fp64_value, 3),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_smooth_point<float> (finest-level Chebyshev-Jacobi sweep, FP16 coefficients, 31 algorithmic "
                      "B/cell), swept over 6 rotating copies of its operands so every launch streams from HBM "
                      "rather than the 256 MB last-level cache",
            "achieved": sm["gbs"],
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": sm["frac"],
            "traffic": traffic_from_profile(),
            "ms_per_launch": sm["ms_per_launch"],
            "algorithmic_bytes_per_launch": sm["algorithmic_bytes"],
            "copy_gbs": kern["copy_1GiB"]["gbs"],
            # the same operands streamed once per cell with no stencil, over the same
            # rotating copies: the rate this array set allows at all
            "stream_floor_ms": kern["smoother_fp32_stream_floor"]["ms_per_launch"],
            "over_stream_floor": round(sm["ms_per_launch"] / kern["smoother_fp32_stream_floor"]["ms_per_launch"], 3),
            "kernels": kern,
        },
        "setup_s": round(mx["setup_s"], 3),
        "set_nvert_ms": round(mx["set_nvert_ms"], 2),
        "cpu_baseline": None,
    }
    if max(divs) > DIV_TOL or max(divs64) > DIV_TOL:
        out["error"] = "max|div u| above 1e-8"
    if not args.single_only:
        out["host_array"] = host_array_block(E, torch, case, max(3, min(args.steps, 10)))
        out["time_step"] = time_step_block(E, torch, case, 5)
    if not args.no_cpu:
        try:
            from oracle import cpu_baseline

            out["cpu_baseline"] = cpu_baseline.measure(case, budget_s=args.cpu_budget,
                                                       reference_leg=args.cpu_ref_cap > 0,
                                                       reference_cap_s=args.cpu_ref_cap)
        except Exception as e:  # the GPU figures stand on their own
            out["cpu_baseline"] = {"error": f"{type(e).__name__}: {e}"}
    if not args.single_only:
        del case
        progress = lambda msg: print(f"bench: {msg}", file=sys.stderr, flush=True)
        progress("towers block")
        out["towers"] = towers_block(E, torch, args.block_steps, min(args.warmup, 3))
        progress("batch block")
        out["batch"] = batch_block(E, torch, gen_batch, args.block_steps)
        del gen_batch
        progress("serve block")
        rates = [float(x) for x in args.block_rates.split(",") if x]
        out["serve"] = serve_block(E, torch, gen_serve, rates, args.block_requests, args.per_case, args.queue_cap)
        del gen_serve
        out["batch_over_plain_fp64"] = out["batch"]["mixed_batched_over_fp64_sequential"]
        out["pregenerate_s"] = round(pre_s, 1)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

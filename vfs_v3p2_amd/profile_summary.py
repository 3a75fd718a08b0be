"""Summarise rocprofv3 output (CSV) of the engine.

    python -m vfs_v3p2_amd.profile_summary <run>_kernel_trace.csv
    python -m vfs_v3p2_amd.profile_summary <run>_kernel_trace.csv --batch
    python -m vfs_v3p2_amd.profile_summary --fetch <fetch>_counter_collection.csv \
        --write <write>_counter_collection.csv --cells N

Trace mode prints, as markdown: per-(kernel, grid) launch counts and durations,
the kernel time of each solve (a solve runs from its RHS kernel to its
divergence kernel), and the per-solve breakdown by kernel for the first and the
last group of solves (the benchmark runs the mixed form first, then FP64).

Batch mode keeps the batched kernels (kb_*, in the multigrid precision given by
--precision), splits them into calls at each batched RHS step (the first launch
of a call), and breaks down the last call by kernel.

Counter mode (--counters a.csv b.csv --kernel NAME --cells N) prints the mean per
dispatch (and per wave, given SQ_WAVES) of every counter those runs collected.

Traffic mode reads one FETCH_SIZE and one WRITE_SIZE pass (KiB per dispatch),
scales each by what the same counter reads for the 1 GiB device copy kernel
(k_copy16: 1 GiB read, 1 GiB written), and prints HBM-side bytes per launch and
per cell for every kernel whose grid covers N cells, plus a machine-readable
line for the finest FP32 smoother sweep.
"""
from __future__ import annotations

import argparse
import collections
import csv
import re


def _short(name: str) -> str:
    n = name.replace("pp::(anonymous namespace)::", "").replace("void ", "")
    return re.sub(r"\(.*", "", n)


def load_trace(path):
    rows = []
    for r in csv.DictReader(open(path)):
        s, e = int(r["Start_Timestamp"]), int(r["End_Timestamp"])
        rows.append(dict(name=_short(r["Kernel_Name"]), s=s, e=e, d=e - s,
                         grid=(int(r["Grid_Size_X"]), int(r["Grid_Size_Y"]), int(r["Grid_Size_Z"])),
                         vgpr=r.get("VGPR_Count", ""), scratch=r.get("Scratch_Size", ""),
                         dispatch=r.get("Dispatch_Id", "")))
    rows.sort(key=lambda r: r["s"])
    return rows


def solves(rows):
    """(first, last, launches) of every solve: from the last RHS launch before a
    Divergence launch to that Divergence (RHS launches of a timing loop in between
    solves start no solve)."""
    out, last_rhs = [], None
    for r in rows:
        if r["name"] == "k_rhs":
            last_rhs = r
        elif r["name"] == "k_divergence" and last_rhs is not None:
            a, b = last_rhs, r
            out.append((a, b, [x for x in rows if x["s"] >= a["s"] and x["e"] <= b["e"]]))
            last_rhs = None
    return out


def _counters(path):
    acc = collections.defaultdict(list)
    for r in csv.DictReader(open(path)):
        acc[(_short(r["Kernel_Name"]), int(r["Grid_Size"]))].append(float(r["Counter_Value"]))
    return {k: (sum(v) / len(v), len(v)) for k, v in acc.items()}


COPY = ("k_copy16", 1 << 26)  # 2^26 work-items x 16 B = 1 GiB read and 1 GiB written
SMOOTHER = "k_smooth_point<float, 0>"


def traffic(fetch_csv, write_csv, cells, top):
    F, W = _counters(fetch_csv), _counters(write_csv)
    if COPY not in F or COPY not in W:
        raise SystemExit("the copy kernel k_copy16 is missing from the counter runs")
    fs = (1 << 30) / (F[COPY][0] * 1024.0)
    ws = (1 << 30) / (W[COPY][0] * 1024.0)
    print(f"Calibration on k_copy16 (1 GiB read + 1 GiB written): FETCH_SIZE reads {F[COPY][0]:.0f} KiB "
          f"-> scale {fs:.4f}; WRITE_SIZE reads {W[COPY][0]:.0f} KiB -> scale {ws:.4f}\n")
    print("| kernel | grid | dispatches | fetched MB | written MB | B/cell |")
    print("|---|---|---|---|---|---|")
    keys = sorted((k for k in F if k[1] == cells), key=lambda k: -(F[k][0] * fs + W.get(k, (0, 0))[0] * ws))
    for k in keys[:top]:
        f = F[k][0] * 1024 * fs
        w = W.get(k, (0.0, 0))[0] * 1024 * ws
        print(f"| {k[0]} | {k[1]} | {F[k][1]} | {f / 1e6:.2f} | {w / 1e6:.2f} | {(f + w) / cells:.1f} |")
    sm = (SMOOTHER, cells)
    if sm in F and sm in W:
        t = F[sm][0] * 1024 * fs + W[sm][0] * 1024 * ws
        print(f"\ntraffic_bytes_per_launch: {t:.0f}")


def counters(paths, kernel, cells):
    """Mean per dispatch of every counter in the given --pmc runs, for the
    dispatches of `kernel` (short name) covering `cells` work-items, and per wave
    when SQ_WAVES is among them."""
    vals = {}
    for p in paths:
        acc = collections.defaultdict(list)
        for r in csv.DictReader(open(p)):
            if _short(r["Kernel_Name"]) == kernel and int(r["Grid_Size"]) == cells:
                acc[r["Counter_Name"]].append(float(r["Counter_Value"]))
        for k, v in acc.items():
            vals[k] = (sum(v) / len(v), len(v))
    waves = vals.get("SQ_WAVES", (0.0, 0))[0]
    print(f"| counter ({kernel}, {cells} work-items) | dispatches | mean per dispatch | per wave |")
    print("|---|---|---|---|")
    for k in sorted(vals):
        m, n = vals[k]
        pw = f"{m / waves:.1f}" if waves else "-"
        print(f"| {k} | {n} | {m:.0f} | {pw} |")


RHS_STEP = "kb_phase<2, "  # the batched RHS step: the first launch of every batched call


def batch(rows, precision, top):
    kb = [r for r in rows if r["name"].startswith("kb_") and (precision in r["name"] or "<" not in r["name"])]
    calls, cur = [], []
    for r in kb:
        if cur and r["name"].startswith(RHS_STEP):
            calls.append(cur)
            cur = []
        cur.append(r)
    if cur:
        calls.append(cur)
    print("## Batched calls (kb_* kernels; a call starts at its batched RHS step)\n")
    print("| call | wall ms (profiled) | kernel ms | launches |")
    print("|---|---|---|---|")
    for q, c in enumerate(calls):
        print(f"| {q} | {(c[-1]['e'] - c[0]['s']) / 1e6:.2f} | {sum(r['d'] for r in c) / 1e6:.2f} | {len(c)} |")
    c = calls[-1]
    agg = collections.defaultdict(float)
    cnt = collections.defaultdict(int)
    for r in c:
        agg[r["name"]] += r["d"]
        cnt[r["name"]] += 1
    tot = sum(agg.values())
    print(f"\n## Last call by kernel: {tot / 1e6:.2f} ms kernel time\n")
    print("| kernel | launches | ms | mean us | share |")
    print("|---|---|---|---|---|")
    for k, v in sorted(agg.items(), key=lambda kv: -kv[1])[:top]:
        print(f"| {k} | {cnt[k]} | {v / 1e6:.2f} | {v / cnt[k] / 1e3:.1f} | {100 * v / tot:.1f}% |")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace", nargs="?")
    ap.add_argument("--batch", action="store_true")
    ap.add_argument("--precision", default="float")
    ap.add_argument("--fetch", default=None)
    ap.add_argument("--write", default=None)
    ap.add_argument("--cells", type=int, default=192 * 96 * 128)
    ap.add_argument("--top", type=int, default=25)
    ap.add_argument("--counters", nargs="*", default=None, help="--pmc counter CSVs (with --kernel)")
    ap.add_argument("--kernel", default=SMOOTHER)
    a = ap.parse_args()
    if a.counters:
        counters(a.counters, a.kernel, a.cells)
        return
    if a.fetch:
        traffic(a.fetch, a.write, a.cells, a.top)
        return
    rows = load_trace(a.trace)
    if a.batch:
        batch(rows, a.precision, a.top)
        return
    by = collections.defaultdict(list)
    for r in rows:
        by[(r["name"], r["grid"])].append(r)
    print("## Kernels by total time (all launches in the trace)\n")
    print("| kernel | grid (work-items) | launches | mean us | min us | total ms | VGPR | scratch |")
    print("|---|---|---|---|---|---|---|---|")
    for (k, g), rs in sorted(by.items(), key=lambda kv: -sum(r["d"] for r in kv[1]))[:a.top]:
        ds = [r["d"] for r in rs]
        print(f"| {k} | {g[0]}x{g[1]}x{g[2]} | {len(ds)} | {sum(ds) / len(ds) / 1e3:.2f} | {min(ds) / 1e3:.2f} | "
              f"{sum(ds) / 1e6:.3f} | {rs[0]['vgpr']} | {rs[0]['scratch']} |")
    sv = solves(rows)
    if sv:
        print("\n## Solves (RHS kernel start to divergence kernel end)\n")
        print("| solve | wall ms (profiled) | kernel ms | launches |")
        print("|---|---|---|---|")
        for q, (s0, s1, seg) in enumerate(sv):
            print(f"| {q} | {(s1['e'] - s0['s']) / 1e6:.3f} | {sum(r['d'] for r in seg) / 1e6:.3f} | {len(seg)} |")
        half = len(sv) // 2
        groups = [("first group", sv[1:half]), ("last group", sv[half + 1:])] if half >= 2 else [("all", sv)]
        for label, chosen in groups:
            if not chosen:
                continue
            agg = collections.defaultdict(float)
            cnt = collections.defaultdict(int)
            for _, _, seg in chosen:
                for r in seg:
                    key = (r["name"], r["grid"])
                    agg[key] += r["d"] / len(chosen)
                    cnt[key] += 1
            tot = sum(agg.values())
            print(f"\n## Per-solve breakdown, {label} of solves ({len(chosen)} solves): {tot / 1e6:.3f} ms kernel time\n")
            print("| kernel | grid | launches/solve | us/solve | share |")
            print("|---|---|---|---|---|")
            for (k, g), v in sorted(agg.items(), key=lambda kv: -kv[1])[:a.top]:
                print(f"| {k} | {g[0]}x{g[1]}x{g[2]} | {cnt[(k, g)] // len(chosen)} | {v / 1e3:.1f} | "
                      f"{100 * v / tot:.1f}% |")


if __name__ == "__main__":
    main()

"""Summarise a rocprofv3 kernel trace of the engine (CSV output).

    python -m vfs_v3p2_amd.profile_summary <run>_kernel_trace.csv [--pmc <run>_counter_collection.csv]

Prints, as markdown: per-(kernel, grid) launch counts and durations, the
per-solve kernel time of each solve (a solve runs from its RHS kernel to its
divergence kernel), and the per-solve breakdown by kernel for the first and the
last group of solves (the benchmark runs the mixed form first, then FP64).
With --pmc, also the mean counter value per dispatch for each (kernel, grid).
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
    starts = [r for r in rows if r["name"] == "k_rhs"]
    ends = [r for r in rows if r["name"] == "k_divergence"]
    out = []
    for a, b in zip(starts, ends):
        seg = [r for r in rows if r["s"] >= a["s"] and r["e"] <= b["e"]]
        out.append((a, b, seg))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("trace")
    ap.add_argument("--pmc", default=None)
    ap.add_argument("--top", type=int, default=25)
    a = ap.parse_args()
    rows = load_trace(a.trace)
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
    if a.pmc:
        acc = collections.defaultdict(list)
        for r in csv.DictReader(open(a.pmc)):
            g = (int(r["Grid_Size"]) if "Grid_Size" in r else 0)
            acc[(_short(r["Kernel_Name"]), r.get("Grid_Size", ""), r["Counter_Name"])].append(
                float(r["Counter_Value"]))
        print("\n## Counters (mean per dispatch)\n")
        print("| kernel | grid size | counter | dispatches | mean |")
        print("|---|---|---|---|---|")
        for (k, g, c), v in sorted(acc.items(), key=lambda kv: -sum(kv[1]))[:a.top]:
            print(f"| {k} | {g} | {c} | {len(v)} | {sum(v) / len(v):.1f} |")


if __name__ == "__main__":
    main()

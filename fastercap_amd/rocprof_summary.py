"""Summaries of rocprofv3 CSV output for this library's kernels.

    python -m fastercap_amd.rocprof_summary stats KERNEL_STATS.csv [--top N] [--group PREFIX ...]
        per-kernel totals (time, calls, average) and, for every PREFIX, the sum over
        all kernels whose name starts with it (template instantiations of one kernel)
    python -m fastercap_amd.rocprof_summary pmc COUNTER_COLLECTION.csv COUNTER PREFIX [--scale F]
        mean / min / max / total of COUNTER per dispatch of the kernels starting with
        PREFIX, optionally multiplied by a calibration factor F
"""
from __future__ import annotations

import argparse
import csv
import sys


def kernel_stats(path: str):
    with open(path) as f:
        rows = list(csv.DictReader(f))
    out = []
    for r in rows:
        out.append({
            "name": r["Name"],
            "calls": int(r["Calls"]),
            "total_ms": float(r["TotalDurationNs"]) / 1e6,
            "avg_us": float(r["AverageNs"]) / 1e3,
            "pct": float(r["Percentage"]),
        })
    out.sort(key=lambda r: -r["total_ms"])
    return out


def group(stats, prefix: str):
    sel = [r for r in stats if r["name"].startswith(prefix)]
    calls = sum(r["calls"] for r in sel)
    tot = sum(r["total_ms"] for r in sel)
    return {"prefix": prefix, "kernels": len(sel), "calls": calls, "total_ms": tot,
            "avg_us": 1e3 * tot / calls if calls else 0.0}


def pmc(path: str, counter: str, prefix: str, scale: float = 1.0):
    with open(path) as f:
        vals = [float(r["Counter_Value"]) * scale for r in csv.DictReader(f)
                if r["Counter_Name"] == counter and r["Kernel_Name"].startswith(prefix)]
    if not vals:
        return {"counter": counter, "prefix": prefix, "dispatches": 0}
    return {"counter": counter, "prefix": prefix, "dispatches": len(vals), "mean": sum(vals) / len(vals),
            "min": min(vals), "max": max(vals), "total": sum(vals), "scale": scale}


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("--top", type=int, default=15)
    a.add_argument("--group", action="append", default=[])
    b = sub.add_parser("pmc")
    b.add_argument("csv")
    b.add_argument("counter")
    b.add_argument("prefix")
    b.add_argument("--scale", type=float, default=1.0)
    args = ap.parse_args(argv)
    if args.cmd == "stats":
        st = kernel_stats(args.csv)
        print("| total ms | % | calls | avg us | kernel |")
        print("|---:|---:|---:|---:|---|")
        for r in st[: args.top]:
            print(f"| {r['total_ms']:.1f} | {r['pct']:.1f} | {r['calls']} | {r['avg_us']:.1f} | `{r['name'][:110]}` |")
        for p in args.group:
            g = group(st, p)
            print(f"\ngroup `{p}*`: {g['kernels']} instantiations, {g['calls']} calls, {g['total_ms']:.1f} ms, "
                  f"average {g['avg_us']:.1f} us per launch")
    else:
        r = pmc(args.csv, args.counter, args.prefix, args.scale)
        print(r)
    return 0


if __name__ == "__main__":
    sys.exit(main())

"""Summaries of rocprofv3 CSV output for this library's kernels.

    python -m fastercap_amd.rocprof_summary stats KERNEL_STATS.csv|RESULTS.db [--top N] [--group PREFIX ...]
        per-kernel totals (time, calls, average) and, for every PREFIX, the sum over
        all kernels whose name starts with it (template instantiations of one kernel)
    python -m fastercap_amd.rocprof_summary pmc COUNTER_COLLECTION.csv COUNTER PREFIX [--scale F]
        mean / min / max / total of COUNTER per dispatch of the kernels starting with
        PREFIX, optionally multiplied by a calibration factor F
    python -m fastercap_amd.rocprof_summary window RESULTS.db --marker NAME --skip S [--steps K] [--gap-ms GAP]
        kernel trace of a rocprofv3 database (default rocpd output) cut to a window:
        the window opens at the (S+1)-th dispatch of kernel NAME (a prefix) that
        starts a burst (the first dispatch more than GAP ms, default 20, after the
        previous dispatch of that kernel) and spans K such bursts; per kernel
        group: device time per step, calls per step, and the window's wall time
        per step against the summed kernel time (the idle share)
"""
from __future__ import annotations

import argparse
import csv
import sys


def kernel_stats(path: str):
    if path.endswith(".db"):
        return _db_kernel_stats(path)
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


def _db_kernel_stats(db: str):
    """The same per-kernel totals from a rocprofv3 database (the default rocpd
    output of --kernel-trace --stats)."""
    import sqlite3

    con = sqlite3.connect(db)
    per = {}
    for name, a, b in con.execute("select name, start, end from kernels"):
        e = per.setdefault(name, [0, 0.0])
        e[0] += 1
        e[1] += (b - a) / 1e6
    tot = sum(v[1] for v in per.values()) or 1.0
    out = [{"name": n, "calls": v[0], "total_ms": v[1], "avg_us": 1e3 * v[1] / v[0], "pct": 100.0 * v[1] / tot}
           for n, v in per.items()]
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


def _short(name: str) -> str:
    """kernel name without the argument list, templates kept"""
    depth, out = 0, []
    for ch in name:
        if ch == "(" and depth == 0:
            break
        if ch == "<":
            depth += 1
        elif ch == ">":
            depth -= 1
        out.append(ch)
    return "".join(out).replace("void ", "")


def window(db: str, marker: str, skip: int, steps: int, gap_ms: float = 20.0):
    import sqlite3

    con = sqlite3.connect(db)
    rows = con.execute("select name, start, end from kernels order by start").fetchall()
    starts = [r[1] for r in rows if r[0].startswith(marker) or _short(r[0]).startswith(marker)]
    bursts, last = [], None
    for t in starts:
        if last is None or t - last > gap_ms * 1e6:
            bursts.append(t)
        last = t
    if len(bursts) < skip + 1:
        raise SystemExit(f"only {len(bursts)} bursts of {marker}")
    t0 = bursts[skip]
    t1 = bursts[skip + steps] if len(bursts) > skip + steps else max(r[2] for r in rows)
    sel = [r for r in rows if t0 <= r[1] < t1]
    per = {}
    for name, a, b in sel:
        k = _short(name)
        e = per.setdefault(k, [0, 0.0])
        e[0] += 1
        e[1] += (b - a) / 1e6
    busy = sum(v[1] for v in per.values())
    wall = (max(r[2] for r in sel) - t0) / 1e6
    n = max(1, min(steps, len(bursts) - skip))
    return {"steps": n, "wall_ms_per_step": wall / n, "kernel_ms_per_step": busy / n,
            "kernels": sorted(((k, v[0] / n, v[1] / n) for k, v in per.items()), key=lambda x: -x[2])}


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
    c = sub.add_parser("window")
    c.add_argument("db")
    c.add_argument("--marker", required=True)
    c.add_argument("--skip", type=int, default=0)
    c.add_argument("--steps", type=int, default=1)
    c.add_argument("--top", type=int, default=40)
    c.add_argument("--gap-ms", type=float, default=20.0)
    args = ap.parse_args(argv)
    if args.cmd == "window":
        w = window(args.db, args.marker, args.skip, args.steps, args.gap_ms)
        print(f"{w['steps']} steps: wall {w['wall_ms_per_step']:.2f} ms per step, kernels "
              f"{w['kernel_ms_per_step']:.2f} ms per step\n")
        print("| ms per step | % of kernel time | calls per step | kernel |")
        print("|---:|---:|---:|---|")
        for k, calls, ms in w["kernels"][: args.top]:
            print(f"| {ms:.3f} | {100 * ms / w['kernel_ms_per_step']:.1f} | {calls:.1f} | `{k[:120]}` |")
        return 0
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

"""Per-kernel duration summary of a rocprofv3 --kernel-trace run, as text:
kernel, launches, mean/min/max duration, total.  Reads either the rocpd
SQLite output (``*_results.db``) or the CSV kernel trace
(``--output-format csv`` -> ``*_kernel_trace.csv``).

    python -m frehdc_amd.rocpd_summary <results.db | kernel_trace.csv> [more ...] [--cmd "..."] [--active-us T]

Several traces (one process per precision form, say) are pooled.

``--active-us T`` adds a second table over the launches longer than T us
only: the PCG's chunked launches include iterations that exit at once after
convergence, and the working launches are the ones the roofline is about.

``--timeline K`` (CSV traces) adds the kernel sequence of the last K steps
of the first trace (a step starts at each k_predictor launch): start,
duration and the gap before each launch, in microseconds.
"""
from __future__ import annotations

import argparse
import collections
import csv
import sqlite3


def durations(path: str):
    """{kernel name: [duration ns, ...]} in launch order."""
    out = collections.defaultdict(list)
    if path.endswith(".csv"):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                out[r["Kernel_Name"]].append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    else:
        con = sqlite3.connect(path)
        for name, d in con.execute("select name, end-start from kernels order by start"):
            out[name].append(int(d))
        con.close()
    return out


def launches(path: str):
    """[(start ns, end ns, kernel name)] of a CSV kernel trace, by start."""
    rows = []
    with open(path, newline="") as f:
        for r in csv.DictReader(f):
            rows.append((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]))
    rows.sort()
    return rows


def timeline(path: str, steps: int) -> str:
    rows = launches(path)
    starts = [q for q, r in enumerate(rows) if "k_predictor" in r[2]]
    sel = starts[-(steps + 1):]
    out = []
    for a, b in zip(sel[:-1], sel[1:]):
        t0 = rows[a][0]
        busy = sum(e - s for s, e, _ in rows[a:b])
        out.append(f"step: {(rows[b][0] - t0) / 1e3:.1f} us from predictor to predictor, {b - a} kernels, "
                   f"{busy / 1e3:.1f} us busy")
        prev = None
        for s, e, n in rows[a:b]:
            gap = (s - prev) / 1e3 if prev is not None else 0.0
            short = n.split("(")[0].replace("void ", "").replace("frehd::", "")
            out.append(f"  {(s - t0) / 1e3:9.1f} {(e - s) / 1e3:9.1f}  gap {gap:7.1f}  {short}")
            prev = e
    return "\n".join(out)


def table(d, min_ns=0):
    rows = []
    for name, v in d.items():
        v = [x for x in v if x > min_ns]
        if v:
            rows.append((name, len(v), sum(v) / len(v), min(v), max(v), sum(v)))
    rows.sort(key=lambda r: -r[5])
    tot = sum(r[5] for r in rows)
    lines = [f"{'kernel':55s} {'launches':>8s} {'mean_us':>10s} {'min_us':>10s} {'max_us':>10s} "
             f"{'total_ms':>10s} {'pct':>6s}"]
    for name, n, avg, mn, mx, s in rows:
        lines.append(f"{name[:55]:55s} {n:8d} {avg / 1e3:10.3f} {mn / 1e3:10.3f} {mx / 1e3:10.3f} "
                     f"{s / 1e6:10.3f} {100 * s / tot:6.1f}")
    return "\n".join(lines)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("paths", nargs="+")
    ap.add_argument("--cmd", default="")
    ap.add_argument("--active-us", type=float, default=0.0)
    ap.add_argument("--timeline", type=int, default=0)
    a = ap.parse_args()
    d = collections.defaultdict(list)
    for p in a.paths:
        for k, v in durations(p).items():
            d[k].extend(v)
    if a.cmd:
        print(f"command: {a.cmd}")
    print("all launches")
    print(table(d))
    if a.active_us > 0:
        print(f"\nlaunches longer than {a.active_us:g} us")
        print(table(d, int(a.active_us * 1e3)))
    if a.timeline > 0 and a.paths[0].endswith(".csv"):
        print(f"\nthe last {a.timeline} steps (start us, duration us, gap us before the launch)")
        print(timeline(a.paths[0], a.timeline))


if __name__ == "__main__":
    main()

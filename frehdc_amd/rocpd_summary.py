"""Per-kernel duration summary of a rocprofv3 --kernel-trace run (the rocpd
SQLite output), as text: kernel, launches, mean/min/max duration, total.

    python -m frehdc_amd.rocpd_summary <results.db> [--cmd "the profiled command"]
"""
from __future__ import annotations

import argparse
import sqlite3


def summarize(db: str):
    con = sqlite3.connect(db)
    rows = con.execute(
        "select name, count(*), avg(end-start), min(end-start), max(end-start), sum(end-start) "
        "from kernels group by name order by sum(end-start) desc").fetchall()
    con.close()
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("db")
    ap.add_argument("--cmd", default="")
    a = ap.parse_args()
    rows = summarize(a.db)
    tot = sum(r[5] for r in rows)
    if a.cmd:
        print(f"command: {a.cmd}")
    print(f"{'kernel':55s} {'launches':>8s} {'mean_us':>10s} {'min_us':>10s} {'max_us':>10s} {'total_ms':>10s} {'pct':>6s}")
    for name, n, avg, mn, mx, s in rows:
        print(f"{name[:55]:55s} {n:8d} {avg / 1e3:10.3f} {mn / 1e3:10.3f} {mx / 1e3:10.3f} {s / 1e6:10.3f} {100 * s / tot:6.1f}")


if __name__ == "__main__":
    main()

"""HBM traffic per kernel dispatch from two rocprofv3 counter passes
(FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by what the same
counter reads for a device-to-device copy of known size (frehdc_amd.copy_probe
run under the same counter).

    python -m frehdc_amd.pmc_summary --fetch <FETCH csv> [...] --write <WRITE csv> [...]
        --fetch-copy <copy FETCH csv> --write-copy <copy WRITE csv> --copy-mib 1024
        [--json out.json] [--cmd "..."]

Dispatches that move less than 10 % of the kernel's largest dispatch are
left out of the means (solver launches past convergence exit at once).
Several --fetch / --write files (one process per precision form, say) are
pooled.
"""
from __future__ import annotations

import argparse
import collections
import csv
import json
import re


def per_dispatch(paths):
    out = collections.defaultdict(list)
    for path in ([paths] if isinstance(paths, str) else paths):
        with open(path, newline="") as f:
            for r in csv.DictReader(f):
                out[r["Kernel_Name"]].append(float(r["Counter_Value"]))
    return out


def copy_factor(path, mib):
    vals = [v for k, vs in per_dispatch(path).items() if "copyBuffer" in k for v in vs]
    if not vals:
        raise SystemExit(f"no copy dispatch in {path}")
    return (mib * 1024.0) / (sum(vals) / len(vals))


def short(name):
    n = name.split("(")[0].replace("void ", "").replace("frehd::", "")
    n = re.sub(r"k_cg1<false(, \d)?>", "k_cg1_f32", n)
    n = re.sub(r"k_cg1<true(, \d)?>", "k_cg1_f32_setup", n)
    n = re.sub(r"k_pcg1<false(, \d)?>", "k_pcg1", n)
    n = re.sub(r"k_pcg1<true(, \d)?>", "k_pcg1_setup", n)
    return n.replace("<double>", "").replace("<float>", "_f32")


def summarize(fetch, write, ffac, wfac):
    F, W = per_dispatch(fetch), per_dispatch(write)
    rows = {}
    for k in F:
        if k not in W or "rocclr" in k:
            continue
        f, w = F[k], W[k]
        tot = [a + b for a, b in zip(f, w)]
        big = max(tot) if tot else 0
        idx = [q for q, t in enumerate(tot) if t >= 0.1 * big] if big > 0 else []
        if not idx:
            continue
        mf = sum(f[q] for q in idx) / len(idx) * ffac
        mw = sum(w[q] for q in idx) / len(idx) * wfac
        rows[short(k)] = {"fetch_bytes": mf * 1024, "write_bytes": mw * 1024,
                          "bytes": (mf + mw) * 1024, "dispatches": len(idx)}
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--fetch", required=True, nargs="+")
    ap.add_argument("--write", required=True, nargs="+")
    ap.add_argument("--fetch-copy", required=True)
    ap.add_argument("--write-copy", required=True)
    ap.add_argument("--copy-mib", type=float, default=1024)
    ap.add_argument("--json")
    ap.add_argument("--cmd", default="")
    a = ap.parse_args()
    ff = copy_factor(a.fetch_copy, a.copy_mib)
    wf = copy_factor(a.write_copy, a.copy_mib)
    rows = summarize(a.fetch, a.write, ff, wf)
    if a.cmd:
        print(f"command: {a.cmd}")
    print(f"calibration: a {a.copy_mib:g} MiB device copy reads FETCH_SIZE x{ff:.3f}, WRITE_SIZE x{wf:.3f} of its bytes")
    print(f"{'kernel':28s} {'dispatches':>10s} {'fetch_MB':>10s} {'write_MB':>10s} {'total_MB':>10s}")
    for k, r in sorted(rows.items(), key=lambda kv: -kv[1]["bytes"]):
        print(f"{k:28s} {r['dispatches']:10d} {r['fetch_bytes'] / 1e6:10.2f} {r['write_bytes'] / 1e6:10.2f} "
              f"{r['bytes'] / 1e6:10.2f}")
    if a.json:
        with open(a.json, "w") as f:
            json.dump({"fetch_factor": ff, "write_factor": wf, "kernels": rows}, f, indent=1)


if __name__ == "__main__":
    main()

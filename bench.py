"""Throughput of the semi-implicit free-surface step on one MI355X.

Workload (default): the wetting/drying marsh, 1024 x 1024 cells, dx = 2 m,
dt = 0.5 s, fBm bed (Hurst 0.7) in [-1.0, +1.5] m, tide 0.6 m / 44712 s on
the YP row, Manning n ~ U(0.02, 0.04) per cell, one scalar, Jacobi PCG to
1e-10, all FP64.  A step is one oneCompleteStep + monitorCFL over the whole
grid; value = interior cells x timed steps / wall time of the timed region.

Prints one JSON line (rank 0).  --gpus above 1 is refused: the build runs on
one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic bytes per interior cell (FP64): one PCG iteration reads the two
# face-coefficient arrays and reads+writes x, r and p
PCG_ITER_BYTES = 8 * 8
# predictor + assembly and update + one scalar (11 + 12 array accesses + 3 for
# the solver start), frozen per cell-step
STEP_BYTES_FP64 = 8 * 26
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scen_builder, n_steps):
    """Time the reference's own step functions (oracle/_ref) on the host, one
    core, on a bounded sample: the same grid for n_steps steps.  Falls back to
    the C restatement (same algorithm, bit-identical results) when the
    reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    scen = scen_builder()
    kind = "reference"
    ntide = scen.cfg["tideLocLengthP"]
    note = ""
    try:
        if not pyoracle.have_reference():
            raise FileNotFoundError
        if ntide > 1000:
            # the reference Config holds at most 1000 tidal cells
            scen.cfg["tideLocLengthP"] = 1000
            scen.tideLocP = scen.tideLocP[:1000].copy()
            note = ", tide on the first 1000 cells of the YP row (reference limit)"
        scen.manningCell = None      # the reference has one global Manning n
        model = pyoracle.Reference(scen)
    except (FileNotFoundError, ValueError, OSError):
        kind = "port"
        scen = scen_builder()
        model = pyoracle.Oracle(scen)
    model.step(1)                    # first step outside the timing
    t0 = time.perf_counter()
    model.step(n_steps)
    dt = time.perf_counter() - t0
    val = scen.N * n_steps / dt
    return {"value": val, "unit": "cell-steps/s", "cores": 1, "kind": kind,
            "sample": f"{scen.nx}x{scen.ny} marsh, steps 2-{n_steps + 1}, Jacobi PCG, FP64, "
                      f"{dt:.1f} s on 1 host core{note}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    from frehdc_amd import cases, device
    import numpy as np

    total = args.warmup + args.steps + 8

    def builder():
        return cases.marsh(n=args.size, Nt=max(total, 2000))

    scen = builder()
    log(f"bench: marsh {args.size}^2, warmup {args.warmup}, steps {args.steps}")
    dev = device.DeviceScenario(scen)
    dev.step(args.warmup)
    it0 = dev.stats()["pcgIterTotal"]
    import ctypes
    t0 = time.perf_counter()
    dev.step(args.steps)             # returns after the stream is synchronised
    t1 = time.perf_counter()
    st = dev.stats()
    iters = (st["pcgIterTotal"] - it0) / args.steps
    sec = t1 - t0
    value = scen.N * args.steps / sec
    log(f"bench: {sec * 1e3 / args.steps:.3f} ms/step, {iters:.1f} PCG iterations/step")

    # roofline of the dominant kernel pair (one PCG iteration) from HIP events
    # on the handle's stream, over a few further steps
    dev.enable_timing(True)
    dev.step(4)
    kt = dev.kernel_times()
    dev.enable_timing(False)
    log("kernel times (ms/launch, launches):", json.dumps(kt))
    t_pq = kt.get("k_pcg_dir", (0, 0))[0]
    t_up = kt.get("k_pcg_res", (0, 0))[0]
    t_iter_ms = t_pq + t_up
    alg_bytes = PCG_ITER_BYTES * scen.N
    achieved = alg_bytes / (t_iter_ms * 1e-3) / 1e9 if t_iter_ms > 0 else 0.0
    roof = {"bound": "hbm", "kernel": "PCG iteration (k_pcg_dir + k_pcg_res)",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": alg_bytes, "ms_per_iteration": round(t_iter_ms, 5)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(builder, args.cpu_steps)
        log("cpu baseline:", json.dumps(cpu))

    out = {
        "metric": "cell-steps/s mixed vs FP64 at eta err<=1e-6 m; p50/p99 scenario latency",
        "value": value, "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": sec * 1e3 / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": f"marsh {args.size}x{args.size}, dx=2 m, dt=0.5 s, plain grid, "
                               "FP64 path, Jacobi PCG eps=1e-10", "cells": scen.N,
                   "pcg_iterations_per_step": round(iters, 2), "precision": "fp64",
                   "subgrid": "not built (plain grid)"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

"""Throughput of the semi-implicit free-surface step on one MI355X.

Workload (default): the wetting/drying marsh, 1024 x 1024 cells, dx = 2 m,
dt = 0.5 s, fBm bed (Hurst 0.7) in [-1.0, +1.5] m, tide 0.6 m / 44712 s on
the YP row, Manning n ~ U(0.02, 0.04) per cell, one scalar, Jacobi PCG to
1e-10.  Forms timed on the same scenario: fp64 (every operation FP64, the
reference's cold start), mixed_cold (FP32 operator and inner PCG, FP64
refinement to the same |b - Ax| < eps |b| test, cold start) and mixed_warm
(the same, refinement started from the current surface).  A step is one
oneCompleteStep + monitorCFL over the whole grid; value = interior cells x
timed steps / wall time of the timed region, of the last mixed form; the
eta / u / v errors of each mixed form against fp64 after the same steps are
reported beside it.

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

# algorithmic bytes per interior cell: one PCG iteration reads the two
# face-coefficient arrays and reads+writes x, r and p (8 accesses), FP64 or
# FP32 (the inner solve of the mixed form)
PCG_ITER_BYTES = 8 * 8
PCG_ITER_BYTES_F32 = 4 * 8
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


def run_form(scen, steps, warmup, **over):
    """Time one precision form: warmup steps, then `steps` timed steps (the
    call returns after the stream is synchronised), then 4 steps with
    per-kernel HIP event timing.  Returns the numbers and the final state."""
    from frehdc_amd import device
    dev = device.DeviceScenario(scen, **over)
    dev.step(warmup)
    it0 = dev.stats()["pcgIterTotal"]
    t0 = time.perf_counter()
    dev.step(steps)
    t1 = time.perf_counter()
    st = dev.stats()
    fields = dev.download(["surf", "uuXP", "vvYP", "depth"])
    dev.enable_timing(True)
    dev.step(4)
    kt = dev.kernel_times()
    dev.enable_timing(False)
    dev.close()
    sec = t1 - t0
    return {"sec": sec, "ms_per_step": sec * 1e3 / steps, "value": scen.N * steps / sec,
            "iters": (st["pcgIterTotal"] - it0) / steps, "sweeps": st["refineLast"],
            "relres": st["relResLast"], "kt": kt, "fields": fields}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,mixed_cold,mixed_warm",
                    help="precision forms to time (the headline value is the last mixed form)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    from frehdc_amd import abi, cases
    import numpy as np

    total = args.warmup + args.steps + 8

    def builder():
        return cases.marsh(n=args.size, Nt=max(total, 2000))

    scen = builder()
    N = scen.N
    forms = {"fp64": {}, "mixed_cold": {"precision": abi.PREC_MIXED, "warmStart": 0},
             "mixed_warm": {"precision": abi.PREC_MIXED, "warmStart": 1}}
    names = [f for f in args.forms.split(",") if f]
    res = {}
    for f in names:
        log(f"bench: marsh {args.size}^2 {f}, warmup {args.warmup}, steps {args.steps}")
        r = run_form(scen, args.steps, args.warmup, **forms[f])
        log(f"bench: {f}: {r['ms_per_step']:.3f} ms/step, {r['iters']:.1f} PCG iterations/step, "
            f"{r['sweeps']} refinement sweeps")
        log(f"kernel times {f} (ms/launch, launches):",
            json.dumps({k: [round(v[0], 5), v[1]] for k, v in r["kt"].items()}))
        res[f] = r

    def err(f, k):
        a, b = res[f]["fields"][k][:N], res["fp64"]["fields"][k][:N]
        return float(np.max(np.abs(a - b)))

    forms_out = {}
    for f, r in res.items():
        o = {"cell_steps_per_s": r["value"], "ms_per_step": r["ms_per_step"],
             "pcg_iterations_per_step": round(r["iters"], 2), "refinement_sweeps": r["sweeps"],
             "rel_residual_last": r["relres"]}
        if f != "fp64" and "fp64" in res:
            o["max_abs_eta_err_m"] = err(f, "surf")
            o["max_abs_u_err"] = err(f, "uuXP")
            o["max_abs_v_err"] = err(f, "vvYP")
            o["wet_mask_mismatches"] = int(np.count_nonzero(
                (r["fields"]["depth"][:N] > 0) != (res["fp64"]["fields"]["depth"][:N] > 0)))
        forms_out[f] = o
    mixed = [f for f in names if f.startswith("mixed")]
    head = mixed[-1] if mixed else names[-1]
    hr = res[head]

    # roofline of the dominant kernel pair of the headline form (one PCG
    # iteration), from HIP events on the handle's stream
    kt = hr["kt"]
    if head == "fp64":
        kd, kr, per_cell, dtype = "k_pcg_dir", "k_pcg_res", PCG_ITER_BYTES, "f64"
    else:
        kd, kr, per_cell, dtype = "k_pcg_dir_f32", "k_pcg_res_f32", PCG_ITER_BYTES_F32, "f32+f64"
    t_iter_ms = kt.get(kd, (0, 0))[0] + kt.get(kr, (0, 0))[0]
    alg_bytes = per_cell * N
    achieved = alg_bytes / (t_iter_ms * 1e-3) / 1e9 if t_iter_ms > 0 else 0.0
    roof = {"bound": "hbm", "kernel": f"PCG iteration ({kd} + {kr})",
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": alg_bytes, "ms_per_iteration": round(t_iter_ms, 5)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(builder, args.cpu_steps)
        log("cpu baseline:", json.dumps(cpu))

    out = {
        "metric": "cell-steps/s mixed vs FP64 at eta err<=1e-6 m; p50/p99 scenario latency",
        "value": hr["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": hr["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": f"marsh {args.size}x{args.size}, dx=2 m, dt=0.5 s, plain grid, "
                               f"Jacobi PCG eps=1e-10, form {head}", "cells": N,
                   "form": head, "forms": forms_out,
                   "subgrid": "plain grid (subgrid path is parity-tested, not benched here)"},
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

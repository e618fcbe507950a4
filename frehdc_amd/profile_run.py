"""A workload stepped under a profiler: the headline marsh (1024^2, subgrid
8x8 from device tables), the large channel (4096^2, subgrid 4x4 on the fly,
three scalars) or the batched ensemble (--members members of 2048^2, plain
grid, one batch), in the given precision forms, each autotuned after 3
steps (or given a fixed solver tiling from its first step), then 5 more
steps and `--steps` steps timed.  Prints one line per form.

    rocprofv3 --kernel-trace --stats ... -- python3 -m frehdc_amd.profile_run --workload marsh --forms fp64,mixed_warm

Every workload here has at least 2^18 cells, so its steps run as streamed
launches (no captured graphs) unless FREHD_GRAPHS=1 in the environment
forces graph replays.
"""
from __future__ import annotations

import argparse
import time

from . import abi, cases
from .device import DeviceScenario

FORMS = {"fp64": {}, "fp64_warm": {"warmStart": 2}, "mixed_cold": {"precision": abi.PREC_MIXED, "warmStart": 0},
         "mixed_warm": {"precision": abi.PREC_MIXED, "warmStart": 2}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", choices=["marsh", "channel", "ensemble"], default="marsh")
    ap.add_argument("--size", type=int, default=0, help="grid side (default 1024 marsh, 4096 channel, "
                                                          "2048 ensemble)")
    ap.add_argument("--members", type=int, default=64, help="ensemble members in the batch")
    ap.add_argument("--forms", default="fp64,mixed_warm")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--tiling", default="",
                    help="H:remap[:cols_per_lane] solver tiling instead of the autotuner (whose trial launches "
                         "would enter a profile's averages): one for every form, or one per form, comma-separated")
    a = ap.parse_args()
    n = a.size or {"marsh": 1024, "channel": 4096, "ensemble": 2048}[a.workload]
    t0 = time.perf_counter()
    if a.workload == "ensemble":
        return ensemble(a, n, t0)
    if a.workload == "marsh":
        sc = cases.marsh_subgrid(n=n, r=8, Nt=a.steps + 40, device_tables=True)
    else:
        sc = cases.channel_subgrid(n=n, r=4, Nt=a.steps + 40)
    print(f"{sc.name}: built in {time.perf_counter() - t0:.1f} s", flush=True)
    forms = [x for x in a.forms.split(",") if x]
    tilings = [x for x in a.tiling.split(",") if x]
    if len(tilings) == 1:
        tilings *= len(forms)
    if tilings and len(tilings) != len(forms):
        raise SystemExit("--tiling: give one H:remap, or one per form")
    for q, f in enumerate(forms):
        d = DeviceScenario(sc, **FORMS[f])
        if tilings:
            H, remap, lanes = ([int(x) for x in tilings[q].split(":")] + [1])[:3]
            if not f.startswith("mixed"):
                lanes = min(lanes, 2)          # k_pcg1 takes 1 or 2 columns per lane
            d.set_pcg_tiling(H, remap, lanes)
            tiling = {"rows_per_tile": H, "xcd_remap": remap, "cols_per_lane": lanes}
            d.step(3)
        else:
            d.step(3)
            tiling = d.autotune(10)
        d.step(5)
        t0 = time.perf_counter()
        d.step(a.steps)
        ms = (time.perf_counter() - t0) * 1e3 / a.steps
        print(f"{sc.name} {f}: {ms:.3f} ms/step, tiling {tiling}", flush=True)
        d.close()


def ensemble(a, n, t0):
    """The batched ensemble: members of one fBm base plus a per-member
    perturbation, one batch per form."""
    import numpy as np
    from .device import Batch
    base = cases.fbm_surface(n, n, 0.7, np.random.default_rng(5), -1.0, 1.5)
    members = [cases.ensemble_member(m, n=n, Nt=a.steps + 40, base=base) for m in range(a.members)]
    print(f"ensemble {a.members} x {n}^2: built in {time.perf_counter() - t0:.1f} s", flush=True)
    forms = [x for x in a.forms.split(",") if x]
    tilings = [x for x in a.tiling.split(",") if x]
    if len(tilings) == 1:
        tilings *= len(forms)
    if tilings and len(tilings) != len(forms):
        raise SystemExit("--tiling: give one H:remap, or one per form")
    for q, f in enumerate(forms):
        devs = [DeviceScenario(sc, **FORMS[f]) for sc in members]
        if tilings:
            H, remap, lanes = ([int(x) for x in tilings[q].split(":")] + [1])[:3]
            if not f.startswith("mixed"):
                lanes = min(lanes, 2)
            for d in devs:
                d.set_pcg_tiling(H, remap, lanes)
            tiling = {"rows_per_tile": H, "xcd_remap": remap, "cols_per_lane": lanes}
        b = Batch(len(devs))
        for d in devs:
            b.add(d)
        b.try_step(3)
        if not tilings:
            tiling = b.autotune(10)
        b.try_step(5)
        t1 = time.perf_counter()
        b.try_step(a.steps)
        ms = (time.perf_counter() - t1) * 1e3 / a.steps
        print(f"ensemble {a.members} x {n}^2 {f}: {ms:.3f} ms/step, tiling {tiling}", flush=True)
        b.close()
        for d in devs:
            d.close()


if __name__ == "__main__":
    main()

"""A workload stepped under a profiler: the headline marsh (1024^2, subgrid
8x8 from device tables) or the large channel (4096^2, subgrid 4x4 on the
fly, three scalars), in the given precision forms, each autotuned after 3
steps (or given a fixed solver tiling from its first step), then 5 more
steps and `--steps` steps timed.  Prints one line per form.

    rocprofv3 --kernel-trace --stats ... -- python3 -m frehdc_amd.profile_run --workload marsh --forms fp64,mixed_warm
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
    ap.add_argument("--workload", choices=["marsh", "channel"], default="marsh")
    ap.add_argument("--size", type=int, default=0, help="grid side (default 1024 marsh, 4096 channel)")
    ap.add_argument("--forms", default="fp64,mixed_warm")
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--tiling", default="",
                    help="H:remap[:cols_per_lane] solver tiling instead of the autotuner (whose trial launches "
                         "would enter a profile's averages): one for every form, or one per form, comma-separated")
    a = ap.parse_args()
    n = a.size or (1024 if a.workload == "marsh" else 4096)
    t0 = time.perf_counter()
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


if __name__ == "__main__":
    main()

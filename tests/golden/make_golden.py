"""Write the golden fixtures: the reference's own step functions (built by
oracle/Makefile into oracle/_ref/libfrehd_ref.so from the reference sources)
run on small synthetic scenarios; inputs and outputs are stored as .npz.

    python tests/golden/make_golden.py

The fixtures are data (inputs and the reference's outputs); the reference is
not needed to use them.
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "oracle")]

from frehdc_amd import cases  # noqa: E402
import pyoracle  # noqa: E402

FIELDS = ["surf", "surfOld", "uuXP", "vvYP", "uuYP", "vvXP", "depth", "depthXP", "depthYP",
          "EnXP", "EnYP", "DragXP", "DragYP", "CDXP", "CDYP", "Fuu", "Fvv", "cellV", "S", "Sm",
          "CFLx", "CFLy"]
SUB_FIELDS = ["V", "Z", "Vx", "Vy", "Zx", "Zy", "Nx", "Oy"]
CHECKPOINTS = [1, 5, 40]


def fast_marsh():
    sc = cases.marsh(n=48, dx=2.0, dt=0.5, Nt=400, amp=0.6)
    sc.manningCell = None            # the reference has one global Manning n
    sc.tideP[:] = cases.tide_series(400, 0.5, 0.6, period=1200.0)
    return sc


def inflow_basin():
    sc = cases.basin(nx=36, ny=32, Nt=100)
    # river inflow on three cells of row 0, with scalar concentration 5
    sc.inflowLoc = np.array([10, 11, 12], dtype=np.int32)
    sc.cfg["inflowLocLength"] = 3
    sc.inflow[:] = 2.0
    sc.cfg["inflowS"] = 5.0
    return sc


def marsh_subgrid():
    return cases.marsh_subgrid(n=12, r=4, Nt=60, period=1200.0, manning_cell=False, dsurf=0.05)


CASES = {
    "marsh12_subgrid4_jacobi": (marsh_subgrid, {}),
    "basin40x36_jacobi": (lambda: cases.basin(nx=40, ny=36, Nt=100), {}),
    "basin40x36_ssor": (lambda: cases.basin(nx=40, ny=36, Nt=100), {"precond": 1}),
    "marsh48_wetdry_jacobi": (fast_marsh, {}),
    "basin36x32_inflow_jacobi": (inflow_basin, {}),
}


def scenario_arrays(sc):
    sub = {}
    if sc.subgrid is not None:
        for k in ("edgeX", "edgeY", "V", "Z", "Vxp", "Vxm", "Vyp", "Vym", "Np", "Nm", "Op", "Om"):
            sub[f"sub_{k}"] = sc.subgrid[k]
    return {**sub,
        "in_bottomZ": sc.bottomZ, "in_tideP": sc.tideP, "in_tideM": sc.tideM,
        "in_inflow": sc.inflow, "in_tideLocP": sc.tideLocP, "in_tideLocM": sc.tideLocM,
        "in_inflowLoc": sc.inflowLoc,
        "in_S0cell": sc.S0cell if sc.S0cell is not None else np.zeros(0),
    }


def main():
    if not pyoracle.have_reference():
        raise SystemExit("oracle/_ref/libfrehd_ref.so is not built (make -C oracle ref)")
    for name, (builder, override) in CASES.items():
        sc = builder()
        cfg = dict(sc.cfg)
        cfg.update(override)
        ref = pyoracle.Reference(sc, **override)
        out = scenario_arrays(sc)
        out["cfg_json"] = np.frombuffer(json.dumps(cfg).encode(), dtype=np.uint8)
        out["surf0"] = np.array([sc.surf0])
        iters, vloss = [], []
        names = FIELDS + (SUB_FIELDS if sc.subgrid is not None else [])
        for n in names:
            out[f"init_{n}"] = ref.field(n).reshape(-1)
        t = 0
        for cp in CHECKPOINTS:
            while t < cp:
                rc = ref.step(1)
                assert rc == 0, (name, t, rc)
                iters.append(ref.pcg_iters)
                vloss.append(ref.vloss())
                t += 1
            for n in names:
                out[f"t{cp}_{n}"] = ref.field(n).reshape(-1)
        out["pcg_iters"] = np.array(iters, dtype=np.int32)
        out["vloss"] = np.array(vloss)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **out)
        print(f"{path}: {os.path.getsize(path) / 1024:.0f} KiB, iterations {iters[:5]}...")


if __name__ == "__main__":
    main()

"""Throughput of the semi-implicit free-surface step on one MI355X.

Workload (default): the wetting/drying marsh, 1024 x 1024 cells, dx = 2 m,
subgrid 8 x 8 pixels per cell (tables generated on the device from the
pixel bed, 250 levels of 0.01 m), dt = 0.5 s, fBm bed (Hurst 0.7) in
[-1.0, +1.5] m, tide 0.6 m / 44712 s on the YP row, Manning n ~
U(0.02, 0.04) per cell, one scalar (0, boundary value 1), Jacobi PCG to
1e-10.  Forms timed on the same scenario: fp64 (every operation FP64, the
reference's cold start), mixed_cold (FP32 operator and inner PCG, FP64
refinement to the same |b - Ax| < eps |b| test, cold start) and mixed_warm
(the same, refinement started from the extrapolated surface 2 eta^n -
eta^(n-1)).  Each form runs the solver-tiling autotuner after 3 steps and
before its warmup.  A step is one oneCompleteStep + monitorCFL over the
whole grid; value = interior cells x timed steps / wall time of the timed
region, of the last mixed form; the eta / u / v errors of each mixed form
against fp64 after the same steps are reported beside it.

Beside it (not the value): the large single domain (4096^2 channel, subgrid
4 x 4 evaluated on the fly, three scalars, river inflow; fp64 and
mixed_warm, with the solver roofline of each), the batched ensemble (64 x
2048^2) and the serving loop (48 requests), and the reference's CPU step on
a bounded sample.

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

# algorithmic bytes per interior cell of one solver iteration.  FP64 Jacobi
# PCG: the two face-coefficient arrays and the diagonal (z = r / |D| and the
# operator's centre term need it every iteration), x, r and p read and
# written: 9 accesses.  The inner FP32 solve of the mixed form runs on the
# Jacobi-scaled operator (unit diagonal): cX, cY, x, r, p: 8 accesses.
PCG_ITER_BYTES = 9 * 8
PCG_ITER_BYTES_F32 = 4 * 8
# the FP64 solve from a nonzero start: r = b - A x_0, R {b, x, D, cX, cY} W r
PCG_START_BYTES = 6 * 8
# predictor + assembly and update + one scalar, frozen per cell-step: FP64,
# 11 + 12 array accesses + 3 for the solver start, 8 B each; the mixed form
# keeps the explicit state at 4 B except eta, b and x (8 B): predictor R
# {eta 8, u, v, zb, zbx, zby} W {b 8, cx, cy, Enx, Eny} = 52 B, update R {x 8,
# eta 8, zb, zbx, zby, Enx, Eny, S} W {eta 8, u, v, S} = 60 B (its solver
# start is the first refinement sweep, counted with the sweeps)
STEP_BYTES_FP64 = 8 * 26
STEP_BYTES_MIXED = 52 + 60
# each scalar beyond the first: read and write its concentration
SCALAR_BYTES = 8 * 2
# one FP64 refinement sweep of the mixed form: residual R {x, b, cX, cY}
# (FP64) + W r (FP32), correction R {x, d} + W x
REFINE_SWEEP_BYTES = 8 * 6 + 4 * 2
HBM_PEAK_GBS = 8000.0


def step_bytes(cfg, mixed, k, m, warm=False):
    """Algorithmic HBM bytes of one cell-step: every array each phase needs,
    read or written once per phase (stencil halos free), at the width the
    form stores it (mixed: 4 B but for eta, b, x).  Subgrid adds the two
    bracketing levels of T tables, the level index (4 B, read and written)
    and the eight face / volume arrays written by the update and read by the
    predictor: T = 10, + 4 minimum face areas when useminA, + 4 minimum
    volumes when they can change a result (useminV = 2, or useminV = 1 with
    useminA: without the minimum areas the volume adjustment never fires,
    and the reference's unconditional reads of those tables are not
    counted).  k solver iterations per step (FP64 72 B, inner FP32 32 B per
    cell), m refinement sweeps, and the FP64 start residual from a nonzero
    start (warm)."""
    w = 4 if mixed else 8
    b = (STEP_BYTES_MIXED if mixed else STEP_BYTES_FP64) + 2 * w * max(cfg.get("nScalar", 1) - 1, 0)
    if cfg.get("useSubgrid", 0):
        minA = cfg.get("useminA", 0) == 1
        minV = cfg.get("useminV", 0) == 2 or (cfg.get("useminV", 0) == 1 and minA)
        T = 10 + (4 if minA else 0) + (4 if minV else 0)
        b += w * (2 * T + 16) + 8
    if mixed:
        b += PCG_ITER_BYTES_F32 * k + REFINE_SWEEP_BYTES * m
    else:
        b += PCG_ITER_BYTES * k + (PCG_START_BYTES if warm else 0)
    return b


def step_roofline(cfg, N, r, mixed, warm=False):
    """The whole step against HBM bandwidth: algorithmic bytes per cell-step
    (step_bytes, with the measured iterations and the refinement sweeps of
    the kernel-timed steps) x cells / the timed ms per step."""
    m = r["kt"].get("k_refine", (0, 0))[1] / 4.0 if mixed else 0.0
    per_cell = step_bytes(cfg, mixed, r["iters"], m, warm)
    achieved = per_cell * N / (r["ms_per_step"] * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": "whole step (all kernels)", "bytes_per_cell_step": round(per_cell, 1),
            "pcg_iterations_per_step": round(r["iters"], 2), "refinement_sweeps_per_step": m,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4)}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(n, r, budget_s, procs=1):
    """Time the reference's own step functions (oracle/_ref, one host core)
    on a bounded sample of the workload: an n x n marsh from the same
    generator (fBm bed, r x r subgrid, tables generated on the device and
    downloaded, uniform Manning n since the reference has no per-cell n),
    stepping until about budget_s seconds have passed; with procs > 1, then
    as many processes at once, each on its own copy of the sample (the host
    cores' aggregate).  Falls back to the C restatement (same algorithm,
    bit-identical results) when the reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from frehdc_amd import cases, device
    scen = cases.marsh_subgrid(n=n, r=r, Nt=4000, device_tables=True, manning_cell=False)
    d = device.DeviceScenario(scen)
    tabs = d.tables()
    d.close()
    sg = {k: np.ascontiguousarray(v.reshape(-1)) for k, v in tabs.items()}
    sg["edgeX"], sg["edgeY"] = scen.subgrid["edgeX"], scen.subgrid["edgeY"]
    scen.subgrid = sg
    kind = "reference"
    try:
        if not pyoracle.have_reference():
            raise FileNotFoundError
        model = pyoracle.Reference(scen)
    except (FileNotFoundError, ValueError, OSError):
        kind = "port"
        model = pyoracle.Oracle(scen)
    model.step(1)                    # first step outside the timing
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < 3000:
        model.step(1)
        steps += 1
    dt = time.perf_counter() - t0
    model.close()
    one = {"value": scen.N * steps / dt, "cores": 1, "steps": steps, "seconds": round(dt, 2)}
    sample = (f"{n}x{n} marsh, subgrid {r}x{r} (Ne={scen.cfg['Ne']}), Jacobi PCG eps 1e-10, FP64, uniform "
              f"Manning n")
    if procs <= 1:
        return {"value": one["value"], "unit": "cell-steps/s", "cores": 1, "kind": kind,
                "sample": f"{sample}, steps 2-{steps + 1}, {dt:.1f} s on 1 host core"}
    many = cpu_processes(scen, kind, procs, budget_s)
    return {"value": many["value"], "unit": "cell-steps/s", "cores": procs, "kind": kind,
            "sample": f"{sample}: {procs} single-threaded processes each stepping its own copy of the sample "
                      f"at once for about {budget_s:.0f} s (no halo exchange, so an upper bound of the "
                      f"reference's MPI strips on as many cores; host CPUs visible: {os.cpu_count()})",
            "single_core": one, "per_process_steps": many["steps"]}


def _cpu_process(path, kind, budget_s, barrier, q):
    """One process of the multi-core CPU baseline: its own copy of the
    sample, stepped for budget_s seconds after every process is ready."""
    import pickle
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    with open(path, "rb") as f:
        scen = pickle.load(f)
    model = pyoracle.Reference(scen) if kind == "reference" else pyoracle.Oracle(scen)
    model.scen = None
    del scen
    model.step(1)
    barrier.wait()
    steps, t0 = 0, time.perf_counter()
    while time.perf_counter() - t0 < budget_s and steps < 3000:
        model.step(1)
        steps += 1
    q.put((steps, time.perf_counter() - t0))
    model.close()


def cpu_processes(scen, kind, procs, budget_s):
    """`procs` processes (spawned: they never touch the GPU) stepping copies
    of the sample together; aggregate cell-steps/s."""
    import dataclasses
    import multiprocessing as mp
    import pickle
    import tempfile
    ctx = mp.get_context("spawn")
    with tempfile.TemporaryDirectory(dir="/tmp") as tmp:
        path = os.path.join(tmp, "sample.pkl")
        with open(path, "wb") as f:
            # the dataclass fields only (building a model caches its ctypes
            # structures on the scenario)
            pickle.dump(dataclasses.replace(scen), f, protocol=5)
        barrier, q = ctx.Barrier(procs), ctx.Queue()
        ps = [ctx.Process(target=_cpu_process, args=(path, kind, budget_s, barrier, q)) for _ in range(procs)]
        for p_ in ps:
            p_.start()
        res = [q.get(timeout=budget_s + 600) for _ in ps]
        for p_ in ps:
            p_.join()
    return {"value": sum(scen.N * st / dt for st, dt in res), "steps": sorted(st for st, _ in res)}


TILING = None     # (H, remap, cols per lane) given on the command line instead of the autotuner


BLOCK_TUNE_STEPS = 2   # steps per candidate of the stencil block-shape tuner (6x that many steps run)


def run_form(scen, steps, warmup, tune=True, spinup=0, job=0, **over):
    """Time one precision form: 3 steps, the solver-tiling autotuner, the
    stencil block-shape tuner (whose steps count as the first of the
    spin-up), `spinup` steps of the job from rest, warmup steps, then `steps` timed
    steps (the call returns after the stream is synchronised); with job > 0,
    then the rest of a `job`-step run, every step after the autotuners timed
    for the whole job's average; then 4 steps with per-kernel HIP event
    timing.  Returns the numbers and the final state."""
    from frehdc_amd import device
    dev = device.DeviceScenario(scen, **over)
    dev.step(3)
    if TILING:
        from frehdc_amd import abi
        lanes = min(TILING[2], 4 if over.get("precision") == abi.PREC_MIXED else 2)
        dev.set_pcg_tiling(TILING[0], TILING[1], lanes)
        tiling = {"rows_per_tile": TILING[0], "xcd_remap": TILING[1], "cols_per_lane": lanes}
    else:
        tiling = dev.autotune(20) if tune else None
    jt = 0.0                     # the job's stepping time after the autotuners
    tuned = 0                    # steps the block-shape tuner ran
    if tune and not TILING and spinup >= 6 * BLOCK_TUNE_STEPS:
        blocks = dev.autotune_blocks(BLOCK_TUNE_STEPS)
        tuned = 6 * BLOCK_TUNE_STEPS
        spinup -= tuned
        if tiling is not None:
            tiling = dict(tiling, stencil_block=[blocks["block_x"], blocks["block_y"]])
    if spinup > 0:
        t0 = time.perf_counter()
        dev.step(spinup)
        jt += time.perf_counter() - t0
    t0 = time.perf_counter()
    dev.step(warmup)
    jt += time.perf_counter() - t0
    it0 = dev.stats()["pcgIterTotal"]
    t0 = time.perf_counter()
    dev.step(steps)
    t1 = time.perf_counter()
    jt += t1 - t0
    st = dev.stats()
    fields = dev.download(["surf", "uuXP", "vvYP", "depth"])
    job_out = None
    rest = job - (3 + tuned + spinup + warmup + steps)
    if job > 0 and rest >= 0:
        jit0 = dev.stats()["pcgIterTotal"]
        t2 = time.perf_counter()
        dev.step(rest)
        jt += time.perf_counter() - t2
        n = job - 3 - tuned
        job_out = {"steps": n, "ms_per_step": jt * 1e3 / n, "cell_steps_per_s": scen.N * n / jt}
    dev.enable_timing(True)
    dev.step(4)
    kt = dev.kernel_times()
    dev.enable_timing(False)
    budget = measure_budget(dev, 10)
    dev.close()
    sec = t1 - t0
    return {"sec": sec, "ms_per_step": sec * 1e3 / steps, "value": scen.N * steps / sec,
            "iters": (st["pcgIterTotal"] - it0) / steps, "sweeps": st["refineLast"],
            "relres": st["relResLast"], "kt": kt, "fields": fields, "tiling": tiling, "budget": budget,
            "job": job_out}


def measure_budget(dev, steps):
    """Conservation over `steps` more steps with the device budgets on
    (flux-form volume and scalar-mass closure per step, summed in
    double-double; accounting changes no result): relative residuals of the
    sum and of the worst step, and what the faces moved (the check is not
    vacuous)."""
    dev.set_budget(True)
    dev.step(steps)
    b = dev.budget()
    dev.set_budget(False)
    return {"steps": b["steps"], "volume_rel_residual": b["volumeRelResidual"],
            "volume_rel_residual_max_step": b["volumeRelResidualMax"], "volume_moved_m3": b["volumeMoved"],
            "mass_rel_residual": b["massRelResidual"], "mass_rel_residual_max_step": b["massRelResidualMax"],
            "mass_moved": b["massMoved"]}


def run_batch(members, steps, warmup, **over):
    """One batch of ensemble members (shared launches): 2 steps, the solver
    tiling and stencil block-shape tuners (6 steps), warmup steps, then
    `steps` timed steps, then 2 steps with per-kernel event timing."""
    from frehdc_amd import abi, device
    devs = [device.DeviceScenario(sc, **over) for sc in members]
    b = device.Batch(len(devs))
    for d in devs:
        b.add(d)
    b.try_step(2)
    tiling = b.autotune(10)
    blocks = b.autotune_blocks(1)          # 6 steps
    tiling = dict(tiling, stencil_block=[blocks["block_x"], blocks["block_y"]])
    b.try_step(warmup)
    t0 = time.perf_counter()
    rc = b.try_step(steps)
    t1 = time.perf_counter()
    if rc not in (0, abi.FREHD_ECFL):
        abi.check(rc)
    halted = sum(1 for d in devs if d.stats()["halted"])
    fields = [{k: v[: sc.N] for k, v in d.download(["surf", "uuXP", "vvYP", "depth"]).items()}
              for d, sc in zip(devs, members)]
    iters = max(d.stats()["pcgIterLast"] for d in devs)
    b.enable_timing(True)
    b.try_step(2)
    kt = b.kernel_times()
    b.close()
    for d in devs:
        d.close()
    cells = sum(sc.N for sc in members)
    return {"sec": t1 - t0, "value": cells * steps / (t1 - t0), "ms_per_step": (t1 - t0) * 1e3 / steps,
            "fields": fields, "iters": iters, "kt": kt, "halted": halted, "tiling": tiling}


def batched_ensemble(n_members, size, steps, warmup):
    """The batched-ensemble workload: members of one fBm base (Hurst 0.7) plus
    a seeded N(0, 0.05 m) perturbation, Manning n ~ U(0.015, 0.04) and tide
    amplitude ~ U(0.3, 0.8) m per member, one scalar, dx = 1 m, dt = 0.5 s,
    plain grid; FP64 and mixed (warm start) forms on the same members."""
    from frehdc_amd import abi, cases
    import numpy as np
    t0 = time.perf_counter()
    base = cases.fbm_surface(size, size, 0.7, np.random.default_rng(5), -1.0, 1.5)
    Nt = warmup + steps + 18
    members = [cases.ensemble_member(m, n=size, Nt=Nt, base=base) for m in range(n_members)]
    log(f"bench: batch of {n_members} x {size}^2 built in {time.perf_counter() - t0:.1f} s")
    out = {}
    ref = None
    # fp64_warm: the FP64 solve from the extrapolated start, the yardstick of
    # how far a change of iterate within the solver tolerance moves the state
    for f, over in (("fp64", {}), ("fp64_warm", {"warmStart": 2}),
                    ("mixed_warm", {"precision": abi.PREC_MIXED, "warmStart": 2})):
        r = run_batch(members, steps, warmup, **over)
        log(f"bench: batch {f}: {r['ms_per_step']:.2f} ms/step, {r['value'] / 1e9:.3f} G cell-steps/s")
        log(f"kernel times batch {f} (ms/launch, launches):",
            json.dumps({k: [round(v[0], 4), v[1]] for k, v in r["kt"].items()}))
        o = {"cell_steps_per_s": r["value"], "ms_per_step": r["ms_per_step"], "pcg_iterations_last": r["iters"],
             "halted_members": r["halted"], "tiling": r["tiling"]}
        if ref is None:
            ref = r["fields"]
        else:
            o.update(field_errors(r["fields"], ref))
        out[f] = o
    return {"workload": f"{n_members} members x {size}x{size}, dx=1 m, dt=0.5 s, plain grid, shared fBm "
                        f"base + N(0, 0.05 m) per member, Manning U(0.015, 0.04), tide U(0.3, 0.8) m, "
                        f"one scalar, {steps} timed steps after {warmup}", "forms": out}


def serving_run(count, subgrid):
    from frehdc_amd import abi, serving
    t0 = time.perf_counter()
    reqs = serving.make_requests(count, 20.0, subgrid_r=subgrid, workers=16)
    log(f"bench: {count} serving requests built in {time.perf_counter() - t0:.1f} s")
    out = {}
    ref = None
    for f, over in (("fp64", {}), ("fp64_warm", {"warmStart": 2}),
                    ("mixed_warm", {"precision": abi.PREC_MIXED, "warmStart": 2})):
        fresh = [serving.Request(rid=r.rid, arrival=r.arrival, steps=r.steps, scen=r.scen) for r in reqs]
        st = serving.Server(budget_bytes=200e9, collect=["surf", "uuXP", "vvYP", "depth"], **over).run(fresh)
        log(f"bench: serving {f}: p50 {st['latency_p50_s']:.2f} s, p99 {st['latency_p99_s']:.2f} s, "
            f"{st['cell_steps_per_s'] / 1e9:.3f} G cell-steps/s")
        # every request's final state, against the FP64 form's
        res = [{k: v[: r.scen.N] for k, v in r.result.items()} for r in sorted(fresh, key=lambda q: q.rid)]
        if ref is None:
            ref = res
        else:
            st.update(field_errors(res, ref))
        out[f] = st
    return {"workload": f"{count} requests, Poisson 20/s, sides log-uniform 64..1024 (multiples of 64), "
                        f"dx=2 m, subgrid {subgrid}x{subgrid}, fBm Hurst U(0.6, 0.9), tide U(0.1, 0.7) m, "
                        f"Manning U(0.015, 0.04), 1-3 scalars, U(200, 2000) steps, HBM budget 200 GB",
            "forms": out}


def field_errors(got, ref):
    """Largest |eta|, |u|, |v| differences and wet-mask mismatches over a
    list of scenarios' final fields, against the FP64 form's."""
    import numpy as np
    e = {"max_abs_eta_err_m": 0.0, "max_abs_u_err": 0.0, "max_abs_v_err": 0.0, "wet_mask_mismatches": 0,
         "scenarios": len(ref), "scenarios_eta_err_over_1e-6_m": 0}
    for a, b in zip(got, ref):
        de = float(np.max(np.abs(a["surf"] - b["surf"])))
        e["max_abs_eta_err_m"] = max(e["max_abs_eta_err_m"], de)
        e["scenarios_eta_err_over_1e-6_m"] += int(de > 1e-6)
        e["max_abs_u_err"] = max(e["max_abs_u_err"], float(np.max(np.abs(a["uuXP"] - b["uuXP"]))))
        e["max_abs_v_err"] = max(e["max_abs_v_err"], float(np.max(np.abs(a["vvYP"] - b["vvYP"]))))
        e["wet_mask_mismatches"] += int(np.count_nonzero((a["depth"] > 0) != (b["depth"] > 0)))
    return e


def traffic_of(kernels, workload):
    """HBM bytes per launch of the kernels, from the committed counter
    summary of the workload (profiles/traffic_<workload>.json, e.g.
    marsh1024_sub8, channel4096_sub4); None when there is none."""
    path = os.path.join(ROOT, "profiles", f"traffic_{workload}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        k = json.load(f)["kernels"]
    if not all(n in k for n in kernels):
        return None
    return int(sum(k[n]["bytes"] for n in kernels))


def solver_roofline(kt, N, mixed):
    """Roofline of one solver iteration from event-timed kernel averages:
    FP64, one working k_pcg1 launch (72 B per cell); mixed, one working k_cg1
    launch of the inner FP32 solve (32 B per cell)."""
    ks, per_cell = (["k_cg1_f32"], PCG_ITER_BYTES_F32) if mixed else (["k_pcg1"], PCG_ITER_BYTES)
    t_ms = sum(kt.get(k, (0, 0))[0] for k in ks)
    alg = per_cell * N
    achieved = alg / (t_ms * 1e-3) / 1e9 if t_ms > 0 else 0.0
    return ks, {"bound": "hbm", "kernel": f"solver iteration ({' + '.join(ks)})", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "bytes_per_launch": alg, "ms_per_iteration": round(t_ms, 5)}


def large_domain(size, r, steps, warmup):
    """The large single domain on one GPU: channel size^2 (dx = 1 m), subgrid
    r x r evaluated on the fly from the pixels, three scalars, tide 0.5 m on
    YP and 50 m^3/s river inflow on row 0; fp64 and mixed_warm timed, the
    solver roofline of each, the mixed form's errors against fp64."""
    from frehdc_amd import abi, cases
    import numpy as np
    t0 = time.perf_counter()
    sc = cases.channel_subgrid(n=size, r=r, Nt=warmup + steps + 20)
    log(f"bench: large domain {sc.name} (Ne={sc.cfg['Ne']}) built in {time.perf_counter() - t0:.1f} s")
    out, res = {}, {}
    for f, over in [("fp64", {}), ("mixed_warm", {"precision": abi.PREC_MIXED, "warmStart": 2})]:
        r_ = run_form(sc, steps, warmup, **over)
        res[f] = r_
        ks, roof = solver_roofline(r_["kt"], sc.N, f != "fp64")
        roof["traffic"] = traffic_of(ks, f"channel{size}_sub{r}")
        out[f] = {"cell_steps_per_s": r_["value"], "ms_per_step": r_["ms_per_step"],
                  "pcg_iterations_per_step": round(r_["iters"], 2), "refinement_sweeps": r_["sweeps"],
                  "rel_residual_last": r_["relres"], "tiling": r_["tiling"], "solver_roofline": roof,
                  "step_roofline": step_roofline(sc.cfg, sc.N, r_, f != "fp64", f != "fp64"),
                  "budget": r_["budget"]}
        log(f"bench: large {f}: {r_['ms_per_step']:.2f} ms/step, {r_['iters']:.1f} iterations/step, "
            f"solver {roof['achieved']:.0f} GB/s ({roof['frac']:.3f} of peak)")
    N = sc.N
    for k, name in (("surf", "max_abs_eta_err_m"), ("uuXP", "max_abs_u_err"), ("vvYP", "max_abs_v_err")):
        out["mixed_warm"][name] = float(np.max(np.abs(res["mixed_warm"]["fields"][k][:N] - res["fp64"]["fields"][k][:N])))
    out["mixed_warm"]["wet_mask_mismatches"] = int(np.count_nonzero(
        (res["mixed_warm"]["fields"]["depth"][:N] > 0) != (res["fp64"]["fields"]["depth"][:N] > 0)))
    return {"workload": f"channel {size}x{size}, dx=1 m, subgrid {r}x{r} on the fly (Ne={sc.cfg['Ne']}), dt=0.5 s, "
                        f"fBm bed (Hurst 0.8) [-1.5, 1] m + meandering channel 2 m deep, 20 cells wide, tide 0.5 m "
                        f"on YP, river 50 m^3/s on row 0, 3 scalars, Jacobi PCG eps=1e-10, {steps} timed steps "
                        f"after {warmup}",
            "cells": N, "forms": out}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--spinup", type=int, default=100,
                    help="steps of the single-domain job from rest before the warmup and the timed steps")
    ap.add_argument("--job", type=int, default=2000,
                    help="steps of the single-domain job whose whole-run average is reported per form (0: none)")
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--subgrid", type=int, default=8, help="subgrid pixels per cell side (0: plain grid)")
    ap.add_argument("--cpu-size", type=int, default=256)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--cpu-procs", type=int, default=16,
                    help="processes of the multi-core CPU baseline (the host cores the GPU's share has)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--forms", default="fp64,fp64_warm,mixed_cold,mixed_warm",
                    help="precision forms to time (the headline value is the last mixed form)")
    ap.add_argument("--batch-members", type=int, default=64)
    ap.add_argument("--batch-size", type=int, default=2048)
    ap.add_argument("--batch-steps", type=int, default=10)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--serve-count", type=int, default=48)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--large-size", type=int, default=4096)
    ap.add_argument("--large-steps", type=int, default=10)
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--tiling", default="", help="H:remap[:cols_per_lane] solver tiling for the single-domain "
                                                 "forms instead of the autotuner (profiled runs: no trial "
                                                 "launches; cols_per_lane: at most 2 for the FP64 forms)")
    args = ap.parse_args()
    global TILING
    if args.tiling:
        TILING = tuple(int(x) for x in args.tiling.split(":"))
        TILING = TILING + (1,) * (3 - len(TILING))
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(this build has no multi-GPU path)", file=sys.stderr)
        sys.exit(2)

    from frehdc_amd import abi, cases
    import numpy as np

    total = max(3 + args.spinup + args.warmup + args.steps, args.job) + 16

    t0 = time.perf_counter()
    if args.subgrid > 0:
        scen = cases.marsh_subgrid(n=args.size, r=args.subgrid, Nt=max(total, 2016), device_tables=True)
    else:
        scen = cases.marsh(n=args.size, Nt=max(total, 2016))
    log(f"bench: scenario built in {time.perf_counter() - t0:.1f} s")
    N = scen.N
    grid = f"subgrid {args.subgrid}x{args.subgrid} (Ne={scen.cfg['Ne']})" if args.subgrid > 0 else "plain grid"
    # fp64: the reference's solve (cold start every step); fp64_warm: the same
    # FP64 solve from the extrapolated start 2 eta^n - eta^(n-1), so that the
    # gain of the start and the gain of the precision are reported apart
    forms = {"fp64": {}, "fp64_warm": {"warmStart": 2}, "mixed_cold": {"precision": abi.PREC_MIXED, "warmStart": 0},
             "mixed_warm": {"precision": abi.PREC_MIXED, "warmStart": 2}}
    names = [f for f in args.forms.split(",") if f]
    res = {}
    for f in names:
        log(f"bench: marsh {args.size}^2 {f}, spin-up {args.spinup}, warmup {args.warmup}, steps {args.steps}")
        r = run_form(scen, args.steps, args.warmup, spinup=args.spinup, job=args.job, **forms[f])
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
             "rel_residual_last": r["relres"], "tiling": r["tiling"], "job": r["job"],
             "step_roofline": step_roofline(scen.cfg, N, r, f.startswith("mixed"), f.endswith("warm")),
             "budget": r["budget"]}
        if f != "fp64" and "fp64" in res:
            o["max_abs_eta_err_m"] = err(f, "surf")
            o["max_abs_u_err"] = err(f, "uuXP")
            o["max_abs_v_err"] = err(f, "vvYP")
            o["wet_mask_mismatches"] = int(np.count_nonzero(
                (r["fields"]["depth"][:N] > 0) != (res["fp64"]["fields"]["depth"][:N] > 0)))
        if f != "fp64" and "fp64" in res:
            o["vs_fp64"] = round(r["value"] / res["fp64"]["value"], 3)
            if r["job"] and res["fp64"]["job"]:
                o["job_vs_fp64"] = round(r["job"]["cell_steps_per_s"] / res["fp64"]["job"]["cell_steps_per_s"], 3)
        forms_out[f] = o
    mixed = [f for f in names if f.startswith("mixed")]
    head = mixed[-1] if mixed else names[-1]
    hr = res[head]

    # roofline of the solver iteration of the headline form, from HIP events
    # on the handle's stream: FP64, the two kernels of one PCG iteration;
    # mixed, the one launch of an inner FP32 iteration (working launches:
    # the set-up and stop-test launches are timed apart)
    dtype = "f32+f64" if head.startswith("mixed") else "f64"
    ks, roof = solver_roofline(hr["kt"], N, head.startswith("mixed"))
    traffic = traffic_of(ks, f"marsh{args.size}_sub{args.subgrid}")
    roof["traffic"] = traffic
    if traffic and roof["ms_per_iteration"] > 0:
        roof["traffic_GBps"] = round(traffic / (roof["ms_per_iteration"] * 1e-3) / 1e9, 1)
    # where the step's time goes in the kernel-timed steps: the solver
    # iteration's share, and the kernel with the largest share (whose bytes
    # have no per-kernel frozen count; the whole step's is step_roofline)
    shares = {k: v[0] * v[1] / 4.0 / hr["ms_per_step"] for k, v in hr["kt"].items()}
    roof["step_share"] = round(sum(shares.get(k, 0.0) for k in ks), 3)
    top = max(shares, key=shares.get)
    roof["largest_kernel"] = {"kernel": top, "ms_per_launch": round(hr["kt"][top][0], 5),
                              "step_share": round(shares[top], 3)}

    large = None
    if not args.no_large and args.large_size > 0:
        large = large_domain(args.large_size, 4, args.large_steps, 5)

    batched = None
    if not args.no_batch and args.batch_members > 0:
        batched = batched_ensemble(args.batch_members, args.batch_size, args.batch_steps, 3)
    served = None
    if not args.no_serve and args.serve_count > 0:
        served = serving_run(args.serve_count, 4)

    cpu = None
    if not args.no_cpu:
        procs = min(args.cpu_procs, len(os.sched_getaffinity(0)))
        cpu = cpu_baseline(args.cpu_size, max(args.subgrid, 1), args.cpu_seconds, procs)
        log("cpu baseline:", json.dumps(cpu))

    out = {
        "metric": "cell-steps/s mixed vs FP64 at eta err<=1e-6 m; p50/p99 scenario latency",
        "value": hr["value"], "unit": "cell-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": hr["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": dtype, "data": "synthetic",
        "config": {"workload": f"marsh {args.size}x{args.size}, dx=2 m, {grid}, dt=0.5 s, fBm bed "
                               f"[-1, 1.5] m, tide 0.6 m / 44712 s, Manning n U(0.02, 0.04) per cell, "
                               f"one scalar, Jacobi PCG eps=1e-10, form {head}; timed: steps "
                               f"{4 + args.spinup + args.warmup}-{3 + args.spinup + args.warmup + args.steps} "
                               f"of a {args.job or 3 + args.spinup + args.warmup + args.steps}-step run from rest "
                               f"(the solver tiling tuned after 3 steps, the stencil block shape within the "
                               f"spin-up; forms.*.job: the average over all of its steps after the tuners)",
                   "cells": N,
                   "form": head, "forms": forms_out},
        "roofline": roof,
        "large": large,
        "cpu_baseline": cpu,
        "batched": batched,
        "serving": served,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

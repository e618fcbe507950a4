"""The benchmark workloads on the device against the CPU oracle, at sizes the
oracle finishes in seconds:

* the large channel (subgrid 4x4, three scalars, river inflow on row 0,
  tide on the YP row) at 64^2, once with level tables made on the host and
  once evaluated on the fly from the pixels;
* the basin on the subgrid model (4x4 pixels per cell);
* the headline marsh at 256^2 with 8x8 pixels per cell, per-cell Manning n
  and the device's own level tables -- the cpu_baseline sample of bench.py;
* the ensemble: four members at 64^2 as one FP64 batch, each against its own
  oracle run, and the mixed-precision batch against the FP64 batch.

Where the device makes the levels (device tables, on the fly) the oracle
runs on the device's tables, downloaded: both then see the same table
values, and every phase but the solve is compared bit for bit.  The solve
sums its dot products in another order than the reference's sequential
loops, so the new surface agrees to the solver tolerance (1e-9 m) and whole
runs to 1e-7, with identical wet/dry masks."""
import dataclasses

import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import Batch, DeviceScenario
import pyoracle as po

from test_gpu_parity import compare
from test_gpu_subgrid import dl, oracle_state

pytestmark = pytest.mark.gpu

SUB = abi.SUB_FIELD_NAMES


def with_device_tables(sc):
    """The scenario with the level tables the device generates (or
    evaluates on the fly) from its pixels, for the oracle."""
    d = DeviceScenario(sc)
    tabs = d.tables()
    d.close()
    sg = {k: np.ascontiguousarray(v.reshape(-1)) for k, v in tabs.items()}
    sg["edgeX"], sg["edgeY"] = sc.subgrid["edgeX"], sc.subgrid["edgeY"]
    cfg = dict(sc.cfg, subgridOnTheFly=0)
    return dataclasses.replace(sc, subgrid=sg, cfg=cfg)


def channel64_tables():
    return cases.channel_subgrid(n=64, r=4, Nt=300, on_the_fly=False)


def channel64_fly():
    return cases.channel_subgrid(n=64, r=4, Nt=300, on_the_fly=True)


def basin32_sub4():
    return cases.basin_subgrid(nx=32, ny=32, r=4, Nt=300)


def marsh256_sub8():
    return cases.marsh_subgrid(n=256, r=8, Nt=60, device_tables=True)


BUILDERS = {"channel64_sub4_tables": channel64_tables, "channel64_sub4_fly": channel64_fly,
            "basin32_sub4": basin32_sub4, "marsh256_sub8": marsh256_sub8}
_made = {}


def pair(name):
    """(device scenario, oracle scenario), built once per session."""
    if name not in _made:
        sc = BUILDERS[name]()
        osc = with_device_tables(sc) if "V" not in sc.subgrid else sc
        _made[name] = (sc, osc)
    return _made[name]


PHASE_CASES = [(n, k) for n in ("channel64_sub4_tables", "channel64_sub4_fly", "basin32_sub4")
               for k in (0, 5, 40)] + [("marsh256_sub8", 0), ("marsh256_sub8", 6)]


@pytest.mark.parametrize("name,k", PHASE_CASES, ids=[f"{n}-k{k}" for n, k in PHASE_CASES])
def test_config_phases_bit_exact(oracle_built, name, k):
    sc, osc = pair(name)
    print(f"case {name}: phases after {k} steps", flush=True)
    o = po.Oracle(osc)
    o.step(k)
    # the device run keeps the scenario's own level source (pixels on the fly)
    d = DeviceScenario(sc, from_state=oracle_state(o, osc))
    d.tt = o.tt
    o.phase(po.PHASE_SOURCE)
    d.phase(0)
    compare(dl(d, sc), oracle_state(o, osc), ["EnXP", "EnYP", "DragXP", "DragYP"], f"{name} k={k} phase0")
    d.phase(1)
    o.phase(po.PHASE_SURFACE)
    got, ref = dl(d, sc), oracle_state(o, osc)
    compare(got, ref, ["surfOld"], "phase1")
    assert np.max(np.abs(got["surf"] - ref["surf"])) <= 1e-9
    d.close()
    d = DeviceScenario(sc, from_state=oracle_state(o, osc))
    d.tt = o.tt
    d.phase(2)
    o.phase(po.PHASE_UPDATE)
    names = [n for n in abi.FIELD_NAMES + SUB + ["ind"] if n not in ("CFLx", "CFLy", "S", "Sm")]
    compare(dl(d, sc), oracle_state(o, osc), names, f"{name} k={k} phase2")
    d.phase(3)
    o.phase(po.PHASE_SCALAR)
    compare(dl(d, sc), oracle_state(o, osc), ["S", "Sm", "surf", "depth"], f"{name} k={k} phase3")
    d.phase(4)
    o.phase(po.PHASE_CFL)
    compare(dl(d, sc), oracle_state(o, osc), ["CFLx", "CFLy"], f"{name} k={k} phase4")
    d.close()


TRAJ = [("channel64_sub4_tables", 200), ("channel64_sub4_fly", 200), ("basin32_sub4", 200),
        ("marsh256_sub8", 12)]


@pytest.mark.parametrize("name,steps", TRAJ, ids=[t[0] for t in TRAJ])
def test_config_trajectory(oracle_built, name, steps):
    sc, osc = pair(name)
    print(f"case {name}: {steps}-step trajectory", flush=True)
    d = DeviceScenario(sc)
    o = po.Oracle(osc)
    N, ns = sc.N, max(1, sc.cfg["nScalar"])
    chunk = min(20, steps)
    worst = {}
    for t in range(steps // chunk):
        d.step(chunk)
        o.step(chunk)
        got = d.download(["surf", "uuXP", "vvYP", "depth", "S", "V"])
        ref = oracle_state(o, osc)
        for n in ("surf", "uuXP", "vvYP", "V"):
            e = float(np.max(np.abs(got[n].reshape(-1)[:N] - ref[n][:N])))
            worst[n] = max(worst.get(n, 0.0), e)
        S, Sr = got["S"].reshape(ns, -1)[:, :N], ref["S"].reshape(ns, -1)[:, :N]
        worst["S"] = max(worst.get("S", 0.0), float(np.max(np.abs(S - Sr))))
        assert np.array_equal(got["depth"][:N] > 0, ref["depth"][:N] > 0), f"wet mask at {(t + 1) * chunk}"
    assert d.stats()["steps"] == steps and not d.stats()["halted"]
    print("max errors", worst, flush=True)
    for n, e in worst.items():
        assert e <= 1e-7, f"{n} error {e}"
    d.close()


def ensemble(n=64, members=4, steps=300, period=1200.0):
    """Members as the ensemble workload makes them (shared base, own
    perturbation, roughness and tide amplitude); the 1200 s tidal period
    wets and dries cells within a short run, the workload's own is 44712 s."""
    base = cases.fbm_surface(n, n, 0.7, np.random.default_rng(5), -1.0, 1.5)
    return [cases.ensemble_member(m, n=n, Nt=steps, base=base, period=period) for m in range(members)]


def batch_of(scs, **over):
    b = Batch(len(scs))
    ds = [DeviceScenario(sc, **over) for sc in scs]
    for d in ds:
        b.add(d)
    return b, ds


def test_ensemble_fp64_batch_against_oracle(oracle_built):
    steps = 120
    scs = ensemble(steps=steps + 20)
    print(f"case ensemble: {len(scs)} members 64^2, FP64 batch vs oracle, {steps} steps", flush=True)
    b, ds = batch_of(scs)
    orc = [po.Oracle(sc) for sc in scs]
    N = scs[0].N
    worst = 0.0
    for t in range(steps // 20):
        b.step(20)
        for o in orc:
            o.step(20)
        for m, o in enumerate(orc):
            got = ds[m].download(["surf", "uuXP", "vvYP", "depth", "S"])
            for n in ("surf", "uuXP", "vvYP"):
                worst = max(worst, float(np.max(np.abs(got[n][:N] - o.field(n)[:N]))))
            worst = max(worst, float(np.max(np.abs(got["S"][0, :N] - o.field("S")[0, :N]))))
            assert np.array_equal(got["depth"][:N] > 0, o.field("depth")[:N] > 0), (m, t)
    print(f"FP64 batch vs oracle {worst:.3g}", flush=True)
    assert worst <= 1e-7
    # members really differ (own bed, roughness and tide)
    assert not np.array_equal(ds[0].field("surf")[:N], ds[1].field("surf")[:N])


def surf_depth(ds):
    return [d.download(["surf", "depth", "uuXP", "vvYP"]) for d in ds]


def test_ensemble_mixed_batch_against_fp64():
    """The mixed-precision batch (extrapolated start) against the FP64 batch:
    at the workload's own tide, eta within 1e-6 m and identical wet/dry
    masks; under the fast tide, where the trajectory itself moves with the
    solver tolerance (threshold rules at the fronts), eta within 10x of what
    the FP64 batch moves when its tolerance tightens to 1e-13."""
    steps = 120
    for period in (cases.TIDE_PERIOD, 1200.0):
        scs = ensemble(steps=steps + 20, period=period)
        print(f"case ensemble: 4 members 64^2, tide period {period:g} s, mixed vs FP64 batch", flush=True)
        runs = {}
        for name, over in (("fp64", {}), ("tight", {"eps": 1e-13}),
                           ("mixed", {"precision": abi.PREC_MIXED, "warmStart": 2})):
            b, ds = batch_of(scs, **over)
            b.step(steps)
            runs[name] = surf_depth(ds)
            for d in ds:
                assert d.stats()["relResLast"] < 1e-10
        N = scs[0].N
        err = {k: max(float(np.max(np.abs(x["surf"][:N] - y["surf"][:N]))) for x, y in zip(runs[k], runs["fp64"]))
               for k in ("tight", "mixed")}
        masks = sum(int(np.count_nonzero((x["depth"][:N] > 0) != (y["depth"][:N] > 0)))
                    for x, y in zip(runs["mixed"], runs["fp64"]))
        print(f"  eta: mixed vs fp64 {err['mixed']:.3g}, fp64 eps 1e-13 vs fp64 {err['tight']:.3g}, "
              f"mask mismatches {masks}", flush=True)
        if period == cases.TIDE_PERIOD:
            assert err["mixed"] <= 1e-6 and masks == 0
        else:
            assert err["mixed"] <= max(1e-6, 10 * err["tight"])

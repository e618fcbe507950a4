"""The solver-tiling autotuner (frehd_dev_autotune / frehd_batch_autotune):
it leaves the model state and the statistics bitwise as they were, returns a
valid tiling, and the run continues as an untuned twin does (the tiling only
changes the order of the solver's dot-product sums)."""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import Batch, DeviceScenario

pytestmark = pytest.mark.gpu

NAMES = ["surf", "surfOld", "uuXP", "vvYP", "depth", "S", "CFLx", "CFLy"]


def snapshot(d):
    return d.download(NAMES), d.stats()


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_MIXED], ids=["fp64", "mixed"])
@pytest.mark.parametrize("make", [lambda: cases.basin(nx=300, ny=70, Nt=60),
                                  lambda: cases.marsh_subgrid(n=32, r=4, Nt=60, period=1200.0, device_tables=True)],
                         ids=["basin300x70", "marsh32_sub4"])
def test_autotune_keeps_state(make, prec):
    sc = make()
    print(f"case autotune {sc.name} precision {prec}", flush=True)
    a = DeviceScenario(sc, precision=prec, warmStart=2)
    b = DeviceScenario(sc, precision=prec, warmStart=2)
    a.step(5)
    b.step(5)
    f0, s0 = snapshot(a)
    best = a.autotune(iters=4)
    f1, s1 = snapshot(a)
    print("  best", best, flush=True)
    for k in f0:
        assert np.array_equal(f0[k], f1[k]), k
    assert s0 == s1
    assert best["rows_per_tile"] >= 1 and best["xcd_remap"] in (0, 1) and best["ms_per_iteration"] > 0
    assert best["cols_per_lane"] in ((1, 2, 4) if prec == abi.PREC_MIXED else (1, 2))
    a.step(10)
    b.step(10)
    fa, fb = a.download(["surf", "uuXP", "vvYP"]), b.download(["surf", "uuXP", "vvYP"])
    N = sc.N
    for k in fa:
        assert np.max(np.abs(fa[k][:N] - fb[k][:N])) <= 1e-8, k
    assert a.stats()["steps"] == b.stats()["steps"]
    a.close()
    b.close()


@pytest.mark.parametrize("warm", [0, 2], ids=["cold", "warm"])
def test_autotune_times_working_iterations(warm):
    # the trial iterations time working launches: the tuned time per
    # iteration is the event-timed k_pcg1 launch's to within a factor of two
    # (trials that ran on an exhausted residual exit at once, a few times
    # faster than a working launch at this size)
    sc = cases.marsh(n=512, Nt=40)
    print(f"case autotune marsh512 fp64, warmStart {warm}", flush=True)
    d = DeviceScenario(sc, warmStart=warm)
    d.step(3)
    best = d.autotune(iters=20)
    d.enable_timing(True)
    d.step(3)
    kt = d.kernel_times()
    d.enable_timing(False)
    print("  best", best, "k_pcg1", kt.get("k_pcg1"), flush=True)
    ms = kt["k_pcg1"][0]
    assert 0.5 * ms <= best["ms_per_iteration"] <= 2.0 * ms
    d.close()


def test_batch_autotune():
    print("case batch autotune, two mixed basins", flush=True)
    scs = [cases.basin(nx=96, ny=64, Nt=40), cases.basin(nx=130, ny=40, Nt=40, seed=3)]
    devs = [DeviceScenario(s, precision=abi.PREC_MIXED) for s in scs]
    b = Batch(2)
    for d in devs:
        b.add(d)
    b.step(3)
    before = [snapshot(d) for d in devs]
    best = b.autotune(iters=4)
    after = [snapshot(d) for d in devs]
    print("  best", best, flush=True)
    for (f0, s0), (f1, s1) in zip(before, after):
        for k in f0:
            assert np.array_equal(f0[k], f1[k]), k
        assert s0 == s1
    assert best["rows_per_tile"] <= 40
    assert best["cols_per_lane"] in (1, 2)          # nx = 130 is not a multiple of 4
    b.step(3)
    for d in devs:
        assert d.stats()["steps"] == 6


def test_batch_lanes_follow_members():
    """A handle joining a batch takes the members' columns per lane, or the
    batch goes back to one column per lane when the newcomer's nx does not
    allow it; either way the batched run equals the solo runs (the same
    tiling there) within the solver tolerance."""
    print("case batch lanes: mixed basins 96x64 (lanes 4), 96x48, 130x40", flush=True)
    scs = [cases.basin(nx=96, ny=64, Nt=40), cases.basin(nx=96, ny=48, Nt=40, seed=4),
           cases.basin(nx=130, ny=40, Nt=40, seed=3)]
    devs = [DeviceScenario(s, precision=abi.PREC_MIXED) for s in scs]
    devs[0].set_pcg_tiling(0, 0, 4)
    solo = []
    for s, lanes, n in zip(scs, (4, 4, 1), (6, 6, 4)):
        d = DeviceScenario(s, precision=abi.PREC_MIXED)
        d.set_pcg_tiling(0, 0, lanes)
        d.step(n)
        solo.append(d.download(["surf", "uuXP"]))
        d.close()
    b = Batch(3)
    b.add(devs[0])
    b.add(devs[1])                  # takes 4
    b.step(2)
    b.add(devs[2])                  # nx = 130: everyone back to 1
    b.step(2)
    b.remove(devs[2])
    devs[2].step(2)
    b.step(2)
    got = [d.download(["surf", "uuXP"]) for d in devs]
    steps = [d.stats()["steps"] for d in devs]
    b.close()
    for d in devs:
        d.close()
    assert steps == [6, 6, 4]
    for g, r, s in zip(got, solo, scs):
        for k in g:
            assert np.max(np.abs(g[k][: s.N] - r[k][: s.N])) <= 1e-8, (s.name, k)

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
    a.step(10)
    b.step(10)
    fa, fb = a.download(["surf", "uuXP", "vvYP"]), b.download(["surf", "uuXP", "vvYP"])
    N = sc.N
    for k in fa:
        assert np.max(np.abs(fa[k][:N] - fb[k][:N])) <= 1e-8, k
    assert a.stats()["steps"] == b.stats()["steps"]
    a.close()
    b.close()


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
    b.step(3)
    for d in devs:
        assert d.stats()["steps"] == 6

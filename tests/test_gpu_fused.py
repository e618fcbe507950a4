"""The fused mixed-precision solve (frehd_dev_set_fused_solve: refinement
sweeps, inner iterations, convergence tests, surface update and flags in one
persistent launch) against the launched solve (one launch per sweep and
iteration).

* Bit for bit, step by step: before every step the launched handle's pattern
  history is reset (frehd_dev_set_fused_solve(-1)), so its step runs every
  inner solve to its own stop test, as the fused solve does; both handles
  then hold bitwise the same state, iteration counts and volume losses after
  each of 40 wet/dry steps -- plain and subgrid, cold, warm and extrapolated
  starts.  The fused handle runs its steps as captured graphs after the
  first few, so both launch forms of the fused solve are covered.
* The plan's grid does not matter: a fused solve capped to a few blocks
  (each walking many tiles) equals the uncapped one bit for bit.
* Batches: with the fused solve every member runs its own solve to its own
  convergence, so a member's mixed-precision result is the same alone and
  in a batch of unequal sizes, bit for bit.
* FP64 handles ignore the setting (the solve stays the reference's).

The fused solve is off by default (a grid barrier costs more than a kernel
boundary on this device, so it is no faster at these sizes); each test
switches it on.
"""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import Batch, DeviceScenario, fused_solve_capacity

pytestmark = pytest.mark.gpu

FIELDS = ["surf", "surfOld", "uuXP", "vvYP", "depth", "S", "cellV", "CDXP", "wtfXP"]
STATS = ("steps", "pcgIterTotal", "pcgIterLast", "refineLast", "relResLast", "maxCFLx", "maxCFLy", "Vloss",
         "halted")


def fast_marsh(n=48, steps=400):
    sc = cases.marsh(n=n, dx=2.0, dt=0.5, Nt=steps, amp=0.6)
    sc.tideP[:] = cases.tide_series(steps, 0.5, 0.6, period=1200.0)
    return sc


SCEN = {
    "marsh48": lambda: fast_marsh(48, 100),
    "marsh64_sub4": lambda: cases.marsh_subgrid(n=64, r=4, Nt=100, period=1200.0),
    "basin40x36": lambda: cases.basin(nx=40, ny=36, Nt=100),
}


def same(a, b, N, names=FIELDS):
    for k in names:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if k == "S":
            x, y = x[..., :N], y[..., :N]
        else:
            x, y = x[:N], y[:N]
        if not np.array_equal(x, y):
            diff = float(np.max(np.abs(x - y)))
            return f"{k} differs (max |d| {diff:.3e})"
    return None


@pytest.mark.parametrize("warm", [0, 1, 2])
@pytest.mark.parametrize("name", sorted(SCEN))
def test_fused_equals_launched_step_by_step(name, warm):
    sc = SCEN[name]()
    over = {"precision": abi.PREC_MIXED, "warmStart": warm}
    F = DeviceScenario(sc, **over)
    L = DeviceScenario(sc, **over)
    try:
        F.set_fused_solve(0)
        for step in range(40):
            L.set_fused_solve(-1)
            L.step(1)
            F.step(1)
            a, b = F.download(FIELDS), L.download(FIELDS)
            msg = same(a, b, sc.N)
            assert msg is None, f"{name} warm={warm} step {step + 1}: {msg}"
            sa, sb = F.stats(), L.stats()
            for k in STATS:
                assert np.array_equal(np.asarray(sa[k]), np.asarray(sb[k])), \
                    f"{name} warm={warm} step {step + 1}: stats {k} {sa[k]} != {sb[k]}"
        assert F.stats()["pcgIterTotal"] > 0
    finally:
        F.close()
        L.close()


def test_fused_grid_does_not_matter():
    sc = cases.marsh_subgrid(n=96, r=4, Nt=80, period=1200.0)
    over = {"precision": abi.PREC_MIXED, "warmStart": 2}
    A = DeviceScenario(sc, **over)
    B = DeviceScenario(sc, **over)
    try:
        A.set_fused_solve(0)
        B.set_fused_solve(2)        # two blocks walk every tile of the member
        A.step(30)
        B.step(30)
        msg = same(A.download(FIELDS), B.download(FIELDS), sc.N)
        assert msg is None, msg
        assert A.stats()["pcgIterTotal"] == B.stats()["pcgIterTotal"]
    finally:
        A.close()
        B.close()


def test_fused_batch_equals_solo():
    scens = [fast_marsh(48, 80), cases.marsh(n=96, Nt=80), fast_marsh(64, 80)]
    over = {"precision": abi.PREC_MIXED, "warmStart": 2}
    solos = []
    for sc in scens:
        d = DeviceScenario(sc, **over)
        d.set_fused_solve(0)
        d.step(30)
        solos.append((d.download(FIELDS), d.stats()))
        d.close()
    devs = [DeviceScenario(sc, **over) for sc in scens]
    b = Batch(4)
    try:
        for d in devs:
            b.add(d)
        b.set_fused_solve(0)
        b.step(30)
        for d, sc, (ref, st) in zip(devs, scens, solos):
            msg = same(d.download(FIELDS), ref, sc.N)
            assert msg is None, f"{sc.name}: {msg}"
            assert d.stats()["pcgIterTotal"] == st["pcgIterTotal"]
    finally:
        b.close()
        for d in devs:
            d.close()


def test_fused_capacity_and_fp64_unaffected():
    assert fused_solve_capacity(4) >= 256
    assert fused_solve_capacity(1) >= fused_solve_capacity(4)
    sc = cases.marsh(n=48, Nt=40)
    A = DeviceScenario(sc)
    B = DeviceScenario(sc)
    try:
        A.set_fused_solve(0)
        B.set_fused_solve(-1)
        A.step(20)
        B.step(20)
        assert same(A.download(FIELDS), B.download(FIELDS), sc.N) is None
    finally:
        A.close()
        B.close()

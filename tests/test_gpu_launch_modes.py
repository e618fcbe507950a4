"""Launch modes of whole steps (FREHD_GRAPHS when a handle or batch is
created: 0 kernel by kernel, 1 captured graphs, 2 streamed launches; unset,
streamed from 2^18 cells up and graphs below).  The three run the same
kernels with the same arguments, so after 40 wet/dry steps every state
field and the solver statistics are bitwise the same -- FP64 and mixed
precision, plain and subgrid, a single handle and a batch of unequal members
(where the shared solver pattern leaves some members' solves open and the
host finishes them) -- and a member that exceeds the CFL limit halts at the
same step whichever the mode.
"""
import os

import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import Batch, DeviceScenario

pytestmark = pytest.mark.gpu

FIELDS = ["surf", "surfOld", "uuXP", "vvYP", "depth", "S", "cellV", "CDXP", "wtfXP"]
STATS = ("steps", "pcgIterTotal", "pcgIterLast", "refineLast", "maxCFLx", "maxCFLy", "halted")


def fast_marsh(n=48, steps=100):
    sc = cases.marsh(n=n, dx=2.0, dt=0.5, Nt=steps, amp=0.6)
    sc.tideP[:] = cases.tide_series(steps, 0.5, 0.6, period=1200.0)
    return sc


SCEN = {
    "marsh48_fp64": (lambda: fast_marsh(48), {}),
    "marsh64_sub4_mixed": (lambda: cases.marsh_subgrid(n=64, r=4, Nt=100, period=1200.0),
                           {"precision": abi.PREC_MIXED, "warmStart": 2}),
    "basin40x36_mixed_cold": (lambda: cases.basin(nx=40, ny=36, Nt=100), {"precision": abi.PREC_MIXED}),
}


def with_mode(mode, make):
    old = os.environ.get("FREHD_GRAPHS")
    os.environ["FREHD_GRAPHS"] = str(mode)
    try:
        return make()
    finally:
        if old is None:
            del os.environ["FREHD_GRAPHS"]
        else:
            os.environ["FREHD_GRAPHS"] = old


def assert_same(a, b, N, label):
    for k in FIELDS:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        x, y = (x[..., :N], y[..., :N]) if k == "S" else (x[:N], y[:N])
        assert np.array_equal(x, y), f"{label}: {k} differs (max |d| {float(np.max(np.abs(x - y))):.3e})"


@pytest.mark.parametrize("name", sorted(SCEN))
def test_modes_bitwise_equal(name):
    make, over = SCEN[name]
    sc = make()
    runs = {}
    for mode in (0, 1, 2):
        d = with_mode(mode, lambda: DeviceScenario(sc, **over))
        try:
            for _ in range(4):
                d.step(10)
            runs[mode] = (d.download(FIELDS), d.stats())
        finally:
            d.close()
    for mode in (1, 2):
        assert_same(runs[mode][0], runs[0][0], sc.N, f"{name} mode {mode}")
        for k in STATS:
            assert runs[mode][1][k] == runs[0][1][k], f"{name} mode {mode}: {k}"
        np.testing.assert_allclose(runs[mode][1]["Vloss"], runs[0][1]["Vloss"], rtol=1e-12, atol=1e-12)


def test_modes_batch_with_open_solves():
    scens = [fast_marsh(48), cases.marsh(n=96, Nt=100), fast_marsh(64)]
    over = {"precision": abi.PREC_MIXED, "warmStart": 2}
    runs = {}
    for mode in (1, 2):
        devs = [with_mode(mode, lambda sc=sc: DeviceScenario(sc, **over)) for sc in scens]
        b = with_mode(mode, lambda: Batch(4))
        try:
            for d in devs:
                b.add(d)
            b.step(40)
            runs[mode] = [(d.download(FIELDS), d.stats()) for d in devs]
        finally:
            b.close()
            for d in devs:
                d.close()
    for (a, sa), (g, sg), sc in zip(runs[2], runs[1], scens):
        assert_same(a, g, sc.N, f"batch member {sc.name}")
        assert sa["pcgIterTotal"] == sg["pcgIterTotal"]


def test_modes_halt_at_the_same_step():
    bad = cases.basin(nx=32, ny=32, Nt=40)
    bad.tideP[:] = 0.0
    bad.tideP[1:] = 1000.0
    for mode in (0, 1, 2):
        d = with_mode(mode, lambda: DeviceScenario(bad))
        try:
            assert d.try_step(20) == abi.FREHD_ECFL
            st = d.stats()
            assert st["halted"] and st["haltStep"] == 1 and st["steps"] == 1, (mode, st)
            assert d.try_step(1) == abi.FREHD_ECFL
        finally:
            d.close()

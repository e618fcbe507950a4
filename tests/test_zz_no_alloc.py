"""frehd_dev_step_noalloc (whole steps that allocate nothing: the kernels'
partial sums in the library's static scratch array, launches one at a time)
against frehd_dev_step on the same scenario, through the C ABI: bit for bit.
Two calls in a row, the second on a smaller grid, so the second call starts
from the scratch contents the first one left; FP64 and mixed precision, a
plain grid and a subgrid, and the entry point mixed with frehd_dev_step on
one handle."""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import DeviceScenario

pytestmark = pytest.mark.gpu

NAMES = ["surf", "surfOld", "uuXP", "vvYP", "uuYP", "vvXP", "depth", "depthXP", "depthYP", "EnXP", "EnYP",
         "DragXP", "DragYP", "CDXP", "CDYP", "Fuu", "Fvv", "cellV", "S", "Sm", "CFLx", "CFLy", "wtfXP", "wtfYP"]


def fast_marsh(n, steps=200):
    sc = cases.marsh(n=n, dx=2.0, dt=0.5, Nt=steps, amp=0.6)
    sc.tideP[:] = cases.tide_series(steps, 0.5, 0.6, period=1200.0)
    return sc


def bitwise(a, b):
    for k in a:
        x, y = np.asarray(a[k]).reshape(-1), np.asarray(b[k]).reshape(-1)
        assert np.array_equal(x.view(np.int64), y.view(np.int64)), k


def noalloc(d, n):
    return abi.check(d.lib.frehd_dev_step_noalloc(d.h, int(n)))


def stats_equal(a, b):
    sa, sb = a.stats(), b.stats()
    for k in ("steps", "pcgIterTotal", "pcgIterLast", "refineLast", "relResLast", "maxCFLx", "maxCFLy", "Vloss"):
        assert sa[k] == sb[k], (k, sa[k], sb[k])


@pytest.mark.parametrize("prec", [abi.PREC_FP64, abi.PREC_MIXED], ids=["fp64", "mixed"])
def test_noalloc_matches_step_two_calls(prec):
    over = {"precision": prec, "warmStart": 2} if prec == abi.PREC_MIXED else {}
    # the first call on the larger grid, the second on the smaller one
    for sc in (fast_marsh(96), cases.basin(nx=40, ny=36, Nt=60)):
        print(f"case {sc.name} precision {prec}: frehd_dev_step_noalloc vs frehd_dev_step, 25 steps", flush=True)
        ref = DeviceScenario(sc, **over)
        ref.step(25)
        got = DeviceScenario(sc, **over)
        noalloc(got, 25)
        bitwise(ref.download(NAMES), got.download(NAMES))
        stats_equal(ref, got)
        ref.close()
        got.close()


def test_noalloc_subgrid_and_mixed_with_step():
    sc = cases.marsh_subgrid(n=32, r=4, Nt=200, period=1200.0)
    names = NAMES + ["V", "Z", "Vx", "Vy", "Zx", "Zy", "Nx", "Oy", "ind"]
    print(f"case {sc.name}: noalloc and step interleaved on one handle vs step alone, 30 steps", flush=True)
    ref = DeviceScenario(sc)
    ref.step(30)
    got = DeviceScenario(sc)
    noalloc(got, 10)
    got.step(10)            # graphs and the handle's own partial buffers
    noalloc(got, 10)
    a, b = ref.download(names), got.download(names)
    for k in names:
        x, y = np.asarray(a[k]).reshape(-1), np.asarray(b[k]).reshape(-1)
        assert np.array_equal(x.view(np.int64) if x.dtype == np.float64 else x,
                              y.view(np.int64) if y.dtype == np.float64 else y), k
    stats_equal(ref, got)
    ref.close()
    got.close()


def test_noalloc_refuses_mid_step_and_batched():
    from frehdc_amd.device import Batch
    sc = cases.basin(nx=40, ny=36, Nt=40)
    print(f"case {sc.name}: noalloc guards", flush=True)
    d = DeviceScenario(sc)
    d.phase(0)
    assert d.lib.frehd_dev_step_noalloc(d.h, 1) == abi.FREHD_ESTATE
    for p in (1, 2, 3, 4):
        d.phase(p)
    noalloc(d, 1)
    b = Batch(1)
    b.add(d)
    assert d.lib.frehd_dev_step_noalloc(d.h, 1) == abi.FREHD_ESTATE
    b.remove(d)
    noalloc(d, 1)
    assert d.stats()["steps"] == 3
    d.close()

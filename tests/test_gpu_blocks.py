"""Block shape of the 2-D stencil kernels (frehd_dev_autotune_blocks,
frehd_batch_autotune_blocks): the tuner runs whole steps under each shape
(64 x 4, 128 x 2, 256 x 1 cells), so a tuned run passes through every shape;
since the shape changes no cell's arithmetic, the tuned run must end bitwise
equal to a run of as many steps under the default shape (every state field,
iteration counts), with only the volume losses' block partial sums in
another order (equal to 1e-12 relative).  Plain, subgrid and mixed-precision
scenarios, and a batch of unequal members against their solo runs.
"""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import Batch, DeviceScenario

pytestmark = pytest.mark.gpu

FIELDS = ["surf", "surfOld", "uuXP", "vvYP", "depth", "S", "cellV", "CDXP", "wtfXP", "EnXP", "DragYP"]


def fast_marsh(n=48, steps=200):
    sc = cases.marsh(n=n, dx=2.0, dt=0.5, Nt=steps, amp=0.6)
    sc.tideP[:] = cases.tide_series(steps, 0.5, 0.6, period=1200.0)
    return sc


SCEN = {
    "marsh48": (lambda: fast_marsh(48, 100), {}),
    "marsh64_sub4": (lambda: cases.marsh_subgrid(n=64, r=4, Nt=100, period=1200.0), {}),
    "basin40x36_mixed": (lambda: cases.basin(nx=40, ny=36, Nt=100), {"precision": abi.PREC_MIXED, "warmStart": 2}),
    "marsh96_sub4_mixed": (lambda: cases.marsh_subgrid(n=96, r=4, Nt=100, period=1200.0),
                           {"precision": abi.PREC_MIXED, "warmStart": 2}),
}


def assert_same(a, b, N, label):
    for k in FIELDS:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        x, y = (x[..., :N], y[..., :N]) if k == "S" else (x[:N], y[:N])
        assert np.array_equal(x, y), f"{label}: {k} differs (max |d| {float(np.max(np.abs(x - y))):.3e})"


@pytest.mark.parametrize("name", sorted(SCEN))
def test_tuned_run_equals_default_shape(name):
    make, over = SCEN[name]
    sc = make()
    A = DeviceScenario(sc, **over)
    B = DeviceScenario(sc, **over)
    try:
        A.step(3)
        B.step(3)
        tuned = B.autotune_blocks(2)          # 12 steps through all three shapes
        assert tuned["block_x"] in (64, 128, 256) and tuned["block_x"] * tuned["block_y"] == 256
        assert tuned["ms_per_step"] > 0
        A.step(12 + 10)
        B.step(10)
        sa, sb = A.stats(), B.stats()
        assert sa["steps"] == sb["steps"] == 25
        assert sa["pcgIterTotal"] == sb["pcgIterTotal"]
        assert_same(A.download(FIELDS), B.download(FIELDS), sc.N, name)
        np.testing.assert_allclose(sb["Vloss"], sa["Vloss"], rtol=1e-12, atol=1e-12)
        assert sa["maxCFLx"] == sb["maxCFLx"] and sa["maxCFLy"] == sb["maxCFLy"]
    finally:
        A.close()
        B.close()


def test_tuned_batch_equals_solo():
    scens = [fast_marsh(48, 60), cases.marsh(n=96, Nt=60), fast_marsh(64, 60)]
    solos = []
    for sc in scens:
        d = DeviceScenario(sc)
        d.step(3 + 6 + 5)
        solos.append((d.download(FIELDS), d.stats()))
        d.close()
    devs = [DeviceScenario(sc) for sc in scens]
    b = Batch(4)
    try:
        for d in devs:
            b.add(d)
        b.step(3)
        tuned = b.autotune_blocks(1)          # 6 steps
        assert tuned["block_x"] in (64, 128, 256)
        b.step(5)
        for d, sc, (ref, st) in zip(devs, scens, solos):
            assert_same(d.download(FIELDS), ref, sc.N, sc.name)
            assert d.stats()["pcgIterTotal"] == st["pcgIterTotal"]
        # a handle joining later takes the batch's shape
        late = DeviceScenario(fast_marsh(40, 60))
        try:
            late.step(2)
            b.add(late)
            b.step(2)
            assert late.stats()["steps"] == 4
        finally:
            b.remove(late)
            late.close()
    finally:
        b.close()
        for d in devs:
            d.close()


def test_autotune_blocks_needs_forcing():
    sc = fast_marsh(48, 20)
    d = DeviceScenario(sc)
    try:
        d.step(3)
        with pytest.raises(abi.FrehdError):
            d.autotune_blocks(5)              # 30 more steps than the 20-sample series
        assert d.stats()["steps"] == 3
    finally:
        d.close()

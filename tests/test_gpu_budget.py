"""Conservation on the device (frehd_dev_set_budget / frehd_dev_budget): the
flux-form volume update (volumeByFlux, nsfunctions.c:373-409, and its
subgrid form, subgrid.c:1227-1265) and each scalar's mass update
(scalar.c:68-274 with the inflow mass) close to 1e-12 relative in FP64:
per step, the change of the domain total minus the flux through the
boundary faces, the inflow and the negative-volume clip, summed in
double-double.  The terms that close the budget are far from negligible
(the faces move water and mass between the cells many orders above the
residual), and accounting changes no result: the state is bit for bit the same
with the budgets on and off."""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import DeviceScenario

pytestmark = pytest.mark.gpu

FIELDS = ["surf", "uuXP", "vvYP", "depth", "S", "Sm", "cellV", "CDXP"]


def inflow_basin():
    sc = cases.basin(nx=64, ny=40, Nt=300)
    sc.inflowLoc = np.array([10, 11, 12], dtype=np.int32)
    sc.cfg["inflowLocLength"] = 3
    sc.inflow[:] = 2.0
    sc.cfg["inflowS"] = 5.0
    return sc


def fast_marsh():
    sc = cases.marsh(n=48, dx=2.0, dt=0.5, Nt=300, amp=0.6)
    sc.tideP[:] = cases.tide_series(300, 0.5, 0.6, period=1200.0)
    return sc


def tvd_marsh():
    sc = fast_marsh()
    sc.cfg["scalarAdv"] = 2
    return sc


CASES = {
    "basin64x40_inflow": (inflow_basin, {}),
    "marsh48": (fast_marsh, {}),
    "marsh48_tvd": (tvd_marsh, {}),
    "marsh32_sub4": (lambda: cases.marsh_subgrid(n=32, r=4, Nt=300, period=1200.0), {}),
    "channel64_sub4_3scalars": (lambda: cases.channel_subgrid(n=64, r=4, Nt=300), {}),
    "marsh48_mixed": (fast_marsh, {"precision": abi.PREC_MIXED, "warmStart": 2}),
}


@pytest.mark.parametrize("name", list(CASES))
def test_budgets_close(name):
    build, over = CASES[name]
    sc = build()
    print(f"case {name}: 150 steps with the budgets on", flush=True)
    d = DeviceScenario(sc, **over)
    d.set_budget(True)
    d.step(150)
    b = d.budget()
    print(f"  volume {b['volume']:.6g} m^3, residual {b['volumeResidual']:.3g} (max step "
          f"{b['volumeResidualMax']:.3g}), boundary {b['volumeBoundary']:.6g}, inflow {b['volumeInflow']:.6g}, "
          f"clip {b['volumeClip']:.3g}; mass residuals {b['massRelResidual']}, boundary {b['massBoundary']}",
          flush=True)
    assert b["on"] and b["steps"] == 150
    assert b["volumeRelResidual"] <= 1e-12 and b["volumeRelResidualMax"] <= 1e-12
    for q in range(len(b["mass"])):
        assert b["massRelResidual"][q] <= 1e-12 and b["massRelResidualMax"][q] <= 1e-12, q
    # the check is not vacuous: the faces moved water and mass between the
    # cells many orders above what is left unbalanced
    assert b["volumeMoved"] > 1e8 * abs(b["volumeResidual"])
    for q in range(len(b["mass"])):
        assert b["massMoved"][q] > 1e8 * abs(b["massResidual"][q]), q
    d.close()


@pytest.mark.parametrize("name", ["basin64x40_inflow", "marsh32_sub4"])
def test_budgets_change_nothing(name):
    build, over = CASES[name]
    sc = build()
    print(f"case {name}: state with the budgets on vs off, 60 steps", flush=True)
    a = DeviceScenario(sc, **over)
    b = DeviceScenario(sc, **over)
    b.set_budget(True)
    a.step(60)
    b.step(60)
    fa, fb = a.download(FIELDS), b.download(FIELDS)
    for k in FIELDS:
        assert np.array_equal(np.asarray(fa[k]).view(np.int64), np.asarray(fb[k]).view(np.int64)), k
    b.set_budget(False)
    b.step(5)
    assert not b.budget()["on"] and b.budget()["steps"] == 0
    a.close()
    b.close()

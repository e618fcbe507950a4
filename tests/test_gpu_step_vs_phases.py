"""A whole step (frehd_dev_step: eager launches for the first steps, then
one captured graph per solver pattern) against the same steps phase by
phase (frehd_dev_phase, the path pinned to the oracle phase by phase in
test_gpu_parity, test_gpu_subgrid and test_gpu_configs), bit for bit, in
every storage form (FP64 cold and warm, the mixed form's FP32 state), on
the plain and the subgrid model, with wind, inflow and tides, and on grids
that are not multiples of any launch tile (300 x 37, 7 x 5)."""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import DeviceScenario

pytestmark = pytest.mark.gpu

CASES = {
    "basin300x37": lambda: cases.basin(nx=300, ny=37, Nt=60),
    "basin7x5": lambda: cases.basin(nx=7, ny=5, Nt=60),
    "marsh64_wind": lambda: cases.with_wind(cases.marsh(n=64, Nt=60)),
    "marsh32_sub4": lambda: cases.marsh_subgrid(n=32, r=4, Nt=60, period=1200.0),
    "channel64_sub4": lambda: cases.channel_subgrid(n=64, r=4, Nt=60, on_the_fly=False),
}

FORMS = {
    "fp64": {},
    "fp64_warm": {"warmStart": 2},
    "mixed": {"precision": abi.PREC_MIXED, "warmStart": 2},
}


def names(sc):
    return list(abi.FIELD_NAMES) + (list(abi.SUB_FIELD_NAMES) if sc.subgrid is not None else [])


@pytest.mark.parametrize("form", list(FORMS))
@pytest.mark.parametrize("case", list(CASES))
def test_step_equals_phases(case, form):
    print(f"case {case} form {form}", flush=True)
    sc = CASES[case]()
    steps = 12          # the step path turns to graphs after its first steps
    whole = DeviceScenario(sc, **FORMS[form])
    whole.step(steps)
    a = whole.download(names(sc))
    ta = whole.tt
    whole.close()
    phased = DeviceScenario(sc, **FORMS[form])
    for _ in range(steps):
        for p in range(5):
            phased.phase(p)
    b = phased.download(names(sc))
    tb = phased.tt
    phased.close()
    assert ta == tb == steps
    for k in a:
        bad = np.flatnonzero(a[k].reshape(-1) != b[k].reshape(-1))
        assert bad.size == 0, (k, bad[:8], a[k].reshape(-1)[bad[:4]], b[k].reshape(-1)[bad[:4]])

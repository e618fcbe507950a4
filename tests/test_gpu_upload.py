"""frehd_dev_upload on a subgrid handle: the uploaded state replaces what the
next step starts from, the subgrid lookup's inputs included.

The lookup (computeSubArea, subgrid.c:55-251) skips a cell whose surface is
bitwise the one it last saw, keeping its outputs.  Those outputs also depend
on the level search index and the face bottoms, so after an upload that
changes them a handle must forget what it saw: stepping it must give bit for
bit what a handle created from the same state gives.  Dry cells keep their
surface at the bed from step to step, so they are the cells where a stale
lookup would survive; resetting the level index to Init's value (Ne / 2,
subgrid.c:2272) exposes them (the walk moves it back to the cell's level)."""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import DeviceScenario

pytestmark = pytest.mark.gpu

NAMES = abi.FIELD_NAMES + abi.SUB_FIELD_NAMES + ["ind"]


def test_upload_forgets_subgrid_lookups():
    sc = cases.marsh_subgrid(n=24, r=4, Nt=100, period=1200.0)
    print(f"case {sc.name}: upload of a new level index and face bottoms", flush=True)
    a = DeviceScenario(sc)
    a.step(10)
    st = a.download(NAMES)
    N, Ne = sc.N, sc.cfg["Ne"]
    dry = st["depth"][:N] == 0
    assert np.any(dry) and np.any(st["ind"][:N][dry] != Ne // 2)
    new = {k: np.array(v, copy=True) for k, v in st.items()}
    new["ind"][:N] = Ne // 2
    rng = np.random.default_rng(3)
    pick = rng.random(N) < 0.3
    new["bottomXP"][:N][pick] -= 0.05
    new["bottomYP"][:N][pick] -= 0.05
    a.upload(new)
    b = DeviceScenario(sc, from_state=new)
    b.tt = a.tt
    a.step(5)
    b.step(5)
    ga, gb = a.download(NAMES), b.download(NAMES)
    for n in NAMES:
        x, y = np.asarray(ga[n]).reshape(-1), np.asarray(gb[n]).reshape(-1)
        if x.dtype == np.float64:
            assert np.array_equal(x.view(np.int64), y.view(np.int64)), n
        else:
            assert np.array_equal(x, y), n
    # the dry cells' index was walked back from Ne / 2
    assert np.any(ga["ind"][:N][dry] != Ne // 2)
    a.close()
    b.close()


def test_mid_step_calls_refused():
    """After phase 0 a handle is inside a step: a whole step, an upload or
    a new step index are refused until the step's phases are done."""
    sc = cases.basin(nx=40, ny=36, Nt=50)
    print(f"case {sc.name}: mid-step guards", flush=True)
    d = DeviceScenario(sc)
    d.phase(0)
    for call in (lambda: d.step(1), lambda: d.upload({"surf": d.field("surf")}),
                 lambda: setattr(d, "tt", 3)):
        with pytest.raises(abi.FrehdError) as e:
            call()
        assert e.value.code == abi.FREHD_ESTATE
    with pytest.raises(abi.FrehdError):
        d.phase(0)                       # would start another step
    for p in (1, 2, 3, 4):
        d.phase(p)
    d.step(1)
    assert d.stats()["steps"] == 2
    d.close()

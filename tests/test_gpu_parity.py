"""Device step vs the CPU oracle (which matches the reference bit for bit).

* Init and every phase except the solve are compared bit for bit: the device
  state is set to the oracle's state, one phase runs on both, and every array
  must agree exactly (the drag coefficient, which goes through pow(), within
  2 ulp).
* The Jacobi-PCG solve sums its dot products in a different order than the
  reference's sequential loops, so the new surface agrees to the solver
  tolerance: |d eta| <= 1e-9 m with eps = 1e-10 on metre-scale water levels.
* Wind stress: the direction's cosine and sine come from the host's libm, as
  in the reference, so the predictor stays bit-exact; on thin faces the
  stress decays through exp() (device and host may differ by an ulp), so
  En there agrees to 1e-14 relative (plus 1e-20 absolute where terms
  cancel).
* Whole runs: eta, u, v, S within 1e-7 after 200 wet/dry steps, wet/dry
  masks identical.
"""
import numpy as np
import pytest

from frehdc_amd import abi, cases
import pyoracle as po

pytestmark = pytest.mark.gpu

ULP_FIELDS = {"CDXP", "CDYP"}


def oracle_state(o):
    out = {}
    for n in abi.FIELD_NAMES:
        a = o.field(n)
        out[n] = np.ascontiguousarray(a).reshape(-1)
    return out


def fast_marsh(n=48, steps=400):
    sc = cases.marsh(n=n, dx=2.0, dt=0.5, Nt=steps, amp=0.6)
    # a short tidal period so cells wet and dry within the run
    sc.tideP[:] = cases.tide_series(steps, 0.5, 0.6, period=1200.0)
    return sc


def tvd_marsh():
    """The wet/dry marsh with TVD superbee scalar advection (scalarAdv = 2)."""
    sc = fast_marsh()
    sc.cfg["scalarAdv"] = 2
    sc.name += "_tvd"
    return sc


def wind_basin():
    """The basin under a gusting, veering wind (deep: no thin-layer faces)."""
    return cases.with_wind(cases.basin(nx=40, ny=36, Nt=200))


def wind_marsh():
    """The wet/dry marsh under wind, with the thin-layer decay on faces
    shallower than hD = 0.3 m."""
    sc = cases.with_wind(fast_marsh(), speed=12.0)
    sc.cfg["hD"] = 0.3
    return sc


def three_scalar_marsh():
    """The wet/dry marsh carrying three scalars in one pass: the boundary
    tracer, a plume, and a patch above 100 that updateScalar's out-of-range
    reset (scalar.c:392-402) removes, lowering those cells' surface and depth
    for the scalars after it."""
    sc = fast_marsh()
    n = sc.nx
    sc.cfg["nScalar"] = 3
    yy, xx = np.mgrid[0:n, 0:n]
    plume = np.exp(-((xx - n / 2) ** 2 + (yy - n / 2) ** 2) / (2 * (n / 8) ** 2))
    patch = np.where((abs(xx - n / 3) < 3) & (abs(yy - n / 2) < 3), 100.5, 5.0)
    sc.S0cell = np.ascontiguousarray(np.concatenate([np.zeros(n * n), plume.reshape(-1), patch.reshape(-1)]))
    sc.name += "_3scalars"
    return sc


def two_tide(sc, amp=0.3, period=1200.0):
    """A second tidal boundary along the first row (bcType = 3): the YM
    boundary's Dirichlet rows, surface values and tidal-face velocities
    (adjustTidalVelocity's second loop, nsfunctions.c:1058-1077, whose faces
    south of the first row are jM ghosts)."""
    nx, steps = sc.cfg["nx"], len(sc.tideP) - 1
    sc.cfg["bcType"] = 3
    sc.cfg["tideLocLengthM"] = nx
    sc.tideLocM = np.arange(nx, dtype=np.int32)
    sc.tideM[:] = cases.tide_series(steps, sc.cfg["dt"], amp, period=period, phase=1.0)
    sc.name += "_ym"
    return sc


def cases_small():
    return [cases.basin(nx=40, ny=36, Nt=200), fast_marsh(), tvd_marsh(), wind_basin(), wind_marsh(),
            three_scalar_marsh(), two_tide(fast_marsh())]


def compare(dev_fields, ref_fields, names, tag, ulp=ULP_FIELDS, rel=None):
    bad = []
    for n in names:
        a = np.asarray(dev_fields[n]).reshape(-1)
        b = np.asarray(ref_fields[n]).reshape(-1)
        m = min(len(a), len(b))
        a, b = a[:m], b[:m]
        if rel and n in rel:
            ok = np.all(np.abs(a - b) <= rel[n] * np.maximum(np.abs(a), np.abs(b)) + 1e-20)
        elif n in ulp:
            ok = np.all(np.abs(a - b) <= 2 * np.spacing(np.maximum(np.abs(a), np.abs(b))))
        elif a.dtype != np.float64 or b.dtype != np.float64:
            ok = np.array_equal(a, b)
        else:
            ok = np.array_equal(a.view(np.int64), b.view(np.int64))
        if not ok:
            idx = np.nonzero(a != b)[0]
            bad.append((n, len(idx), idx[:4].tolist(), a[idx[:2]].tolist(), b[idx[:2]].tolist()))
    assert not bad, f"{tag}: mismatches {bad}"


@pytest.fixture(scope="module")
def dev_mod():
    from frehdc_amd import device
    if device.device_count() < 1:
        pytest.skip("no GPU")
    return device


@pytest.mark.parametrize("sc", cases_small(), ids=lambda s: s.name)
def test_init_bit_exact(dev_mod, oracle_built, sc):
    print(f"case {sc.name}: init", flush=True)
    d = dev_mod.DeviceScenario(sc)
    o = po.Oracle(sc)
    compare(d.download(), oracle_state(o), abi.FIELD_NAMES, "init")


@pytest.mark.parametrize("k", [0, 3, 25])
@pytest.mark.parametrize("sc", cases_small(), ids=lambda s: s.name)
def test_phases_bit_exact(dev_mod, oracle_built, sc, k):
    print(f"case {sc.name}: phases after {k} steps", flush=True)
    o = po.Oracle(sc)
    o.step(k)
    # phase 0: predictor
    d = dev_mod.DeviceScenario(sc, from_state=oracle_state(o))
    d.tt = o.tt
    o.phase(po.PHASE_SOURCE)
    d.phase(0)
    thin_wind = sc.cfg.get("useWind") == 1 and sc.cfg.get("useThinLayer") == 1
    compare(d.download(), oracle_state(o), ["EnXP", "EnYP", "DragXP", "DragYP", "uuXP", "vvYP"],
            f"{sc.name} k={k} phase0", rel={"EnXP": 1e-14, "EnYP": 1e-14} if thin_wind else None)
    # phase 1: solve; surface to solver tolerance, the rest exact
    d.phase(1)
    o.phase(po.PHASE_SURFACE)
    got = d.download(["surf", "surfOld"])
    ref = oracle_state(o)
    compare(got, ref, ["surfOld"], "phase1")
    err = np.max(np.abs(got["surf"] - ref["surf"]))
    assert err <= 1e-9, f"phase1 surf error {err}"
    # phases 2-4 from the oracle's mid-step state
    d = dev_mod.DeviceScenario(sc, from_state=oracle_state(o))
    d.tt = o.tt
    d.phase(2)
    o.phase(po.PHASE_UPDATE)
    names = [n for n in abi.FIELD_NAMES if n not in ("CFLx", "CFLy", "S", "Sm")]
    compare(d.download(), oracle_state(o), names, f"{sc.name} k={k} phase2")
    d.phase(3)
    o.phase(po.PHASE_SCALAR)
    compare(d.download(), oracle_state(o), ["S", "Sm", "surf", "depth"], f"{sc.name} k={k} phase3")
    d.phase(4)
    o.phase(po.PHASE_CFL)
    compare(d.download(), oracle_state(o), ["CFLx", "CFLy"], f"{sc.name} k={k} phase4")


@pytest.mark.parametrize("sc", cases_small(), ids=lambda s: s.name)
def test_trajectory(dev_mod, oracle_built, sc):
    steps = 200
    print(f"case {sc.name}: {steps}-step trajectory", flush=True)
    d = dev_mod.DeviceScenario(sc)
    o = po.Oracle(sc)
    worst = {}
    for t in range(steps // 20):
        d.step(20)
        o.step(20)
        got = d.download(["surf", "uuXP", "vvYP", "depth", "S"])
        ref = oracle_state(o)
        N = sc.N
        for n in ("surf", "uuXP", "vvYP", "S"):
            e = float(np.max(np.abs(got[n].reshape(-1)[:N] - ref[n][:N])))
            worst[n] = max(worst.get(n, 0.0), e)
        wet_d = got["depth"][:N] > 0
        wet_o = ref["depth"][:N] > 0
        assert np.array_equal(wet_d, wet_o), f"wet mask mismatch at step {(t + 1) * 20}: {np.sum(wet_d != wet_o)} cells"
    print("max errors", worst, "iterations", d.stats()["pcgIterLast"], o.pcg_iters)
    for n, e in worst.items():
        assert e <= 1e-7, f"{n} error {e}"


def test_deterministic(dev_mod):
    sc = fast_marsh()
    runs = []
    for _ in range(2):
        d = dev_mod.DeviceScenario(sc)
        d.step(50)
        runs.append(d.download(["surf", "uuXP", "vvYP", "S"]))
    for n in runs[0]:
        assert np.array_equal(runs[0][n].view(np.int64), runs[1][n].view(np.int64)), n


def test_cfl_halt(dev_mod):
    sc = cases.basin(nx=32, ny=32, Nt=10)
    sc.tideP[:] = 0.0
    sc.tideP[1:] = 1000.0        # a 1 km jump at the tidal row drives CFL far above 20
    d = dev_mod.DeviceScenario(sc)
    rc = d.try_step(5)
    assert rc == abi.FREHD_ECFL
    assert d.tt == 1
    # the handle stays halted: later steps are refused
    assert d.try_step(1) == abi.FREHD_ECFL

"""Subgrid bathymetry on the device vs the CPU oracle (which matches the
reference's subgrid.c bit for bit): Init, every phase and whole runs.

Bit-exact except the PCG sums (tolerance as for the plain grid) and pow()
in the drag update (2 ulp)."""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import DeviceScenario
import pyoracle as po

from test_gpu_parity import compare

pytestmark = pytest.mark.gpu

SUB = abi.SUB_FIELD_NAMES


def oracle_state(o, sc):
    out = {}
    for n in abi.FIELD_NAMES + SUB:
        out[n] = np.ascontiguousarray(o.field(n) if n in abi.FIELD_NAMES else o.raw(n)).reshape(-1)
    out["ind"] = o.ind()
    return out


def tvd_sub():
    sc = cases.marsh_subgrid(n=28, r=4, Nt=300, period=1200.0)
    sc.cfg["scalarAdv"] = 2
    sc.name += "_tvd"
    return sc


def wind_sub():
    """Subgrid marsh under wind (windTerm also runs on the subgrid path); no
    thin layer, so the predictor is bit-exact (no exp)."""
    sc = cases.with_wind(cases.marsh_subgrid(n=28, r=4, Nt=300, period=1200.0), speed=12.0)
    sc.cfg["useThinLayer"] = 0
    return sc


def volp_sub():
    """Volp 2013 subgrid drag (useSubDrag = 2) with the CD cap active."""
    sc = cases.with_subdrag(cases.marsh_subgrid(n=28, r=4, Nt=300, period=1200.0), 2)
    sc.cfg["CDmax"] = 0.02
    return sc


def casas_sub():
    """Casas 2010 subgrid drag (useSubDrag = 4)."""
    return cases.with_subdrag(cases.marsh_subgrid(n=28, r=4, Nt=300, period=1200.0, manning_cell=False), 4)


def minarea_sub(useminV):
    """Minimum face areas (useminA = 1, beta 0.5) with the effective volumes
    of useminV = 0 (none), 1 (pow-based reduction factor) or 2 (the minimum
    volumes)."""
    return cases.with_minarea(cases.marsh_subgrid(n=28, r=4, Nt=300, period=1200.0), 0.5, useminV)


def generated_sub(**kw):
    """The generator's block-checked tables (minimum areas and volumes, drag
    tables) from the marsh's own pixels."""
    return cases.marsh_subgrid_generated(n=28, r=4, Nt=300, period=1200.0, **kw)


def ym_sub():
    """The subgrid marsh with a second tidal boundary along its first row
    (bcType = 3): the YM loop of adjustTidalVelocitySub (subgrid.c:1291-1332),
    whose faces south of the first row are jM ghosts."""
    sc = cases.marsh_subgrid(n=32, r=4, Nt=300, period=1200.0)
    nx = sc.cfg["nx"]
    sc.cfg["bcType"] = 3
    sc.cfg["tideLocLengthM"] = nx
    sc.tideLocM = np.arange(nx, dtype=np.int32)
    sc.tideM[:] = cases.tide_series(len(sc.tideP) - 1, sc.cfg["dt"], 0.6, period=900.0)
    sc.name += "_ym"
    return sc


def sub_cases():
    return [ym_sub(), cases.marsh_subgrid(n=32, r=4, Nt=300, period=1200.0),
            cases.marsh_subgrid(n=24, r=8, Nt=300, period=1200.0, manning_cell=False), tvd_sub(), wind_sub(),
            volp_sub(), casas_sub(), minarea_sub(0), minarea_sub(2), generated_sub(useminV=2),
            generated_sub(useminA=0, useminV=0, subdrag=2), generated_sub(useminA=0, useminV=0, subdrag=4,
                                                                        manning_cell=False)]


def test_minarea_effective_volume_pow(dev_mod, oracle_built):
    """useminA = 1 with useminV = 1: the volume reduction factor goes through
    pow(1 - r, 2), which the device and the host's libm round differently in
    the last place (glibc's pow(x, 2) is not x*x either, 1 in ~1200 draws), so
    Vx, Vy agree to 4 ulp after Init and whole runs to the trajectory
    tolerance, wet/dry masks identical."""
    sc = minarea_sub(1)
    print(f"case {sc.name}: init and 200-step trajectory", flush=True)
    d = dev_mod.DeviceScenario(sc)
    o = po.Oracle(sc)
    got, ref = dl(d, sc), oracle_state(o, sc)
    N = sc.N
    for n in ("Vx", "Vy"):
        a, b = got[n][:N], ref[n][:N]
        assert np.all(np.abs(a - b) <= 4 * np.spacing(np.maximum(np.abs(a), np.abs(b)))), n
    worst = {}
    for t in range(10):
        d.step(20)
        o.step(20)
        got, ref = d.download(["surf", "uuXP", "vvYP", "depth", "V"]), oracle_state(o, sc)
        for n in ("surf", "uuXP", "vvYP", "V"):
            worst[n] = max(worst.get(n, 0.0), float(np.max(np.abs(got[n].reshape(-1)[:N] - ref[n][:N]))))
        assert np.array_equal(got["depth"][:N] > 0, ref["depth"][:N] > 0), f"wet mask at {(t + 1) * 20}"
    print("max errors", worst)
    for n, e in worst.items():
        assert e <= 1e-7, f"{n} error {e}"


@pytest.fixture(scope="module")
def dev_mod():
    from frehdc_amd import device
    if device.device_count() < 1:
        pytest.skip("no GPU")
    return device


def dl(d, sc):
    return d.download(abi.FIELD_NAMES + SUB + ["ind"])


@pytest.mark.parametrize("sc", sub_cases(), ids=lambda s: s.name)
def test_subgrid_init_bit_exact(dev_mod, oracle_built, sc):
    print(f"case {sc.name}: init", flush=True)
    d = dev_mod.DeviceScenario(sc)
    o = po.Oracle(sc)
    compare(dl(d, sc), oracle_state(o, sc), abi.FIELD_NAMES + SUB + ["ind"], "init")


@pytest.mark.parametrize("k", [0, 5, 40])
@pytest.mark.parametrize("sc", sub_cases(), ids=lambda s: s.name)
def test_subgrid_phases_bit_exact(dev_mod, oracle_built, sc, k):
    print(f"case {sc.name}: phases after {k} steps", flush=True)
    o = po.Oracle(sc)
    o.step(k)
    d = dev_mod.DeviceScenario(sc, from_state=oracle_state(o, sc))
    d.tt = o.tt
    o.phase(po.PHASE_SOURCE)
    d.phase(0)
    compare(dl(d, sc), oracle_state(o, sc), ["EnXP", "EnYP", "DragXP", "DragYP"], f"k={k} phase0")
    d.phase(1)
    o.phase(po.PHASE_SURFACE)
    got, ref = dl(d, sc), oracle_state(o, sc)
    compare(got, ref, ["surfOld"], "phase1")
    assert np.max(np.abs(got["surf"] - ref["surf"])) <= 1e-9
    d = dev_mod.DeviceScenario(sc, from_state=oracle_state(o, sc))
    d.tt = o.tt
    d.phase(2)
    o.phase(po.PHASE_UPDATE)
    names = [n for n in abi.FIELD_NAMES + SUB + ["ind"] if n not in ("CFLx", "CFLy", "S", "Sm")]
    compare(dl(d, sc), oracle_state(o, sc), names, f"k={k} phase2")
    d.phase(3)
    o.phase(po.PHASE_SCALAR)
    compare(dl(d, sc), oracle_state(o, sc), ["S", "Sm", "surf", "depth"], f"k={k} phase3")
    d.phase(4)
    o.phase(po.PHASE_CFL)
    compare(dl(d, sc), oracle_state(o, sc), ["CFLx", "CFLy"], f"k={k} phase4")


@pytest.mark.parametrize("sc", sub_cases(), ids=lambda s: s.name)
def test_subgrid_trajectory(dev_mod, oracle_built, sc):
    steps = 200
    print(f"case {sc.name}: {steps}-step trajectory", flush=True)
    d = dev_mod.DeviceScenario(sc)
    o = po.Oracle(sc)
    worst = {}
    for t in range(steps // 20):
        d.step(20)
        o.step(20)
        got = d.download(["surf", "uuXP", "vvYP", "depth", "S", "V"])
        ref = oracle_state(o, sc)
        N = sc.N
        for n in ("surf", "uuXP", "vvYP", "S", "V"):
            e = float(np.max(np.abs(got[n].reshape(-1)[:N] - ref[n][:N])))
            worst[n] = max(worst.get(n, 0.0), e)
        assert np.array_equal(got["depth"][:N] > 0, ref["depth"][:N] > 0), f"wet mask at {(t + 1) * 20}"
    print("max errors", worst)
    for n, e in worst.items():
        assert e <= 1e-7, f"{n} error {e}"


@pytest.mark.parametrize("r", [1, 4, 8])
def test_device_tables_match_host(r):
    """Tables generated on the device from the fine bathymetry agree with the
    host generator (pinned by tests/test_subgrid_tables.py) to rounding: the
    sums run in a different order."""
    print(f"case device tables r={r}", flush=True)
    host = cases.marsh_subgrid(n=24, r=r, Nt=50, manning_cell=False)
    dev_sc = cases.marsh_subgrid(n=24, r=r, Nt=50, manning_cell=False, device_tables=True)
    d = DeviceScenario(dev_sc)
    t = d.tables()
    d.close()
    for k, v in t.items():
        ref = host.subgrid[k].reshape(v.shape)
        scale = max(1.0, float(np.max(np.abs(ref))))
        err = float(np.max(np.abs(v - ref)))
        assert err <= 1e-13 * scale, (k, err)


def test_device_tables_run_matches_host_tables():
    print("case marsh32_sub4 device vs host tables, 100 steps", flush=True)
    host = cases.marsh_subgrid(n=32, r=4, Nt=200, period=1200.0)
    dev_sc = cases.marsh_subgrid(n=32, r=4, Nt=200, period=1200.0, device_tables=True)
    a = DeviceScenario(host)
    b = DeviceScenario(dev_sc)
    a.step(100)
    b.step(100)
    fa = a.download(["surf", "uuXP", "vvYP"])
    fb = b.download(["surf", "uuXP", "vvYP"])
    N = host.N
    for k in fa:
        assert np.max(np.abs(fa[k][:N] - fb[k][:N])) <= 1e-9, k


@pytest.mark.parametrize("n,r,prec", [(40, 1, abi.PREC_FP64), (36, 3, abi.PREC_FP64), (32, 4, abi.PREC_MIXED),
                                      (24, 8, abi.PREC_FP64)])
def test_on_the_fly_bitwise_equal_to_tables(n, r, prec):
    """The on-the-fly lookup (pixels kept, two levels evaluated per lookup)
    gives bitwise the state of the stored device tables: both evaluate the
    levels with the same function.  Whole runs over the wet/dry cycle, every
    state array and the level index compared."""
    print(f"case on-the-fly marsh{n}_sub{r} precision {prec}, 150 steps", flush=True)
    a = DeviceScenario(cases.marsh_subgrid(n=n, r=r, Nt=300, period=1200.0, device_tables=True), precision=prec)
    b = DeviceScenario(cases.marsh_subgrid(n=n, r=r, Nt=300, period=1200.0, on_the_fly=True), precision=prec)
    assert a._cfg.subgridOnTheFly == 0 and b._cfg.subgridOnTheFly == 1
    for t in range(3):
        a.step(50)
        b.step(50)
        fa, fb = dl(a, None), dl(b, None)
        for k in fa:
            assert np.array_equal(fa[k], fb[k]), f"{k} differs after {(t + 1) * 50} steps"
    ta, tb = a.tables(), b.tables()
    for k in ta:
        assert np.array_equal(ta[k], tb[k]), k
    a.close()
    b.close()


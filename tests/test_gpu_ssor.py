"""SSOR-preconditioned CG on the device (k_ssor: the reference's
SSORPrecond with omega = 1, precond.c:48-73, swept diagonal by diagonal)
against the CPU oracle (pinned bit for bit to the reference's SSOR solve)
and against the reference's own SSOR run (tests/golden/basin40x36_ssor.npz).

* The preconditioned vector itself: with one CG iteration (maxIter = 1) the
  solve returns x = alpha z with z = M^-1 b, so the device's surface over
  the oracle's is one constant (their alphas differ in the last places: the
  dot products are summed in another order) on every cell the surface
  update does not clip -- a wrong sweep order or coefficient changes z
  cell by cell.  Tidal Dirichlet rows, wet/dry marsh rows and isolated dry
  subgrid rows (identity rows) are all in the cases.
* Whole solves: the same iteration count as the oracle's SSOR-PCG and the
  surface to the solver tolerance (1e-9 m).
* The reference's run: surface within 1e-9 m of the golden states at steps
  1, 5 and 40, wet/dry masks identical, iteration counts equal.  The
  Jacobi-preconditioned device path against the same SSOR run: both stop at
  |r| < eps |b| (eps = 1e-10) but at different iterates, so the surfaces
  differ by a few 1e-9 m (4.2e-9 m by step 40 here), held to 1e-8 m."""
import os

import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import DeviceScenario
import pyoracle as po

from test_golden import load_fixture
import test_gpu_parity
import test_gpu_subgrid

pytestmark = pytest.mark.gpu


def oracle_state(o, sc):
    if sc.subgrid is not None:
        return test_gpu_subgrid.oracle_state(o, sc)
    return test_gpu_parity.oracle_state(o)


HERE = os.path.dirname(os.path.abspath(__file__))
SSOR = abi.PRECOND_SSOR


def fast_marsh():
    sc = cases.marsh(n=48, dx=2.0, dt=0.5, Nt=200, amp=0.6)
    sc.tideP[:] = cases.tide_series(200, 0.5, 0.6, period=1200.0)
    return sc


BUILD = {"basin40x36": lambda: cases.basin(nx=40, ny=36, Nt=100),
         "marsh48": fast_marsh,
         "marsh24_sub4": lambda: cases.marsh_subgrid(n=24, r=4, Nt=200, period=1200.0),
         "basin96x80": lambda: cases.basin(nx=96, ny=80, Nt=100)}   # beyond the LDS vector: z in HBM


@pytest.mark.parametrize("name", list(BUILD))
def test_ssor_vector_is_the_references(oracle_built, name):
    sc = BUILD[name]()
    print(f"case {name}: SSOR, one CG iteration from Init", flush=True)
    o = po.Oracle(sc, precond=SSOR, maxIter=1)
    d = DeviceScenario(sc, from_state=oracle_state(o, sc) if sc.subgrid is not None else None,
                       precond=SSOR, maxIter=1)
    o.phase(po.PHASE_SOURCE)
    o.phase(po.PHASE_SURFACE)
    d.phase(0)
    d.phase(1)
    N = sc.N
    got, ref = d.field("surf")[:N], o.field("surf")[:N]
    zb = o.field("bottomZ")[:N]
    tidal = np.zeros(N, bool)
    tidal[sc.tideLocP] = True
    free = (got > zb) & (ref > zb) & ~tidal & (np.abs(ref) > 1e-3)
    assert np.count_nonzero(free) > N // 4
    ratio = got[free] / ref[free]
    spread = float(ratio.max() - ratio.min())
    print(f"  ratio {ratio.mean():.17g}, spread {spread:.3g} over {np.count_nonzero(free)} cells", flush=True)
    assert spread <= 8 * np.finfo(np.float64).eps
    d.close()


@pytest.mark.parametrize("name", list(BUILD))
def test_ssor_steps_match_oracle(oracle_built, name):
    sc = BUILD[name]()
    print(f"case {name}: SSOR-PCG steps vs the oracle", flush=True)
    o = po.Oracle(sc, precond=SSOR)
    d = DeviceScenario(sc, precond=SSOR)
    N = sc.N
    worst = 0.0
    for t in range(30):
        d.step(1)
        o.step(1)
        assert abs(d.stats()["pcgIterLast"] - o.pcg_iters) <= 1, (t, d.stats()["pcgIterLast"], o.pcg_iters)
        got, ref = d.download(["surf", "depth"]), oracle_state(o, sc)
        worst = max(worst, float(np.max(np.abs(got["surf"][:N] - ref["surf"][:N]))))
        assert np.array_equal(got["depth"][:N] > 0, ref["depth"][:N] > 0), t
    print(f"  max |d eta| {worst:.3g}, iterations {o.pcg_iters}", flush=True)
    assert worst <= 1e-9
    d.close()


@pytest.mark.parametrize("precond", [SSOR, abi.PRECOND_JACOBI], ids=["ssor", "jacobi"])
def test_against_reference_ssor_run(precond):
    sc, z = load_fixture(os.path.join(HERE, "golden", "basin40x36_ssor.npz"))
    print(f"case basin40x36_ssor golden, device precond {precond}", flush=True)
    d = DeviceScenario(sc, precond=precond)
    N = sc.N
    t, iters, worst = 0, [], {}
    for cp in (1, 5, 40):
        while t < cp:
            d.step(1)
            iters.append(d.stats()["pcgIterLast"])
            t += 1
        got = d.download(["surf", "depth", "uuXP", "vvYP"])
        e = float(np.max(np.abs(got["surf"][:N] - z[f"t{cp}_surf"][:N])))
        worst[cp] = e
        assert e <= (1e-9 if precond == SSOR else 1e-8), (cp, e)
        assert np.array_equal(got["depth"][:N] > 0, z[f"t{cp}_depth"][:N] > 0), cp
    print(f"  max |d eta| vs the reference's SSOR run at steps 1, 5, 40: {worst}; iterations {iters[:5]}...",
          flush=True)
    if precond == SSOR:
        ref = z["pcg_iters"][:40].tolist()
        assert all(abs(a - b) <= 1 for a, b in zip(iters, ref)), (iters, ref)
    d.close()


def test_ssor_refused_for_mixed_precision():
    sc = cases.basin(nx=40, ny=36, Nt=10)
    with pytest.raises(abi.FrehdError) as e:
        DeviceScenario(sc, precond=SSOR, precision=abi.PREC_MIXED)
    assert e.value.code == abi.FREHD_EINVAL

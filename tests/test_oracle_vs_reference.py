"""The CPU oracle against the reference's own functions, live: every phase of
every step bit for bit (runs where oracle/_ref/libfrehd_ref.so is built, i.e.
next to the reference checkout)."""
import numpy as np
import pytest

from frehdc_amd import cases

FIELDS = ["surf", "surfOld", "uuXP", "vvYP", "uuYP", "vvXP", "depth", "depthXP", "depthYP",
          "EnXP", "EnYP", "DragXP", "DragYP", "CDXP", "CDYP", "Fuu", "Fvv", "cellV", "S",
          "CFLx", "CFLy", "wtfXP", "wtfYP", "bottomZ", "bottomXP", "bottomYP", "GnCt", "GnXP",
          "GnXM", "GnYP", "GnYM"]
INTERIOR = {"uuYP", "vvXP", "depth", "cellV", "GnCt", "GnXP", "GnXM", "GnYP", "GnYM", "Vxp",
            "Vxm", "Vyp", "Vym", "Np", "Nm", "Op", "Om"}


def marsh_fast():
    sc = cases.marsh(n=40, dx=2.0, dt=0.5, Nt=200, amp=0.6)
    sc.manningCell = None
    sc.tideP[:] = cases.tide_series(200, 0.5, 0.6, period=1200.0)
    return sc


def marsh_sub():
    return cases.marsh_subgrid(n=20, r=4, Nt=200, period=1200.0, manning_cell=False)


def gen_sub(**kw):
    return cases.marsh_subgrid_generated(n=20, r=4, Nt=200, period=1200.0, manning_cell=False, **kw)


def two_tide(sc, amp=0.3):
    """A second tidal boundary along the first row (bcType = 3): the YM
    Dirichlet rows and surface values and adjustTidalVelocity's second loop
    (nsfunctions.c:1058-1077)."""
    nx, steps = sc.cfg["nx"], len(sc.tideP) - 1
    sc.cfg["bcType"] = 3
    sc.cfg["tideLocLengthM"] = nx
    sc.tideLocM = np.arange(nx, dtype=np.int32)
    sc.tideM[:] = cases.tide_series(steps, sc.cfg["dt"], amp, period=1200.0, phase=1.0)
    return sc


SUB_FIELDS = ["V", "Z", "Vx", "Vy", "Zx", "Zy", "Nx", "Oy", "Vxp", "Vxm", "Vyp", "Vym", "Np",
              "Nm", "Op", "Om", "minVx", "minVy"]

CASES = [
    ("marsh_subgrid4_jacobi", marsh_sub, {}),
    ("marsh_subgrid4_ssor", marsh_sub, {"precond": 1}),
    ("basin_jacobi", lambda: cases.basin(nx=30, ny=26, Nt=40), {}),
    ("basin_ssor", lambda: cases.basin(nx=30, ny=26, Nt=40), {"precond": 1}),
    ("marsh_jacobi", marsh_fast, {}),
    # two tidal boundaries (bcType 3: tideP on the last row, tideM on the first)
    ("marsh_ym", lambda: two_tide(marsh_fast()), {}),
    ("marsh_subgrid4_ym", lambda: two_tide(marsh_sub()), {}),
    ("basin_constCD_thin", lambda: cases.basin(nx=28, ny=24, Nt=40), {"CDnotN": 1, "hD": 2.6}),
    # TVD superbee scalar advection (scalar.c:99-255)
    ("basin_tvd", lambda: cases.basin(nx=30, ny=26, Nt=40), {"scalarAdv": 2}),
    ("marsh_tvd", marsh_fast, {"scalarAdv": 2}),
    ("marsh_subgrid4_tvd", marsh_sub, {"scalarAdv": 2}),
    # wind stress (nsfunctions.c:609-648), with the thin-layer decay on faces
    # shallower than hD
    ("basin_wind", lambda: cases.with_wind(cases.basin(nx=30, ny=26, Nt=40)), {}),
    ("marsh_wind", lambda: cases.with_wind(marsh_fast()), {"hD": 0.3}),
    ("marsh_subgrid4_wind", lambda: cases.with_wind(marsh_sub()), {}),
    # subgrid drag models (updateCDSub, subgrid.c:1481-1519): Volp 2013 with
    # the CD cap active on shallow levels, Casas 2010 (both model numbers)
    ("marsh_subgrid4_volp", lambda: cases.with_subdrag(marsh_sub(), 2), {"CDmax": 0.02}),
    ("marsh_subgrid4_casas3", lambda: cases.with_subdrag(marsh_sub(), 3), {}),
    ("marsh_subgrid4_casas4", lambda: cases.with_subdrag(marsh_sub(), 4), {"useThinLayer": 0}),
    # minimum face areas and effective volumes (combineSubArea, subgrid.c:360-414)
    ("marsh_subgrid4_minA_V1", lambda: cases.with_minarea(marsh_sub(), 0.5, 1), {}),
    ("marsh_subgrid4_minA_V0", lambda: cases.with_minarea(marsh_sub(), 0.8, 0), {}),
    ("marsh_subgrid4_minA_V2", lambda: cases.with_minarea(marsh_sub(), 0.3, 2), {}),
    ("marsh_subgrid4_minV2", lambda: cases.with_minarea(marsh_sub(), 0.0, 2), {"useminA": 0}),
    # the generator's own tables (block check, minimum areas and volumes,
    # drag tables) in the reference's step
    ("marsh_subgrid4_gen_minA_V1", lambda: gen_sub(useminV=1), {}),
    ("marsh_subgrid4_gen_minA_V2", lambda: gen_sub(useminV=2), {}),
    ("marsh_subgrid4_gen_volp", lambda: gen_sub(useminA=0, useminV=0, subdrag=2), {}),
    ("marsh_subgrid4_gen_casas", lambda: gen_sub(useminA=0, useminV=0, subdrag=3), {}),
]


@pytest.mark.parametrize("name,builder,override", CASES, ids=[c[0] for c in CASES])
def test_every_phase_bit_exact(oracle_built, name, builder, override):
    import pyoracle
    if not pyoracle.have_reference():
        pytest.skip("reference build absent")
    sc = builder()
    o = pyoracle.Oracle(sc, **override)
    r = pyoracle.Reference(sc, **override)
    steps = 40 if name.startswith("marsh") else 12
    names = FIELDS + (SUB_FIELDS if sc.subgrid is not None else [])
    for t in range(steps):
        for p in range(5):
            assert o.phase(p) == r.phase(p)
            if sc.subgrid is not None:
                assert np.array_equal(o.ind(), r.ind()), (name, t, p, "ind")
            for n in names:
                a = o.raw(n)
                b = r.raw(n)
                m = sc.N if n in INTERIOR else sc.Ng
                assert np.array_equal(a[:m].view(np.int64), b[:m].view(np.int64)), (name, t, p, n)
            zo, zr = o.z(), r.z()
            assert np.array_equal(zo.view(np.int64), zr.view(np.int64)), (name, t, p, "z")
        assert o.pcg_iters == r.pcg_iters
    assert np.array_equal(o.vloss(), r.vloss())

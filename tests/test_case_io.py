"""Input decks (include/frehd_io.h, frehdc_amd/case.py) against the
reference's own readers, on its shipped sample deck (FrehdC_v42/data.dat
and input/, copied to tests/golden/sample_deck/ by
tests/golden/make_sample_deck.py):

* settings: the derived values ReadUserSettings computes
  (configuration.c:18-158) for that deck;
* boundary series: readOneBC's output (initialize.c:400-460, via
  oracle/_ref) recorded in tests/golden/sample_deck/reference_readers.npz, bit for
  bit;
* bed and offset: ReadBathymetry + InitBathymetry's, bit for bit;
* with the reference present: the writer's files equal WriteFile's byte for
  byte, the subgrid tables equal initAllSubVar's, and the loaded deck run
  through the reference's own step functions equals the C restatement.
"""
import os
import shutil
import time

import numpy as np
import pytest

from frehdc_amd import abi
from frehdc_amd.case import Case

HERE = os.path.dirname(os.path.abspath(__file__))
DECK = os.path.join(HERE, "golden", "sample_deck")
GOLD = os.path.join(HERE, "golden", "sample_deck", "reference_readers.npz")
REFDIR = "/root/reference/FrehdC_v42"


@pytest.fixture(autouse=True)
def cst6():
    old = os.environ.get("TZ")
    os.environ["TZ"] = "CST6"
    time.tzset()
    yield
    if old is None:
        del os.environ["TZ"]
    else:
        os.environ["TZ"] = old
    time.tzset()


@pytest.fixture(scope="module")
def lib_built():
    try:
        abi.load_library()
    except OSError as e:
        pytest.skip(f"device library not built: {e}")


def test_sample_settings(lib_built):
    c = Case(os.path.join(DECK, "data.dat"), DECK)
    k, i = c.cfg, c.info
    assert (k.nx, k.ny, k.dx, k.dy, k.dt) == (30, 124, 16.0, 16.0, 2.0)
    assert i.Nt == k.Nt == 43200                # 1 day * 86400 / 2 s
    assert i.OutItvl == 900                     # 0.5 h * 3600 / 2 s
    assert i.tNStart == 1370844000.0            # 2013-06-10 00:00 CST6 = the tide files' first stamp
    assert k.bcType == 1 and k.tideLocLengthP == 10 and k.tideLocLengthM == 1
    tide = np.ctypeslib.as_array(c.forcing.tideLocP, shape=(10,))
    assert list(tide) == list(range(5 + 30 * 123, 15 + 30 * 123))   # first cell 5 on the top row
    inflow = np.ctypeslib.as_array(c.forcing.inflowLoc, shape=(1,))
    assert list(inflow) == [1]                  # j*NX + i - 1 for i = 2, j = 0
    assert (k.g, k.NUx, k.manningN, k.CDnotN, k.useThinLayer, k.hD, k.CDmax) == (9.8066, 1e-4, 0.03, 0, 1, 0.1, 1.0)
    assert (k.minDepth, k.wtfh, k.CFLl, k.CFLh, k.eps, k.maxIter) == (1e-5, 1e-5, 0.5, 0.7, 1e-8, 10000000)
    assert (k.useScalar, k.scalarAdv, k.useConstTidePS, k.tidePS, k.inflowS) == (1, 1, 1, 20.0, 0.0)
    assert (c.ic.surf0, c.ic.U0, c.ic.V0, c.ic.S0) == (0.2, 0.0, 0.0, 10.0)
    assert (i.savesurface, i.saveuu, i.savevv, i.savedepth, i.savescalar, i.saveCD) == (1, 1, 1, 1, 1, 0)
    assert k.useSubgrid == 0


def test_sample_series_and_bed_match_reference_readers(lib_built):
    c = Case(os.path.join(DECK, "data.dat"), DECK)
    sc = c.scenario()
    g = np.load(GOLD)
    for name in ("tideP", "inflow"):
        ours = getattr(sc, name)
        assert len(ours) == 43201
        assert np.array_equal(ours[:43200].view(np.int64), g[name].view(np.int64)), name
    assert np.array_equal(sc.bottomZ.view(np.int64), g["allZ"].view(np.int64))
    assert sc.cfg["offset"] == g["offset"][0]


def test_writer_format(lib_built, tmp_path):
    deck = tmp_path / "deck"
    shutil.copytree(DECK, deck)
    c = Case(str(deck / "data.dat"), str(deck))
    N = 30 * 124
    Ng = 32 * 126
    surf = np.zeros(Ng)
    surf[:N] = np.linspace(-1, 2, N) + c.info.offset
    c.write_arrays(7, {"surf": surf})
    text = (deck / "output" / "surfaceZ_7.dat").read_text().splitlines()
    assert len(text) == N
    assert text[0] == "%4.4f " % -1.0 and text[-1] == "%4.4f " % 2.0


def have_ref():
    import pyoracle
    return os.path.isdir(REFDIR) and pyoracle.have_reference()


@pytest.mark.skipif(not have_ref(), reason="reference checkout / oracle/_ref not present")
def test_writer_bytes_equal_reference(lib_built, tmp_path):
    import ctypes as C
    import pyoracle
    deck = tmp_path / "deck"
    shutil.copytree(DECK, deck)
    c = Case(str(deck / "data.dat"), str(deck))
    N, Ng = 30 * 124, 32 * 126
    rng = np.random.default_rng(4)
    u = rng.normal(0, 0.3, Ng)
    c.write_arrays(3, {"uuXP": u})
    ref = tmp_path / "ref/"
    ref.mkdir()
    lib = C.CDLL(pyoracle.REF_LIB)
    lib.ref_io_write((str(ref) + "/").encode(), b"uVelocity", 3,
                     np.ascontiguousarray(u[:N]).ctypes.data_as(C.POINTER(C.c_double)), N)
    assert (deck / "output" / "uVelocity_3.dat").read_bytes() == (ref / "uVelocity_3.dat").read_bytes()


@pytest.mark.skipif(not have_ref(), reason="reference checkout / oracle/_ref not present")
def test_subgrid_tables_equal_reference(lib_built, tmp_path):
    """The shipped subgrid data (input/subdata/, 50 levels) with useSubgrid
    switched on: our loader's tables equal initAllSubVar's."""
    import ctypes as C
    import pyoracle
    deck = tmp_path / "deck"
    deck.mkdir()
    os.symlink(os.path.join(REFDIR, "input"), deck / "input")
    lines = open(os.path.join(REFDIR, "data.dat")).read().splitlines()
    out = []
    for ln in lines:
        if ln.strip().endswith("# If subgrid model is used"):
            ln = "1;" + ln[2:]
        if ln.strip().endswith("# If effective volume is used"):
            ln = "1;" + ln[2:]
        out.append(ln)
    (deck / "data.dat").write_text("\n".join(out) + "\n")
    c = Case(str(deck / "data.dat"), str(deck))
    k = c.cfg
    assert k.useSubgrid == 1 and k.Ne == 50
    N = k.nx * k.ny
    sc = c.scenario()
    lib = C.CDLL(pyoracle.REF_LIB)
    P = C.POINTER(C.c_double)
    outs = [np.zeros(N * 50) for _ in range(14)]
    arr = (P * 14)(*[o.ctypes.data_as(P) for o in outs])
    ne = C.c_int()
    cwd = os.getcwd()
    os.chdir(deck)
    try:
        lib.ref_io_subtables(b"input/", b"subdata/", k.nx, k.ny, C.c_double(0.25), C.c_double(0.75),
                             C.c_double(0.01), C.c_double(k.offset), 1, arr, C.byref(ne))
    finally:
        os.chdir(cwd)
    assert ne.value == 50
    for t, name in enumerate(abi.SUB_TABLES):
        assert np.array_equal(sc.subgrid[name].view(np.int64), outs[t].view(np.int64)), name


@pytest.mark.skipif(not have_ref(), reason="reference checkout / oracle/_ref not present")
def test_loaded_deck_runs_like_reference(lib_built, oracle_built):
    import pyoracle
    c = Case(os.path.join(DECK, "data.dat"), DECK)
    sc = c.scenario()
    o = pyoracle.Oracle(sc)
    r = pyoracle.Reference(sc)
    for _ in range(30):
        assert o.step(1) == 0 and r.step(1) == 0
    for n in ("surf", "uuXP", "vvYP", "S", "depth"):
        a, b = o.field(n).reshape(-1), r.field(n).reshape(-1)
        m = min(len(a), len(b))
        assert np.array_equal(a[:m].view(np.int64), b[:m].view(np.int64)), n


def wind_deck(tmp_path):
    """The sample deck with the wind switched on (north angle 15 deg) and
    windspd.dat / winddir.dat in the tide files' format (copies of tideP.dat
    with the direction scaled, 721 samples each as data.dat says)."""
    deck = tmp_path / "wdeck"
    shutil.copytree(DECK, deck)
    out = []
    for ln in (deck / "data.dat").read_text().splitlines():
        if ln.strip().endswith("# If wind boundary condition exists"):
            ln = "1;" + ln[2:]
        elif ln.strip().endswith("# North angle"):
            ln = "15;" + ln[3:]
        out.append(ln)
    (deck / "data.dat").write_text("\n".join(out) + "\n")
    shutil.copy(deck / "input" / "tideP.dat", deck / "input" / "windspd.dat")
    rows = [ln.split() for ln in (deck / "input" / "tideP.dat").read_text().splitlines() if ln.strip()]
    (deck / "input" / "winddir.dat").write_text("".join(f"{t} {float(v) * 900.0:.6f}\n" for t, v in rows))
    return deck


def test_wind_deck_settings_and_series(lib_built, tmp_path):
    """Wind settings come from data.dat entries 32-35 and 77-79
    (configuration.c:63-66, 117-119); the wind series go through the same
    reader as the tides (readBC, initialize.c:362-366), so a copy of tideP.dat
    read as windspd.dat gives tideP's series bit for bit."""
    deck = wind_deck(tmp_path)
    c = Case(str(deck / "data.dat"), str(deck))
    k = c.cfg
    assert (k.useWind, k.northAngle, k.CwT, k.rhoa, k.Cw) == (1, 15.0, 10.0, 1.25, 0.0013)
    sc = c.scenario()
    assert np.array_equal(sc.windspd.view(np.int64), sc.tideP.view(np.int64))
    # the direction file holds 900 x tide to three more digits: the same
    # interpolation of values scaled by 900 (to the file's rounding)
    assert np.allclose(sc.winddir, 900.0 * sc.tideP, rtol=1e-6, atol=1e-6)
    assert np.ptp(sc.winddir) > 10.0


@pytest.mark.skipif(not have_ref(), reason="reference checkout / oracle/_ref not present")
def test_wind_deck_runs_like_reference(lib_built, oracle_built, tmp_path):
    import pyoracle
    deck = wind_deck(tmp_path)
    c = Case(str(deck / "data.dat"), str(deck))
    sc = c.scenario()
    o = pyoracle.Oracle(sc)
    r = pyoracle.Reference(sc)
    for _ in range(30):
        assert o.step(1) == 0 and r.step(1) == 0
    for n in ("surf", "uuXP", "vvYP", "EnXP", "EnYP", "S", "depth"):
        a, b = o.field(n).reshape(-1), r.field(n).reshape(-1)
        m = min(len(a), len(b))
        assert np.array_equal(a[:m].view(np.int64), b[:m].view(np.int64)), n


def subgrid_deck(tmp_path, subdrag=0, minA=None):
    """The reference's own deck with its shipped subgrid data (input/subdata/,
    50 levels) switched on, optionally with a subgrid drag model or the
    minimum face areas (minA = beta)."""
    deck = tmp_path / f"sdeck{subdrag}_{minA}"
    deck.mkdir()
    os.symlink(os.path.join(REFDIR, "input"), deck / "input")
    out = []
    for ln in open(os.path.join(REFDIR, "data.dat")).read().splitlines():
        if ln.strip().endswith("# If subgrid model is used"):
            ln = "1;" + ln[2:]
        if ln.strip().endswith("# If effective volume is used"):
            ln = "1;" + ln[2:]
        if subdrag and ln.strip().endswith("# The subgrid drag model used"):
            ln = f"{subdrag};" + ln[2:]
        if minA is not None and ln.strip().endswith("# If minimum face area is used"):
            ln = "1;" + ln[2:]
        if minA is not None and ln.strip().endswith("# Face area correction factor"):
            ln = f"{minA};" + ln[len(ln.split(';')[0]) + 1:]
        out.append(ln)
    (deck / "data.dat").write_text("\n".join(out) + "\n")
    return deck


@pytest.mark.skipif(not have_ref(), reason="reference checkout / oracle/_ref not present")
def test_volp_tables_equal_reference(lib_built, tmp_path):
    """useSubDrag = 2 on the shipped subgrid data: our loader's YhX / YhY
    tables equal initAllSubVar's (subgrid.c:2206-2216), and the loaded deck
    runs through the reference's own step functions like the C restatement,
    bit for bit."""
    import ctypes as C
    import pyoracle
    deck = subgrid_deck(tmp_path, 2)
    c = Case(str(deck / "data.dat"), str(deck))
    k = c.cfg
    assert k.useSubgrid == 1 and k.useSubDrag == 2 and k.Ne == 50
    N = k.nx * k.ny
    sc = c.scenario()
    lib = C.CDLL(pyoracle.REF_LIB)
    P = C.POINTER(C.c_double)
    outs = [np.zeros(N * 50) for _ in range(7)]
    arr = (P * 7)(*[o.ctypes.data_as(P) for o in outs])
    ne = C.c_int()
    cwd = os.getcwd()
    os.chdir(deck)
    try:
        lib.ref_io_extratables(b"input/", b"subdata/", k.nx, k.ny, C.c_double(0.25), C.c_double(0.75),
                               C.c_double(0.01), C.c_double(k.offset), 2, 0, arr, C.byref(ne))
    finally:
        os.chdir(cwd)
    assert ne.value == 50
    for t, name in enumerate(("YhX", "YhY")):
        assert np.array_equal(sc.subgrid[name].view(np.int64), outs[t].view(np.int64)), name
    assert np.any(sc.subgrid["YhX"] > 0)
    o = pyoracle.Oracle(sc)
    r = pyoracle.Reference(sc)
    for _ in range(20):
        assert o.step(1) == 0 and r.step(1) == 0
    for n in ("surf", "uuXP", "vvYP", "CDXP", "CDYP", "depth"):
        a, b = o.field(n).reshape(-1), r.field(n).reshape(-1)
        m = min(len(a), len(b))
        assert np.array_equal(a[:m].view(np.int64), b[:m].view(np.int64)), n


@pytest.mark.skipif(not have_ref(), reason="reference checkout / oracle/_ref not present")
def test_minarea_tables_and_run_equal_reference(lib_built, tmp_path):
    """useminA = 1 (beta 0.5) with useminV = 1 on the shipped subgrid data:
    our loader's Nmin/Nmax/Omin/Omax tables equal initAllSubVar's
    (subgrid.c:2160-2190), and the deck runs through the reference's own step
    functions like the C restatement, bit for bit."""
    import ctypes as C
    import pyoracle
    deck = subgrid_deck(tmp_path, 0, 0.5)
    c = Case(str(deck / "data.dat"), str(deck))
    k = c.cfg
    assert (k.useSubgrid, k.useminA, k.useminV, k.beta) == (1, 1, 1, 0.5)
    N = k.nx * k.ny
    sc = c.scenario()
    lib = C.CDLL(pyoracle.REF_LIB)
    P = C.POINTER(C.c_double)
    outs = [np.zeros(N * 50) for _ in range(7)]
    arr = (P * 7)(*[o.ctypes.data_as(P) for o in outs])
    ne = C.c_int()
    cwd = os.getcwd()
    os.chdir(deck)
    try:
        lib.ref_io_extratables(b"input/", b"subdata/", k.nx, k.ny, C.c_double(0.25), C.c_double(0.75),
                               C.c_double(0.01), C.c_double(k.offset), 0, 1, arr, C.byref(ne))
    finally:
        os.chdir(cwd)
    for t, name in enumerate(abi.MINA_TABLES):
        assert np.array_equal(sc.subgrid[name].view(np.int64), outs[3 + t].view(np.int64)), name
    assert np.any(sc.subgrid["Nmin"] > 0)
    o = pyoracle.Oracle(sc)
    r = pyoracle.Reference(sc)
    for _ in range(20):
        assert o.step(1) == 0 and r.step(1) == 0
    for n in ("surf", "uuXP", "vvYP", "depth"):
        a, b = o.field(n).reshape(-1), r.field(n).reshape(-1)
        m = min(len(a), len(b))
        assert np.array_equal(a[:m].view(np.int64), b[:m].view(np.int64)), n
    for n in ("Vx", "Vy", "Nx", "Oy"):
        a, b = o.raw(n), r.raw(n)
        assert np.array_equal(a[:N].view(np.int64), b[:N].view(np.int64)), n

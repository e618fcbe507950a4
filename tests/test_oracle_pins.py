"""Pins of the oracle's two extensions of the reference, against the
reference's own functions (oracle/_ref, built next to the reference
checkout):

* per-cell Manning n.  The reference has one Config.manningN (updateCD,
  nsfunctions.c:1170-1189; updateCDSub, subgrid.c:1520-1535; thinLayerDrag,
  nsfunctions.c:691-727).  The oracle runs with n drawn per cell from two
  values n1, n2; two reference runs hold n1 and n2 everywhere.  Between
  steps the reference runs take the oracle's drag coefficients, so all three
  walk the same trajectory; after every phase every array must agree bit
  for bit, except that the drag coefficients of updateData must equal the
  n1 run's on the n1 cells and the n2 run's on the n2 cells.  An index
  mistake (a neighbour's n, a transposed index) fails on the cells where
  the two values meet.
* several scalars.  The reference carries one (scalar.c:30-56).  The oracle
  runs the scalar pass once per scalar, in order.  Without an out-of-range
  reset the hydrodynamics do not see the scalars, so each scalar must equal
  a one-scalar reference run from its own initial field, bit for bit.  With
  a reset (S > 100 or S < 0: S = 0 and eta -= depth, scalar.c:400-412) in
  the second scalar only, the hydrodynamics and that scalar must follow the
  one-scalar run of the second field, the first scalar its own run up to
  the step of the reset, and the reset cells must lose exactly their depth.
"""
import dataclasses

import numpy as np
import pytest

from frehdc_amd import cases

HYDRO = ["surf", "surfOld", "uuXP", "vvYP", "uuYP", "vvXP", "depth", "depthXP", "depthYP",
         "EnXP", "EnYP", "DragXP", "DragYP", "Fuu", "Fvv", "cellV", "CFLx", "CFLy", "wtfXP", "wtfYP"]
SUB = ["V", "Z", "Vx", "Vy", "Zx", "Zy", "Nx", "Oy"]
INTERIOR = {"uuYP", "vvXP", "depth", "cellV"}


def need_reference():
    import pyoracle
    if not pyoracle.have_reference():
        pytest.skip("reference build absent")
    return pyoracle


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


def same(a, b, m):
    return np.array_equal(bits(a[:m]), bits(b[:m]))


def marsh_plain():
    sc = cases.marsh(n=40, dx=2.0, dt=0.5, Nt=200, amp=0.6)
    sc.tideP[:] = cases.tide_series(200, 0.5, 0.6, period=1200.0)
    return sc


def marsh_sub():
    return cases.marsh_subgrid(n=20, r=4, Nt=200, period=1200.0, manning_cell=False)


N1, N2 = 0.02, 0.035


def two_valued_manning(sc, seed=7):
    rng = np.random.default_rng(seed)
    cls = rng.random(sc.N) < 0.5
    return np.where(cls, N1, N2), cls


@pytest.mark.parametrize("builder", [marsh_plain, marsh_sub], ids=["marsh40", "marsh20_sub4"])
def test_per_cell_manning_matches_uniform_reference_runs(oracle_built, builder):
    po = need_reference()
    sc = builder()
    sc.manningCell, cls = two_valued_manning(sc)
    sc.manningCell = np.ascontiguousarray(sc.manningCell)
    o = po.Oracle(sc)
    plain = dataclasses.replace(sc, manningCell=None)
    r1 = po.Reference(plain, manningN=N1)
    r2 = po.Reference(plain, manningN=N2)
    N, Ng = sc.N, sc.Ng
    names = HYDRO + (SUB if sc.subgrid is not None else [])

    def check_cd(stage):
        # every interior face coefficient from its own cell's n
        for f in ("CDXP", "CDYP"):
            a, b1, b2 = o.raw(f)[:N], r1.raw(f)[:N], r2.raw(f)[:N]
            assert np.array_equal(bits(a)[cls], bits(b1)[cls]), (stage, f, "n1 cells")
            assert np.array_equal(bits(a)[~cls], bits(b2)[~cls]), (stage, f, "n2 cells")
        # the two classes really differ somewhere (the check is not vacuous)
        wet = o.raw("depth")[:N] > 0
        assert np.any(bits(r1.raw("CDXP")[:N])[wet] != bits(r2.raw("CDXP")[:N])[wet]), stage

    def sync_cd():
        for f in ("CDXP", "CDYP"):
            v = o.raw(f)
            r1.poke(f, v)
            r2.poke(f, v)

    check_cd("init")          # Init's thinLayerDrag
    sync_cd()
    mixed_faces = 0
    for t in range(30):
        for p in range(5):
            assert o.phase(p) == r1.phase(p) == r2.phase(p)
            for r in (r1, r2):
                for n in names:
                    m = N if n in INTERIOR else Ng
                    assert same(o.raw(n), r.raw(n), m), (t, p, n)
                assert same(o.raw("S"), r.raw("S"), Ng), (t, p, "S")
            if p == 2:
                check_cd((t, p))
                d = o.raw("depth")[:N] > 0
                mixed_faces += int(np.count_nonzero(d))
                sync_cd()
        assert o.pcg_iters == r1.pcg_iters == r2.pcg_iters
    assert mixed_faces > 0


S_NAMES = ["zero", "ramp", "patch"]


def scalar_fields(sc, big_patch=False):
    N, nx, ny = sc.N, sc.nx, sc.ny
    jj, ii = np.divmod(np.arange(N), nx)
    ramp = 35.0 * jj / (ny - 1)
    patch = ((np.abs(ii - nx // 2) < nx // 6) & (np.abs(jj - ny // 2) < ny // 6)).astype(np.float64)
    if big_patch:
        patch = patch * 150.0         # out of range: the reset fires on the wet patch cells
    return [np.zeros(N), ramp, patch]


def single(sc, field):
    return dataclasses.replace(sc, S0cell=np.ascontiguousarray(field), cfg=dict(sc.cfg, nScalar=1))


@pytest.mark.parametrize("builder", [marsh_plain, marsh_sub], ids=["marsh40", "marsh20_sub4"])
def test_scalars_match_one_scalar_reference_runs(oracle_built, builder):
    po = need_reference()
    sc = builder()
    sc.manningCell = None
    fields = scalar_fields(sc)
    multi = dataclasses.replace(sc, S0cell=np.ascontiguousarray(np.concatenate(fields)),
                                cfg=dict(sc.cfg, nScalar=3))
    o = po.Oracle(multi)
    refs = [po.Reference(single(sc, f)) for f in fields]
    N, Ng = sc.N, sc.Ng
    names = HYDRO + (SUB if sc.subgrid is not None else [])
    for t in range(30):
        for p in range(5):
            rcs = {o.phase(p)} | {r.phase(p) for r in refs}
            assert len(rcs) == 1
            S, Sm = o.raw("S").reshape(3, Ng), o.raw("Sm").reshape(3, Ng)
            for q, r in enumerate(refs):
                for n in names:
                    m = N if n in INTERIOR else Ng
                    assert same(o.raw(n), r.raw(n), m), (t, p, n, q)
                assert same(S[q], r.raw("S"), Ng), (t, p, "S", S_NAMES[q])
                assert same(Sm[q], r.raw("Sm"), N), (t, p, "Sm", S_NAMES[q])
    # the scalars really differ (each pass read its own field)
    S = o.raw("S").reshape(3, Ng)[:, :N]
    assert not np.array_equal(S[1], S[2]) and np.any(S[1] != 0)


def test_scalar_reset_chain(oracle_built):
    po = need_reference()
    sc = marsh_plain()
    sc.manningCell = None
    f = scalar_fields(sc, big_patch=True)
    first, second = f[1], f[2]          # benign ramp, then the out-of-range patch
    multi = dataclasses.replace(sc, S0cell=np.ascontiguousarray(np.concatenate([first, second])),
                                cfg=dict(sc.cfg, nScalar=2))
    o = po.Oracle(multi)
    r1 = po.Reference(single(sc, first))
    r2 = po.Reference(single(sc, second))
    N, Ng = sc.N, sc.Ng
    reset_step = None
    total_reset = 0
    for t in range(12):
        for p in range(5):
            if p == 3:
                surf0, depth0 = o.raw("surf")[:N].copy(), o.raw("depth")[:N].copy()
            assert o.phase(p) == r1.phase(p) == r2.phase(p)
            S = o.raw("S").reshape(2, Ng)
            # hydrodynamics and the resetting scalar follow its own run
            for n in HYDRO:
                m = N if n in INTERIOR else Ng
                assert same(o.raw(n), r2.raw(n), m), (t, p, n)
            assert same(S[1], r2.raw("S"), Ng), (t, p, "S second")
            if reset_step is None:
                # before the first reset the benign scalar follows its run too
                assert same(S[0], r1.raw("S"), Ng), (t, p, "S first")
            if p == 3:
                surf1, depth1 = o.raw("surf")[:N], o.raw("depth")[:N]
                hit = bits(surf1) != bits(surf0)
                if np.any(hit):
                    # eta dropped by exactly the depth, which is now zero
                    assert np.array_equal(bits(surf1[hit]), bits(surf0[hit] - depth0[hit])), t
                    assert np.all(depth1[hit] == 0) and np.all(S[1][:N][hit] == 0)
                    total_reset += int(np.count_nonzero(hit))
                    if reset_step is None:
                        reset_step = t
                        # the benign scalar ran before the reset in this step
                        assert same(S[0], r1.raw("S"), Ng), (t, "S first at the reset step")
    assert reset_step == 0 and total_reset > 0

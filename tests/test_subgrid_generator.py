"""The subgrid table generator with the reference's block check
(frehd_subgrid_generate, include/frehd_subgrid.h; preProcessing/
script_CreateSubAreaWithBlockV16.m:144-460 with ComputeFaceAreaWithBlockCheck.m)
and its text output.

The CPU oracle (oracle/subgrid_blockcheck.py, the script restated statement
by statement) is pinned first, by what does not depend on it:

* flat beds: every table has a closed form in h - z (face areas r dx d,
  minimum areas r d -- the script sums depths without the pixel width --,
  maximum areas the face areas, minimum volumes the half volumes);
* without any blocking (every cell one wet cluster touching its four faces)
  the face tables equal the plain generator's (frehdc_amd/subgrid.py,
  pinned by brute force), and checkBlock = 0 equals it everywhere;
* constructed beds worked out by hand: a dry ridge along a face, a cell
  whose small pool is cleaned away (its face then carries nothing although
  its pixels are wet), a channel with a contraction inside the cell (the
  maximum area falls to the contraction's and the minimum volume follows).

Then the native generator equals the oracle exactly on random beds (beds and
levels on a 1/64 m grid: every sum exact, so the order of summation does not
enter), with the drag-model tables to 1e-13, and the text files read back
equal the tables to the "%f" rounding.
"""
import os
import sys

import numpy as np
import pytest

from frehdc_amd import subgrid as sgm

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import subgrid_blockcheck as ob  # noqa: E402

TEN = ("V", "Z", "Vxp", "Vxm", "Vyp", "Vym", "Np", "Nm", "Op", "Om")


def level(t, name, k, N):
    return t[name][k * N:(k + 1) * N]


@pytest.mark.parametrize("r,dxf", [(4, 1.0), (8, 0.5)])
def test_oracle_flat_bed_closed_forms(r, dxf):
    nx, ny = 3, 4
    z0 = 0.25
    fine = np.full((ny * r, nx * r), z0)
    Ne, surfmin, dsurf = 6, 0.0, 0.125
    t = ob.generate(fine, r, dxf, surfmin, dsurf, Ne, check_block=True, min_volume=True)
    N = nx * ny
    dA = dxf * dxf
    for k in range(Ne):
        h = surfmin + k * dsurf
        d = max(h - z0, 0.0)
        assert np.all(level(t, "V", k, N) == dA * r * r * d)
        assert np.all(level(t, "Z", k, N) == (dA * r * r if d > 0 else 0.0))
        for q in ("Vxp", "Vxm", "Vyp", "Vym"):
            assert np.all(level(t, q, k, N) == dA * r * (r // 2) * d)
        for q in ("Np", "Nm", "Op", "Om"):
            assert np.all(level(t, q, k, N) == r * dxf * d), (q, k)
        # minimum areas: r pixel depths, no pixel width; maximum areas: the
        # face areas (the wider neighbourhood's minimum sits at its ends)
        for q in ("Nmin", "Omin"):
            assert np.all(level(t, q, k, N) == r * d)
        assert np.all(level(t, "Nmax", k, N) == level(t, "Np", k, N))
        assert np.all(level(t, "Omax", k, N) == level(t, "Op", k, N))
        for q, v in (("minVxp", "Vxp"), ("minVxm", "Vxm"), ("minVyp", "Vyp"), ("minVym", "Vym")):
            assert np.all(level(t, q, k, N) == level(t, v, k, N)), (q, k)
    assert np.all(t["bottom"] == z0) and np.all(t["bottomXP"] == z0) and np.all(t["bottomYP"] == z0)


def test_oracle_plain_faces_equal_the_plain_generator():
    rng = np.random.default_rng(3)
    r, nx, ny = 4, 3, 3
    fine = rng.integers(0, 32, (ny * r, nx * r)) / 64.0
    Ne, surfmin, dsurf = 10, 0.0, 1 / 16
    plain = sgm.tables(fine, r, 1.0, surfmin, dsurf, Ne)
    t0 = ob.generate(fine, r, 1.0, surfmin, dsurf, Ne, check_block=False, min_volume=False)
    for q in TEN:
        assert np.array_equal(t0[q], plain[q]), q
    # above every pixel nothing is blocked: the block check changes nothing
    t1 = ob.generate(fine, r, 1.0, 0.75, 1 / 16, 3, check_block=True, min_volume=False)
    p1 = sgm.tables(fine, r, 1.0, 0.75, 1 / 16, 3)
    for q in TEN:
        assert np.array_equal(t1[q], p1[q]), q


def test_oracle_dry_ridge_blocks_the_face():
    r, nx, ny = 4, 3, 2
    fine = np.zeros((ny * r, nx * r))
    fine[:, r - 1] = 1.0                      # the last pixel column of the first cells (x = 3)
    N = nx * ny
    t = ob.generate(fine, r, 1.0, 0.5, 1.0, 2, check_block=True, min_volume=False)
    # h = 0.5: the ridge is dry; the first cells' east faces and the second
    # cells' west faces carry nothing, the others r dx h
    for jj in range(ny):
        assert t["Np"][jj * nx + 0] == 0 and t["Nm"][jj * nx + 1] == 0
        assert t["Np"][jj * nx + 1] == r * 0.5 and t["Nm"][jj * nx + 2] == r * 0.5
    # the ridge cells' wet volume excludes the ridge column
    assert t["V"][0] == (r - 1) * r * 0.5
    # h = 1.5: everything wet, the face takes the shallower side: the ridge
    assert np.all(level(t, "Np", 1, N).reshape(ny, nx)[:, 0] == r * 0.5)
    assert np.all(level(t, "Np", 1, N).reshape(ny, nx)[:, 1] == r * 1.5)


def test_oracle_small_pool_is_cleaned_away():
    # cell (1, 1) of 3 x 3 (r = 8): a dry column at y = 5 splits it into a
    # 5-column pool and a 2-column pool; the smaller one touches the cell's
    # north face, which therefore carries nothing although its pixels are
    # wet, and the north neighbour's south face neither
    r, nx, ny = 8, 3, 3
    fine = np.zeros((ny * r, nx * r))
    fine[r + 5, r:2 * r] = 2.0                # y index r + 5 across the cell's x range
    t = ob.generate(fine, r, 1.0, 1.0, 1.0, 1, check_block=True, min_volume=False)
    c, n = 1 * nx + 1, 2 * nx + 1
    assert t["Op"][c] == 0 and t["Om"][n] == 0
    # without the block check the face is wet across its r pixels
    t0 = ob.generate(fine, r, 1.0, 1.0, 1.0, 1, check_block=False, min_volume=False)
    assert t0["Op"][c] == r * 1.0 and t0["Om"][n] == r * 1.0
    # the volume counts both pools; the east face keeps the kept pool's 5
    # pixels (7 wet ones without the block check: all but the dry one)
    assert t["V"][c] == (r * r - r) * 1.0
    assert t["Np"][c] == 5.0 and t0["Np"][c] == (r - 1) * 1.0


def test_oracle_contraction_sets_maximum_area_and_minimum_volume():
    # an x-channel 6 pixels wide through the middle row of cells, narrowing to
    # 2 pixels at x = 13 inside cell (1, 1) (its face is at x = 15)
    r, nx, ny = 8, 4, 3
    fine = np.full((ny * r, nx * r), 3.0)
    fine[r + 1:r + 7, :] = 0.0
    fine[r + 1:r + 3, 13] = 3.0
    fine[r + 5:r + 7, 13] = 3.0
    h = 1.0
    t = ob.generate(fine, r, 1.0, h, 1.0, 1, check_block=True, min_volume=True)
    c = 1 * nx + 1
    assert t["Np"][c] == 6 * h
    assert t["Nmin"][c] == 2 * h
    assert t["Nmax"][c] == 2 * h               # contraction: Nmax falls to Nmin
    Dx = r * 1.0
    assert t["minVxp"][c] == 0.5 * Dx * (2 * h)
    assert t["minVxm"][c + 1] == 0.5 * Dx * (2 * h)
    # the neighbouring face on the same channel keeps its area as maximum
    assert t["Nmax"][1 * nx + 2] == t["Np"][1 * nx + 2] == 6 * h


@pytest.mark.parametrize("seed", range(8))
def test_native_equals_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    r = [2, 4, 6, 8][seed % 4]
    nx, ny = 2 + seed % 3, 2 + (seed // 2) % 3
    fine = rng.integers(0, 64, (ny * r, nx * r)) / 64.0
    if seed % 2:
        fine = np.round(fine * 4) / 4              # plateaus: ties, blocked faces, cleaned pools
    kw = dict(check_block=seed % 3 != 2, min_volume=True, yh=True, cn=True)
    a = ob.generate(fine, r, 0.5, 0.0, 1 / 16, 18, **kw)
    b = sgm.generate(fine, r, 0.5, 0.0, 1 / 16, 18, threads=3, **kw)
    assert set(a) == set(b)
    for q in a:
        if q in ("YhX", "YhY", "Cn"):
            np.testing.assert_allclose(b[q], a[q], rtol=1e-13, atol=0, err_msg=q)
        else:
            assert np.array_equal(a[q], b[q]), q


def test_text_files_read_back(tmp_path):
    rng = np.random.default_rng(7)
    r, nx, ny, Ne = 4, 3, 2, 5
    fine = rng.integers(0, 64, (ny * r, nx * r)) / 64.0
    t = sgm.generate(fine, r, 1.0, 0.0, 0.125, Ne)
    folder = str(tmp_path / "subdata")
    sgm.write_text(folder, t, fine.shape, r, 1.0, 0.0, 0.125, Ne)
    for q, v in t.items():
        back = np.loadtxt(os.path.join(folder, f"{q}.dat"))
        assert back.shape == v.shape, q
        # "%f": six decimals, a half unit of the last one at most (ties such
        # as 1.0828125 land on it exactly)
        assert np.max(np.abs(back - v)) <= 5e-7 * (1 + 1e-6), q


def test_geometry_equals_the_plain_generator():
    # cell bottoms and face bottoms (script V16 :152-172) as the plain
    # generator computes them (the coarse bed and edgeX / edgeY the step reads)
    rng = np.random.default_rng(11)
    r, nx, ny = 4, 5, 3
    fine = rng.normal(0.0, 0.5, (ny * r, nx * r))
    t = sgm.generate(fine, r, 1.0, 0.0, 0.1, 2)
    bottom, edgeX, edgeY = sgm.coarse_geometry(fine, r)
    assert np.array_equal(t["bottom"], bottom.reshape(-1))
    assert np.array_equal(t["bottomXP"], edgeX.reshape(-1))
    assert np.array_equal(t["bottomYP"], edgeY.reshape(-1))

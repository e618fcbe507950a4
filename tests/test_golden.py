"""The CPU oracle against the golden fixtures: outputs of the reference's own
step functions (tests/golden/make_golden.py) on small scenarios, compared
bit for bit -- state after Init and after steps 1, 5 and 40, the PCG
iteration count of every step and the Vloss accumulators."""
import glob
import json
import os

import numpy as np
import pytest

from frehdc_amd import cases

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load_fixture(path):
    z = np.load(path, allow_pickle=False)
    cfg = json.loads(bytes(z["cfg_json"]).decode())
    S0 = z["in_S0cell"]
    sc = cases.Scenario(
        cfg=cfg, bottomZ=z["in_bottomZ"].copy(), tideP=z["in_tideP"].copy(),
        tideM=z["in_tideM"].copy(), inflow=z["in_inflow"].copy(),
        tidalPS=np.zeros_like(z["in_tideP"]), tidalMS=np.zeros_like(z["in_tideP"]),
        tideLocP=z["in_tideLocP"].astype(np.int32), tideLocM=z["in_tideLocM"].astype(np.int32),
        inflowLoc=z["in_inflowLoc"].astype(np.int32), manningCell=None,
        surf0=float(z["surf0"][0]), S0cell=S0.copy() if S0.size else None,
        name=os.path.basename(path)[:-4])
    if "sub_V" in z.files:
        sg = {k[4:]: z[k].copy() for k in z.files if k.startswith("sub_")}
        for m, src in (("minVxp", "Vxp"), ("minVxm", "Vxm"), ("minVyp", "Vyp"), ("minVym", "Vym")):
            sg[m] = sg[src]
        sc.subgrid = sg
    return sc, z


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.int64)


@pytest.mark.parametrize("path", FIXTURES, ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_matches_reference_fixture(oracle_built, path):
    import pyoracle
    sc, z = load_fixture(path)
    o = pyoracle.Oracle(sc)
    names = sorted({k.split("_", 1)[1] for k in z.files if k.startswith("t1_")})

    def check(tag):
        for n in names:
            ref = z[f"{tag}_{n}"]
            got = o.raw(n)[: len(ref)]
            assert np.array_equal(bits(got), bits(ref)), f"{sc.name} {tag} {n}"

    check("init")
    iters, vloss = [], []
    t = 0
    for cp in sorted(int(k[1:].split("_")[0]) for k in z.files if k.startswith("t") and k.endswith("_surf")):
        while t < cp:
            assert o.step(1) == 0
            iters.append(o.pcg_iters)
            vloss.append(o.vloss())
            t += 1
        check(f"t{cp}")
    assert iters == z["pcg_iters"].tolist()
    assert np.array_equal(bits(np.array(vloss)), bits(z["vloss"]))


def test_fixtures_present():
    assert len(FIXTURES) >= 4

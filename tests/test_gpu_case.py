"""The reference's shipped sample deck (tests/golden/sample_deck) through the
deck reader and the device time loop (frehdc_amd.case.run_case): the output
files of steps 0 and 900 (the deck's first output interval) hold the values
of the C restatement of the reference run on the same loaded inputs, to the
4 decimals the files carry (FP64 parity is 1e-9 m; a value within 1e-9 of a
rounding boundary may print one unit apart)."""
import os
import shutil
import time

import numpy as np
import pytest

import pyoracle
from frehdc_amd.case import Case, run_case

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DECK = os.path.join(HERE, "golden", "sample_deck")


def read(path):
    return np.array([float(x) for x in open(path).read().split()])


def test_sample_deck_on_device(tmp_path, oracle_built):
    os.environ["TZ"] = "CST6"
    time.tzset()
    deck = tmp_path / "deck"
    shutil.copytree(DECK, deck)
    print("case sample deck 30x124, 900 steps on the device", flush=True)
    dev = run_case(str(deck / "data.dat"), str(deck), steps=900)
    assert dev.tt == 900
    out = deck / "output"
    names = ["surfaceZ", "uVelocity", "vVelocity", "depth", "scalar"]
    for n in names:
        assert (out / f"{n}_0.dat").exists() and (out / f"{n}_900.dat").exists(), n
    # the C restatement on the same inputs, written by the same writer
    ref_dir = tmp_path / "oracle"
    shutil.copytree(DECK, ref_dir)
    case = Case(str(ref_dir / "data.dat"), str(ref_dir))
    sc = case.scenario()
    o = pyoracle.Oracle(sc)
    assert o.step(900) == 0
    fields = {"surf": o.field("surf").reshape(-1), "uuXP": o.field("uuXP").reshape(-1),
              "vvYP": o.field("vvYP").reshape(-1), "depth": o.field("depth").reshape(-1),
              "S": o.field("S").reshape(-1)}
    case.write_arrays(900, fields)
    for n in names:
        a, b = read(out / f"{n}_900.dat"), read(ref_dir / "output" / f"{n}_900.dat")
        assert a.shape == b.shape == (30 * 124,)
        assert np.max(np.abs(a - b)) <= 1.0001e-4, n
        assert np.mean(a == b) > 0.999, n
    # and at full precision through the handle
    got = dev.download(["surf"])["surf"][: sc.N]
    assert np.max(np.abs(got - fields["surf"][: sc.N])) <= 1e-9

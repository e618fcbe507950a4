"""bench.py's frozen algorithmic byte counts per cell-step, against the
closed forms they restate: plain FP64 208 + 72k B (26 accesses of the
explicit phases, 9 per Jacobi PCG iteration: cX, cY, D read, x, r, p read
and written), subgrid FP64 with ten tables 504 + 72k B, the FP64 solve from
a nonzero start 48 B more (r = b - A x_0); the mixed form's explicit state at
4 B but for eta, b, x (112 B plain, 4 (2T + 16) + 8 B more for the subgrid),
32 B per inner iteration and 56 B per refinement sweep; 16 B (FP64) or 8 B
(mixed) per extra scalar; and the roofline arithmetic (bytes x cells /
time)."""
import os
import sys

import pytest

sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import bench  # noqa: E402

PLAIN = {"nScalar": 1, "useSubgrid": 0}
SUB = {"nScalar": 1, "useSubgrid": 1, "useminV": 0, "useminA": 0}


def test_plain_fp64():
    assert bench.step_bytes(PLAIN, False, 18, 0) == 208 + 72 * 18
    assert bench.step_bytes(PLAIN, False, 0, 0) == 208
    # the warm start's residual pass
    assert bench.step_bytes(PLAIN, False, 2, 0, warm=True) == 208 + 144 + 48


def test_subgrid_fp64_tables():
    # ten tables: 208 + 8 (2*10 + 16) + 8 = 504, + 72 per iteration
    assert bench.step_bytes(SUB, False, 3, 0) == 504 + 216
    # minimum volumes that can act (useminV = 2, or 1 with the minimum
    # areas): T = 14 -> 568 + 72k
    assert bench.step_bytes(dict(SUB, useminV=2), False, 5, 0) == 568 + 360
    # useminV = 1 without useminA: the adjustment never fires, T = 10
    assert bench.step_bytes(dict(SUB, useminV=1), False, 5, 0) == 504 + 360
    # useminA with useminV = 1: T = 18
    assert bench.step_bytes(dict(SUB, useminV=1, useminA=1), False, 0, 0) == 208 + 8 * (36 + 16) + 8


def test_mixed_and_scalars():
    # predictor 52 B + update 60 B at the mixed widths
    assert bench.step_bytes(PLAIN, True, 10, 2) == 112 + 320 + 112
    # subgrid, ten tables at 4 B: 112 + 4 (2*10 + 16) + 8 = 264
    assert bench.step_bytes(SUB, True, 0, 0) == 264
    assert bench.step_bytes(dict(PLAIN, nScalar=3), False, 0, 0) == 208 + 32
    assert bench.step_bytes(dict(PLAIN, nScalar=3), True, 0, 0) == 112 + 16


def test_step_roofline_arithmetic():
    r = {"kt": {"k_refine": (0.02, 8)}, "iters": 4.0, "ms_per_step": 0.5}
    roof = bench.step_roofline(PLAIN, 1 << 20, r, True)
    per_cell = 112 + 32 * 4 + 56 * 2          # 8 sweeps over the 4 kernel-timed steps
    assert roof["bytes_per_cell_step"] == per_cell
    assert roof["achieved"] == pytest.approx(per_cell * (1 << 20) / 0.5e-3 / 1e9, rel=1e-4)
    assert roof["frac"] == pytest.approx(roof["achieved"] / 8000.0, rel=1e-3)

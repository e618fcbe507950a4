"""Batches: several scenarios advanced by shared kernel launches
(frehd_batch_*) give the same results as each scenario alone.

* FP64, plain and subgrid: bit for bit (each member's kernels, partial sums
  and solver iterations are its own; the launch only covers more members);
* mixed precision: within 1e-8 m (the launch pattern of the refinement is
  the batch's, so an inner solve may stop at a different iteration; both
  runs meet the same FP64 residual test);
* membership changes between steps (continuous batching): members that join
  late or leave early still match their solo runs bit for bit;
* a member that halts (CFL > 20) stops while the others go on.
"""
import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import Batch, DeviceScenario

pytestmark = pytest.mark.gpu

FIELDS = ["surf", "uuXP", "vvYP", "depth", "S", "cellV", "CDXP", "wtfXP"]


def fast_marsh(n=48, steps=400):
    sc = cases.marsh(n=n, dx=2.0, dt=0.5, Nt=steps, amp=0.6)
    sc.tideP[:] = cases.tide_series(steps, 0.5, 0.6, period=1200.0)
    return sc


def inflow_basin():
    sc = cases.basin(nx=64, ny=40, Nt=300)
    sc.inflowLoc = np.array([10, 11, 12], dtype=np.int32)
    sc.cfg["inflowLocLength"] = 3
    sc.inflow[:] = 2.0
    sc.cfg["inflowS"] = 5.0
    return sc


def solo(sc, steps, **over):
    d = DeviceScenario(sc, **over)
    d.step(steps)
    out = d.download(FIELDS)
    d.close()
    return out


def same(a, b, tol=0.0):
    for k in a:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        if tol == 0.0:
            assert np.array_equal(x.view(np.int64), y.view(np.int64)), k
        else:
            assert np.max(np.abs(x - y)) <= tol, (k, float(np.max(np.abs(x - y))))


def test_batch_plain_bitwise():
    print("case batch of 3 plain scenarios, 60 steps", flush=True)
    scs = [cases.basin(nx=40, ny=36, Nt=200), fast_marsh(), inflow_basin()]
    ref = [solo(sc, 60) for sc in scs]
    b = Batch(4)
    devs = [DeviceScenario(sc) for sc in scs]
    for d in devs:
        b.add(d)
    assert len(b) == 3
    b.step(60)
    for d, r in zip(devs, ref):
        same(d.download(FIELDS), r)
        assert d.tt == 60


def test_batch_subgrid_bitwise():
    print("case batch of 2 subgrid scenarios, 40 steps", flush=True)
    scs = [cases.marsh_subgrid(n=32, r=4, Nt=200, period=1200.0),
           cases.marsh_subgrid(n=24, r=8, Nt=200, period=1200.0, manning_cell=False)]
    fields = FIELDS + ["V", "Nx", "Oy", "ind"]
    ref = []
    for sc in scs:
        d = DeviceScenario(sc)
        d.step(40)
        ref.append(d.download(fields))
        d.close()
    b = Batch(2)
    devs = [DeviceScenario(sc) for sc in scs]
    for d in devs:
        b.add(d)
    b.step(40)
    for d, r in zip(devs, ref):
        got = d.download(fields)
        for k in fields:
            if k == "ind":
                assert np.array_equal(got[k], r[k])
            else:
                assert np.array_equal(got[k].view(np.int64), r[k].view(np.int64)), k


def test_batch_mixed_within_tolerance():
    print("case batch of 2 mixed-precision scenarios, 80 steps", flush=True)
    scs = [cases.basin(nx=40, ny=36, Nt=200), cases.marsh(n=96, Nt=200)]
    over = dict(precision=abi.PREC_MIXED, warmStart=2)
    ref = [solo(sc, 80, **over) for sc in scs]
    b = Batch(2)
    devs = [DeviceScenario(sc, **over) for sc in scs]
    for d in devs:
        b.add(d)
    b.step(80)
    for d, r, sc in zip(devs, ref, scs):
        got = d.download(["surf"])
        N = sc.N
        assert np.max(np.abs(got["surf"][:N] - r["surf"][:N])) <= 1e-8
        assert d.stats()["relResLast"] < sc.cfg["eps"]


def test_batch_membership_changes():
    """A joins at 0 and leaves at 30; B joins at 10 and runs to 50; C joins at
    30.  Every member ends bit-identical to its solo run of the same length."""
    print("case continuous batching, joins and leaves", flush=True)
    sa, sb, sc_ = cases.basin(nx=40, ny=36, Nt=200), fast_marsh(), inflow_basin()
    ra, rb, rc = solo(sa, 30), solo(sb, 40), solo(sc_, 20)
    b = Batch(3)
    A, B, C = DeviceScenario(sa), DeviceScenario(sb), DeviceScenario(sc_)
    b.add(A)
    b.step(10)
    b.add(B)
    b.step(20)
    b.remove(A)
    b.add(C)
    b.step(20)
    assert A.tt == 30 and B.tt == 40 and C.tt == 20
    same(A.download(FIELDS), ra)
    same(B.download(FIELDS), rb)
    same(C.download(FIELDS), rc)


def test_batch_rejects_mixed_kinds():
    b = Batch(2)
    b.add(DeviceScenario(cases.basin(nx=40, ny=36, Nt=50)))
    sub = DeviceScenario(cases.marsh_subgrid(n=24, r=4, Nt=50))
    with pytest.raises(abi.FrehdError):
        b.add(sub)
    assert len(b) == 1


def test_batch_member_halts_others_continue():
    print("case halting member in a batch", flush=True)
    bad = cases.basin(nx=32, ny=32, Nt=40)
    bad.tideP[:] = 0.0
    bad.tideP[1:] = 1000.0
    good = cases.basin(nx=40, ny=36, Nt=200)
    ref = solo(good, 20)
    b = Batch(2)
    G, X = DeviceScenario(good), DeviceScenario(bad)
    b.add(G)
    b.add(X)
    assert b.try_step(20) == abi.FREHD_ECFL
    assert X.tt == 1
    assert G.tt == 20
    same(G.download(FIELDS), ref)

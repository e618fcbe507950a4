"""Serving loop (frehdc_amd/serving.py): requests of uneven grid size and
run length, arriving over time, share batched launches; every result equals
the same scenario run alone, bit for bit (FP64), and every request finishes
with its own step count."""
import numpy as np
import pytest

from frehdc_amd import abi, serving
from frehdc_amd.device import DeviceScenario

pytestmark = pytest.mark.gpu


def test_serving_results_equal_solo_runs():
    print("case serving 10 requests, sizes 64..256, steps 20..80", flush=True)
    reqs = serving.make_requests(10, rate=200.0, seed=11, lo=64, hi=256, steps_lo=20, steps_hi=80)
    fields = ["surf", "uuXP", "vvYP", "S"]
    srv = serving.Server(budget_bytes=2e9, quantum=7, collect=fields)
    st = srv.run(reqs)
    print(st)
    assert st["requests"] == 10 and st["latency_p99_s"] >= st["latency_p50_s"] > 0
    for r in reqs:
        d = DeviceScenario(r.scen)
        d.step(r.steps)
        solo = d.download(fields)
        d.close()
        for k in fields:
            assert np.array_equal(solo[k].view(np.int64), r.result[k].view(np.int64)), (r.rid, k)


def test_serving_warm_fp64_equals_solo_runs():
    # warm FP64 starts vary their iteration counts from step to step and
    # member to member, so the shared pattern leaves solves open that the
    # host finishes, with several buckets stepping from their own threads;
    # the FP64 iterates do not depend on that, so each result is still the
    # solo run's, bit for bit, with every step taken
    print("case serving 16 requests, warm FP64, sizes 64..512, steps 40..160", flush=True)
    reqs = serving.make_requests(16, rate=400.0, seed=21, lo=64, hi=512, steps_lo=40, steps_hi=160, subgrid_r=4)
    srv = serving.Server(budget_bytes=20e9, quantum=16, collect=["surf", "depth"], warmStart=2)
    srv.run(reqs)
    for r in reqs:
        d = DeviceScenario(r.scen, warmStart=2)
        d.step(r.steps)
        assert d.stats()["steps"] == r.steps
        solo = d.download(["surf", "depth"])
        d.close()
        for k in ("surf", "depth"):
            assert np.array_equal(solo[k], r.result[k]), (r.rid, k, float(np.max(np.abs(solo[k] - r.result[k]))))


def test_serving_budget_limits_residency():
    print("case serving under a tight HBM budget", flush=True)
    reqs = serving.make_requests(6, rate=1000.0, seed=12, lo=128, hi=128, steps_lo=10, steps_hi=10)
    one = serving.footprint(reqs[0].scen)
    srv = serving.Server(budget_bytes=2.5 * one, quantum=5)
    st = srv.run(reqs)
    assert st["peak_resident_scenarios"] <= 2
    assert st["peak_resident_bytes"] <= 2.5 * one


def test_serving_mixed_precision():
    print("case serving, mixed precision", flush=True)
    reqs = serving.make_requests(6, rate=500.0, seed=13, lo=64, hi=192, steps_lo=20, steps_hi=40)
    srv = serving.Server(budget_bytes=2e9, quantum=10, collect=["surf"], precision=abi.PREC_MIXED, warmStart=1)
    srv.run(reqs)
    for r in reqs:
        d = DeviceScenario(r.scen)
        d.step(r.steps)
        ref = d.download(["surf"])["surf"]
        d.close()
        assert np.max(np.abs(ref - r.result["surf"])) <= 1e-6, r.rid

"""Steps whose solve the launch pattern leaves open, one after another.

FREHD_PATTERN_CAP (read when a handle or batch is created) caps the solver
launches a step enqueues ahead, so nearly every solve is left open and the
host finishes it (eager iterations, then the gather that gives the kernels
after the solve their go-ahead, then those kernels).  In the captured and
streamed launch modes the host does not wait for that gather; a gather it
does not wait for writes no flags to the host, or it could land after the
host has cleared the flags for the next step and stand in for that step's
(which then looked closed and lost its update).  Here every step of a
multi-step call is left open, in every launch mode: each run must take all
its steps and equal the kernel-by-kernel run bit for bit -- and, for the
FP64 solve, whose iterates do not depend on where the pattern cuts it, the
uncapped run too; single handles and a batch of unequal members.
"""
import os

import numpy as np
import pytest

from frehdc_amd import abi, cases
from frehdc_amd.device import Batch, DeviceScenario

pytestmark = pytest.mark.gpu

FIELDS = ["surf", "surfOld", "uuXP", "vvYP", "depth", "S", "cellV"]


def fast_marsh(n=48, steps=120):
    sc = cases.marsh(n=n, dx=2.0, dt=0.5, Nt=steps, amp=0.6)
    sc.tideP[:] = cases.tide_series(steps, 0.5, 0.6, period=1200.0)
    return sc


SCEN = {
    "marsh48_fp64": (lambda: fast_marsh(48), {}),
    "marsh48_fp64_warm": (lambda: fast_marsh(48), {"warmStart": 2}),
    "marsh64_sub4_mixed": (lambda: cases.marsh_subgrid(n=64, r=4, Nt=120, period=1200.0),
                           {"precision": abi.PREC_MIXED, "warmStart": 2}),
}


def make(env, fn):
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        return fn()
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v


def same(a, b, N):
    for k in FIELDS:
        x, y = np.asarray(a[k]), np.asarray(b[k])
        x, y = (x[..., :N], y[..., :N]) if k == "S" else (x[:N], y[:N])
        if not np.array_equal(x, y):
            return f"{k} differs (max |d| {float(np.max(np.abs(x - y))):.3e})"
    return None


@pytest.mark.parametrize("name", sorted(SCEN))
def test_open_solves_every_step(name):
    mk, over = SCEN[name]
    sc = mk()
    runs = {}
    for mode in ("0", "1", "2"):
        d = make({"FREHD_GRAPHS": mode, "FREHD_PATTERN_CAP": "1"}, lambda: DeviceScenario(sc, **over))
        try:
            d.step(40)                 # one call: the host goes on to the next step right away
            d.step(20)
            runs[mode] = (d.download(FIELDS), d.stats())
        finally:
            d.close()
    for mode in ("1", "2"):
        st = runs[mode][1]
        assert st["steps"] == 60, f"{name} mode {mode}: {st['steps']} steps taken of 60"
        msg = same(runs[mode][0], runs["0"][0], sc.N)
        assert msg is None, f"{name} mode {mode}: {msg}"
        assert st["pcgIterTotal"] == runs["0"][1]["pcgIterTotal"]
    if "precision" not in over:
        # the FP64 iterates do not depend on where the pattern stops
        d = make({"FREHD_GRAPHS": "2"}, lambda: DeviceScenario(sc, **over))
        try:
            d.step(60)
            msg = same(d.download(FIELDS), runs["0"][0], sc.N)
            assert msg is None, f"{name} uncapped: {msg}"
        finally:
            d.close()


@pytest.mark.parametrize("mode", ["1", "2"])
def test_open_solves_in_a_batch(mode):
    scens = [fast_marsh(48), cases.marsh(n=96, Nt=120), fast_marsh(64)]
    over = {"warmStart": 2}
    solos = []
    for sc in scens:
        d = DeviceScenario(sc, **over)
        d.step(50)
        solos.append((d.download(FIELDS), d.stats()))
        d.close()
    env = {"FREHD_GRAPHS": mode, "FREHD_PATTERN_CAP": "2"}
    devs = [make(env, lambda sc=sc: DeviceScenario(sc, **over)) for sc in scens]
    b = make(env, lambda: Batch(4))
    try:
        for d in devs:
            b.add(d)
        b.step(50)
        for d, sc, (ref, st) in zip(devs, scens, solos):
            got = d.stats()
            assert got["steps"] == 50, f"{sc.name}: {got['steps']} steps taken of 50"
            msg = same(d.download(FIELDS), ref, sc.N)
            assert msg is None, f"{sc.name}: {msg}"
            assert got["pcgIterTotal"] == st["pcgIterTotal"]
    finally:
        b.close()
        for d in devs:
            d.close()


def test_open_solves_concurrent_handles():
    # two handles stepped from two host threads (as the serving loop steps its
    # batches): each one's gathers run behind the other's kernels as well
    import threading
    scens = [(cases.marsh_subgrid(n=64, r=4, Nt=120, period=1200.0), {"warmStart": 2}),
             (cases.marsh(n=96, Nt=120), {"warmStart": 2}), (fast_marsh(64), {})]
    refs = []
    for sc, over in scens:
        d = make({"FREHD_GRAPHS": "0"}, lambda: DeviceScenario(sc, **over))
        d.step(80)
        refs.append(d.download(FIELDS))
        d.close()
    env = {"FREHD_GRAPHS": "2", "FREHD_PATTERN_CAP": "1"}
    devs = [make(env, lambda sc=sc, over=over: DeviceScenario(sc, **over)) for sc, over in scens]
    rcs = [None] * len(devs)

    def run(q):
        rcs[q] = devs[q].try_step(80)

    try:
        ts = [threading.Thread(target=run, args=(q,)) for q in range(len(devs))]
        for t in ts:
            t.start()
        for t in ts:
            t.join()
        for (sc, _), d, ref, rc in zip(scens, devs, refs, rcs):
            assert rc == 0
            assert d.stats()["steps"] == 80, f"{sc.name}: {d.stats()['steps']} steps taken of 80"
            msg = same(d.download(FIELDS), ref, sc.N)
            assert msg is None, f"{sc.name}: {msg}"
    finally:
        for d in devs:
            d.close()

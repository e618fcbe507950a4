"""The serving loop's host logic on the CPU (frehdc_amd/serving.py), with the
device handles replaced by stand-ins that count steps: admission against
the HBM budget in arrival order, size buckets, scheduling rounds that never
step a request past its run, departures that free budget and slots, and
the halt path.  Footprints come from the real library (frehd_dev_footprint
runs without a device)."""
import types

import numpy as np
import pytest

from frehdc_amd import abi, cases, serving


class FakeDev:
    live = []

    def __init__(self, scen, **over):
        self.scen = scen
        self.steps = 0
        self.halted = False
        self.closed = False
        self.halt_at = getattr(scen, "halt_at", None)
        FakeDev.live.append(self)

    def stats(self):
        return {"halted": self.halted}

    def download(self, names):
        return {n: np.zeros(1) for n in names}

    def close(self):
        self.closed = True


class FakeBatch:
    log = []

    def __init__(self, capacity):
        self.capacity = capacity
        self.members = []
        self.lib = types.SimpleNamespace(frehd_last_error=lambda: b"")

    def add(self, d):
        assert len(self.members) < self.capacity
        self.members.append(d)

    def remove(self, d):
        self.members.remove(d)

    def try_step(self, k):
        FakeBatch.log.append((id(self), k, [m.scen.name for m in self.members]))
        rc = 0
        for m in self.members:
            for _ in range(k):
                if m.halted:
                    break
                m.steps += 1
                if m.halt_at is not None and m.steps >= m.halt_at:
                    m.halted = True
            if m.halted:
                rc = abi.FREHD_ECFL
        return rc

    def close(self):
        pass


@pytest.fixture
def fake_device(monkeypatch):
    FakeDev.live = []
    FakeBatch.log = []
    monkeypatch.setattr(serving, "DeviceScenario", FakeDev)
    monkeypatch.setattr(serving, "Batch", FakeBatch)


def request(rid, side, steps, arrival=0.0, halt_at=None):
    sc = cases.basin(nx=side, ny=side, Nt=steps + 1)
    sc.name = f"r{rid}"
    if halt_at is not None:
        sc.halt_at = halt_at
    return serving.Request(rid=rid, arrival=arrival, steps=steps, scen=sc)


def test_every_request_runs_exactly_its_steps(fake_device):
    reqs = [request(q, side, steps) for q, (side, steps) in
            enumerate([(64, 37), (100, 50), (200, 16), (64, 3), (300, 41), (128, 16)])]
    st = serving.Server(budget_bytes=1e12, quantum=16).run(reqs)
    assert st["requests"] == 6
    for r, d in zip(reqs, FakeDev.live):
        assert d.steps == r.steps, r.rid          # never past its run
        assert d.closed
        assert r.completed >= r.admitted >= r.arrival
    # a round advances each bucket by at most the quantum and by at most the
    # fewest steps any member still needs
    assert max(k for _, k, _ in FakeBatch.log) <= 16
    assert st["cell_steps"] == sum(r.scen.N * r.steps for r in reqs)


def test_buckets_group_by_side(fake_device):
    reqs = [request(0, 64, 5), request(1, 128, 5), request(2, 192, 5), request(3, 256, 5), request(4, 320, 5)]
    srv = serving.Server(budget_bytes=1e12, quantum=5, buckets=(128, 256, 384))
    assert [srv.bucket_of(s) for s in (64, 128, 129, 256, 257, 384, 999)] == [128, 128, 256, 256, 384, 384, 384]
    srv.run(reqs)
    groups = {tuple(sorted(names)) for _, _, names in FakeBatch.log}
    assert groups == {("r0", "r1"), ("r2", "r3"), ("r4",)}


def test_budget_admits_in_arrival_order(fake_device):
    reqs = [request(q, 64, 10 * (q + 1)) for q in range(5)]
    one = serving.footprint(reqs[0].scen)
    # room for two at a time: the third waits at the head of the queue until
    # one leaves, and nobody overtakes it
    st = serving.Server(budget_bytes=2.5 * one, quantum=4).run(reqs)
    assert st["peak_resident_bytes"] <= 2.5 * one
    assert st["peak_resident_scenarios"] == 2
    admitted = [r.admitted for r in reqs]
    assert admitted == sorted(admitted)
    assert all(d.steps == r.steps for r, d in zip(reqs, FakeDev.live))


def test_capacity_limits_a_bucket(fake_device):
    reqs = [request(q, 64, 8) for q in range(5)]
    st = serving.Server(budget_bytes=1e12, quantum=8, capacity=2).run(reqs)
    assert st["peak_resident_scenarios"] == 2
    assert all(len(names) <= 2 for _, _, names in FakeBatch.log)


def test_oversized_request_is_rejected(fake_device):
    r = request(0, 64, 5)
    with pytest.raises(ValueError, match="budget"):
        serving.Server(budget_bytes=serving.footprint(r.scen) / 2).run([r])


def test_halted_request_leaves(fake_device):
    reqs = [request(0, 64, 40, halt_at=6), request(1, 64, 12)]
    st = serving.Server(budget_bytes=1e12, quantum=4).run(reqs)
    d0, d1 = FakeDev.live
    assert d0.halted and d0.closed and d0.steps == 6
    assert d1.steps == 12 and not d1.halted
    assert st["requests"] == 2

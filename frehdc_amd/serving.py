"""Serving loop: scenario runs of uneven grid size arrive over time, are
admitted against an HBM budget, stay resident on the device and are advanced
in shared launches (continuous batching), and leave when their run is done.

* Size buckets: requests are grouped by grid side (<= 128, 256, 384, ...,
  1024, 2048, 4096); each bucket is one device batch (frehd_batch_*), so a
  launch covers every resident scenario of similar size and the surplus
  blocks of the smaller ones stay bounded (finer buckets measured no
  faster: the launches of the concurrent buckets overlap on the device).
* Admission: first come, first served, while the device footprint of the
  resident scenarios (frehd_dev_footprint) stays within the budget and the
  bucket has a free slot; a request that does not fit waits at the head of
  the queue.
* Scheduling rounds: every non-empty bucket advances by
  min(quantum, fewest steps any of its members still needs), so nobody
  overshoots its run; finished scenarios are downloaded (optionally),
  removed and destroyed, which frees budget and slots; arrivals are admitted
  between rounds, so new work joins running batches.
* Reported: per-request latency (completion - arrival) p50/p99, queue wait,
  aggregate cell-steps/s over the makespan, peak resident bytes and members.
* Subgrid requests keep their pixels on the device and evaluate the two
  levels bracketing the surface at each lookup (--tables: level tables in
  HBM instead, 20 KB per cell at 250 levels, which the budget admits only a
  few dozen of at a time).

    python -m frehdc_amd.serving --count 512 --rate 20 --subgrid 4 --forms fp64,mixed_warm
"""
from __future__ import annotations

import argparse
import collections
import json
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, cases
from .device import Batch, DeviceScenario


@dataclass
class Request:
    rid: int
    arrival: float
    steps: int
    scen: object
    footprint: int = 0
    admitted: float = -1.0
    completed: float = -1.0
    done_steps: int = 0
    result: dict | None = None
    dev: DeviceScenario | None = field(default=None, repr=False)


def footprint(scen, **over) -> int:
    lib = abi.load_library()
    cfg = scen.c_config(**over)
    import ctypes as C
    b = C.c_int64()
    abi.check(lib.frehd_dev_footprint(C.byref(cfg), C.byref(b)))
    return int(b.value)


class Server:
    def __init__(self, budget_bytes=200e9, buckets=(128, 256, 384, 512, 640, 768, 896, 1024, 2048, 4096), quantum=16,
                 capacity=128, collect=None, progress=0.0, **over):
        self.budget = float(budget_bytes)
        self.edges = list(buckets)
        self.quantum = int(quantum)
        self.capacity = int(capacity)
        self.collect = collect          # field names to download at completion, or None
        self.progress = float(progress)  # seconds between progress lines on stdout (0: none)
        self.over = over                # config overrides for every request (precision, ...)
        self.batches = {}
        self.members = collections.defaultdict(list)

    def bucket_of(self, side):
        for e in self.edges:
            if side <= e:
                return e
        return self.edges[-1]

    def run(self, reqs):
        """Serve `reqs` (sorted by arrival, seconds from the start); returns
        the statistics.  Wall-clock driven: arrivals are honoured in real
        time."""
        for r in reqs:
            r.footprint = footprint(r.scen, **self.over)
            if r.footprint > self.budget:
                raise ValueError(f"request {r.rid} needs {r.footprint} B, more than the budget")
        from concurrent.futures import ThreadPoolExecutor
        self.pool = ThreadPoolExecutor(max_workers=len(self.edges))
        queue = collections.deque()
        used = 0
        peak_used = 0
        peak_members = 0
        nxt = 0
        active = 0
        rounds = 0
        done_count = 0
        t0 = time.perf_counter()
        next_report = self.progress

        def now():
            return time.perf_counter() - t0

        while nxt < len(reqs) or queue or active:
            t = now()
            while nxt < len(reqs) and reqs[nxt].arrival <= t:
                queue.append(reqs[nxt])
                nxt += 1
            while queue:
                r = queue[0]
                key = self.bucket_of(max(r.scen.cfg["nx"], r.scen.cfg["ny"]))
                if used + r.footprint > self.budget or len(self.members[key]) >= self.capacity:
                    break
                queue.popleft()
                r.dev = DeviceScenario(r.scen, **self.over)
                if key not in self.batches:
                    self.batches[key] = Batch(self.capacity)
                self.batches[key].add(r.dev)
                self.members[key].append(r)
                r.admitted = now()
                used += r.footprint
                active += 1
            peak_used = max(peak_used, used)
            peak_members = max(peak_members, active)
            if not active:
                if nxt < len(reqs):
                    time.sleep(max(0.0, reqs[nxt].arrival - now()))
                continue
            rounds += 1
            # every bucket advances in its own host thread (each batch has its
            # own stream; the library releases the interpreter while it runs),
            # so the buckets' launches and synchronisations overlap
            work = []
            for key, mem in self.members.items():
                if mem:
                    k = min(self.quantum, min(r.steps - r.done_steps for r in mem))
                    work.append((key, mem, k))
            futs = [(key, mem, k, self.pool.submit(_step_in_thread, self.batches[key], k)) for key, mem, k in work]
            for key, mem, k, fut in futs:
                rc, msg = fut.result()
                b = self.batches[key]
                if rc not in (0, abi.FREHD_ECFL):
                    raise abi.FrehdError(rc, msg)
                for r in mem:
                    r.done_steps += k
                halted = {id(r) for r in mem if rc == abi.FREHD_ECFL and r.dev.stats()["halted"]}
                done = [r for r in mem if r.done_steps >= r.steps or id(r) in halted]
                for r in done:
                    r.completed = now()
                    if self.collect:
                        r.result = r.dev.download(self.collect)
                    b.remove(r.dev)
                    r.dev.close()
                    r.dev = None
                    mem.remove(r)
                    used -= r.footprint
                    active -= 1
                    done_count += 1
            if self.progress > 0 and now() >= next_report:
                print(f"# t={now():.0f}s served {done_count}/{len(reqs)}, resident {active}, queued {len(queue)}, "
                      f"arrived {nxt}, resident bytes {used / 1e9:.1f} GB", flush=True)
                next_report += self.progress
        self.pool.shutdown()
        for b in self.batches.values():
            b.close()
        self.batches = {}
        lat = np.array([r.completed - r.arrival for r in reqs])
        wait = np.array([r.admitted - r.arrival for r in reqs])
        work = sum(r.scen.N * r.steps for r in reqs)
        span = max(r.completed for r in reqs) - min(r.arrival for r in reqs)
        return {
            "requests": len(reqs),
            "latency_p50_s": float(np.percentile(lat, 50)),
            "latency_p99_s": float(np.percentile(lat, 99)),
            "queue_wait_p50_s": float(np.percentile(wait, 50)),
            "queue_wait_p99_s": float(np.percentile(wait, 99)),
            "cell_steps": int(work),
            "makespan_s": float(span),
            "cell_steps_per_s": float(work / span),
            "peak_resident_bytes": int(peak_used),
            "peak_resident_scenarios": int(peak_members),
            "rounds": rounds,
        }


def _step_in_thread(batch, k):
    """Step a batch; the library's error message is per thread, so read it
    in the thread that made the call."""
    rc = batch.try_step(k)
    return rc, (batch.lib.frehd_last_error().decode() if rc else "")


def make_requests(count, rate, seed=6, lo=64, hi=1024, steps_lo=200, steps_hi=2000, subgrid_r=0, workers=1,
                  on_the_fly=False):
    """Build the request scenarios up front (their inputs exist before they
    arrive); `workers` > 1 builds them in a process pool."""
    spec = cases.serving_requests(count=count, rate=rate, seed=seed, lo=lo, hi=hi, steps_lo=steps_lo,
                                  steps_hi=steps_hi, subgrid_r=subgrid_r, on_the_fly=on_the_fly)
    if workers > 1:
        # spawned workers: a fork of a process whose HIP runtime is up would
        # carry its threads' state, and the pool is shut down (its processes
        # joined) before this returns
        import multiprocessing as mp
        from concurrent.futures import ProcessPoolExecutor
        with ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn")) as ex:
            scens = list(ex.map(_build, [(count, rate, seed, lo, hi, steps_lo, steps_hi, subgrid_r, on_the_fly, q)
                                         for q in range(count)]))
    else:
        scens = [b() for (_, _, _, b) in spec]
    return [Request(rid=q, arrival=a, steps=st, scen=sc) for q, ((a, st, _, _), sc) in enumerate(zip(spec, scens))]


def _build(args):
    count, rate, seed, lo, hi, steps_lo, steps_hi, subgrid_r, on_the_fly, q = args
    spec = cases.serving_requests(count=count, rate=rate, seed=seed, lo=lo, hi=hi, steps_lo=steps_lo,
                                  steps_hi=steps_hi, subgrid_r=subgrid_r, on_the_fly=on_the_fly)
    return spec[q][3]()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=512)
    ap.add_argument("--rate", type=float, default=20.0)
    ap.add_argument("--lo", type=int, default=64)
    ap.add_argument("--hi", type=int, default=1024)
    ap.add_argument("--steps-lo", type=int, default=200)
    ap.add_argument("--steps-hi", type=int, default=2000)
    ap.add_argument("--subgrid", type=int, default=4)
    ap.add_argument("--budget-gb", type=float, default=200.0)
    ap.add_argument("--quantum", type=int, default=16)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--forms", default="fp64,mixed_warm")
    ap.add_argument("--progress", type=float, default=15.0, help="seconds between progress lines (0: none)")
    ap.add_argument("--buckets", default="128,256,384,512,640,768,896,1024,2048,4096",
                    help="bucket edges (grid side), comma-separated")
    ap.add_argument("--tables", action="store_true",
                    help="subgrid level tables in HBM (20 KB per cell at 250 levels) instead of evaluating the "
                         "levels from the pixels at each lookup")
    a = ap.parse_args()
    t0 = time.perf_counter()
    reqs = make_requests(a.count, a.rate, lo=a.lo, hi=a.hi, steps_lo=a.steps_lo, steps_hi=a.steps_hi,
                         subgrid_r=a.subgrid, workers=a.workers, on_the_fly=not a.tables)
    print(f"# built {len(reqs)} requests in {time.perf_counter() - t0:.1f} s", flush=True)
    forms = {"fp64": {}, "fp64_warm": {"warmStart": 2}, "mixed_warm": {"precision": abi.PREC_MIXED, "warmStart": 2},
             "mixed_cold": {"precision": abi.PREC_MIXED, "warmStart": 0}}
    out = {}
    for f in a.forms.split(","):
        fresh = [Request(rid=r.rid, arrival=r.arrival, steps=r.steps, scen=r.scen) for r in reqs]
        st = Server(budget_bytes=a.budget_gb * 1e9, quantum=a.quantum, progress=a.progress,
                    buckets=tuple(int(x) for x in a.buckets.split(",")), **forms[f]).run(fresh)
        out[f] = st
        print(json.dumps({f: st}), flush=True)
    print(json.dumps({"serving": out, "count": a.count, "rate": a.rate, "subgrid": a.subgrid}), flush=True)


if __name__ == "__main__":
    main()

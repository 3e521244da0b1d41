"""Request loop over the continuous-batching solve engine (one GPU).

Extraction requests name a resident job (refined mesh + link structure already
uploaded and filled in HBM).  The loop admits waiting requests while the device
has room, both in concurrent solve slots and in HBM bytes (the job's resident
footprint counted once, plus each solve's own footprint from
`fcg_job_footprint`).  It advances every admitted solve together, one
`fcg_engine_step` per iteration, so requests arriving while others run join the
next step.  Each result carries its latency from arrival to completion.

There is no reference counterpart: FasterCap solves one geometry per process.
The per-request work is `CSolveCap::Solve` (Solver/SolveCapacitance.cpp:1138-1425).
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field

import numpy as np

from . import FCG_PREC_MIXED, Device, SolveResult

HBM_BYTES = 288 * (1 << 30)


@dataclass
class Request:
    rid: int
    job: int
    arrival: float  # seconds, on the loop's clock
    start: float = -1.0
    finish: float = -1.0
    result: SolveResult | None = None


@dataclass
class ServeReport:
    completed: int
    elapsed_s: float
    solves_per_s: float
    latency_ms: dict
    queue_ms: dict
    max_active: int
    steps: int
    peak_solve_bytes: int  # admitted solve bytes (footprint estimates) at their peak
    peak_engine_bytes: int  # device bytes the engine's active solves held at their peak
    resident_bytes: int  # resident bytes of the registered jobs
    requests: list = field(default_factory=list)


def _pct(x, p):
    if not x:
        return float("nan")
    return float(np.percentile(np.asarray(x), p))


class Server:
    """Admission by slots and HBM footprint over one engine."""

    def __init__(self, dev: Device, tol: float = 1e-4, precision: int = FCG_PREC_MIXED, max_active: int = 256,
                 hbm_budget: int = int(0.9 * HBM_BYTES), max_iter: int = 1000):
        self.dev = dev
        self.engine = dev.engine(tol=tol, precision=precision, max_iter=max_iter)
        self.precision = precision
        self.max_active = max_active
        self.hbm_budget = hbm_budget
        self.max_iter = max_iter
        self.resident = {}  # job -> resident bytes
        self.solve_bytes = {}  # job -> bytes of one solve

    def add_job(self, job: int):
        """Register a filled, resident job and its footprints."""
        res, sol = self.dev.footprint(job, self.precision, iters=min(self.max_iter, 200))
        self.resident[job] = res
        self.solve_bytes[job] = sol

    def run(self, arrivals, duration_s: float | None = None) -> ServeReport:
        """Serve `arrivals` = [(arrival_s, job), ...] sorted by time.  Arrival times
        are honoured on the wall clock: a request is visible to admission only once
        its time has passed.  Stops when every request completed (or after
        duration_s with whatever completed)."""
        waiting = collections.deque()
        pending = collections.deque(Request(i, j, t) for i, (t, j) in enumerate(arrivals))
        active = {}
        done = []
        resident_total = sum(self.resident.values())
        active_bytes = 0
        peak = 0
        peak_dev = 0
        steps = 0
        max_act = 0
        t0 = time.perf_counter()
        while pending or waiting or active:
            now = time.perf_counter() - t0
            if duration_s is not None and now > duration_s:
                break
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission: slots and HBM
            while waiting and len(active) < self.max_active:
                r = waiting[0]
                need = self.solve_bytes[r.job]
                if active and resident_total + active_bytes + need > self.hbm_budget:
                    break
                waiting.popleft()
                r.start = time.perf_counter() - t0
                self.engine.submit(r.job, r.rid)
                active[r.rid] = r
                active_bytes += need
            peak = max(peak, active_bytes)
            max_act = max(max_act, len(active))
            if not active:
                # idle until the next arrival
                if pending:
                    time.sleep(max(0.0, min(pending[0].arrival - (time.perf_counter() - t0), 0.001)))
                continue
            self.engine.step()
            steps += 1
            peak_dev = max(peak_dev, self.engine.bytes())
            while True:
                got = self.engine.poll()
                if got is None:
                    break
                tag, job, res = got
                r = active.pop(tag)
                r.finish = time.perf_counter() - t0
                r.result = res
                active_bytes -= self.solve_bytes[r.job]
                done.append(r)
        elapsed = time.perf_counter() - t0
        lat = [1e3 * (r.finish - r.arrival) for r in done]
        que = [1e3 * (r.start - r.arrival) for r in done]
        return ServeReport(
            completed=len(done),
            elapsed_s=elapsed,
            solves_per_s=len(done) / elapsed if elapsed > 0 else 0.0,
            latency_ms={"p50": _pct(lat, 50), "p99": _pct(lat, 99), "mean": float(np.mean(lat)) if lat else 0.0},
            queue_ms={"p50": _pct(que, 50), "p99": _pct(que, 99)},
            max_active=max_act,
            steps=steps,
            peak_solve_bytes=peak,
            peak_engine_bytes=peak_dev,
            resident_bytes=resident_total,
            requests=done,
        )

    def close(self):
        self.engine.close()


def poisson_arrivals(rate_per_s: float, count: int, jobs, seed: int, weights=None):
    """`count` arrivals of a Poisson process with the given rate, each naming a job
    drawn from `jobs` (with optional weights)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate_per_s, size=count)
    t = np.cumsum(gaps)
    p = None if weights is None else np.asarray(weights, float) / np.sum(weights)
    pick = rng.choice(len(jobs), size=count, p=p)
    return [(float(t[i]), int(jobs[pick[i]])) for i in range(count)]

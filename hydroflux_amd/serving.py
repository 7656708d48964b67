"""Continuous-batching serving loop for flood-scenario requests on one GPU.

Requests arrive over time (e.g. Poisson). Each is admitted into the resident
batch of its kind (FP32 precision form for plain requests, Dual for requests
carrying forward-mode tangents) as soon as a slot and enough HBM are free;
its state stays on the device until its clock reaches its end time. The GPU
advances every resident scenario together in chunks of steps; the next
chunk is enqueued as soon as one ends, and the host retires finished
scenarios (download, free) and admits waiting ones while it runs, so new
work overlaps with work already on the device.

Reported: per-request latency-to-completion (completion - arrival, wall
clock), its p50/p99, and the aggregate cell-updates/s over the run.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional

import numpy as np

from .batch import Batch, BatchFull, ScenarioSpec, footprint


@dataclass
class Request:
    rid: int
    arrival: float          # seconds after the start of the run
    spec: ScenarioSpec
    kind: str = "f32"       # "f32" or "dual"
    # filled by the server
    admitted: Optional[float] = None
    completed: Optional[float] = None
    steps: int = 0
    sim_time: float = 0.0
    status: int = 0
    result: Optional[np.ndarray] = None
    tangent: Optional[np.ndarray] = None

    @property
    def latency(self):
        return None if self.completed is None else self.completed - self.arrival


@dataclass
class ServeReport:
    requests: List[Request]
    wall_s: float
    cell_updates: float
    latencies: np.ndarray = field(default=None)

    @property
    def cell_updates_per_s(self):
        return self.cell_updates / self.wall_s

    def percentile(self, q):
        return float(np.percentile(self.latencies, q))

    def summary(self):
        return {
            "requests": len(self.requests),
            "wall_s": self.wall_s,
            "cell_updates_per_s": self.cell_updates_per_s,
            "cell_updates_per_s_by_kind": {k: v / self.wall_s for k, v in getattr(self, "by_kind", {}).items()},
            "latency_p50_s": self.percentile(50),
            "latency_p99_s": self.percentile(99),
            "latency_mean_s": float(np.mean(self.latencies)),
            "failed": int(sum(1 for r in self.requests if r.status != 0)),
        }


def poisson_arrivals(n, rate, seed=0):
    """n arrival times of a Poisson process with `rate` requests/s (rate=inf: all at 0)."""
    if not np.isfinite(rate):
        return np.zeros(n)
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))


class Server:
    def __init__(self, capacity_bytes: Dict[str, int], max_slots: int = 512, chunk_steps: int = 32,
                 keep_results: bool = True, autotune_steps: int = 0, tile_config: Optional[Dict[str, int]] = None):
        self.batches = {k: Batch(k, capacity_bytes=v, max_slots=max_slots) for k, v in capacity_bytes.items()}
        self.chunk = int(chunk_steps)
        self.keep = keep_results
        self.autotune_steps = autotune_steps
        for k, ci in (tile_config or {}).items():
            self.batches[k].set_config(ci)

    def close(self):
        for b in self.batches.values():
            b.close()

    def run(self, requests: List[Request], on_done: Optional[Callable[[Request], None]] = None) -> ServeReport:
        """Serve the requests to completion. on_done(request) is called as
        each one completes (its result fields set, before it leaves the batch).

        Each batch (FP32, Dual) runs chunks on its own stream; the host
        serves whichever batch has finished its chunk, without waiting on
        the other: remove the slots that finished a round earlier (their
        slots and HBM are reused only once no enqueued launch refers to
        them); read the clocks; enqueue the next chunk at once, so the device
        works while the host downloads the results of the slots that just
        finished (a finished slot's steps are no-ops, its state is final)
        and admits waiting requests into free slots and memory, which join
        at the next chunk."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: Dict[str, List[Request]] = {k: [] for k in self.batches}
        resident: Dict[str, Dict[int, Request]] = {k: {} for k in self.batches}
        retiring: Dict[str, List[int]] = {k: [] for k in self.batches}
        in_flight = {k: False for k in self.batches}
        cell_updates = {k: 0.0 for k in self.batches}
        i = 0
        t0 = time.perf_counter()
        tuned = set()

        def admit(now):
            nonlocal i
            while i < len(pending) and pending[i].arrival <= now:
                waiting[pending[i].kind].append(pending[i])
                i += 1
            for k, q in waiting.items():  # by HBM footprint, FIFO per kind
                while q:
                    r = q[0]
                    try:
                        slot = self.batches[k].add(r.spec)
                    except BatchFull:
                        break
                    r.admitted = time.perf_counter() - t0
                    resident[k][slot] = r
                    q.pop(0)

        while i < len(pending) or any(waiting[k] or resident[k] or retiring[k] for k in self.batches):
            admit(time.perf_counter() - t0)
            if not any(in_flight.values()) and not any(
                    len(resident[k]) > len(retiring[k]) for k in self.batches):
                for k in self.batches:  # nothing running: drop what finished
                    for slot in retiring[k]:
                        self.batches[k].remove(slot)
                        del resident[k][slot]
                    retiring[k] = []
                if i < len(pending):  # idle until the next arrival
                    time.sleep(max(0.0, min(0.01, pending[i].arrival - (time.perf_counter() - t0))))
                continue
            for k in self.batches:  # start every idle batch that has work
                b = self.batches[k]
                if not in_flight[k] and len(resident[k]) > len(retiring[k]):
                    if self.autotune_steps and k not in tuned and len(resident[k]) > 1:
                        b.autotune(self.autotune_steps)
                        tuned.add(k)
                    b.step(self.chunk)
                    in_flight[k] = True
            progressed = False
            for k in self.batches:
                b = self.batches[k]
                if not in_flight[k]:  # idle since its last wait: nothing refers to its finished slots
                    for slot in retiring[k]:
                        b.remove(slot)
                        del resident[k][slot]
                    retiring[k] = []
                    continue
                if not b.idle():  # another batch may finish first: serve whichever does
                    continue
                progressed = True
                in_flight[k] = False
                for slot in retiring[k]:  # no enqueued launch refers to them now
                    b.remove(slot)
                    del resident[k][slot]
                retiring[k] = []
                t, steps, status, done, used = b.poll()
                finished = [slot for slot in resident[k] if done[slot]]
                if len(resident[k]) > len(finished):
                    b.step(self.chunk)  # the next chunk runs during the host work below
                    in_flight[k] = True
                for slot in finished:
                    r = resident[k][slot]
                    r.completed = time.perf_counter() - t0
                    r.steps = int(steps[slot])
                    r.sim_time = float(t[slot])
                    r.status = int(status[slot])
                    if self.keep:
                        out = b.download(slot, r.spec.shape, finished=True)
                        if k == "dual":
                            r.result, r.tangent = out
                        else:
                            r.result = out
                    cell_updates[k] += r.spec.n_cells * r.steps
                    if on_done is not None:
                        on_done(r)
                    retiring[k].append(slot)
                admit(time.perf_counter() - t0)
            if not progressed and any(in_flight.values()):
                time.sleep(2e-4)
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in requests])
        rep = ServeReport(requests, wall, sum(cell_updates.values()), lat)
        rep.by_kind = {k: v for k, v in cell_updates.items()}
        return rep


def request_footprint(req: Request) -> int:
    nr, nc = req.spec.shape
    return footprint(req.kind, nr, nc, max(1, len(req.spec.sources)))

"""Serving: rollout requests of different sizes and lengths, admitted as they
arrive and advanced together on the GPU (continuous batching).

Every active request is one system of a segmented batch
(``jaxdem_amd.batching``), or steps alone with ``System.step`` while it is
the only one; its state, cell list and tangential-spring history stay in
device memory between step chunks.  At each chunk boundary
finished requests leave and arrived requests are admitted (first come first
served, or least work first) while their device footprint fits the memory
budget and the batch holds fewer than ``max_active`` requests.  A chunk runs
until the next request finishes, so no request runs past its own step
count, and for about ``chunk_ms`` of GPU time (estimated from the previous
chunks' time per step; ``chunk`` steps for the first): arrivals wait at most
one chunk, while the per-chunk costs (packing the systems' arrays, the
entry and exit conversions of the batched step, a host synchronisation) stay
a small fraction of each chunk.  Because a system advances in a batch exactly as it
does alone, each result is bit-identical to running the request by itself.

    server = Server(device, hbm_budget=...)
    done = server.run(requests)          # requests with arrival times
    stats = latency_stats(done)
"""

from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Any, Callable, Sequence

import numpy as np
import torch

from .batching import Batch
from .system import System


@dataclass
class Request:
    """One rollout: ``steps`` time steps of (state, system) (System.step)."""

    rid: int
    state: Any
    system: Any
    steps: int
    arrival: float = 0.0  # seconds after the start of Server.run
    # filled by the server
    admitted: float | None = None
    finished: float | None = None
    result: tuple | None = None
    done_steps: int = 0

    @property
    def n(self) -> int:
        return int(self.state.N)

    @property
    def latency(self) -> float:
        return float(self.finished - self.arrival)


def footprint_bytes(state, system) -> int:
    """Device bytes a request holds while active: the caller-layout state and
    neighbour list plus the working set of the batched step (double-buffered
    float4 fields, list, springs and sort scratch)."""
    N, D = state.N, state.dim
    M = int(getattr(system.collider, "max_neighbors", 0))
    caller = N * (4 * (3 * D + 3 + 4 + 3 + 3) + 4 * 4) + N * M * (4 + 4 * D) + 4 * N * max((M + 31) // 32, 1)
    work = N * (16 * 13 + 4 * 12) + N * M * (4 * 2 + 16 * 2) + 8 * N * max((M + 63) // 64, 1) * 2
    return int(caller + work)


class Server:
    """Continuous batching of rollout requests.

    ``policy`` picks which waiting request is admitted next when a place in
    the batch is free: "fcfs" the earliest arrival, "sjf" the least work
    (particles x steps; ties by arrival).  ``max_active`` caps the requests
    in the batch: with a cap, requests share the GPU with a few others
    instead of with every arrival, and "sjf" lets short requests pass long
    ones in the queue."""

    def __init__(self, device, hbm_budget: int | None = None, chunk: int = 50, max_active: int = 4096,
                 clock: Callable[[], float] = time.perf_counter, chunk_ms: float = 20.0, policy: str = "fcfs",
                 launch_mode: int = 2, spec_min_gap: int = 0):
        if policy not in ("fcfs", "sjf"):
            raise ValueError(f"unknown admission policy {policy!r} (fcfs, sjf)")
        if int(max_active) < 1:
            raise ValueError("max_active must be >= 1")
        self.policy = policy
        self.device = torch.device(device)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.8 * free)
        self.hbm_budget = int(hbm_budget)
        self.chunk = int(chunk)
        self.chunk_ms = float(chunk_ms)
        self.max_active = int(max_active)
        self.clock = clock
        # a served batch holds a few systems that rebuild rarely: speculative
        # chunks launch fewer kernels than a rebuild chain after every step
        self.launch_mode = int(launch_mode)
        self.spec_min_gap = int(spec_min_gap)
        self.log: list = []

    def run(self, requests: Sequence[Request], realtime: bool = True) -> list:
        """Serve ``requests`` (their ``arrival`` times are seconds after the
        start).  With ``realtime`` the server waits for arrivals on the wall
        clock; otherwise every request is present from the start.  Returns
        copies of the requests with their results and timestamps (the
        requests passed in are left as they were, so they can be served
        again)."""
        pending = sorted((Request(r.rid, r.state, r.system, r.steps, r.arrival) for r in requests),
                         key=lambda r: (r.arrival, r.rid))
        queue: list = []  # arrived, not admitted yet, in arrival order
        active: list = []  # in the batch, in batch order
        batch: Batch | None = None
        t0 = self.clock()
        now = lambda: self.clock() - t0  # noqa: E731
        used = 0
        done: list = []
        single = None  # (state, system) of a request that runs alone
        ms_per_step = None  # of the current composition, from the last chunk
        while pending or queue or active:
            t = now()
            while pending and (not realtime or pending[0].arrival <= t):
                queue.append(pending.pop(0))
            admit = []
            while queue and len(active) + len(admit) < self.max_active:
                k = 0
                if self.policy == "sjf":
                    k = min(range(len(queue)), key=lambda i: (queue[i].n * queue[i].steps, queue[i].arrival,
                                                             queue[i].rid))
                if queue[k].steps <= 0:
                    # nothing to run: the result is the request's own state
                    r = queue.pop(k)
                    r.admitted = r.finished = now()
                    r.result = (r.state, r.system)
                    done.append(r)
                    continue
                fp = footprint_bytes(queue[k].state, queue[k].system)
                if used + fp > self.hbm_budget and (active or admit):
                    break
                r = queue.pop(k)
                r.admitted = now()
                r.__dict__["_fp"] = fp
                used += fp
                admit.append(r)
            if admit:
                pairs = self._pairs(batch, single) + [(r.state, r.system) for r in admit]
                active = active + admit
                batch, single = self._hold(pairs)
                ms_per_step = None
            if not active:
                if pending and realtime:
                    time.sleep(max(0.0, min(pending[0].arrival - now(), 0.01)))
                continue
            want = self.chunk if ms_per_step is None else max(1, int(self.chunk_ms / max(ms_per_step, 1e-6)))
            n = min(want, min(r.steps - r.done_steps for r in active))
            t_c = self.clock()
            if single is not None:
                single = System.step(*single, n=n)
            else:
                batch = batch.step(n)
            torch.cuda.synchronize(self.device)
            ms_per_step = (self.clock() - t_c) * 1e3 / n
            for r in active:
                r.done_steps += n
            self.log.append((now(), len(active), int(batch.n_particles if batch is not None else single[0].N), n))
            if any(r.done_steps >= r.steps for r in active):
                cur = self._pairs(batch, single)
                t_f = now()
                keep = []
                for r, pair in zip(active, cur):
                    if r.done_steps >= r.steps:
                        r.result = pair
                        r.finished = t_f
                        used -= r.__dict__.pop("_fp")
                        done.append(r)
                    else:
                        keep.append((r, pair))
                active = [r for r, _ in keep]
                batch, single = self._hold([p for _, p in keep])
                ms_per_step = None
        return done

    @staticmethod
    def _pairs(batch, single) -> list:
        if single is not None:
            return [single]
        return batch.unpack() if batch is not None else []

    def _hold(self, pairs: list) -> tuple:
        """(batch, single) holding the active requests: one request steps
        alone with System.step (a batch of one only adds the segmented
        kernels' costs), several in one batch."""
        if not pairs:
            return None, None
        if len(pairs) == 1:
            return None, pairs[0]
        batch = Batch.pack(pairs)
        batch.launch_mode, batch.spec_min_gap = self.launch_mode, self.spec_min_gap
        return batch, None


def run_one_at_a_time(requests: Sequence[Request], device, realtime: bool = True,
                      clock: Callable[[], float] = time.perf_counter) -> list:
    """The same requests served one after another with System.step, in
    arrival order (the plain kernel, no batching)."""
    out = []
    t0 = clock()
    for r0 in sorted(requests, key=lambda r: (r.arrival, r.rid)):
        r = Request(r0.rid, r0.state, r0.system, r0.steps, r0.arrival)
        if realtime:
            wait = r.arrival - (clock() - t0)
            if wait > 0:
                time.sleep(wait)
        r.admitted = clock() - t0
        st, sy = System.step(r.state, r.system, n=r.steps)
        torch.cuda.synchronize(torch.device(device))
        r.finished = clock() - t0
        r.result = (st, sy)
        r.done_steps = r.steps
        out.append(r)
    return out


def latency_stats(done: Sequence[Request]) -> dict:
    """p50 / p99 latency (arrival to result), makespan and aggregate
    particle-steps/s."""
    lat = np.array([r.latency for r in done])
    work = sum(r.n * r.steps for r in done)
    span = max(r.finished for r in done) - min(r.arrival for r in done)
    return {
        "requests": len(done),
        "p50_latency_s": float(np.percentile(lat, 50)),
        "p99_latency_s": float(np.percentile(lat, 99)),
        "mean_latency_s": float(lat.mean()),
        "makespan_s": float(span),
        "particle_steps": int(work),
        "particle_steps_per_s": float(work / span),
    }


__all__ = ["Request", "Server", "footprint_bytes", "latency_stats", "run_one_at_a_time"]

"""Serving simulation requests with continuous batching.

Requests (automaton, size, step count, seed) arrive over time. The server admits them in
arrival order while the HBM budget and the slot count allow, keeps their grids on the device
(generated there from the seed), and runs *ticks*: one `cellato_hip_multi` call that advances
every live job by up to `tick_passes` launches of its own depth T (min(tick_passes x T, steps
left) generations). Jobs that arrive while others are running join at the next tick;
finished jobs leave and free their memory. Completion time is the GPU time at which a job's last tick finished (HIP events
against an event recorded at start), so latency = completion - arrival.

`serve_one_at_a_time` is the comparison point: the same arrivals served in order, each alone
with the plain simple traverser (one launch per generation).
"""
import collections
import ctypes
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from .planes import DeviceGrid
from .rules import all_rules, rule as get_rule


@dataclass
class Request:
    id: int
    rule: str
    width: int
    height: int
    steps: int
    arrival: float           # seconds after the start of serving
    seed: int
    bytes: int = 0           # HBM the job holds while live (two buffers)

    @property
    def cell_updates(self):
        return self.width * self.height * self.steps


def serving_workload(n=64, rate=2.0, seed=7, side_range=(128, 8192), steps_range=(64, 2048), rules=None):
    """Poisson arrivals at `rate` per second; automaton uniform over all ten rules; square
    side log-uniform in side_range rounded to a multiple of 64; steps uniform in steps_range."""
    rng = np.random.Generator(np.random.PCG64(seed))
    names = rules or [r.name for r in all_rules()]
    t = 0.0
    reqs = []
    lo, hi = math.log(side_range[0]), math.log(side_range[1])
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        name = names[int(rng.integers(0, len(names)))]
        side = min(max(int(round(math.exp(rng.uniform(lo, hi)) / 64)) * 64, side_range[0]), side_range[1])
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        r = get_rule(name)
        reqs.append(Request(i, name, side, side, steps, t, seed * 1000003 + i, 2 * r.planes * side * side // 8))
    return reqs


@dataclass
class _Live:
    req: Request
    grid: DeviceGrid
    T: int
    chunk: int
    done: int = 0


@dataclass
class ServeReport:
    completions: dict = field(default_factory=dict)     # request id -> completion time (s)
    latencies: dict = field(default_factory=dict)       # request id -> completion - arrival (s)
    admitted_at: dict = field(default_factory=dict)
    ticks: int = 0
    max_live: int = 0
    peak_bytes: int = 0
    refused_budget: int = 0
    makespan: float = 0.0
    cell_updates: int = 0

    def summary(self):
        lat = np.array(sorted(self.latencies.values())) if self.latencies else np.zeros(1)
        return {"requests": len(self.latencies), "p50_s": float(np.percentile(lat, 50)),
                "p99_s": float(np.percentile(lat, 99)), "mean_s": float(lat.mean()),
                "gcups": self.cell_updates / self.makespan / 1e9 if self.makespan > 0 else 0.0,
                "ticks": self.ticks, "max_live": self.max_live, "peak_GB": self.peak_bytes / 1e9,
                "makespan_s": self.makespan}


def admit(waiting, n_live, used, budget, max_live):
    """Admission policy: take requests from the front of `waiting` (arrival order) while a
    slot is free and the HBM budget holds their two buffers; a request that could never fit
    the budget is refused. Head-of-line: a request that fits later blocks those behind it.
    Returns (admitted, refused count, new bytes in use) and pops them from `waiting`."""
    admitted, refused = [], 0
    while waiting and n_live + len(admitted) < max_live:
        r = waiting[0]
        if r.bytes > budget:
            waiting.popleft()
            refused += 1
            continue
        if used + r.bytes > budget:
            break
        waiting.popleft()
        admitted.append(r)
        used += r.bytes
    return admitted, refused, used


class GpuEngine:
    """The device side of the serving loop: ticks go to `cellato_hip_multi` on the current
    stream, completion times come from HIP events against an event recorded at start."""

    def __init__(self):
        import torch
        self.torch = torch
        self.stream = torch.cuda.current_stream()
        self.start_ev = None

    def start(self):
        self.torch.cuda.synchronize()
        self.start_ev = self.torch.cuda.Event(enable_timing=True)
        self.start_ev.record(self.stream)

    def tick(self, arr, n):
        """Advance the n grids of `arr` (cellato_batch_grid array); returns result_in_b per grid."""
        in_b = (ctypes.c_int * n)()
        _lib.check(_lib.lib().cellato_hip_multi(arr, n, 0, ctypes.c_void_p(self.stream.cuda_stream), in_b))
        return list(in_b)

    def record(self):
        ev = self.torch.cuda.Event(enable_timing=True)
        ev.record(self.stream)
        return ev

    def done(self, ev):
        return ev.query()

    def wait(self, ev):
        ev.synchronize()

    def seconds(self, ev):
        """GPU time of a recorded event after start(), in seconds."""
        return self.start_ev.elapsed_time(ev) / 1e3


class Server:
    """Continuous batching of simulation requests under an HBM budget.

    A job's two buffers count against the budget from its admission until the tick that
    finishes it has retired on the GPU (its grid is referenced by the launches in flight until
    then); only then are its bytes given to new requests. `allocate(request)` makes a job's
    grid and `release(grid)` drops it (a test may pass fakes for both, together with a fake
    `engine`, to follow the bytes held)."""

    def __init__(self, hbm_budget=200e9, max_live=512, tuner=None, default_T=4, keep_results=False,
                 allocate=None, tick_passes=4, release=None, engine=None):
        self.budget, self.max_live = hbm_budget, max_live
        self.tick_passes = tick_passes
        self.tuner, self.default_T = tuner, default_T
        self.keep_results = keep_results
        self.results = {}       # request id -> final DeviceGrid (keep_results)
        self.allocate = allocate or (lambda r: DeviceGrid(get_rule(r.rule), r.width, r.height, 0))
        self.release = release or (lambda grid: None)
        self.engine = engine

    def _config(self, r):
        if self.tuner is not None:
            return self.tuner.pick(r.rule, r.width, r.height)
        return self.default_T, 0

    def run(self, requests, clock=time.perf_counter, sleep=time.sleep):
        eng = self.engine or GpuEngine()
        rep = ServeReport()
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        waiting = collections.deque()
        live = []
        used = 0                # bytes of every admitted job whose last tick has not retired
        eng.start()
        t0 = clock()
        inflight = []           # (event, [live jobs finishing with this tick])
        while pending or waiting or live or inflight:
            now = clock() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(pending.popleft())
            # admission in arrival order, by slots and HBM budget
            admitted, refused, used = admit(waiting, len(live), used, self.budget, self.max_live)
            rep.refused_budget += refused
            for r in admitted:
                g = self.allocate(r)
                g.fill_random(r.seed)
                T, chunk = self._config(r)
                live.append(_Live(r, g, T, chunk))
                rep.admitted_at[r.id] = now
            rep.max_live = max(rep.max_live, len(live))
            rep.peak_bytes = max(rep.peak_bytes, used)
            if live:
                arr = (_lib.BatchGrid * len(live))()
                for i, j in enumerate(live):
                    g = arr[i]
                    g.rule, g.layout, g.width, g.height = j.grid.rule.id, 0, j.req.width, j.req.height
                    g.steps = min(j.T * self.tick_passes, j.req.steps - j.done)
                    g.t0, g.temporal_steps, g.chunk_rows = j.done, j.T, j.chunk
                    for p, ptr in enumerate(j.grid.ptrs(j.grid.cur)):
                        g.a[p] = ptr
                    for p, ptr in enumerate(j.grid.ptrs(1 - j.grid.cur)):
                        g.b[p] = ptr
                in_b = eng.tick(arr, len(live))
                rep.ticks += 1
                finished, still = [], []
                for i, j in enumerate(live):
                    if in_b[i]:
                        j.grid.cur = 1 - j.grid.cur
                    j.done += arr[i].steps
                    (finished if j.done >= j.req.steps else still).append(j)
                inflight.append((eng.record(), finished))
                live = still
            # retire ticks whose GPU work is done (keep at most two in flight); a finished job's
            # bytes return to the budget here, once the launches that use its grid are over
            while inflight and (len(inflight) > 2 or eng.done(inflight[0][0]) or not live):
                ev, fin = inflight.pop(0)
                eng.wait(ev)
                t_done = eng.seconds(ev)
                for j in fin:
                    rep.completions[j.req.id] = t_done
                    rep.latencies[j.req.id] = t_done - j.req.arrival
                    rep.cell_updates += j.req.cell_updates
                    used -= j.req.bytes
                    if self.keep_results:
                        self.results[j.req.id] = j.grid
                    else:
                        self.release(j.grid)
                    j.grid = None
            if not live and not waiting and pending and not inflight:
                sleep(max(0.0, min(0.01, pending[0].arrival - (clock() - t0))))
        rep.makespan = max(rep.completions.values()) - min(r.arrival for r in requests) if rep.completions else 0.0
        return rep


def serve_one_at_a_time(requests, keep_results=False, clock=time.perf_counter, sleep=time.sleep, time_limit=None):
    """Baseline: requests served in arrival order, one at a time, each with the plain simple
    traverser. Returns (ServeReport, results)."""
    import torch
    rep = ServeReport()
    results = {}
    stream = torch.cuda.current_stream()
    start_ev = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = clock()
    start_ev.record(stream)
    for r in sorted(requests, key=lambda r: r.arrival):
        gpu_now = start_ev.elapsed_time(_record(torch, stream)) / 1e3
        if r.arrival > gpu_now:
            sleep(r.arrival - gpu_now)
        g = DeviceGrid(get_rule(r.rule), r.width, r.height, 0)
        g.fill_random(r.seed)
        in_b = _lib.simple(g.rule.id, 0, g.ptrs(0), g.ptrs(1), r.width, r.height, 0, r.steps, 0, stream.cuda_stream)
        g.cur = int(in_b)
        ev = _record(torch, stream)
        t_done = start_ev.elapsed_time(ev) / 1e3
        rep.completions[r.id] = t_done
        rep.latencies[r.id] = t_done - r.arrival
        rep.cell_updates += r.cell_updates
        if keep_results:
            results[r.id] = g
        if time_limit is not None and clock() - t0 > time_limit:
            break
    done = [r for r in requests if r.id in rep.completions]
    rep.makespan = max(rep.completions.values()) - min(r.arrival for r in done) if done else 0.0
    return rep, results


def _record(torch, stream):
    ev = torch.cuda.Event(enable_timing=True)
    ev.record(stream)
    ev.synchronize()
    return ev

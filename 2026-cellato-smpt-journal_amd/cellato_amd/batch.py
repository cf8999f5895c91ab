"""Many independent grids advanced by one hipGraph launch (C ABI `cellato_batch_*`).

`Batch(jobs)` puts every job's grid in HBM (two buffers each), plans the launch schedule
once (per-grid depth and chunk, from the job or from the autotuner) and captures it into a
graph; `run()` replays that graph. `plain_loop(jobs)` is the comparison point: one
`cellato_hip_simple` call per grid, one kernel launch per generation.
"""
import ctypes
import math
from dataclasses import dataclass, field

import numpy as np

from . import _lib, grid as hg
from .planes import DeviceGrid
from .rules import rule as get_rule, iid_values


@dataclass
class Job:
    rule: str
    width: int
    height: int
    steps: int
    t0: int = 0
    temporal_steps: int = 0      # 0: ask the tuner
    chunk_rows: int = 0
    seed: int = 0
    values: np.ndarray = field(default=None, repr=False)   # initial state values (host), or
    on_device: bool = False      # True: seeded i.i.d. state generated on the device (fill_random)

    @property
    def cells(self):
        return self.width * self.height

    def initial_values(self):
        if self.values is None:
            self.values = iid_values(self.rule, self.width, self.height, self.seed)
        return self.values


BATCH_RULES = ["game-of-life", "brian", "forest-fire", "traffic", "maze"]


def batch_workload(n=512, seed=2026, rules=BATCH_RULES, side_range=(64, 4096), steps_range=(16, 1024)):
    """Batching configuration: automata uniform over `rules`, square sides log-uniform in
    side_range rounded to a multiple of 64, steps uniform in steps_range, one seed."""
    rng = np.random.Generator(np.random.PCG64(seed))
    jobs = []
    lo, hi = math.log(side_range[0]), math.log(side_range[1])
    for i in range(n):
        name = rules[int(rng.integers(0, len(rules)))]
        side = int(round(math.exp(rng.uniform(lo, hi)) / 64)) * 64
        side = min(max(side, side_range[0]), side_range[1])
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        jobs.append(Job(name, side, side, steps, seed=seed * 100003 + i, on_device=True))
    return jobs


def _init(d, j):
    if j.on_device:
        d.fill_random(j.seed)
    else:
        d.upload(hg.pack(d.rule.index_of(j.initial_values()), d.rule.planes, hg.LINEAR))


def plan(jobs, max_waves=0):
    """The schedule `Batch(jobs)` would build (host only; no GPU needed): start round of each
    grid's chain of passes, round count, launches, tasks. Every job needs its depth set."""
    arr = (_lib.BatchGrid * len(jobs))()
    fake = [0x1000 * (i + 1) for i in range(5)]      # plane pointers are not dereferenced
    for i, j in enumerate(jobs):
        g = arr[i]
        g.rule, g.layout, g.width, g.height = get_rule(j.rule).id, hg.LINEAR, j.width, j.height
        g.steps, g.t0, g.temporal_steps, g.chunk_rows = j.steps, j.t0, j.temporal_steps, j.chunk_rows
        for p in range(5):
            g.a[p] = fake[p]
            g.b[p] = fake[p] + 0x100000
    starts = (ctypes.c_int * max(1, len(jobs)))()
    r, l, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    _lib.check(_lib.lib().cellato_batch_plan(arr, len(jobs), max_waves, starts, ctypes.byref(r), ctypes.byref(l),
                                             ctypes.byref(t)))
    return {"start_round": list(starts[:len(jobs)]), "rounds": r.value, "launches": l.value, "tasks": t.value}


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


class Batch:
    def __init__(self, jobs, tuner=None, max_waves=0):
        self.jobs = list(jobs)
        self.grids = []
        arr = (_lib.BatchGrid * len(self.jobs))()
        # depths: the job's own, else the tuner's choice measured on the batch itself (4
        # without a tuner)
        auto = [j for j in self.jobs if j.temporal_steps <= 0]
        picked = {}
        if tuner is not None and auto:
            depths, _, _ = tuner.batch_depths(auto)
            picked = {id(j): T for j, T in zip(auto, depths)}
        self.depths = []
        for i, j in enumerate(self.jobs):
            r = get_rule(j.rule)
            T, chunk = (j.temporal_steps, j.chunk_rows) if j.temporal_steps > 0 else (picked.get(id(j), 4), 0)
            self.depths.append(T)
            d = DeviceGrid(r, j.width, j.height, hg.LINEAR)
            _init(d, j)
            self.grids.append(d)
            g = arr[i]
            g.rule, g.layout, g.width, g.height = r.id, hg.LINEAR, j.width, j.height
            g.steps, g.t0, g.temporal_steps, g.chunk_rows = j.steps, j.t0, T, chunk
            for p, ptr in enumerate(d.ptrs(0)):
                g.a[p] = ptr
            for p, ptr in enumerate(d.ptrs(1)):
                g.b[p] = ptr
        self._arr = arr
        self.handle = ctypes.c_void_p()
        _lib.check(_lib.lib().cellato_batch_create(arr, len(self.jobs), max_waves, ctypes.byref(self.handle)))

    @property
    def cell_updates(self):
        return sum(j.cells * j.steps for j in self.jobs)

    def stats(self):
        r, l, t = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        _lib.check(_lib.lib().cellato_batch_stats(self.handle, ctypes.byref(r), ctypes.byref(l), ctypes.byref(t)))
        return {"rounds": r.value, "launches": l.value, "tasks": t.value}

    def reset(self):
        """Put every grid's initial state back into its first buffer (before a replay)."""
        for j, d in zip(self.jobs, self.grids):
            d.cur = 0
            _init(d, j)

    def run(self, stream=None):
        _lib.check(_lib.lib().cellato_batch_launch(self.handle, ctypes.c_void_p(stream if stream is not None else _stream())))

    def _settle(self, i):
        d = self.grids[i]
        d.cur = _lib.check(_lib.lib().cellato_batch_result_in_b(self.handle, i))
        return d

    def result(self, i):
        import torch
        torch.cuda.synchronize()
        d = self._settle(i)
        return d.rule.value_of(hg.unpack(d.download(), d.width, d.height, hg.LINEAR))

    def checksum(self, i):
        """16-tile checksum of grid i's result, computed on the device."""
        return self._settle(i).checksum_device()

    def close(self):
        if self.handle:
            _lib.lib().cellato_batch_destroy(self.handle)
            self.handle = ctypes.c_void_p()
        self.grids = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlainLoop:
    """The comparison point: the jobs' grids in HBM, advanced one grid after the other with
    `cellato_hip_simple` (one launch per generation) or, with `temporal_steps`, with
    `cellato_hip_temporal`."""

    def __init__(self, jobs):
        self.jobs = list(jobs)
        self.grids = []
        for j in self.jobs:
            r = get_rule(j.rule)
            d = DeviceGrid(r, j.width, j.height, hg.LINEAR)
            _init(d, j)
            self.grids.append(d)

    def reset(self):
        for j, d in zip(self.jobs, self.grids):
            d.cur = 0
            _init(d, j)

    def run(self, temporal_steps=0, tuner=None):
        st = _stream()
        for j, d in zip(self.jobs, self.grids):
            if temporal_steps or tuner:
                T, chunk = tuner.pick(j.rule, j.width, j.height) if tuner else (temporal_steps, 0)
                in_b = _lib.temporal(d.rule.id, 0, d.ptrs(0), d.ptrs(1), j.width, j.height, j.t0, j.steps, T, chunk, st)
            else:
                in_b = _lib.simple(d.rule.id, 0, d.ptrs(0), d.ptrs(1), j.width, j.height, j.t0, j.steps, 0, st)
            d.cur = int(in_b)

    def result(self, i):
        import torch
        torch.cuda.synchronize()
        d = self.grids[i]
        return d.rule.value_of(hg.unpack(d.download(), d.width, d.height, hg.LINEAR))

    def checksum(self, i):
        return self.grids[i].checksum_device()

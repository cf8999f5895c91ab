"""Autotuner: the launch shape per (automaton, size): temporal depth T and rows per task
(the tile shape: a wave's chunk of rows, 0 = the planner's choice, csrc/plan.hpp auto_chunk).

A candidate (T, chunk) is measured as throughput, the way a batch uses it: enough copies of a
grid of the bucket's size to fill the GPU, advanced by one graph launch (`Batch`), timed with
HIP events over a few replays. Sizes are bucketed by powers of two of width and height.
Without a cost model every candidate of a bucket is measured; with one
(`cellato_amd.costmodel.CostModel`) only the `budget` candidates it ranks best are measured.
Results are cached in a JSON file.
"""
import json
import math
from pathlib import Path

import numpy as np

from .rules import rule as get_rule

DEPTHS = (1, 2, 3, 4, 5, 6, 8)
CHUNKS = (0, 32, 128)          # rows per task: the planner's choice, short and tall tiles
BATCH_DEPTHS = (2, 4, 6, 8)
CANDIDATES = [(T, c) for T in DEPTHS for c in CHUNKS]


def bucket(side):
    return int(2 ** max(6, min(15, round(math.log2(max(side, 1))))))


_KERNELS = {}


def kernel_shape(rule_name, width, T):
    """(depth the pass runs at, registers per lane) of the kernel a pass of up to T generations
    over rows `width` cells wide launches (cellato_hip_kernel_info; needs the GPU)."""
    from . import _lib
    wx = width // 64
    key = (rule_name, 0 if wx < 64 else 1 if wx == 64 else 2, T)
    if key not in _KERNELS:
        eff, _, regs = _lib.kernel_info(get_rule(rule_name).id, 0, width, T)
        _KERNELS[key] = (eff, regs)
    return _KERNELS[key]


def features(rule_name, width, height, T, chunk=0):
    """Inputs of the cost model: what the rule is made of (planes and the kinds of nodes of
    its AST, not its identity), what the compiler made of it at this depth (the kernel's
    registers per lane, the waves per SIMD they allow, the depth the pass really runs at),
    so that the model says something about rules it has not seen; the grid size, the depth,
    the rows per task, and the lane mapping the size implies."""
    r = get_rule(rule_name)
    wx = width // 64
    eff, regs = kernel_shape(rule_name, width, T)
    waves = min(8, 512 // max(regs, 1))
    return [r.planes / 5.0, math.log2(width) / 15.0, math.log2(height) / 15.0, T / 8.0, 1.0 / T,
            (T / 8.0) ** 2, r.planes * T / 40.0,
            float(r.has("moore")), float(r.has("von_neumann")), float(r.has("margolus")),
            float(r.has("neighbor")), float(r.has("dynamic_target")), float(r.has("alternating")),
            1.0 if wx < 64 else 0.0, 1.0 if wx == 64 else 0.0,
            regs / 512.0, waves / 8.0, eff / 8.0,
            0.0 if chunk == 0 else math.log2(chunk) / 10.0, 1.0 if chunk == 0 else 0.0]


class Tuner:
    def __init__(self, cache_path=None, model=None, budget=3, target_cells=1 << 27, steps=24, replays=2, measure=True):
        self.cache_path = Path(cache_path) if cache_path else None
        self.cache = {}
        if self.cache_path and self.cache_path.exists():
            self.cache = json.loads(self.cache_path.read_text())
        self.model, self.budget = model, budget
        self.target_cells, self.steps, self.measure_enabled = target_cells, steps, measure
        self.replays = replays
        self.measurements = 0

    @staticmethod
    def key(rule_name, width, height):
        return f"{rule_name}/{bucket(width)}x{bucket(height)}"

    def pick(self, rule_name, width, height):
        k = self.key(rule_name, width, height)
        if k not in self.cache:
            if not self.measure_enabled:
                return 4, 0
            self.cache[k] = self.tune(rule_name, bucket(width), bucket(height))
            self.save()
        e = self.cache[k]
        return int(e["T"]), int(e["chunk"])

    def save(self):
        if self.cache_path:
            self.cache_path.parent.mkdir(parents=True, exist_ok=True)
            self.cache_path.write_text(json.dumps(self.cache, indent=1, sort_keys=True))

    def ranked_candidates(self, rule_name, width, height):
        if self.model is None:
            return list(CANDIDATES)
        X = np.array([features(rule_name, width, height, T, c) for T, c in CANDIDATES], np.float32)
        pred = self.model.predict(X)
        order = np.argsort(pred)
        return [CANDIDATES[i] for i in order[: self.budget]]

    def measure(self, rule_name, width, height, T, chunk):
        """GCUPS of `copies` grids of this size advanced `steps` generations in one batch,
        timed over `replays` graph launches (the grids keep evolving; only time is read)."""
        import torch
        from .batch import Batch, Job
        copies = max(1, min(2048, self.target_cells // (width * height)))
        jobs = [Job(rule_name, width, height, self.steps, temporal_steps=T, chunk_rows=chunk, seed=i, on_device=True)
                for i in range(copies)]
        b = Batch(jobs)
        b.run()                      # warm-up replay
        torch.cuda.synchronize()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(self.replays):
            b.run()
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop)
        b.close()
        self.measurements += 1
        return copies * width * height * self.steps * self.replays / (ms * 1e-3) / 1e9

    def ns_per_cell(self, rule_name, width, height, T):
        """Time per cell update at depth T for this bucket: measured when the bucket's tuning
        measured T, else the cost model's prediction (tuning the bucket first if needed)."""
        self.pick(rule_name, width, height)
        e = self.cache.get(self.key(rule_name, width, height))
        best = [m["gcups"] for m in (e or {}).get("measured", []) if m["T"] == T and m["gcups"] > 0]
        if best:
            return 1.0 / max(best)
        if self.model is not None:
            x = np.array([features(rule_name, bucket(width), bucket(height), T)], np.float32)
            return float(np.exp(self.model.predict(x)[0]))
        return float("inf")

    def automaton_depths(self, jobs):
        """One depth per automaton: the depth with the least predicted time over that
        automaton's jobs (cell updates x time per cell update of each job's bucket)."""
        by_rule = {}
        for j in jobs:
            by_rule.setdefault(j.rule, []).append(j)
        best = {}
        for name, js in by_rule.items():
            cost = {T: sum(j.cells * j.steps * self.ns_per_cell(name, j.width, j.height, T) for j in js)
                    for T in DEPTHS}
            best[name] = min(DEPTHS, key=lambda T: (cost[T], -T))
        return [best[j.rule] for j in jobs]

    def batch_candidates(self, jobs):
        """Depth assignments worth timing on a batch: one depth for all grids (BATCH_DEPTHS),
        one per automaton, and the per-(automaton, size) picks."""
        cands = {f"T{T}": [T] * len(jobs) for T in BATCH_DEPTHS}
        cands["per_automaton"] = self.automaton_depths(jobs)
        cands["per_automaton_and_size"] = [self.pick(j.rule, j.width, j.height)[0] for j in jobs]
        return cands

    def batch_chunks(self, jobs):
        """Rows per task of each job at its bucket's tuned tile shape (0 = the planner's)."""
        return [self.pick(j.rule, j.width, j.height)[1] for j in jobs]

    def batch_depths(self, jobs):
        """Depth per job of a batch, measured on the batch itself: every candidate of
        `batch_candidates` is built (`Batch`) and replayed once after a warm-up replay, and
        the fastest is kept. A batch's time is set by its rounds (one launch per automaton and
        depth in each round, each paying a pass's pipeline fill), which per-bucket
        measurements do not see, so the batch is timed whole. Returns (depths, name of the
        winner, {candidate: GCUPS})."""
        import torch
        from .batch import Batch
        saved = [(j.temporal_steps, j.chunk_rows) for j in jobs]
        cu = sum(j.cells * j.steps for j in jobs)
        cands = self.batch_candidates(jobs)
        chunks = {"per_automaton_and_size": self.batch_chunks(jobs)}   # the buckets' tuned tile shapes
        out = {}
        try:
            for name, depths in cands.items():
                for j, T, c in zip(jobs, depths, chunks.get(name, [0] * len(jobs))):
                    j.temporal_steps, j.chunk_rows = T, c
                b = Batch(jobs)
                b.run()
                torch.cuda.synchronize()
                start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                start.record()
                b.run()
                stop.record()
                torch.cuda.synchronize()
                out[name] = cu / (start.elapsed_time(stop) * 1e-3) / 1e9
                b.close()
                del b
                self.measurements += 1
        finally:
            for j, (t, c) in zip(jobs, saved):
                j.temporal_steps, j.chunk_rows = t, c
        best = max(out, key=out.get)
        if best in chunks:
            for j, c in zip(jobs, chunks[best]):
                j.chunk_rows = c
        return cands[best], best, out

    def tune(self, rule_name, width, height):
        results = []
        for T, chunk in self.ranked_candidates(rule_name, width, height):
            results.append({"T": T, "chunk": chunk, "gcups": self.measure(rule_name, width, height, T, chunk)})
        best = max(results, key=lambda r: r["gcups"])
        return {"T": best["T"], "chunk": best["chunk"], "gcups": best["gcups"], "measured": results}


SIDES = (64, 128, 256, 512, 1024, 2048, 4096, 8192)


def collect_dataset(path, rules=None, sides=SIDES, log=print):
    """Measure every candidate of every (rule, square side) bucket and write the samples
    (features and measured ns per cell update) to `path` as JSON."""
    from .rules import all_rules
    names = rules or [r.name for r in all_rules()]
    tuner = Tuner(measure=True)
    rows = []
    for name in names:
        for side in sides:
            for T, chunk in CANDIDATES:
                g = tuner.measure(name, side, side, T, chunk)
                rows.append({"rule": name, "width": side, "height": side, "T": T, "chunk": chunk, "gcups": g,
                             "ns_per_cell": 1.0 / g})
            best = max((r for r in rows if r["rule"] == name and r["width"] == side), key=lambda r: r["gcups"])
            log(f"{name} {side}: best T={best['T']} chunk={best['chunk']} {best['gcups']:.0f} GCUPS")
    Path(path).write_text(json.dumps({"device": "MI355X", "steps": tuner.steps, "replays": tuner.replays,
                                      "target_cells": tuner.target_cells, "samples": rows}))
    return rows


if __name__ == "__main__":
    import sys
    if len(sys.argv) >= 3 and sys.argv[1] == "collect":
        collect_dataset(sys.argv[2], rules=sys.argv[3].split(",") if len(sys.argv) > 3 else None)

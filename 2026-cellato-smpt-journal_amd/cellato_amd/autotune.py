"""Autotuner: the temporal depth T per (automaton, size).

The rows per task are not tuned here: the launch planner picks them per launch round from
the round's total work (csrc/plan.hpp, auto_chunk). A depth is measured as throughput, the
way a batch uses it: enough copies of a grid of the bucket's size to fill the GPU, advanced
by one graph launch (`Batch`), timed with HIP events over a few replays. Sizes are bucketed
by powers of two of width and height. Without a cost model every depth of a bucket is
measured; with one (`cellato_amd.costmodel.CostModel`) only the `budget` depths it ranks
best are measured. Results are cached in a JSON file.
"""
import json
import math
from pathlib import Path

import numpy as np

from .rules import rule as get_rule

DEPTHS = (1, 2, 3, 4, 5, 6, 8)
CANDIDATES = [(T, 0) for T in DEPTHS]


def bucket(side):
    return int(2 ** max(6, min(15, round(math.log2(max(side, 1))))))


def features(rule_name, width, height, T, chunk=0):
    """Inputs of the cost model: what the rule is made of (planes and the kinds of nodes of
    its AST, not its identity, so that the model says something about rules it has not
    seen), the grid size, the depth, and the lane mapping the size implies."""
    r = get_rule(rule_name)
    wx = width // 64
    return [r.planes / 5.0, math.log2(width) / 15.0, math.log2(height) / 15.0, T / 8.0, 1.0 / T,
            (T / 8.0) ** 2, r.planes * T / 40.0,
            float(r.has("moore")), float(r.has("von_neumann")), float(r.has("margolus")),
            float(r.has("neighbor")), float(r.has("dynamic_target")), float(r.has("alternating")),
            1.0 if wx < 64 else 0.0, 1.0 if wx == 64 else 0.0]


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

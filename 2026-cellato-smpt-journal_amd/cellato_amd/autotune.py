"""Autotuner: the temporal depth T and the rows per task (`chunk_rows`) per (automaton, size).

A configuration is measured as throughput, the way a batch uses it: enough copies of a grid
of the bucket's size to fill the GPU, advanced by one graph launch (`Batch`), timed with
HIP events. Sizes are bucketed by powers of two of width and height. Without a cost model
every candidate of a bucket is measured; with one (`cellato_amd.costmodel.CostModel`) only the
`budget` candidates it ranks best are measured. Results are cached in a JSON file.
"""
import json
import math
from pathlib import Path

import numpy as np

from .rules import rule as get_rule

DEPTHS = (1, 2, 3, 4, 5, 6, 8)
CHUNKS = (0, 8, 32, 64)
CANDIDATES = [(T, c) for T in DEPTHS for c in CHUNKS]


def bucket(side):
    return int(2 ** max(6, min(15, round(math.log2(max(side, 1))))))


def features(rule_name, width, height, T, chunk):
    """Inputs of the cost model: rule one-hot, planes, log sizes, depth, log chunk, lane mapping."""
    r = get_rule(rule_name)
    onehot = [0.0] * 10
    onehot[r.id] = 1.0
    wx = width // 64
    eff_chunk = chunk if chunk > 0 else max(16, 4 * T)
    return onehot + [r.planes / 5.0, math.log2(width) / 15.0, math.log2(height) / 15.0, T / 8.0,
                     math.log2(eff_chunk) / 8.0, 1.0 if wx < 64 else 0.0, 1.0 if wx == 64 else 0.0,
                     (2.0 * T / eff_chunk), 1.0 / T]


class Tuner:
    def __init__(self, cache_path=None, model=None, budget=4, target_cells=1 << 26, steps=24, measure=True):
        self.cache_path = Path(cache_path) if cache_path else None
        self.cache = {}
        if self.cache_path and self.cache_path.exists():
            self.cache = json.loads(self.cache_path.read_text())
        self.model, self.budget = model, budget
        self.target_cells, self.steps, self.measure_enabled = target_cells, steps, measure
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
        """GCUPS of `copies` grids of this size advanced `steps` generations in one batch."""
        import torch
        from .batch import Batch, Job
        copies = max(1, min(512, self.target_cells // (width * height)))
        jobs = [Job(rule_name, width, height, self.steps, temporal_steps=T, chunk_rows=chunk, seed=i)
                for i in range(copies)]
        b = Batch(jobs)
        b.run()                      # warm-up replay
        torch.cuda.synchronize()
        b.reset()
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        b.run()
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop)
        b.close()
        self.measurements += 1
        return copies * width * height * self.steps / (ms * 1e-3) / 1e9

    def tune(self, rule_name, width, height):
        results = []
        for T, chunk in self.ranked_candidates(rule_name, width, height):
            results.append({"T": T, "chunk": chunk, "gcups": self.measure(rule_name, width, height, T, chunk)})
        best = max(results, key=lambda r: r["gcups"])
        return {"T": best["T"], "chunk": best["chunk"], "gcups": best["gcups"], "measured": results}

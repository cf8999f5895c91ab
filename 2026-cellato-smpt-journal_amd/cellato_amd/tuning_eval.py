"""What the learned cost model gains: on the measured dataset (every launch shape (T, rows
per task) of every (automaton, size) bucket), compare picking shapes with and without the
network.

Held-out evaluation: the network is trained on all buckets except one automaton
(holdout="rule") or one grid size (holdout="size") and ranks the candidates of the held-out
buckets; a strategy "measures" its chosen candidates (reads them from the dataset) and keeps
the best. Reported per strategy: mean fraction of the bucket's best throughput reached with
`budget` measurements, against random picks of the same number of candidates, a fixed shape
(T=4, the planner's rows per task), and measuring all candidates (fraction 1 by definition)."""
import json
from pathlib import Path

import numpy as np

from .autotune import features

DATA = Path(__file__).parent / "data" / "depth_tuning_mi355x.json"


def load(path=DATA):
    rows = json.loads(Path(path).read_text())["samples"]
    table = {}
    for r in rows:
        table.setdefault((r["rule"], r["width"], r["height"]), {})[(r["T"], r.get("chunk", 0))] = r["gcups"]
    return rows, table


def evaluate(budget=2, epochs=600, seed=0, hidden=(64, 64), path=DATA, holdout="rule"):
    from .costmodel import CostModel
    rows, table = load(path)
    part = (lambda rule, w: rule) if holdout in ("rule", "automaton") else (lambda rule, w: w)
    folds = sorted({part(k[0], k[1]) for k in table})
    rng = np.random.default_rng(seed)
    res = {"model_topk": [], "model_top1": [], "default_T4": [], "random_k": []}
    losses = []
    for held in folds:
        train = [r for r in rows if part(r["rule"], r["width"]) != held]
        X = np.array([features(r["rule"], r["width"], r["height"], r["T"], r.get("chunk", 0)) for r in train],
                     np.float32)
        y = np.log(np.array([r["ns_per_cell"] for r in train], np.float32))
        m = CostModel(X.shape[1], hidden=hidden, seed=seed)
        hist = m.fit(X, y, epochs=epochs)
        losses.append(hist[-1])
        for key, by_c in table.items():
            if part(key[0], key[1]) != held:
                continue
            best = max(by_c.values())
            cands = sorted(by_c)
            pred = m.predict(np.array([features(key[0], key[1], key[2], T, c) for T, c in cands], np.float32))
            order = [cands[i] for i in np.argsort(pred)]
            res["model_topk"].append(max(by_c[c] for c in order[:budget]) / best)
            res["model_top1"].append(by_c[order[0]] / best)
            res["default_T4"].append(by_c[(4, 0)] / best)
            idx = np.arange(len(cands))
            picks = [max(by_c[cands[i]] for i in rng.choice(idx, size=budget, replace=False)) for _ in range(64)]
            res["random_k"].append(float(np.mean(picks)) / best)
    summary = {k: float(np.mean(v)) for k, v in res.items()}
    summary["budget"] = budget
    summary["holdout"] = holdout
    summary["exhaustive_measurements_per_bucket"] = len(next(iter(table.values())))
    summary["final_train_loss_mean"] = float(np.mean(losses))
    return summary

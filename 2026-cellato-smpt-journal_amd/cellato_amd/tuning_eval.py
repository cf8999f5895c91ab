"""What the learned cost model gains: on the measured dataset (every depth of every
(automaton, size) bucket), compare picking depths with and without the network.

Held-out evaluation: the network is trained on all buckets except one automaton
(holdout="rule") or one grid size (holdout="size") and ranks the depths of the held-out
buckets; a strategy "measures" its chosen depths (reads them from the dataset) and keeps the
best. Reported per strategy: mean fraction of the bucket's best throughput reached with
`budget` measurements, against random picks of the same number of depths, a fixed depth
(T=4), and measuring all depths (fraction 1 by definition, len(DEPTHS) measurements)."""
import json
from pathlib import Path

import numpy as np

from .autotune import DEPTHS, features

DATA = Path(__file__).parent / "data" / "depth_tuning_mi355x.json"


def load(path=DATA):
    rows = json.loads(Path(path).read_text())["samples"]
    table = {}
    for r in rows:
        table.setdefault((r["rule"], r["width"], r["height"]), {})[r["T"]] = r["gcups"]
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
        X = np.array([features(r["rule"], r["width"], r["height"], r["T"]) for r in train], np.float32)
        y = np.log(np.array([r["ns_per_cell"] for r in train], np.float32))
        m = CostModel(X.shape[1], hidden=hidden, seed=seed)
        hist = m.fit(X, y, epochs=epochs)
        losses.append(hist[-1])
        for key, by_T in table.items():
            if part(key[0], key[1]) != held:
                continue
            best = max(by_T.values())
            Ts = [T for T in DEPTHS if T in by_T]
            pred = m.predict(np.array([features(key[0], key[1], key[2], T) for T in Ts], np.float32))
            order = [Ts[i] for i in np.argsort(pred)]
            res["model_topk"].append(max(by_T[T] for T in order[:budget]) / best)
            res["model_top1"].append(by_T[order[0]] / best)
            res["default_T4"].append(by_T[4] / best)
            picks = [max(by_T[int(T)] for T in rng.choice(Ts, size=budget, replace=False)) for _ in range(64)]
            res["random_k"].append(float(np.mean(picks)) / best)
    summary = {k: float(np.mean(v)) for k, v in res.items()}
    summary["budget"] = budget
    summary["holdout"] = holdout
    summary["exhaustive_measurements_per_bucket"] = len(DEPTHS)
    summary["final_train_loss_mean"] = float(np.mean(losses))
    return summary

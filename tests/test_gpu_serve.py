"""Serving with continuous batching: every served result equals the same request run alone
(plain simple traverser) and the CPU oracle on the host restatement of its initial state;
latencies are non-negative and every request completes."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def test_served_results_equal_alone():
    import torch
    from cellato_amd.serve import Server, serving_workload, serve_one_at_a_time
    reqs = serving_workload(n=24, rate=200.0, seed=3, side_range=(64, 512), steps_range=(8, 64))
    srv = Server(keep_results=True, default_T=3)
    rep = srv.run(reqs)
    assert len(rep.latencies) == len(reqs) and min(rep.latencies.values()) >= 0
    base, alone = serve_one_at_a_time(reqs, keep_results=True)
    for r in reqs:
        a = srv.results[r.id].to_plane_grid().to_standard()
        b = alone[r.id].to_plane_grid().to_standard()
        assert np.array_equal(a, b), r
    from cellato_amd.rules import device_random_values
    for r in reqs[:6]:
        init = device_random_values(r.rule, r.width, r.height, r.seed)
        want = O.run_standard(r.rule, init, r.steps)
        assert np.array_equal(srv.results[r.id].to_plane_grid().to_standard(), want), r
    s = rep.summary()
    assert s["requests"] == 24 and s["p99_s"] >= s["p50_s"] >= 0


def test_tight_budget_still_completes():
    from cellato_amd.serve import Server, serving_workload
    reqs = serving_workload(n=12, rate=500.0, seed=4, side_range=(64, 256), steps_range=(8, 32))
    budget = 2 * max(r.bytes for r in reqs)
    srv = Server(hbm_budget=budget, default_T=2)
    rep = srv.run(reqs)
    assert len(rep.latencies) == 12 and rep.peak_bytes <= budget

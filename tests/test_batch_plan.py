"""Batch schedules planned on the host (no GPU): every grid's chain of passes sits in
consecutive rounds inside the schedule, the round count is the longest chain, and placing
the chains levels the rounds' work compared with starting every grid in round 0."""
import math

import pytest

from cellato_amd.batch import Job, batch_workload, plan


def loads(jobs, starts, R):
    out = [0] * R
    for j, s in zip(jobs, starts):
        done, r = 0, 0
        while done < j.steps:
            T = min(j.temporal_steps, j.steps - done)
            out[s + r] += j.cells * T
            done += T
            r += 1
    return out


@pytest.mark.parametrize("T", [1, 4, 8])
def test_chains_fit_and_rounds_are_level(T):
    jobs = batch_workload(256, seed=11)
    for j in jobs:
        j.temporal_steps = T
    p = plan(jobs)
    passes = [math.ceil(j.steps / T) for j in jobs]
    assert p["rounds"] == max(passes)
    for s, n in zip(p["start_round"], passes):
        assert 0 <= s and s + n <= p["rounds"]
    packed = loads(jobs, p["start_round"], p["rounds"])
    unpacked = loads(jobs, [0] * len(jobs), p["rounds"])
    assert sum(packed) == sum(unpacked)
    assert max(packed) <= 0.75 * max(unpacked)
    mean = sum(packed) / p["rounds"]
    assert max(packed) <= 1.25 * mean


def test_single_chain_and_equal_chains():
    p = plan([Job("game-of-life", 128, 64, 40, temporal_steps=4)])
    assert p["start_round"] == [0] and p["rounds"] == 10
    # four equal grids of 3 passes and one of 12: the short chains spread over the rounds
    jobs = [Job("maze", 256, 256, 12, temporal_steps=4) for _ in range(4)] + [Job("maze", 64, 64, 48, temporal_steps=4)]
    p = plan(jobs)
    assert p["rounds"] == 12
    assert sorted(p["start_round"][:4]) == [0, 3, 6, 9]


def test_tasks_and_launches_do_not_depend_on_placement_count():
    jobs = batch_workload(64, seed=3)
    for j in jobs:
        j.temporal_steps = 2
    p = plan(jobs)
    assert p["launches"] >= p["rounds"]
    assert p["tasks"] >= sum(math.ceil(j.steps / 2) for j in jobs)

"""Batches: many grids of different automata, sizes, step counts, start steps and depths in
one graph launch; every grid equals the CPU oracle, alone and in the mix; replays of the
graph give the same results."""
import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

ALL = ["game-of-life", "wire", "forest-fire", "excitable", "brian", "maze", "traffic", "fluid", "critters", "cyclic"]


def small_jobs(seed=3, n=40):
    from cellato_amd.batch import Job
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        w = int(rng.choice([64, 128, 192, 640, 4096, 4160, 8192]))
        h = int(rng.integers(1, 90))
        jobs.append(Job(ALL[i % 10], w, h, int(rng.integers(0, 40)), t0=int(rng.integers(0, 3)),
                        temporal_steps=int(rng.integers(1, 9)), chunk_rows=int(rng.choice([0, 5, 16])), seed=i))
    return jobs


def test_batch_matches_oracle_and_replays():
    from cellato_amd.batch import Batch
    jobs = small_jobs()
    b = Batch(jobs)
    want = [O.run_standard(j.rule, j.initial_values(), j.steps, t0=j.t0) for j in jobs]
    for replay in range(3):
        if replay:
            b.reset()
        b.run()
        for i, j in enumerate(jobs):
            assert np.array_equal(b.result(i), want[i]), (replay, i, j)
    st = b.stats()
    assert st["launches"] >= st["rounds"] > 0
    b.close()


def test_alone_equals_in_the_mix():
    from cellato_amd.batch import Batch
    jobs = small_jobs(seed=5, n=24)
    mixed = Batch(jobs)
    mixed.run()
    for i, j in enumerate(jobs[:6]):
        alone = Batch([j])
        alone.run()
        assert np.array_equal(alone.result(0), mixed.result(i)), j
        alone.close()
    mixed.close()


def test_batch_rejects_bad_grid():
    from cellato_amd import _lib
    from cellato_amd.batch import Batch, Job
    with pytest.raises(_lib.CellatoError) as e:
        Batch([Job("maze", 64, 8, 3, temporal_steps=12)])
    assert e.value.code == _lib.E_TEMPORAL

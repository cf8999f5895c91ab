"""The benchmark workloads at their real shapes, on the HIP path, against the CPU oracle and
against the plain per-grid path:

- Game of Life 1024^2, Bernoulli(0.5), 128 steps: simple and temporal traversers, both
  layouts, equal to the standard evaluator (the first workload compares the standard and
  bit_planes evaluators on the CPU; here the GPU path joins them);
- Game of Life and WireWorld 2048^2 subsets of the single-grid workload (simple, depths
  1/2/4/8) and the multi-plane rules (Cyclic, Greenberg-Hastings, HPP, Critters; depths 2/4/8),
  64 steps, equal to the standard evaluator;
- the 512-grid batch (five automata, sides 64..4096, 16..1024 steps): every grid's 16-tile
  checksum equals the plain per-grid loop's, under one depth for all grids and under mixed
  per-(automaton, size) depths, and the 8 smallest grids equal the standard evaluator;
- the serving workload (256 Poisson requests over the ten automata, sides 128..8192, 64..2048
  steps, 200 GB budget): every served result equals the one-at-a-time result and the 4
  smallest equal the standard evaluator.
"""
import math
import sys

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import run_gpu

pytestmark = pytest.mark.gpu


def note(msg):
    print(f"[gpu] {msg}", file=sys.stderr, flush=True)


@pytest.mark.parametrize("layout", [0, 1])
def test_gol_1024_128_steps(layout):
    import cellato_amd as C
    v = C.iid_values("game-of-life", 1024, 1024, 2026)
    assert 0.49 < v.mean() < 0.51
    want = O.run_standard("game-of-life", v, 128)
    assert np.array_equal(O.run_bitplanes("game-of-life", v, 128), want)
    for kind, T in (("simple", 1), ("temporal", 1), ("temporal", 4), ("temporal", 8)):
        got = run_gpu("game-of-life", v, layout, kind, T, 128)
        assert np.array_equal(got, want), (kind, T)


SUBSETS = [("game-of-life", [("simple", 1), ("temporal", 1), ("temporal", 2), ("temporal", 4), ("temporal", 8)]),
           ("wire", [("simple", 1), ("temporal", 1), ("temporal", 2), ("temporal", 4), ("temporal", 8)]),
           ("cyclic", [("temporal", 2), ("temporal", 4), ("temporal", 8)]),
           ("excitable", [("temporal", 2), ("temporal", 4), ("temporal", 8)]),
           ("fluid", [("temporal", 2), ("temporal", 4), ("temporal", 8)]),
           ("critters", [("temporal", 2), ("temporal", 4), ("temporal", 8)])]


@pytest.mark.parametrize("rule,variants", SUBSETS, ids=[s[0] for s in SUBSETS])
def test_2048_subsets(rule, variants):
    import cellato_amd as C
    v = C.iid_values(rule, 2048, 2048, 11)
    want = O.run_standard(rule, v, 64)
    for kind, T in variants:
        assert np.array_equal(run_gpu(rule, v, 0, kind, T, 64), want), (kind, T)


def mixed_depths(jobs):
    """A depth per (automaton, size bucket), spread over 1..8, so that one launch round mixes
    passes of several depths."""
    names = sorted({j.rule for j in jobs})
    return [1 + (names.index(j.rule) * 3 + int(math.log2(j.width))) % 8 for j in jobs]


def test_batch_512_grids():
    import torch
    from cellato_amd.batch import Batch, PlainLoop, batch_workload
    from cellato_amd.rules import device_random_values
    jobs = batch_workload(512, seed=2026)
    assert len({j.rule for j in jobs}) == 5 and all(64 <= j.width <= 4096 and 16 <= j.steps <= 1024 for j in jobs)
    note("plain loop, 512 grids")
    pl = PlainLoop(jobs)
    pl.run()
    plain = [pl.checksum(i) for i in range(len(jobs))]
    del pl
    torch.cuda.empty_cache()
    small = sorted(range(len(jobs)), key=lambda i: jobs[i].cells * jobs[i].steps)[:8]
    for label, depths in (("T6", [6] * len(jobs)), ("mixed", mixed_depths(jobs))):
        note(f"batch of 512 grids, depths {label}")
        for j, T in zip(jobs, depths):
            j.temporal_steps = T
        b = Batch(jobs)
        b.run()
        sums = [b.checksum(i) for i in range(len(jobs))]
        bad = [i for i in range(len(jobs)) if sums[i] != plain[i]]
        assert not bad, (label, [(jobs[i].rule, jobs[i].width, jobs[i].steps, jobs[i].temporal_steps) for i in bad[:5]])
        for i in small:
            j = jobs[i]
            want = O.run_standard(j.rule, device_random_values(j.rule, j.width, j.height, j.seed), j.steps)
            assert np.array_equal(b.result(i), want), (label, j)
        b.close()
        del b
        torch.cuda.empty_cache()


def test_serving_256_requests():
    import torch
    from cellato_amd.rules import device_random_values
    from cellato_amd.serve import Server, serve_one_at_a_time, serving_workload
    reqs = serving_workload(n=256, rate=2000.0, seed=7)
    assert len({r.rule for r in reqs}) == 10
    assert min(r.width for r in reqs) < 512 and max(r.width for r in reqs) > 4096
    note("serving 256 requests, continuous batching")
    srv = Server(hbm_budget=200e9, default_T=4, tick_passes=16, keep_results=True)
    rep = srv.run(reqs)
    assert len(rep.latencies) == 256 and min(rep.latencies.values()) >= 0
    served = {i: g.checksum_device() for i, g in srv.results.items()}
    for r in sorted(reqs, key=lambda r: r.cell_updates)[:4]:
        want = O.run_standard(r.rule, device_random_values(r.rule, r.width, r.height, r.seed), r.steps)
        assert np.array_equal(srv.results[r.id].to_plane_grid().to_standard(), want), r
    srv.results.clear()
    torch.cuda.empty_cache()
    note("serving 256 requests one at a time")
    _, alone = serve_one_at_a_time(reqs, keep_results=True)
    assert len(alone) == 256
    bad = [r for r in reqs if alone[r.id].checksum_device() != served[r.id]]
    assert not bad, bad[:3]


ALL = ["game-of-life", "wire", "forest-fire", "excitable", "brian", "maze", "traffic", "fluid", "critters", "cyclic"]


@pytest.mark.parametrize("rule", ALL)
def test_no_reachable_kernel_uses_scratch(rule):
    """Every kernel the planner can launch (both layouts, the three lane mappings, the tiled
    planes processed as rows at 4096 and 16384 cells wide, the simple traverser and every
    depth, and the batch kernels) keeps its registers out of scratch memory; depths whose
    kernel would spill run at a shallower depth."""
    from cellato_amd import _lib
    from cellato_amd.rules import rule as get_rule
    rid = get_rule(rule).id
    for layout, widths in ((0, (448, 4096, 16384)), (1, (448, 512, 4096, 16384))):
        for width in widths:
            for T in range(0, 9):
                eff, priv, regs = _lib.kernel_info(rid, layout, width, T)
                assert priv == 0, (layout, width, T, priv, regs)
                assert (eff == 0) if T == 0 else (1 <= eff <= T), (layout, width, T, eff)
    for T in range(1, 9):
        eff, priv, regs = _lib.kernel_info(rid, 0, 4096, T, batch=True)
        assert priv == 0 and 1 <= eff <= T, (T, eff, priv, regs)
    if rule == "cyclic":
        assert _lib.kernel_info(rid, 0, 16384, 8)[0] == 4 and _lib.kernel_info(rid, 1, 16384, 8)[0] == 8

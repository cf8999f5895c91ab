"""Bit-exactness of the HIP traversers against the CPU oracle (standard evaluator), for all
ten automata, both layouts, the simple traverser and the temporal traverser at several
depths, over grid shapes that exercise every lane mapping (rows narrower than a wave, exactly
one wave, several full waves, a partial last wave), odd start steps, steps that are not a
multiple of the depth, and very short tori."""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from gpu_util import run_gpu

pytestmark = pytest.mark.gpu

RULES = ["game-of-life", "wire", "forest-fire", "excitable", "brian", "maze", "traffic", "fluid", "critters", "cyclic"]
LINEAR_SHAPES = [(192, 70), (4096, 40), (8192, 50), (8320, 33), (4160, 20), (16384, 24), (7936, 12), (128, 1), (64, 3)]
# 4104: segments of 62 block columns and a tail of narrow lane groups. Widths of at least 64
# columns of 64 cells run the temporal traverser of rules with at most two planes as rows:
# 4096 one ring of 64 columns (also 8 rows high, a single block row), 4160, 8320 and 12352
# segments of 62 columns and a tail (3, 6 and 7 columns)
TILED_SHAPES = [(40, 24), (512, 16), (1024, 24), (1040, 32), (520, 40), (8, 8), (4096, 40), (4160, 16), (8320, 24),
                (4104, 16), (4096, 8), (12352, 56)]
TRAVERSERS = [("simple", 1), ("temporal", 1), ("temporal", 2), ("temporal", 3), ("temporal", 5), ("temporal", 8)]


def iid(rule, w, h, seed):
    rng = np.random.default_rng(seed)
    st = O.states(rule)
    return st[rng.integers(0, len(st), size=(h, w))].astype(np.int32)


@pytest.mark.parametrize("layout", [0, 1], ids=["linear", "tiled"])
@pytest.mark.parametrize("rule", RULES)
def test_matches_standard_evaluator(rule, layout):
    shapes = LINEAR_SHAPES if layout == 0 else TILED_SHAPES
    for si, (w, h) in enumerate(shapes):
        v = iid(rule, w, h, 100 + si)
        for t0 in (0, 3):
            want = O.run_standard(rule, v, 11, t0=t0)
            for kind, T in TRAVERSERS:
                got = run_gpu(rule, v, layout, kind, T, 11, t0)
                assert np.array_equal(got, want), f"{rule} {w}x{h} {kind} T={T} t0={t0}"


@pytest.mark.parametrize("rule", ["game-of-life", "critters", "cyclic", "traffic"])
def test_chunk_sizes(rule):
    v = iid(rule, 8320, 45, 9)
    want = O.run_standard(rule, v, 16, t0=1)
    for chunk in (1, 2, 3, 7, 64):
        for kind, T in [("simple", 1), ("temporal", 4), ("temporal", 8)]:
            assert np.array_equal(run_gpu(rule, v, 0, kind, T, 16, 1, chunk), want), (chunk, kind, T)
    # the tiled layout: chunks of block rows
    v = iid(rule, 8320, 48, 10)
    want = O.run_standard(rule, v, 16, t0=1)
    for chunk in (1, 2, 5):
        for T in (2, 4, 8):
            assert np.array_equal(run_gpu(rule, v, 1, "temporal", T, 16, 1, chunk), want), ("tiled", chunk, T)


@pytest.mark.parametrize("rule", RULES)
def test_2048_subset_run(rule):
    # the bit-exactness subset of the 16384^2 benchmark: 2048^2, against the CPU bit_planes
    # oracle (itself equal to the standard evaluator, see test_oracle_pinning)
    import cellato_amd
    v = cellato_amd.iid_values(rule, 2048, 2048, 5)
    want = O.run_bitplanes(rule, v, 96)
    for layout, kind, T in [(0, "simple", 1), (0, "temporal", 4), (0, "temporal", 8), (1, "simple", 1), (1, "temporal", 6)]:
        assert np.array_equal(run_gpu(rule, v, layout, kind, T, 96), want), (layout, kind, T)


GOLD = json.loads((Path(__file__).parent / "golden" / "reference_checksums.json").read_text())


@pytest.mark.parametrize("rule", RULES)
def test_published_checksum_16384(rule):
    g = next(e for e in GOLD if e["automaton"] == rule and e["x_size"] == e["y_size"] == 16384 and e["steps"] == 128)
    init = O.init_reference(rule, 16384, 16384, g["seed"])
    for layout, kind, T in [(0, "simple", 1), (0, "temporal", 8), (1, "temporal", 4)]:
        out = run_gpu(rule, init, layout, kind, T, 128)
        assert O.checksum(out) == g["checksum"], (g["source"], layout, kind, T)


@pytest.mark.parametrize("rule", ["game-of-life", "wire"])
def test_large_grid_paths_agree(rule):
    """Past the Infinity Cache and past 2^28-byte plane offsets (32768^2: 128 MiB per plane):
    the simple traverser (64-column segments, loaded outside columns) and the temporal one at
    depths 3 and 8 (62-column segments with halo lanes, tail lane groups) give the same grid
    after 16 generations from the same device-made state; depth 1 and the simple traverser
    are pinned to the oracle above, so the three agreeing pins the large-grid runs too."""
    import torch
    import cellato_amd as C
    from cellato_amd import _lib
    from cellato_amd.planes import DeviceGrid
    r = C.rule(rule)
    n, steps = 32768, 16
    stream = torch.cuda.current_stream().cuda_stream
    results = []
    for kind, T in (("simple", 1), ("temporal", 3), ("temporal", 8)):
        dev = DeviceGrid(r, n, n, 0)
        dev.fill_random(5)
        if kind == "simple":
            in_b = _lib.simple(r.id, 0, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), n, n, 0, steps, 0, stream)
        else:
            in_b = _lib.temporal(r.id, 0, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), n, n, 0, steps, T, 0, stream)
        if in_b:
            dev.cur = 1 - dev.cur
        torch.cuda.synchronize()
        results.append(dev.buf[dev.cur].clone())
        del dev
        torch.cuda.empty_cache()
    assert bool(results[0].any()), f"{rule}: empty grid after {steps} generations"
    assert torch.equal(results[0], results[1]), f"{rule}: simple != temporal T=3 at {n}^2"
    assert torch.equal(results[0], results[2]), f"{rule}: simple != temporal T=8 at {n}^2"


@pytest.mark.parametrize("layout", [0, 1], ids=["linear", "tiled"])
def test_largest_plane_paths_agree(layout):
    """A plane of 4 GiB - 512 KiB, the largest the 32-bit buffer offsets address (Game of Life
    262144 x 131056 cells; offsets past 2^31, rows of 4096 words): the simple traverser and
    the temporal one at depths 4 and 8 give the same grid after 8 generations from the same
    device-made state (16 GiB of device memory for the two grids being compared)."""
    import torch
    import cellato_amd as C
    from cellato_amd import _lib
    from cellato_amd.planes import DeviceGrid
    r = C.rule("game-of-life")
    W, H, steps = 262144, 131056, 8
    assert W * H // 8 == 2**32 - 2**19
    stream = torch.cuda.current_stream().cuda_stream
    ref = None
    for kind, T in (("simple", 1), ("temporal", 4), ("temporal", 8)):
        print(f"[gpu] largest plane layout={layout} {kind} T={T}", flush=True)
        dev = DeviceGrid(r, W, H, layout)
        dev.fill_random(9)
        if kind == "simple":
            in_b = _lib.simple(r.id, layout, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), W, H, 0, steps, 0, stream)
        else:
            in_b = _lib.temporal(r.id, layout, dev.ptrs(dev.cur), dev.ptrs(1 - dev.cur), W, H, 0, steps, T, 0, stream)
        if in_b:
            dev.cur = 1 - dev.cur
        torch.cuda.synchronize()
        out = dev.buf[dev.cur]
        if ref is None:
            ref = out.clone()
            assert bool(ref.any()), "empty grid"
        else:
            assert torch.equal(ref, out), f"layout {layout}: simple != temporal T={T}"
        del dev, out
        torch.cuda.empty_cache()
    del ref
    torch.cuda.empty_cache()

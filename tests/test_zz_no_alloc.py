"""cellato_hip_multi_static (tables in a static device array, no allocation) against
cellato_hip_multi on the same grids: two calls in a row, the second on fewer and smaller grids
continuing from the first call's results; every grid's result buffer equal bit for bit, and
the smallest grids equal to the CPU oracle."""
import ctypes
import sys

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu


def make(jobs):
    from cellato_amd.planes import DeviceGrid
    from cellato_amd.rules import rule as get_rule
    grids = []
    for name, w, h, seed in jobs:
        g = DeviceGrid(get_rule(name), w, h, 0)
        g.fill_random(seed)
        grids.append(g)
    return grids


def call(fn, grids, steps, t0, T, stream):
    from cellato_amd import _lib
    arr = (_lib.BatchGrid * len(grids))()
    for i, g in enumerate(grids):
        b = arr[i]
        b.rule, b.layout, b.width, b.height = g.rule.id, 0, g.width, g.height
        b.steps, b.t0, b.temporal_steps, b.chunk_rows = steps[i], t0[i], T[i], 0
        for p, ptr in enumerate(g.ptrs(g.cur)):
            b.a[p] = ptr
        for p, ptr in enumerate(g.ptrs(1 - g.cur)):
            b.b[p] = ptr
    in_b = (ctypes.c_int * len(grids))()
    _lib.check(fn(arr, len(grids), 0, ctypes.c_void_p(stream), in_b))
    for i, g in enumerate(grids):
        if in_b[i]:
            g.cur = 1 - g.cur


def test_static_tables_match_multi_on_two_calls():
    import torch
    from cellato_amd import _lib
    L = _lib.lib()
    L.cellato_hip_multi_static.argtypes = L.cellato_hip_multi.argtypes
    stream = torch.cuda.current_stream().cuda_stream
    jobs = [("game-of-life", 4096, 512, 1), ("wire", 2048, 300, 2), ("cyclic", 640, 200, 3), ("critters", 128, 96, 4),
            ("traffic", 4160, 64, 5), ("fluid", 64, 77, 6), ("excitable", 192, 40, 7), ("brian", 8192, 128, 8)]
    T = [4, 3, 8, 5, 2, 7, 1, 6]
    steps1 = [37, 20, 16, 33, 9, 14, 5, 24]
    a, b = make(jobs), make(jobs)
    init = [g.buf[g.cur].cpu().numpy().copy() for g in a]
    print("[gpu] multi vs multi_static, call 1 (8 grids)", file=sys.stderr, flush=True)
    call(L.cellato_hip_multi, a, steps1, [0] * 8, T, stream)
    call(L.cellato_hip_multi_static, b, steps1, [0] * 8, T, stream)
    torch.cuda.synchronize()
    for i in range(8):
        assert torch.equal(a[i].buf[a[i].cur], b[i].buf[b[i].cur]), ("call 1", jobs[i])
    # the second call: fewer, smaller grids (the first four cut down to their top rows are new
    # grids; the last two continue from call 1), other depths
    keep = [6, 7]
    a2, b2 = make([(n, w, min(h, 64), s + 10) for n, w, h, s in jobs[:3]]) + [a[i] for i in keep], \
        make([(n, w, min(h, 64), s + 10) for n, w, h, s in jobs[:3]]) + [b[i] for i in keep]
    init2 = [g.buf[g.cur].cpu().numpy().copy() for g in a2[:3]]
    steps2 = [11, 6, 3, 4, 7]
    t02 = [0, 0, 0, steps1[6], steps1[7]]
    T2 = [2, 5, 4, 3, 8]
    print("[gpu] multi vs multi_static, call 2 (5 smaller grids)", file=sys.stderr, flush=True)
    call(L.cellato_hip_multi, a2, steps2, t02, T2, stream)
    call(L.cellato_hip_multi_static, b2, steps2, t02, T2, stream)
    torch.cuda.synchronize()
    for i in range(5):
        assert torch.equal(a2[i].buf[a2[i].cur], b2[i].buf[b2[i].cur]), ("call 2", i)
    # the oracle on the small new grids of call 2 and on two grids of call 1
    from cellato_amd import grid as hg
    for g, planes, st in [(b2[2], init2[2], steps2[2]), (b[3], init[3], steps1[3]), (b[5], init[5], steps1[5])]:
        start = g.rule.value_of(hg.unpack(planes.view(np.uint64), g.width, g.height, 0))
        want = O.run_standard(g.rule.name, start, st)
        got = g.rule.value_of(hg.unpack(g.download(), g.width, g.height, 0))
        assert np.array_equal(got, want), g.rule.name

"""Helpers of the GPU tests: run a grid through the C ABI and bring it back as state values."""
import sys

import numpy as np

from cellato_amd import _lib, grid as hg
from cellato_amd.planes import DeviceGrid
from cellato_amd.rules import rule as get_rule


def run_gpu(rule, values, layout, kind, T, steps, t0=0, chunk=0):
    import torch
    r = get_rule(rule)
    H, W = values.shape
    print(f"[gpu] {rule} layout={layout} {W}x{H} {kind} T={T} steps={steps} t0={t0} chunk={chunk}",
          file=sys.stderr, flush=True)
    d = DeviceGrid(r, W, H, layout)
    d.upload(hg.pack(r.index_of(values), r.planes, layout))
    stream = torch.cuda.current_stream().cuda_stream
    if kind == "simple":
        in_b = _lib.simple(r.id, layout, d.ptrs(0), d.ptrs(1), W, H, t0, steps, chunk, stream)
    else:
        in_b = _lib.temporal(r.id, layout, d.ptrs(0), d.ptrs(1), W, H, t0, steps, T, chunk, stream)
    torch.cuda.synchronize()
    d.cur = int(in_b)
    return r.value_of(hg.unpack(d.download(), W, H, layout))

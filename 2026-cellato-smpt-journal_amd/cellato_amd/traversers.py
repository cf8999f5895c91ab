"""Traversers over the HIP library, mirroring the reference's CUDA traversers
(include/traversers/cuda/simple.hpp:25-106 and temporal.hpp:24-150): `init(grid, params)`
puts the grid on the device, `run(steps, callback=None)` advances it (with a callback the
state is copied to the host after every launch, the reference's VERBOSE mode), and
`fetch_result()` returns the host grid and releases the device buffers.
"""
from dataclasses import dataclass

import numpy as np

from . import _lib
from .planes import BitArrayGrid, DeviceGrid, PlaneGrid


@dataclass
class RunParams:
    """The reference's run parameters (include/experiments/run_params.hpp:10-43) that the
    GPU traversers read, plus `chunk_rows` (rows a wave streams per task, 0 = default)."""
    automaton: str = "game-of-life"
    device: str = "HIP"
    traverser: str = "simple"
    evaluator: str = "bit_planes"
    layout: str = "bit_planes"
    x_size: int = 0
    y_size: int = 0
    steps: int = 0
    word_size: int = 64
    temporal_steps: int = 0
    temporal_tile_size_y: int = 0
    rounds: int = 1
    warmup_rounds: int = 0
    seed: int = 42
    chunk_rows: int = 0


def _stream():
    import torch
    return torch.cuda.current_stream().cuda_stream


class _Base:
    is_CUDA = True   # name kept from the reference: "runs on the GPU"

    def __init__(self):
        self._dev = None
        self._t = 0

    def init(self, grid: PlaneGrid, params: RunParams):
        if params.word_size not in (0, 64):
            raise RuntimeError("the HIP traversers use 64-bit words")
        self._params = params
        self._dev = DeviceGrid(grid.rule, grid.width, grid.height, grid.layout)
        self._dev.upload(grid.planes)
        self._t = 0

    def run(self, steps, callback=None):
        if self._dev is None:
            raise RuntimeError("run() before init()")
        if callback is None:
            self._advance(steps)
            return
        callback(0, self._dev.to_plane_grid())
        for k in range(steps):
            self._advance(1)
            callback(k + 1, self._dev.to_plane_grid())

    def fetch_result(self) -> PlaneGrid:
        import torch
        torch.cuda.synchronize()
        g = self._dev.to_plane_grid()
        self._dev = None
        return g

    def _advance(self, steps):
        d = self._dev
        in_b = self._launch(d.rule.id, d.layout, d.ptrs(d.cur), d.ptrs(1 - d.cur), d.width, d.height, self._t, steps)
        if in_b:
            d.cur = 1 - d.cur
        self._t += steps


class SimpleTraverser(_Base):
    """One generation per kernel launch (reference: traversers/cuda/simple.cu:45-87)."""

    def _launch(self, rule, layout, a, b, w, h, t0, steps):
        return _lib.simple(rule, layout, a, b, w, h, t0, steps, self._params.chunk_rows, _stream())


class TemporalTraverser(_Base):
    """`temporal_steps` generations per launch (reference: traversers/cuda/temporal.cu:139-272)."""

    def init(self, grid, params):
        if not 1 <= params.temporal_steps <= _lib.lib().cellato_max_temporal_steps():
            raise RuntimeError("Invalid temporal steps for temporal traverser.")
        super().init(grid, params)

    def _launch(self, rule, layout, a, b, w, h, t0, steps):
        return _lib.temporal(rule, layout, a, b, w, h, t0, steps, self._params.temporal_steps,
                             self._params.chunk_rows, _stream())


class BitArrayTraverser:
    """The simple traverser over the bit_array layout (reference: evaluators::bit_array under
    traversers/cuda/simple.cu): one generation per launch, every cell evaluated on its own."""
    is_CUDA = True

    def __init__(self):
        self._buf = None
        self._t = 0

    def init(self, grid: BitArrayGrid, params: RunParams):
        import torch
        if params.word_size not in (0, grid.word_bits):
            raise RuntimeError("word_size does not match the grid's word type")
        self._grid = grid
        dt = torch.int64 if grid.word_bits == 64 else torch.int32
        host = np.ascontiguousarray(grid.words).view(np.int64 if grid.word_bits == 64 else np.int32)
        a = torch.from_numpy(host.copy()).to("cuda")
        self._buf = [a, torch.zeros_like(a, dtype=dt)]
        self._cur = 0
        self._t = 0

    def run(self, steps, callback=None):
        if self._buf is None:
            raise RuntimeError("run() before init()")
        if callback is None:
            self._advance(steps)
            return
        callback(0, self._host())
        for k in range(steps):
            self._advance(1)
            callback(k + 1, self._host())

    def fetch_result(self) -> BitArrayGrid:
        import torch
        torch.cuda.synchronize()
        g = self._host()
        self._buf = None
        return g

    def _host(self):
        g = self._grid
        words = self._buf[self._cur].cpu().numpy().view(np.uint64 if g.word_bits == 64 else np.uint32).copy()
        return BitArrayGrid(g.rule, g.width, g.height, words, g.word_bits)

    def _advance(self, steps):
        g = self._grid
        in_b = _lib.bit_array(g.rule.id, g.word_bits, self._buf[self._cur].data_ptr(),
                              self._buf[1 - self._cur].data_ptr(), g.width, g.height, self._t, steps, _stream())
        if in_b:
            self._cur = 1 - self._cur
        self._t += steps

"""Bit-plane grids on the host and on the GPU.

`PlaneGrid` mirrors the reference's bit_planes::grid / tiled_bit_planes::grid
(include/memory/bit_planes_grid.hpp:30-271, tiled_bit_planes_grid.hpp:30-289): P planes of
64-bit words, built from a standard grid of state values, convertible back, with the
reference's checksum. `DeviceGrid` holds the two ping-pong buffers of a grid in HBM (one
torch allocation, planes as views), the way the reference traversers own two device grids.
"""
import numpy as np

from . import grid as hg
from .rules import rule as get_rule


class PlaneGrid:
    def __init__(self, rule, width, height, planes, layout=hg.LINEAR):
        self.rule = get_rule(rule) if not hasattr(rule, "planes") else rule
        self.width, self.height, self.layout = width, height, layout
        self.planes = planes   # (P, rows, words) uint64

    @classmethod
    def from_standard(cls, rule, values, layout=hg.LINEAR):
        r = get_rule(rule) if not hasattr(rule, "planes") else rule
        v = np.asarray(values)
        H, W = v.shape
        return cls(r, W, H, hg.pack(r.index_of(v), r.planes, layout), layout)

    def to_standard(self):
        return self.rule.value_of(hg.unpack(self.planes, self.width, self.height, self.layout))

    def x_size_physical(self):
        return self.planes.shape[2]

    def y_size_physical(self):
        return self.planes.shape[1]

    def get_checksum(self):
        return hg.checksum_str(self.to_standard())


class DeviceGrid:
    """Two device buffers (a, b) of P planes each; `cur` says which one holds the state."""

    def __init__(self, rule, width, height, layout=hg.LINEAR, device="cuda"):
        import torch
        self.rule = get_rule(rule) if not hasattr(rule, "planes") else rule
        self.width, self.height, self.layout = width, height, layout
        rows, words = hg.plane_shape(layout, width, height)
        self.shape = (self.rule.planes, rows, words)
        self.buf = torch.zeros((2,) + self.shape, dtype=torch.int64, device=device)
        self.cur = 0

    @property
    def nbytes_state(self):
        return int(np.prod(self.shape)) * 8

    def ptrs(self, which):
        return [self.buf[which, p].data_ptr() for p in range(self.shape[0])]

    def upload(self, planes):
        import torch
        self.buf[self.cur].copy_(torch.from_numpy(np.ascontiguousarray(planes).view(np.int64)))

    def download(self):
        return self.buf[self.cur].cpu().numpy().view(np.uint64).copy()

    def to_plane_grid(self):
        return PlaneGrid(self.rule, self.width, self.height, self.download(), self.layout)

    # ---- on-device conversions (cellato_hip_pack / unpack / checksum / fill_random) ----
    def _stream(self):
        import torch
        return torch.cuda.current_stream().cuda_stream

    def fill_random(self, seed, probs=None):
        """Seeded i.i.d. initial state written on the device into the current buffer."""
        import ctypes
        from . import _lib
        from .rules import iid_index_probs
        probs = iid_index_probs(self.rule.name) if probs is None else probs
        arr = (ctypes.c_double * len(probs))(*probs)
        _lib.check(_lib.lib().cellato_hip_fill_random(self.rule.id, self.layout, _lib.plane_array(self.ptrs(self.cur)),
                                                      self.width, self.height, seed, arr, len(probs), self._stream()))

    def pack_values(self, values):
        """values: torch int32 tensor (H, W) of state values on the device -> current buffer."""
        import torch
        from . import _lib
        v = values.contiguous()
        bad = torch.zeros(1, dtype=torch.int32, device=v.device)
        _lib.check(_lib.lib().cellato_hip_pack(self.rule.id, self.layout, v.data_ptr(), self.width, self.height,
                                               _lib.plane_array(self.ptrs(self.cur)), bad.data_ptr(), self._stream()))
        if int(bad.item()):
            raise ValueError(f"{self.rule.name}: value not in the state dictionary")

    def unpack_values(self):
        """Current buffer -> torch int32 tensor (H, W) of state values on the device."""
        import torch
        from . import _lib
        out = torch.empty((self.height, self.width), dtype=torch.int32, device=self.buf.device)
        _lib.check(_lib.lib().cellato_hip_unpack(self.rule.id, self.layout, _lib.plane_array(self.ptrs(self.cur)),
                                                 self.width, self.height, out.data_ptr(), self._stream()))
        return out

    def checksum_device(self):
        """The reference's 16-tile checksum, computed on the device."""
        import torch
        from . import _lib
        out = torch.zeros(16, dtype=torch.int64, device=self.buf.device)
        _lib.check(_lib.lib().cellato_hip_checksum(self.rule.id, self.layout, _lib.plane_array(self.ptrs(self.cur)),
                                                   self.width, self.height, out.data_ptr(), self._stream()))
        return [int(x) for x in out.cpu().tolist()]


class BitArrayGrid:
    """The reference's bit_array::grid (include/memory/bit_array_grid.hpp:85-321): every cell's
    dictionary index in P bits, word_bits // P cells per word (32- or 64-bit words), row-major."""

    def __init__(self, rule, width, height, words, word_bits=32):
        self.rule = get_rule(rule) if not hasattr(rule, "planes") else rule
        self.width, self.height, self.word_bits = width, height, word_bits
        self.words = words   # (H, W / cells_per_word) uint32 / uint64

    @property
    def cells_per_word(self):
        return self.word_bits // self.rule.planes

    @classmethod
    def from_standard(cls, rule, values, word_bits=32):
        r = get_rule(rule) if not hasattr(rule, "planes") else rule
        v = np.asarray(values)
        H, W = v.shape
        return cls(r, W, H, hg.pack_bit_array(r.index_of(v), r.planes, word_bits), word_bits)

    def to_standard(self):
        return self.rule.value_of(hg.unpack_bit_array(self.words, self.width, self.height, self.rule.planes,
                                                      self.word_bits))

    def x_size_physical(self):
        return self.width // self.cells_per_word

    def y_size_physical(self):
        return self.height

    def get_checksum(self):
        return hg.checksum_str(self.to_standard())

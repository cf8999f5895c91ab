"""Host-side grids: packing of dictionary indices into 64-bit bit planes (linear and tiled
layouts, as include/memory/bit_planes_grid.hpp:234-270 and tiled_bit_planes_grid.hpp:243-288
define them) and into bit_array words (include/memory/bit_array_grid.hpp:293-312), unpacking,
and the reference's 16-tile checksum (include/memory/standard_grid.hpp:179-213)."""
import numpy as np

LINEAR, TILED = 0, 1


def plane_shape(layout, width, height):
    """(rows, words per row) of one plane."""
    if layout == LINEAR:
        if width % 64:
            raise ValueError("`x_size` must be a multiple of `word store bits` (64)")
        return height, width // 64
    if width % 8 or height % 8:
        raise ValueError("tiled layout needs width and height multiples of 8")
    return height // 8, width // 8


def pack(indices, planes, layout=LINEAR):
    """indices: (H, W) dictionary indices -> (P, rows, words) uint64."""
    idx = np.ascontiguousarray(indices)
    H, W = idx.shape
    rows, words = plane_shape(layout, W, H)
    out = np.empty((planes, rows, words), np.uint64)
    for b in range(planes):
        bit = ((idx >> b) & 1).astype(np.uint8)
        if layout == LINEAR:
            cells = bit.reshape(H, words, 64)
        else:
            cells = bit.reshape(rows, 8, words, 8).transpose(0, 2, 1, 3).reshape(rows, words, 64)
        out[b] = np.packbits(cells, axis=-1, bitorder="little").view("<u8").reshape(rows, words)
    return out


def unpack(planes_arr, width, height, layout=LINEAR):
    """(P, rows, words) uint64 -> (H, W) dictionary indices (int32)."""
    P, rows, words = planes_arr.shape
    idx = np.zeros((height, width), np.int32)
    for b in range(P):
        bits = np.unpackbits(np.ascontiguousarray(planes_arr[b]).view(np.uint8).reshape(rows, words, 8),
                             axis=-1, bitorder="little")
        if layout == LINEAR:
            cells = bits.reshape(height, width)
        else:
            cells = bits.reshape(rows, words, 8, 8).transpose(0, 2, 1, 3).reshape(height, width)
        idx |= cells.astype(np.int32) << b
    return idx


def pack_bit_array(indices, planes, word_bits=64):
    """indices: (H, W) dictionary indices -> (H, W / cpw) words (uint32 or uint64), cpw =
    word_bits // planes cells per word, cell x of a row at bits [(x % cpw) * P, + P) of word
    x // cpw (rows start words: W must be a multiple of cpw)."""
    idx = np.ascontiguousarray(indices)
    H, W = idx.shape
    cpw = word_bits // planes
    if W % cpw:
        raise ValueError(f"Grid x_size must be divisible by cells_per_word. x_size={W}, cells_per_word={cpw}")
    dt = np.uint64 if word_bits == 64 else np.uint32
    cells = idx.reshape(H, W // cpw, cpw).astype(np.uint64) & np.uint64((1 << planes) - 1)
    shifts = (np.arange(cpw, dtype=np.uint64) * np.uint64(planes))
    return np.bitwise_or.reduce(cells << shifts, axis=-1).astype(dt)


def unpack_bit_array(words, width, height, planes, word_bits=64):
    """(H, W / cpw) words -> (H, W) dictionary indices (int32)."""
    cpw = word_bits // planes
    w = np.ascontiguousarray(words).astype(np.uint64).reshape(height, width // cpw, 1)
    shifts = (np.arange(cpw, dtype=np.uint64) * np.uint64(planes))
    return ((w >> shifts) & np.uint64((1 << planes) - 1)).astype(np.int32).reshape(height, width)


def checksum(values):
    """16 tile sums of state values over a 4x4 tiling of ceil(W/4) x ceil(H/4) cells."""
    v = np.asarray(values)
    H, W = v.shape
    tw, th = (W + 3) // 4, (H + 3) // 4
    out = []
    for ty in range(4):
        for tx in range(4):
            out.append(int(v[ty * th:min(ty * th + th, H), tx * tw:min(tx * tw + tw, W)].sum(dtype=np.int64)))
    return out


def checksum_str(values):
    return "-".join(str(c) for c in checksum(values))

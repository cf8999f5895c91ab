"""Bit order of the bit_array layout (include/memory/bit_array_grid.hpp:293-312 of the reference:
cell i of the row-major grid in word i / cpw at bits [(i % cpw) * P, + P), cpw = word bits / P),
pinned by words written out by hand for the reference's own grid tests
(include/tests/bit_array_grid.hpp:63-95, 134-176: a 3-state dictionary, 2 bits per cell,
32-bit words of 16 cells), and the product's packer (cellato_amd.grid) against the oracle's
restatement on random grids of every plane count and word size."""
import numpy as np
import pytest

import oracle_lib as O
from cellato_amd import grid as hg

DEAD, ALIVE, DYING = 0, 1, 2   # dictionary indices of the reference test's TestCellState


def test_cell_access_pattern():
    # bit_array_grid.hpp:72-80: 5 x 16 grid, (0,0) ALIVE, (2,0) DYING, (3,1) ALIVE
    W, H = 16, 5
    idx = np.zeros((H, W), np.int32)
    idx[0, 0], idx[0, 2], idx[1, 3] = ALIVE, DYING, ALIVE
    want = np.zeros(H, np.uint64)
    want[0] = (ALIVE << 0) | (DYING << 4)
    want[1] = ALIVE << 6
    for words in (hg.pack_bit_array(idx, 2, 32).reshape(-1), O.pack_bit_array(idx, 2, 32)):
        assert [int(x) for x in words] == [int(x) for x in want]
    assert hg.pack_bit_array(idx, 2, 32).shape == (H, 1)   # x_size_physical = 16 / 16


def test_simple_pattern_words():
    # bit_array_grid.hpp:140-158: the first eight cells of two rows
    W, H = 16, 2
    idx = np.zeros((H, W), np.int32)
    idx[0, :8] = [DEAD, ALIVE, DYING, DEAD, ALIVE, DEAD, ALIVE, DYING]
    idx[1, :8] = [DYING, ALIVE, DEAD, DEAD, ALIVE, ALIVE, DYING, DEAD]
    w0 = sum(int(v) << (2 * i) for i, v in enumerate(idx[0, :8]))
    w1 = sum(int(v) << (2 * i) for i, v in enumerate(idx[1, :8]))
    assert (w0, w1) == (0b10_01_00_01_00_10_01_00, 0b00_10_01_01_00_00_01_10)
    words = hg.pack_bit_array(idx, 2, 32)
    assert [int(x) for x in words.reshape(-1)] == [w0, w1]
    assert np.array_equal(hg.unpack_bit_array(words, W, H, 2, 32), idx)


def test_width_must_fill_words():
    with pytest.raises(ValueError, match="cells_per_word"):
        hg.pack_bit_array(np.zeros((2, 20), np.int32), 2, 32)


@pytest.mark.parametrize("P", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("word_bits", [32, 64])
def test_product_packer_matches_oracle(P, word_bits):
    rng = np.random.default_rng(P * 100 + word_bits)
    cpw = word_bits // P
    for k, h in [(1, 3), (5, 7)]:
        idx = rng.integers(0, 1 << P, size=(h, cpw * k)).astype(np.int32)
        a = hg.pack_bit_array(idx, P, word_bits)
        assert a.dtype == (np.uint64 if word_bits == 64 else np.uint32)
        assert np.array_equal(a.reshape(-1).astype(np.uint64), O.pack_bit_array(idx, P, word_bits))
        assert np.array_equal(hg.unpack_bit_array(a, cpw * k, h, P, word_bits), idx)
        assert np.array_equal(O.unpack_bit_array(O.pack_bit_array(idx, P, word_bits), cpw * k, h, P, word_bits), idx)
        # every word holds cpw cells; with 3 or 5 planes the top bits of a word stay 0
        if (word_bits % P) != 0:
            top = np.uint64(((1 << word_bits) - 1) ^ ((1 << (cpw * P)) - 1))
            assert not np.any(a.astype(np.uint64) & top)

"""The C-ABI library loads, exports every symbol include/cellato_hip.h declares, reports
the same rule registry as the oracle, and rejects invalid arguments before touching the GPU."""
import ctypes
import re
from pathlib import Path

import pytest

import oracle_lib as O
from cellato_amd import _lib

HEADER = Path(__file__).resolve().parent.parent / "include" / "cellato_hip.h"


def declared():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(cellato_\w+)\s*\(", text)))


def test_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 8
    for n in names:
        assert hasattr(L, n), n


def test_registry_matches_oracle():
    L = _lib.lib()
    assert L.cellato_rule_count() == O.lib().oracle_rule_count() == 10
    for name, i in O.rule_ids().items():
        assert L.cellato_rule_name(i).decode() == name
        assert L.cellato_rule_find(name.encode()) == i
        assert L.cellato_rule_planes(i) == O.planes(name)
        buf = (ctypes.c_int32 * 32)()
        n = L.cellato_rule_states(i, buf, 32)
        assert list(buf[:n]) == list(O.states(name))
    assert L.cellato_rule_find(b"no-such-rule") == _lib.E_RULE
    assert L.cellato_rule_name(10) is None
    assert L.cellato_max_temporal_steps() >= 8


FAKE = [0x1000 * (i + 1) for i in range(5)]   # never dereferenced: every case fails validation


@pytest.mark.parametrize("args,code", [
    (dict(rule=11), _lib.E_RULE),
    (dict(layout=7), _lib.E_LAYOUT),
    (dict(width=100), _lib.E_GEOMETRY),
    (dict(width=0), _lib.E_GEOMETRY),
    (dict(layout=1, width=64, height=12), _lib.E_GEOMETRY),
    (dict(steps=-1), _lib.E_ARG),
    (dict(a=[0]), _lib.E_ARG),
])
def test_invalid_arguments_are_rejected(args, code):
    p = dict(rule=0, layout=0, a=FAKE, b=FAKE, width=128, height=16, steps=3)
    p.update(args)
    with pytest.raises(_lib.CellatoError) as e:
        _lib.simple(p["rule"], p["layout"], p["a"], p["b"], p["width"], p["height"], 0, p["steps"])
    assert e.value.code == code
    assert _lib.lib().cellato_hip_last_error().decode()


def test_temporal_depth_is_checked():
    for T in (0, 9, -2):
        with pytest.raises(_lib.CellatoError) as e:
            _lib.temporal(0, 0, FAKE, FAKE, 128, 16, 0, 4, T)
        assert e.value.code == _lib.E_TEMPORAL


def test_zero_steps_launch_nothing():
    assert _lib.simple(0, 0, FAKE, FAKE, 128, 16, 0, 0) is False
    assert _lib.temporal(9, 1, FAKE, FAKE, 64, 16, 0, 0, 4) is False


@pytest.mark.parametrize("T", [0, 1, 2, 4, 8])
def test_plan_fills_every_simd_in_whole_unrolled_blocks(T):
    # automatic rows per task at 16384^2: every task walks whole unrolled blocks of rows, the
    # tasks fill 2 or 3 blocks of 4 waves per CU (256 CUs), and with halo-lane segments the
    # 8 columns left over after four 62-column segments go to groups of 10 lanes, 6 per wave
    import math
    chunk, tasks, blocks = _lib.plan(0, 0, 16384, 16384, T)
    U = 6 if T <= 2 else 3
    steps = chunk + 2 * max(T, 1)
    assert steps % U == 0
    assert blocks == math.ceil(tasks / 4)
    assert math.ceil(blocks / 256) in (2, 3)
    if T >= 1:
        assert tasks == 4 * math.ceil(16384 / chunk) + math.ceil(16384 / (6 * chunk))
    else:
        assert tasks == 4 * math.ceil(16384 / chunk)
    # an explicit chunk is taken as given
    assert _lib.plan(0, 0, 16384, 16384, T, 7)[0] == 7


def test_plan_tiled_temporal_passes_as_rows_or_blocks():
    # Game of Life at depth 4 takes the tiled-as-rows kernels: 16384 cells are 256 columns of
    # 64 cells, four 62-column halo-lane segments and 8 columns in groups of 10 lanes (6 per
    # wave), each task a whole number of block rows (chunks reported in block rows)
    import math
    chunk, tasks, blocks = _lib.plan(0, 1, 16384, 16384, 4)
    assert tasks == 4 * math.ceil(16384 / (8 * chunk)) + math.ceil(16384 / (6 * 8 * chunk))
    assert blocks == math.ceil(tasks / 4)
    assert _lib.plan(0, 1, 16384, 16384, 4, 5)[0] == 5
    # below its crossover depth, for WireWorld (no such kernels) and for rows of fewer than 64
    # columns of 64 cells: 8x8 blocks, 2048 block columns = 33 segments + 2 columns in groups
    # of 4 lanes (16 per wave)
    for rule, T, width in [(0, 2, 16384), (1, 4, 16384), (0, 4, 2048)]:
        chunk, tasks, _ = _lib.plan(rule, 1, width, 16384, T)
        wx = width // 8
        nseg, left = wx // 62, wx % 62
        want = nseg * math.ceil(2048 / chunk)
        if left:
            want += math.ceil(2048 / ((64 // (left + 2)) * chunk))
        assert tasks == want, (rule, T, width)


def test_plan_rejects_what_the_traversers_reject():
    for args, code in [((11, 0, 128, 16, 0), _lib.E_RULE), ((0, 3, 128, 16, 0), _lib.E_LAYOUT),
                       ((0, 0, 100, 16, 0), _lib.E_GEOMETRY), ((0, 0, 128, 16, 9), _lib.E_TEMPORAL)]:
        with pytest.raises(_lib.CellatoError) as e:
            _lib.plan(*args)
        assert e.value.code == code


def test_multi_static_refuses_tables_beyond_its_scratch():
    """The allocation-free batch call checks its table size on the host before touching the
    GPU: 4M one-row tasks (32 MiB of task table) exceed its 8 MiB static array."""
    L = _lib.lib()
    arr = (_lib.BatchGrid * 1)()
    g = arr[0]
    g.rule, g.layout, g.width, g.height, g.steps, g.t0, g.temporal_steps, g.chunk_rows = 0, 0, 4096, 1 << 22, 1, 0, 1, 1
    g.a[0], g.b[0] = FAKE[0], FAKE[1]
    in_b = (ctypes.c_int * 1)()
    L.cellato_hip_multi_static.argtypes = L.cellato_hip_multi.argtypes
    rc = L.cellato_hip_multi_static(arr, 1, 0, None, in_b)
    assert rc == _lib.E_ARG
    assert b"cellato_hip_multi_static's scratch" in L.cellato_hip_last_error()


def test_bit_array_rejects_widths_that_do_not_fill_words():
    """cellato_hip_bit_array checks its geometry on the host: a row must be whole words, words
    are 32 or 64 bits."""
    with pytest.raises(_lib.CellatoError) as e:
        _lib.bit_array(3, 64, FAKE[0], FAKE[1], 16384, 8, 0, 1)   # 3 planes: 21 cells per 64-bit word
    assert e.value.code == _lib.E_GEOMETRY and "cells_per_word" in str(e.value)
    with pytest.raises(_lib.CellatoError) as e:
        _lib.bit_array(0, 16, FAKE[0], FAKE[1], 64, 8, 0, 1)
    assert e.value.code == _lib.E_ARG
    with pytest.raises(_lib.CellatoError) as e:
        _lib.bit_array(11, 64, FAKE[0], FAKE[1], 64, 8, 0, 1)
    assert e.value.code == _lib.E_RULE

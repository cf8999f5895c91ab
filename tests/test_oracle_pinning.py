"""Pins the CPU oracle before it is trusted as the checker.

- the reference's published 16-tile checksums (results/*/_h100.csv, _a100.csv,
  grid-search-results.csv, extracted to tests/golden/reference_checksums.json), reproduced
  from the reference's own initial states by both oracle evaluators;
- the two oracle evaluators (standard and bit_planes) agree on random grids of every rule,
  including the 1024^2 Game-of-Life equality of the first benchmark configuration;
- textbook formulations of the rules (numpy, whole-grid) and conservation laws.
"""
import json
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

GOLD = json.loads((Path(__file__).parent / "golden" / "reference_checksums.json").read_text())
RULES = ["game-of-life", "wire", "forest-fire", "excitable", "brian", "maze", "traffic", "fluid", "critters", "cyclic"]


def golden(rule, w, h, steps, src="_h100"):
    for e in GOLD:
        if e["automaton"] == rule and (e["x_size"], e["y_size"], e["steps"]) == (w, h, steps) and src in e["source"]:
            return e
    raise KeyError((rule, w, h, steps))


@pytest.mark.parametrize("rule", RULES)
def test_bitplanes_oracle_reproduces_published_checksum(rule):
    g = golden(rule, 16384, 16384, 128)
    init = O.init_reference(rule, 16384, 16384, g["seed"])
    out = O.run_bitplanes(rule, init, 128)
    assert O.checksum(out) == g["checksum"], g["source"]


@pytest.mark.slow
def test_standard_oracle_reproduces_published_checksum():
    g = golden("game-of-life", 3840, 4292, 1020, "grid-search")
    init = O.init_reference("game-of-life", 3840, 4292, g["seed"])
    assert O.checksum(O.run_standard("game-of-life", init, 1020)) == g["checksum"], g["source"]


def test_golden_fixture_groups_are_consistent():
    # the same (automaton, size, steps) group carries one checksum in every file
    seen = {}
    for e in GOLD:
        k = (e["automaton"], e["x_size"], e["y_size"], e["steps"], e["seed"])
        assert seen.setdefault(k, e["checksum"]) == e["checksum"]
    assert len(GOLD) > 1000


def iid(rule, w, h, seed):
    rng = np.random.default_rng(seed)
    st = O.states(rule)
    return st[rng.integers(0, len(st), size=(h, w))].astype(np.int32)


@pytest.mark.parametrize("rule", RULES)
@pytest.mark.parametrize("t0", [0, 1])
def test_standard_equals_bitplanes(rule, t0):
    v = iid(rule, 320, 97, 7)
    a = O.run_standard(rule, v, 37, t0=t0)
    b = O.run_bitplanes(rule, v, 37, t0=t0)
    np.testing.assert_array_equal(a, b)


def test_first_benchmark_configuration_gol_1024():
    rng = np.random.default_rng(2026)
    v = (rng.random((1024, 1024)) < 0.5).astype(np.int32)
    np.testing.assert_array_equal(O.run_standard("game-of-life", v, 128), O.run_bitplanes("game-of-life", v, 128))


def test_bitplanes_u8_words_match_u64_words():
    # the reference's unit tests run the bit_planes evaluator on uint8_t words
    for rule in ["game-of-life", "critters", "wire", "fluid"]:
        v = iid(rule, 128, 40, 3)
        P = O.planes(rule)
        idx = O.to_index(rule, v)
        u8 = np.zeros((P, 40, 16), np.uint8)
        for b in range(P):
            u8[b] = np.packbits(((idx >> b) & 1).astype(np.uint8).reshape(40, 16, 8), axis=-1, bitorder="little")[..., 0]
        assert O.lib().oracle_run_bitplanes_u8(O.rule_ids()[rule], u8.reshape(-1), 128, 40, 0, 9) == 0
        out = np.zeros((40, 128), np.int32)
        for b in range(P):
            out |= np.unpackbits(u8[b][..., None], axis=-1, bitorder="little").reshape(40, 128).astype(np.int32) << b
        np.testing.assert_array_equal(O.states(rule)[out], O.run_standard(rule, v, 9))


# ---- textbook rules on whole grids ---------------------------------------------------------
def roll(a, dx, dy):
    """value of the cell at (x + dx, y + dy) on the torus"""
    return np.roll(np.roll(a, -dy, axis=0), -dx, axis=1)


MOORE = [(dx, dy) for dy in (-1, 0, 1) for dx in (-1, 0, 1) if (dx, dy) != (0, 0)]


def textbook(rule, a, t):
    if rule in ("game-of-life", "maze", "brian", "wire", "excitable", "cyclic"):
        def count(s):
            return sum((roll(a, dx, dy) == s).astype(int) for dx, dy in MOORE)
    if rule == "game-of-life":           # B3/S23
        n = count(1)
        return np.where((n == 3) | ((a == 1) & (n == 2)), 1, 0)
    if rule == "maze":                   # B3/S12345
        n = count(1)
        return np.where((n == 3) | ((a == 1) & (n < 6)), 1, 0)
    if rule == "brian":                  # dead(0) -> alive(2) on exactly 2 alive; alive -> dying(1) -> dead
        n = count(2)
        return np.where(a == 0, np.where(n == 2, 2, 0), np.where(a == 2, 1, 0))
    if rule == "wire":                   # head(1) -> tail(2) -> conductor(3) -> head on 1-2 heads
        n = count(1)
        return np.select([a == 1, a == 2, a == 3], [2, 3, np.where((n == 1) | (n == 2), 1, 3)], 0)
    if rule == "excitable":              # 0 -> 1 on any excited neighbour; k -> k+1; 7 -> 0
        return np.where(a == 0, np.where(count(1) > 0, 1, 0), (a + 1) % 8)
    if rule == "cyclic":
        succ = (a + 1) % 32
        hit = sum((roll(a, dx, dy) == succ).astype(int) for dx, dy in MOORE) > 0
        return np.where(hit, succ, a)
    if rule == "forest-fire":            # values: empty 0, tree 1, ash 2, fire 3
        burning = sum((roll(a, dx, dy) == 3).astype(int) for dx, dy in [(0, 1), (0, -1), (1, 0), (-1, 0)]) > 0
        return np.select([a == 3, a == 2, a == 1], [2, np.where(burning, 2, 0), np.where(burning, 3, 1)], 0)
    if rule == "traffic":                # even t: red(1) moves right; odd t: blue(2) moves down
        car, fwd, back = (1, (1, 0), (-1, 0)) if t % 2 == 0 else (2, (0, 1), (0, -1))
        ahead, behind = roll(a, *fwd), roll(a, *back)
        return np.where(a == car, np.where(ahead == 0, 0, car), np.where(a == 0, np.where(behind == car, car, 0), a))
    if rule == "fluid":                  # the reference's HPP bit convention (src/fluid/algorithm.hpp)
        up, down, left, right = roll(a, 0, -1), roll(a, 0, 1), roll(a, -1, 0), roll(a, 1, 0)
        vin = (up & 1) | (down & 2)
        hin = (left & 4) | (right & 8)
        vcol = ((up & 1) > 0) & ((down & 2) > 0) & (hin == 0)
        hcol = ((left & 4) > 0) & ((right & 8) > 0) & (vin == 0)
        return np.where(vcol, 12, vin) | np.where(hcol, 3, hin)
    if rule == "critters":               # Margolus 2x2 blocks at even (t=0) / odd offset
        o = t % 2
        b = np.roll(np.roll(a, -o, 0), -o, 1)
        H, W = b.shape
        blk = b.reshape(H // 2, 2, W // 2, 2)
        n = blk.sum(axis=(1, 3), keepdims=True)
        rot = blk[:, ::-1, :, ::-1]
        new = np.where(n == 3, 1 - rot, np.where(n == 2, blk, 1 - blk)).reshape(H, W)
        return np.roll(np.roll(new, o, 0), o, 1)
    raise KeyError(rule)


@pytest.mark.parametrize("rule", RULES)
def test_oracle_matches_textbook_rule(rule):
    v = iid(rule, 64, 48, 11)
    for t in range(4):
        nxt = O.run_standard(rule, v, 1, t0=t)
        np.testing.assert_array_equal(nxt, textbook(rule, v, t))
        v = nxt


def test_conservation_laws():
    v = iid("fluid", 128, 64, 5)
    bits = lambda a: sum(((a >> b) & 1).sum() for b in range(4))
    assert bits(O.run_standard("fluid", v, 25)) == bits(v)           # HPP conserves particles
    v = iid("traffic", 128, 64, 5)
    out = O.run_standard("traffic", v, 25)
    assert (out == 1).sum() == (v == 1).sum() and (out == 2).sum() == (v == 2).sum()
    v = iid("wire", 128, 64, 5)
    out = O.run_standard("wire", v, 25)                              # WireWorld keeps empty cells
    assert ((out == 0) == (v == 0)).all()


def test_known_patterns():
    z = np.zeros((16, 16), np.int32)
    blinker = z.copy(); blinker[5, 4:7] = 1
    assert not (O.run_standard("game-of-life", blinker, 1) == blinker).all()
    np.testing.assert_array_equal(O.run_standard("game-of-life", blinker, 2), blinker)
    glider = z.copy()
    for y, x in [(0, 1), (1, 2), (2, 0), (2, 1), (2, 2)]:
        glider[y, x] = 1
    np.testing.assert_array_equal(O.run_standard("game-of-life", glider, 4), np.roll(glider, (1, 1), (0, 1)))
    # Critters: an empty block has count 0 and toggles, so an all-dead grid turns all-alive
    np.testing.assert_array_equal(O.run_standard("critters", z, 1), 1 - z)


def test_pack_layouts_match_package(oracle):
    import cellato_amd.grid as hg
    rng = np.random.default_rng(1)
    idx = rng.integers(0, 8, size=(24, 192)).astype(np.int32)
    for layout, fpack in [(hg.LINEAR, O.lib().oracle_pack_linear), (hg.TILED, O.lib().oracle_pack_tiled)]:
        ref = np.zeros(3 * 24 * 192 // (64 if layout == hg.LINEAR else 64), np.uint64)
        fpack(np.ascontiguousarray(idx).reshape(-1), 192, 24, 3, ref)
        mine = hg.pack(idx, 3, layout)
        np.testing.assert_array_equal(mine.reshape(-1), ref)
        np.testing.assert_array_equal(hg.unpack(mine, 192, 24, layout), idx)
    # bit positions: linear cell x -> bit x % 64; tiled (x, y) -> bit (y % 8) * 8 + x % 8
    one = np.zeros((8, 128), np.int32); one[3, 70] = 1
    assert hg.pack(one, 1, hg.LINEAR)[0, 3, 1] == np.uint64(1 << 6)
    assert hg.pack(one, 1, hg.TILED)[0, 0, 8] == np.uint64(1 << (3 * 8 + 6))

"""The bit_array layout on the GPU (cellato_hip_bit_array): every cell packed as its P-bit
dictionary index, word_bits / P cells per word, one generation per launch.

- all ten automata, 32- and 64-bit words, against the standard evaluator on random grids
  (rows of one word, so that a word is its own left and right neighbour on the torus; one row;
  several words; odd start steps for the rules that read the step's parity);
- the reference's published checksums of its bit_array runs (results/H100/_h100.csv, reference
  initial states, seed 42, 128 steps), for every automaton and both word sizes, at the sizes the
  reference ran (16384^2, and 16320x16448 / 16128x16644 where 16384 is not a multiple of the
  cells per word).
"""
import json
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O

pytestmark = pytest.mark.gpu

RULES = ["game-of-life", "wire", "forest-fire", "excitable", "brian", "maze", "traffic", "fluid", "critters", "cyclic"]
GOLD = json.loads((Path(__file__).parent / "golden" / "reference_checksums.json").read_text())


def run_bit_array(rule, values, word_bits, steps, t0=0):
    """values (H, W) state values -> state values after `steps` generations on the GPU."""
    import torch
    from cellato_amd import _lib
    from cellato_amd.rules import rule as get_rule
    r = get_rule(rule)
    H, W = values.shape
    print(f"[gpu] bit_array {rule} {W}x{H} word_bits={word_bits} steps={steps} t0={t0}", file=sys.stderr, flush=True)
    words = O.pack_bit_array(O.to_index(rule, values), r.planes, word_bits)
    dt = torch.int64 if word_bits == 64 else torch.int32
    host = torch.from_numpy(words.view(np.int64) if word_bits == 64 else words.astype(np.uint32).view(np.int32))
    a = host.to("cuda")
    b = torch.zeros_like(a)
    assert a.dtype == dt
    in_b = _lib.bit_array(r.id, word_bits, a.data_ptr(), b.data_ptr(), W, H, t0, steps,
                          torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    out = (b if in_b else a).cpu().numpy()
    out = out.view(np.uint64) if word_bits == 64 else out.view(np.uint32).astype(np.uint64)
    return O.states(rule)[O.unpack_bit_array(out, W, H, r.planes, word_bits)]


def iid(rule, w, h, seed):
    rng = np.random.default_rng(seed)
    st = O.states(rule)
    return st[rng.integers(0, len(st), size=(h, w))].astype(np.int32)


@pytest.mark.parametrize("word_bits", [32, 64])
@pytest.mark.parametrize("rule", RULES)
def test_matches_standard_evaluator(rule, word_bits):
    cpw = word_bits // O.planes(rule)
    for si, (k, h) in enumerate([(1, 7), (3, 37), (9, 1), (2, 2), (40, 13)]):
        v = iid(rule, cpw * k, h, 300 + si)
        for t0 in (0, 1):
            want = O.run_standard(rule, v, 9, t0=t0)
            assert np.array_equal(run_bit_array(rule, v, word_bits, 9, t0), want), (rule, word_bits, cpw * k, h, t0)


@pytest.mark.parametrize("rule", RULES)
def test_published_checksums(rule):
    P = O.planes(rule)
    done = 0
    inits = {}
    for word_bits in (32, 64):
        cpw = word_bits // P
        sizes = [(16384, 16384), (16320, 16448), (16128, 16644)]
        W, H = next((w, h) for w, h in sizes if w % cpw == 0 and any(
            e["automaton"] == rule and (e["x_size"], e["y_size"], e["steps"]) == (w, h, 128) for e in GOLD))
        g = next(e for e in GOLD if e["automaton"] == rule and (e["x_size"], e["y_size"], e["steps"]) == (W, H, 128))
        if (W, H) not in inits:
            inits[(W, H)] = O.init_reference(rule, W, H, g["seed"])
        out = run_bit_array(rule, inits[(W, H)], word_bits, 128)
        assert O.checksum(out) == g["checksum"], (g["source"], word_bits)
        done += 1
    assert done == 2

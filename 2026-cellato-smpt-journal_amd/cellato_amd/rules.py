"""Registry of the ten automata as the HIP library reports them (name, planes, states),
plus the seeded i.i.d. initial-state distributions used by the benchmarks.

Names and dictionary order follow the reference's src/<automaton>/config.hpp.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib


TRAITS = {"moore": 1, "von_neumann": 2, "margolus": 4, "neighbor": 8, "dynamic_target": 16,
          "arithmetic": 32, "alternating": 64}


@dataclass(frozen=True)
class Rule:
    id: int
    name: str
    planes: int
    states: tuple  # enum value of each dictionary index
    traits: int = 0  # rule_trait flags derived from the rule's AST at compile time

    def has(self, trait):
        return bool(self.traits & TRAITS[trait])

    def index_of(self, values):
        """Dictionary indices of an array of state values (enum values)."""
        lut = np.full(max(self.states) + 1, -1, np.int32)
        lut[list(self.states)] = np.arange(len(self.states), dtype=np.int32)
        idx = lut[np.asarray(values)]
        if (idx < 0).any():
            raise ValueError(f"{self.name}: value not in the state dictionary")
        return idx

    def value_of(self, indices):
        return np.asarray(self.states, np.int32)[np.asarray(indices)]


def _load():
    L = _lib.lib()
    out = []
    for r in range(L.cellato_rule_count()):
        buf = (ctypes.c_int32 * 32)()
        n = L.cellato_rule_states(r, buf, 32)
        out.append(Rule(r, L.cellato_rule_name(r).decode(), L.cellato_rule_planes(r), tuple(buf[:n]),
                        L.cellato_rule_traits(r)))
    return out


_RULES = None


def all_rules():
    global _RULES
    if _RULES is None:
        _RULES = _load()
    return _RULES


def rule(name_or_id):
    for r in all_rules():
        if r.name == name_or_id or r.id == name_or_id:
            return r
    raise KeyError(f"unknown automaton {name_or_id!r}")


# Seeded i.i.d. initial states: (state value, probability) per automaton. Densities follow
# the benchmark configurations; forest fire uses the reference's own 20/79/1/0
# (src/fire/data_init.hpp:16-21). Fluid and cyclic are drawn per bit / uniformly below.
IID = {
    "game-of-life": [(1, 0.5), (0, 0.5)],
    "critters": [(1, 0.5), (0, 0.5)],
    "wire": [(0, 0.80), (3, 0.15), (1, 0.03), (2, 0.02)],          # empty, conductor, head, tail
    "excitable": [(0, 0.90), (1, 0.05)] + [(v, 0.05 / 6) for v in range(2, 8)],
    "brian": [(0, 0.80), (2, 0.10), (1, 0.10)],                     # dead, alive(on), dying
    "forest-fire": [(0, 0.20), (1, 0.79), (3, 0.01), (2, 0.00)],    # empty, tree, fire(=3), ash(=2)
    "traffic": [(0, 0.6), (1, 0.2), (2, 0.2)],
    "maze": [(1, 0.1), (0, 0.9)],
}


def iid_values(name, width, height, seed):
    """Seeded i.i.d. grid of state values (int32, shape (height, width))."""
    rng = np.random.Generator(np.random.PCG64(seed))
    if name == "fluid":       # four independent direction bits, each Bernoulli(0.25)
        bits = rng.random((4, height, width)) < 0.25
        return (bits[0] | (bits[1] << 1) | (bits[2] << 2) | (bits[3] << 3)).astype(np.int32)
    if name == "cyclic":      # uniform over the 32 states
        return rng.integers(0, 32, size=(height, width), dtype=np.int32)
    vals, probs = zip(*IID[name])
    cum = np.cumsum(probs)
    u = rng.random((height, width))
    idx = np.searchsorted(cum, u, side="right").clip(max=len(vals) - 1)
    return np.asarray(vals, np.int32)[idx]


def iid_index_probs(name):
    """The i.i.d. distribution of `iid_values` as probabilities per dictionary index (for the
    device generator `DeviceGrid.fill_random`)."""
    r = rule(name)
    probs = [0.0] * len(r.states)
    if name == "fluid":
        for v in range(16):
            k = bin(v).count("1")
            probs[r.states.index(v)] = 0.25 ** k * 0.75 ** (4 - k)
    elif name == "cyclic":
        probs = [1.0 / 32] * 32
    else:
        for v, p in IID[name]:
            probs[r.states.index(v)] += p
    return probs


def mix64(z):
    """splitmix64 finaliser on numpy uint64 arrays (the device generator's hash)."""
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = z + np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def device_random_values(name, width, height, seed, probs=None):
    """Host restatement of the device generator (cellato_hip_fill_random): the state values
    it writes for this rule, size and seed."""
    r = rule(name)
    probs = iid_index_probs(name) if probs is None else probs
    total = sum(probs)
    cum, acc = [], 0.0
    for p in probs:
        acc += p / total
        t = acc * 4294967296.0
        cum.append(0xFFFFFFFF if t >= 4294967295.0 else int(t))
    cum[-1] = 0xFFFFFFFF
    cell = np.arange(width * height, dtype=np.uint64)
    u = (mix64(np.uint64(seed) ^ mix64(cell)) >> np.uint64(32)).astype(np.uint64)
    idx = np.searchsorted(np.asarray(cum[:-1], np.uint64), u, side="right")
    return np.asarray(r.states, np.int32)[idx].reshape(height, width)

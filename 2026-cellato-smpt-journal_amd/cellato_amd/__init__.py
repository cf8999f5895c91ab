"""MI355X (gfx950) bit-plane cellular-automaton step: Python host side over the C ABI of
include/cellato_hip.h (libcellato_hip.so)."""
from . import _lib, grid, rules
from ._lib import CellatoError, LAYOUT_LINEAR, LAYOUT_TILED
from .rules import Rule, all_rules, rule, iid_values
from .planes import BitArrayGrid, PlaneGrid, DeviceGrid
from .traversers import BitArrayTraverser, RunParams, SimpleTraverser, TemporalTraverser

__all__ = ["CellatoError", "LAYOUT_LINEAR", "LAYOUT_TILED", "Rule", "all_rules", "rule", "iid_values",
           "BitArrayGrid", "PlaneGrid", "DeviceGrid", "BitArrayTraverser", "RunParams", "SimpleTraverser",
           "TemporalTraverser", "grid", "rules"]

"""ctypes access to oracle/liboracle.so, the CPU restatement of the reference used as the
checker by the tests (never by the product)."""
import ctypes
import subprocess
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORACLE_DIR = REPO / "oracle"
LIB = ORACLE_DIR / "liboracle.so"

_I32 = np.ctypeslib.ndpointer(np.int32, flags="C")
_U64 = np.ctypeslib.ndpointer(np.uint64, flags="C")
_U8 = np.ctypeslib.ndpointer(np.uint8, flags="C")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
        L = ctypes.CDLL(str(LIB))
        L.oracle_rule_name.restype = ctypes.c_char_p
        L.oracle_rule_states.argtypes = [ctypes.c_int, _I32]
        L.oracle_init_reference.argtypes = [ctypes.c_int] * 4 + [_I32]
        L.oracle_run_standard.argtypes = [ctypes.c_int, _I32] + [ctypes.c_int] * 5
        L.oracle_run_bitplanes.argtypes = [ctypes.c_int, _U64] + [ctypes.c_int] * 5
        L.oracle_run_bitplanes_u8.argtypes = [ctypes.c_int, _U8] + [ctypes.c_int] * 4
        L.oracle_checksum.argtypes = [_I32, ctypes.c_int, ctypes.c_int, _U64]
        for f in ("oracle_pack_linear", "oracle_pack_tiled"):
            getattr(L, f).argtypes = [_I32, ctypes.c_int, ctypes.c_int, ctypes.c_int, _U64]
        for f in ("oracle_unpack_linear", "oracle_unpack_tiled"):
            getattr(L, f).argtypes = [_U64, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I32]
        L.oracle_pack_bit_array.argtypes = [_I32] + [ctypes.c_int] * 4 + [_U64]
        L.oracle_unpack_bit_array.argtypes = [_U64] + [ctypes.c_int] * 4 + [_I32]
        _lib = L
    return _lib


def rule_ids():
    return {lib().oracle_rule_name(i).decode(): i for i in range(lib().oracle_rule_count())}


def planes(rule):
    return lib().oracle_rule_planes(rule_ids()[rule])


def states(rule):
    buf = np.zeros(32, np.int32)
    n = lib().oracle_rule_states(rule_ids()[rule], buf)
    return buf[:n].copy()


def to_index(rule, values):
    st = states(rule)
    lut = np.full(int(st.max()) + 1, -1, np.int32)
    lut[st] = np.arange(len(st))
    return lut[values]


def init_reference(rule, width, height, seed=42):
    cells = np.zeros(width * height, np.int32)
    assert lib().oracle_init_reference(rule_ids()[rule], width, height, seed, cells) == 0
    return cells.reshape(height, width)


def run_standard(rule, values, steps, t0=0, threads=8):
    """Standard evaluator on a grid of state values; returns the new grid."""
    v = np.ascontiguousarray(values, np.int32).copy()
    H, W = v.shape
    assert lib().oracle_run_standard(rule_ids()[rule], v.reshape(-1), W, H, t0, steps, threads) == 0
    return v


def run_bitplanes(rule, values, steps, t0=0, threads=8):
    """The reference's bit_planes evaluator on the CPU (linear layout, 64-bit words)."""
    v = np.ascontiguousarray(values, np.int32)
    H, W = v.shape
    P = planes(rule)
    idx = np.ascontiguousarray(to_index(rule, v), np.int32)
    pl = np.zeros(P * H * (W // 64), np.uint64)
    lib().oracle_pack_linear(idx.reshape(-1), W, H, P, pl)
    assert lib().oracle_run_bitplanes(rule_ids()[rule], pl, W, H, t0, steps, threads) == 0
    out = np.zeros(W * H, np.int32)
    lib().oracle_unpack_linear(pl, W, H, P, out)
    return states(rule)[out].reshape(H, W)


def checksum(values):
    v = np.ascontiguousarray(values, np.int32)
    H, W = v.shape
    out = np.zeros(16, np.uint64)
    lib().oracle_checksum(v.reshape(-1), W, H, out)
    return [int(x) for x in out]


def pack_bit_array(indices, P, word_bits):
    """The bit_array words of a grid of dictionary indices (one word per uint64 slot)."""
    idx = np.ascontiguousarray(indices, np.int32)
    H, W = idx.shape
    cpw = word_bits // P
    out = np.zeros((W * H + cpw - 1) // cpw, np.uint64)
    lib().oracle_pack_bit_array(idx.reshape(-1), W, H, P, word_bits, out)
    return out


def unpack_bit_array(words, W, H, P, word_bits):
    out = np.zeros(W * H, np.int32)
    lib().oracle_unpack_bit_array(np.ascontiguousarray(words, np.uint64), W, H, P, word_bits, out)
    return out.reshape(H, W)

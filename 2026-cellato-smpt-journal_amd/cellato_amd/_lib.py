"""ctypes binding of the C ABI in include/cellato_hip.h (libcellato_hip.so, built in-tree).

The library is the only compute path: if it is missing this module raises instead of
falling back to anything else.
"""
import ctypes
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent.parent
LIB_PATH = PKG_DIR / "libcellato_hip.so"

OK = 0
E_RULE, E_LAYOUT, E_GEOMETRY, E_TEMPORAL, E_ARG, E_HIP = -1, -2, -3, -4, -5, -6
LAYOUT_LINEAR, LAYOUT_TILED = 0, 1

_P64 = ctypes.POINTER(ctypes.c_uint64)
_PP64 = ctypes.POINTER(_P64)


class BatchGrid(ctypes.Structure):
    """cellato_batch_grid of include/cellato_hip.h"""
    _fields_ = [("rule", ctypes.c_int), ("layout", ctypes.c_int), ("width", ctypes.c_int), ("height", ctypes.c_int),
                ("steps", ctypes.c_int), ("t0", ctypes.c_int), ("temporal_steps", ctypes.c_int),
                ("chunk_rows", ctypes.c_int), ("a", ctypes.c_void_p * 5), ("b", ctypes.c_void_p * 5)]


def plan(rule_id, layout, width, height, temporal_steps=0, chunk_rows=0):
    """The launch one pass would make (host only): (rows per task, tasks, blocks)."""
    c, t, b = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().cellato_hip_plan(rule_id, layout, width, height, temporal_steps, chunk_rows, ctypes.byref(c),
                                 ctypes.byref(t), ctypes.byref(b)))
    return c.value, t.value, b.value


def kernel_info(rule_id, layout, width, temporal_steps=0, batch=False):
    """The kernel a pass would launch (needs a device): (depth it runs at, scratch bytes per
    lane, registers per lane). temporal_steps 0 = the simple traverser."""
    t, p, r = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    check(lib().cellato_hip_kernel_info(rule_id, layout, width, temporal_steps, int(batch), ctypes.byref(t),
                                        ctypes.byref(p), ctypes.byref(r)))
    return t.value, p.value, r.value


class CellatoError(RuntimeError):
    """A failed C-ABI call: `code` is the negative status, the message comes from the library."""

    def __init__(self, code, message):
        super().__init__(f"{message} (status {code})")
        self.code = code


_lib = None


def lib():
    """Load libcellato_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        # torch ships its own libamdhip64.so.7. Load it first, so that the library's dependency
        # on libamdhip64.so.7 resolves to that same runtime: two HIP runtimes in one process
        # would not share streams or device state.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                              "(make -C 2026-cellato-smpt-journal_amd/csrc)")
        L = ctypes.CDLL(str(LIB_PATH))
        L.cellato_rule_count.restype = ctypes.c_int
        L.cellato_rule_name.restype = ctypes.c_char_p
        L.cellato_rule_name.argtypes = [ctypes.c_int]
        L.cellato_rule_find.argtypes = [ctypes.c_char_p]
        L.cellato_rule_planes.argtypes = [ctypes.c_int]
        L.cellato_rule_states.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32), ctypes.c_int]
        L.cellato_max_temporal_steps.restype = ctypes.c_int
        L.cellato_rule_traits.argtypes = [ctypes.c_int]
        L.cellato_hip_last_error.restype = ctypes.c_char_p
        L.cellato_hip_simple.argtypes = [ctypes.c_int, ctypes.c_int, _PP64, _PP64, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_int)]
        L.cellato_hip_temporal.argtypes = [ctypes.c_int, ctypes.c_int, _PP64, _PP64, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_int)]
        L.cellato_batch_create.argtypes = [ctypes.POINTER(BatchGrid), ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p)]
        L.cellato_batch_launch.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        L.cellato_batch_result_in_b.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.cellato_batch_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                          ctypes.POINTER(ctypes.c_longlong)]
        L.cellato_batch_destroy.argtypes = [ctypes.c_void_p]
        L.cellato_hip_plan.argtypes = [ctypes.c_int] * 6 + [ctypes.POINTER(ctypes.c_int)] * 3
        L.cellato_hip_kernel_info.argtypes = [ctypes.c_int] * 5 + [ctypes.POINTER(ctypes.c_int)] * 3
        L.cellato_hip_bit_array.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p] + \
            [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        L.cellato_batch_plan.argtypes = [ctypes.POINTER(BatchGrid), ctypes.c_int, ctypes.c_int,
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                         ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_longlong)]
        L.cellato_hip_multi.argtypes = [ctypes.POINTER(BatchGrid), ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_int)]
        L.cellato_hip_pack.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                       _PP64, ctypes.c_void_p, ctypes.c_void_p]
        L.cellato_hip_unpack.argtypes = [ctypes.c_int, ctypes.c_int, _PP64, ctypes.c_int, ctypes.c_int,
                                         ctypes.c_void_p, ctypes.c_void_p]
        L.cellato_hip_checksum.argtypes = [ctypes.c_int, ctypes.c_int, _PP64, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_void_p, ctypes.c_void_p]
        L.cellato_hip_fill_random.argtypes = [ctypes.c_int, ctypes.c_int, _PP64, ctypes.c_int, ctypes.c_int,
                                              ctypes.c_ulonglong, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                                              ctypes.c_void_p]
        vp = ctypes.c_void_p
        L.cellato_mlp_linear_forward.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                                 ctypes.c_int, vp]
        L.cellato_mlp_linear_backward.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, ctypes.c_int, ctypes.c_int,
                                                  ctypes.c_int, ctypes.c_int, vp]
        L.cellato_mlp_mse.argtypes = [vp, vp, vp, vp, ctypes.c_int, vp]
        L.cellato_mlp_adam.argtypes = [vp, vp, vp, vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                       ctypes.c_float, ctypes.c_int, vp]
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise CellatoError(rc, lib().cellato_hip_last_error().decode())
    return rc


def plane_array(ptrs):
    """Host array of device plane pointers (uint64_t* const*)."""
    arr = (_P64 * len(ptrs))(*[ctypes.cast(ctypes.c_void_p(int(p)), _P64) for p in ptrs])
    return arr


def simple(rule, layout, ptrs_a, ptrs_b, width, height, t0, steps, chunk_rows=0, stream=None):
    """cellato_hip_simple; returns True when the result is in buffer b."""
    res = ctypes.c_int(-1)
    check(lib().cellato_hip_simple(rule, layout, plane_array(ptrs_a), plane_array(ptrs_b), width, height, t0,
                                   steps, chunk_rows, ctypes.c_void_p(stream or 0), ctypes.byref(res)))
    return bool(res.value)


def temporal(rule, layout, ptrs_a, ptrs_b, width, height, t0, steps, temporal_steps, chunk_rows=0, stream=None):
    """cellato_hip_temporal; returns True when the result is in buffer b."""
    res = ctypes.c_int(-1)
    check(lib().cellato_hip_temporal(rule, layout, plane_array(ptrs_a), plane_array(ptrs_b), width, height, t0,
                                     steps, temporal_steps, chunk_rows, ctypes.c_void_p(stream or 0),
                                     ctypes.byref(res)))
    return bool(res.value)


def bit_array(rule, word_bits, ptr_a, ptr_b, width, height, t0, steps, stream=None):
    """cellato_hip_bit_array (one device array of words per buffer); True when the result is in b."""
    res = ctypes.c_int(-1)
    check(lib().cellato_hip_bit_array(rule, word_bits, ctypes.c_void_p(int(ptr_a)), ctypes.c_void_p(int(ptr_b)), width,
                                      height, t0, steps, ctypes.c_void_p(stream or 0), ctypes.byref(res)))
    return bool(res.value)

"""ctypes mirror of include/xg.h (the C ABI) and the loader of libxg.so.

The structs here must match include/xg.h field for field; a test checks the
sizes against the library's own view of them.
"""

from __future__ import annotations

import ctypes as C
import os

MAX_RANK = 8
MAX_NODES = 96
MAX_LEAVES = 16
MAX_CONSTS = 32

# status codes
OK = 0
BROADCAST_ERROR = 1
BAD_AXES = 2
UNSUPPORTED = 3
INVALID_ARGUMENT = 4
TYPE_ERROR = 5
DEVICE_ERROR = 6
OUT_OF_MEMORY = 7
BUSY = 8

# dtypes
BOOL, INT8, UINT8, INT16, UINT16, INT32, UINT32, INT64, UINT64, FLOAT32, FLOAT64, BFLOAT16 = range(12)
DTYPE_NAMES = ["bool", "int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64",
               "float32", "float64", "bfloat16"]
DTYPE_SIZE = [1, 1, 1, 2, 2, 4, 4, 8, 8, 4, 8, 2]

# opcodes
OP = dict(
    LEAF=0, CONST=1, NEGATE=2, IDENTITY=3, LOGICAL_NOT=4, BITWISE_NOT=5, CAST=6,
    PLUS=10, MINUS=11, MULTIPLIES=12, DIVIDES=13, MODULUS=14, LOGICAL_AND=15, LOGICAL_OR=16,
    BITWISE_AND=17, BITWISE_OR=18, BITWISE_XOR=19, LEFT_SHIFT=20, RIGHT_SHIFT=21,
    LESS=22, LESS_EQUAL=23, GREATER=24, GREATER_EQUAL=25, EQUAL_TO=26, NOT_EQUAL_TO=27,
    MAXIMUM=28, MINIMUM=29, POW=30, FMOD=31, ATAN2=32, HYPOT=33,
    WHERE=40, CLIP=41,
    ABS=50, FABS=51, EXP=52, EXP2=53, EXPM1=54, LOG=55, LOG2=56, LOG10=57, LOG1P=58, SQRT=59,
    CBRT=60, SIN=61, COS=62, TAN=63, TANH=64, SINH=65, COSH=66, ASIN=67, ACOS=68, ATAN=69,
    ERF=70, FLOOR=71, CEIL=72, TRUNC=73, ROUND=74, GATHER=80,
)

# reductions / scans
RED = dict(sum=0, prod=1, amax=2, amin=3, mean=4, norm_sq=5, norm_l2=6, norm_l1=7, norm_linf=8, norm_l0=9,
           variance=10, stddev=11, minmax=12)
SCAN = dict(cumsum=0, cumprod=1)
REDUCE_IMMEDIATE = 1  # xg_reduce_desc.flags: xt::evaluation_strategy::immediate
AXIS_NONE = -1000


class Tensor(C.Structure):
    _fields_ = [
        ("data", C.c_void_p),
        ("dtype", C.c_int32),
        ("ndim", C.c_int32),
        ("offset", C.c_int64),
        ("shape", C.c_int64 * MAX_RANK),
        ("strides", C.c_int64 * MAX_RANK),
    ]


class _ScalarValue(C.Union):
    _fields_ = [("i", C.c_int64), ("u", C.c_uint64), ("f", C.c_double)]


class Scalar(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("pad_", C.c_int32), ("v", _ScalarValue)]


class Node(C.Structure):
    _fields_ = [("op", C.c_int32), ("nargs", C.c_int32), ("args", C.c_int32 * 3), ("index", C.c_int32)]


class Expr(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32),
        ("n_leaves", C.c_int32),
        ("n_consts", C.c_int32),
        ("flags", C.c_int32),
        ("nodes", C.POINTER(Node)),
        ("leaves", C.POINTER(Tensor)),
        ("consts", C.POINTER(Scalar)),
    ]


class ReduceDesc(C.Structure):
    _fields_ = [
        ("op", C.c_int32),
        ("acc_dtype", C.c_int32),
        ("n_axes", C.c_int32),
        ("axes", C.c_int32 * MAX_RANK),
        ("keep_dims", C.c_int32),
        ("has_initial", C.c_int32),
        ("initial", Scalar),
        ("flags", C.c_int32),
        ("ddof", C.c_int32),
    ]


class ScanDesc(C.Structure):
    _fields_ = [("op", C.c_int32), ("acc_dtype", C.c_int32), ("axis", C.c_int32), ("flags", C.c_int32)]


class BatchItem(C.Structure):
    _fields_ = [("in_", C.c_void_p), ("out", C.c_void_p), ("rows", C.c_int64), ("cols", C.c_int64),
                ("in_row_stride", C.c_int64), ("out_row_stride", C.c_int64)]


class BatchTuning(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("task_elems", C.c_int32),
                ("dyn_percent", C.c_int32)]


class Request(C.Structure):
    _fields_ = [("tpl", C.c_int32), ("dtype", C.c_int32), ("shape", C.c_int64 * 4), ("in_", C.c_void_p),
                ("params", C.c_void_p * 4), ("out", C.c_void_p), ("tag", C.c_uint64), ("arrival_ns", C.c_int64),
                ("expr", C.c_void_p), ("out_tensor", C.c_void_p), ("reduce", C.c_void_p)]


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_int64), ("streams", C.c_int32), ("max_batch", C.c_int32),
                ("queue_capacity", C.c_int32), ("plain", C.c_int32)]


class Completion(C.Structure):
    _fields_ = [("tag", C.c_uint64), ("status", C.c_int32), ("batch_size", C.c_int32), ("arrival_ns", C.c_int64),
                ("admit_ns", C.c_int64), ("launch_ns", C.c_int64), ("done_ns", C.c_int64)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("launches", C.c_int64), ("inflight_bytes", C.c_int64), ("peak_inflight_bytes", C.c_int64),
                ("queued", C.c_int64), ("admitted", C.c_int64)]


HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libxg.so")

# entry points declared in include/xg.h: name -> (restype, argtypes)
_P = C.POINTER
EXPORTS = {
    "xg_version": (C.c_char_p, []),
    "xg_last_error_message": (C.c_char_p, []),
    "xg_init": (C.c_int, []),
    "xg_expr_info": (C.c_int, [_P(Expr), _P(C.c_int32), _P(C.c_int32), _P(C.c_int64)]),
    "xg_assign": (C.c_int, [_P(Expr), _P(Tensor), C.c_void_p]),
    "xg_reduce_info": (C.c_int, [_P(Expr), _P(ReduceDesc), _P(C.c_int32), _P(C.c_int32), _P(C.c_int64)]),
    "xg_reduce": (C.c_int, [_P(Expr), _P(ReduceDesc), _P(Tensor), C.c_void_p]),
    "xg_assign_reduce2d": (C.c_int, [_P(Expr), _P(Tensor), _P(ReduceDesc), _P(Tensor), _P(ReduceDesc), _P(Tensor),
                                     C.c_void_p]),
    "xg_reduce_noalloc": (C.c_int, [_P(Expr), _P(ReduceDesc), _P(Tensor), C.c_void_p]),
    "xg_reduce_noalloc_max_bytes": (C.c_size_t, []),
    "xg_scan_info": (C.c_int, [_P(Expr), _P(ScanDesc), _P(C.c_int32), _P(C.c_int32), _P(C.c_int64)]),
    "xg_scan": (C.c_int, [_P(Expr), _P(ScanDesc), _P(Tensor), C.c_void_p]),
    "xg_assign_indexed": (C.c_int, [_P(Expr), _P(Expr), _P(Tensor), C.c_void_p]),
    "xg_argwhere": (C.c_int, [_P(Expr), C.c_int32, _P(Tensor), C.c_void_p]),
    "xg_npy_info": (C.c_int, [C.c_char_p, _P(C.c_int32), _P(C.c_int32), _P(C.c_int64), _P(C.c_int32)]),
    "xg_npy_load": (C.c_int, [C.c_char_p, _P(Tensor), C.c_void_p]),
    "xg_npy_dump": (C.c_int, [C.c_char_p, _P(Tensor), C.c_void_p]),
    "xg_plan_assign": (C.c_int, [_P(Expr), _P(Tensor), _P(C.c_void_p)]),
    "xg_plan_assign_reduce2d": (C.c_int, [_P(Expr), _P(Tensor), _P(ReduceDesc), _P(Tensor), _P(ReduceDesc),
                                          _P(Tensor), _P(C.c_void_p)]),
    "xg_plan_reduce": (C.c_int, [_P(Expr), _P(ReduceDesc), _P(Tensor), _P(C.c_void_p)]),
    "xg_plan_scan": (C.c_int, [_P(Expr), _P(ScanDesc), _P(Tensor), _P(C.c_void_p)]),
    "xg_plan_workspace_size": (C.c_size_t, [C.c_void_p]),
    "xg_plan_execute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "xg_plan_destroy": (None, [C.c_void_p]),
    "xg_batch_create": (C.c_int, [C.c_int32, C.c_int32, _P(BatchItem), C.c_int64, _P(C.c_void_p)]),
    "xg_batch_autotune": (C.c_int, [C.c_void_p, C.c_void_p]),
    "xg_batch_get_tuning": (C.c_int, [C.c_void_p, _P(BatchTuning)]),
    "xg_batch_set_tuning": (C.c_int, [C.c_void_p, _P(BatchTuning)]),
    "xg_batch_launch": (C.c_int, [C.c_void_p, C.c_void_p]),
    "xg_batch_execute": (C.c_int, [C.c_void_p, C.c_void_p]),
    "xg_batch_info": (C.c_int, [C.c_void_p, _P(C.c_int64), _P(C.c_int64)]),
    "xg_batch_destroy": (None, [C.c_void_p]),
    "xg_plan_execute_many": (C.c_int, [_P(C.c_void_p), C.c_int64, C.c_void_p]),
    "xg_server_now_ns": (C.c_int64, []),
    "xg_server_create": (C.c_int, [_P(ServerConfig), _P(C.c_void_p)]),
    "xg_server_submit": (C.c_int, [C.c_void_p, _P(Request)]),
    "xg_server_poll": (C.c_int, [C.c_void_p, _P(Completion), C.c_int32, _P(C.c_int32)]),
    "xg_server_drain": (C.c_int, [C.c_void_p, C.c_int64]),
    "xg_server_stats_get": (C.c_int, [C.c_void_p, _P(ServerStats)]),
    "xg_server_destroy": (None, [C.c_void_p]),
    "xg_server_hold": (C.c_int, [C.c_void_p, C.c_int32]),
}

# diagnostic entry points (not part of include/xg.h)
DEBUG_EXPORTS = {
    "xg_last_kernel": (C.c_char_p, []),
    "xg_debug_force_dynamic": (None, [C.c_int]),
    "xg_debug_units_per_thread": (None, [C.c_int]),
    "xg_debug_nontemporal": (None, [C.c_int, C.c_int]),
    "xg_debug_fill_granules": (None, [C.c_int]),
    "xg_debug_jit": (None, [C.c_int]),
    "xg_debug_reduce_waves": (None, [C.c_int]),
    "xg_debug_cols_split": (None, [C.c_int, C.c_int]),
    "xg_debug_scan_blocks_per_cu": (None, [C.c_int]),
    "xg_debug_scan_single_pass": (None, [C.c_int]),
    "xg_debug_scan_flat_single_pass": (None, [C.c_int]),
    "xg_debug_jit_stats": (C.c_char_p, [C.POINTER(C.c_int64)]),
    "xg_debug_assign_kernel": (C.c_int, [_P(Expr), _P(Tensor), C.c_char_p, C.c_int32]),
    "xg_debug_assign_pieces": (C.c_int, [_P(Expr), _P(Tensor), _P(C.c_int64), _P(C.c_int32), _P(C.c_int64)]),
    "xg_debug_reduce_pieces": (C.c_int, [_P(Expr), _P(ReduceDesc), _P(Tensor), _P(C.c_int64), _P(C.c_int32),
                                         _P(C.c_int64), _P(C.c_int32)]),
}

_lib = None


def lib():
    """Load libxg.so (raises if it has not been built: there is no fallback)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'` "
                "(the evaluator has no CPU fallback)")
        # torch (the allocator and stream provider) brings its own HIP runtime;
        # it must be loaded before this library's, or the library's runtime
        # finds no device once torch has initialised the GPU (measured on MI355X)
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in {**EXPORTS, **DEBUG_EXPORTS}.items():
            f = getattr(L, name, None)
            if f is None:
                continue
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib

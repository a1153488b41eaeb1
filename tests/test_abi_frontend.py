"""CPU checks of the C ABI and the host side (no GPU needed).

* libxg.so loads and exports every entry point include/xg.h declares;
* the ctypes mirror of the structs has the C layout;
* the library's own type / shape derivation (planner) agrees with the
  oracle's on random expression trees, reductions and scans, including
  which ones are errors;
* the front end's view arithmetic agrees with numpy's.
"""

import ctypes as C
import os
import re
import subprocess
import tempfile

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd import _abi as A
from xtensor_amd.core import _Placeholder

from oracle.pyoracle import Oracle, OracleError

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "xg.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(xg_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = A.lib()
    names = declared_functions()
    assert len(names) >= 9
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.xg_version().decode().startswith("xtensor_amd")


def test_struct_layout_matches_header():
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "xg.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(xg_tensor), sizeof(xg_scalar), sizeof(xg_node),
         sizeof(xg_expr), sizeof(xg_reduce_desc), sizeof(xg_scan_desc));
  printf("%zu %zu %zu\n", offsetof(xg_reduce_desc, initial), offsetof(xg_expr, nodes), offsetof(xg_tensor, shape));
  printf("%zu %zu %zu\n", sizeof(xg_batch_item), sizeof(xg_batch_tuning), offsetof(xg_batch_item, out_row_stride));
  printf("%zu %zu %zu %zu %zu %zu\n", sizeof(xg_request), offsetof(xg_request, out), sizeof(xg_server_config),
         sizeof(xg_completion), offsetof(xg_completion, done_ns), sizeof(xg_server_stats));
  return 0;
}
'''
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(src)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    sizes = [int(x) for x in out]
    mine = [C.sizeof(A.Tensor), C.sizeof(A.Scalar), C.sizeof(A.Node), C.sizeof(A.Expr),
            C.sizeof(A.ReduceDesc), C.sizeof(A.ScanDesc),
            A.ReduceDesc.initial.offset, A.Expr.nodes.offset, A.Tensor.shape.offset,
            C.sizeof(A.BatchItem), C.sizeof(A.BatchTuning), A.BatchItem.out_row_stride.offset,
            C.sizeof(A.Request), A.Request.out.offset, C.sizeof(A.ServerConfig),
            C.sizeof(A.Completion), A.Completion.done_ns.offset, C.sizeof(A.ServerStats)]
    assert sizes == mine


def test_server_config_errors_without_device():
    """The configuration is validated before any device call."""
    lib = A.lib()
    h = C.c_void_p()
    assert lib.xg_server_create(None, C.byref(h)) == A.INVALID_ARGUMENT
    for cfg in (A.ServerConfig(0, 4, 8, 16, 0), A.ServerConfig(1 << 30, 0, 8, 16, 0),
                A.ServerConfig(1 << 30, 17, 8, 16, 0), A.ServerConfig(1 << 30, 4, 0, 16, 0),
                A.ServerConfig(1 << 30, 4, 8, 0, 0)):
        assert lib.xg_server_create(C.byref(cfg), C.byref(h)) == A.INVALID_ARGUMENT
    assert lib.xg_server_submit(None, None) == A.INVALID_ARGUMENT
    assert lib.xg_plan_execute_many(None, 2, None) == A.INVALID_ARGUMENT


def test_batch_argument_errors_without_device():
    """Op and dtype are validated before any device call."""
    lib = A.lib()
    h = C.c_void_p()
    items = (A.BatchItem * 1)()
    assert lib.xg_batch_create(7, A.FLOAT32, items, 1, C.byref(h)) == A.INVALID_ARGUMENT
    assert lib.xg_batch_create(0, A.INT32, items, 1, C.byref(h)) == A.TYPE_ERROR
    assert b"dtype" in lib.xg_last_error_message()
    assert lib.xg_batch_create(0, A.FLOAT32, None, 3, C.byref(h)) == A.INVALID_ARGUMENT
    assert lib.xg_batch_execute(None, None) == A.INVALID_ARGUMENT


# ---------------------------------------------------------------- type fuzz
DTYPES = ["bool", "int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64", "float32", "float64"]
UNARY = ["NEGATE", "IDENTITY", "LOGICAL_NOT", "BITWISE_NOT", "ABS", "FABS", "EXP", "SQRT", "LOG", "FLOOR", "ROUND",
         "TANH", "CBRT"]
BINARY = ["PLUS", "MINUS", "MULTIPLIES", "DIVIDES", "MODULUS", "LOGICAL_AND", "LOGICAL_OR", "BITWISE_AND", "BITWISE_OR",
          "BITWISE_XOR", "LEFT_SHIFT", "RIGHT_SHIFT", "LESS", "LESS_EQUAL", "GREATER", "GREATER_EQUAL", "EQUAL_TO",
          "NOT_EQUAL_TO", "MAXIMUM", "MINIMUM", "POW", "FMOD", "ATAN2", "HYPOT"]
TERNARY = ["WHERE", "CLIP"]
SHAPES = [(), (1,), (3,), (4, 3), (1, 3), (4, 1), (2, 4, 3), (2, 1, 1), (0,), (0, 3), (5, 3)]


def random_expr(rng, depth, leaves):
    if depth == 0 or rng.random() < 0.25:
        if rng.random() < 0.2:
            dt = DTYPES[rng.integers(len(DTYPES))]
            return xt.scalar(1, dt)
        dt = DTYPES[rng.integers(len(DTYPES))]
        shp = SHAPES[rng.integers(len(SHAPES))]
        a = _Placeholder(xt.dtype_code(dt), shp)
        leaves.append(a)
        return a
    k = rng.random()
    if k < 0.3:
        op = UNARY[rng.integers(len(UNARY))]
        return xt.Func(A.OP[op], random_expr(rng, depth - 1, leaves))
    if k < 0.35:
        return xt.cast(random_expr(rng, depth - 1, leaves), DTYPES[rng.integers(len(DTYPES))])
    if k < 0.9:
        op = BINARY[rng.integers(len(BINARY))]
        return xt.Func(A.OP[op], random_expr(rng, depth - 1, leaves), random_expr(rng, depth - 1, leaves))
    op = TERNARY[rng.integers(len(TERNARY))]
    return xt.Func(A.OP[op], *[random_expr(rng, depth - 1, leaves) for _ in range(3)])


def lib_info(e):
    try:
        return ("ok",) + xt.expr_info(e)
    except xt.BroadcastError:
        return ("broadcast",)
    except TypeError:
        return ("type",)


def oracle_info(e, leaves):
    o = Oracle({a: np.zeros(a.shape, np.uint8) for a in leaves})
    try:
        return ("ok",) + o.info(e)
    except OracleError as err:
        return {A.BROADCAST_ERROR: ("broadcast",), A.TYPE_ERROR: ("type",)}.get(err.code, ("other", err.code))


@pytest.mark.parametrize("seed", range(40))
def test_expression_types_match_oracle(seed):
    rng = np.random.default_rng(seed)
    for _ in range(25):
        leaves = []
        e = random_expr(rng, 3, leaves)
        if not isinstance(e, xt.Func):
            continue
        assert lib_info(e) == oracle_info(e, leaves)


REDUCERS = ["sum", "prod", "amax", "amin", "mean", "norm_sq", "norm_l2", "norm_l1", "norm_linf", "norm_l0"]


@pytest.mark.parametrize("seed", range(20))
def test_reduce_and_scan_types_match_oracle(seed):
    rng = np.random.default_rng(100 + seed)
    for _ in range(30):
        dt = DTYPES[rng.integers(len(DTYPES))]
        shp = [(2, 3, 4), (5,), (0, 3), (3, 1, 2), ()][rng.integers(5)]
        a = _Placeholder(xt.dtype_code(dt), shp)
        kind = REDUCERS[rng.integers(len(REDUCERS))]
        nd = len(shp)
        axes = None if rng.random() < 0.3 else sorted(set(int(x) for x in rng.integers(-nd if nd else 0, max(nd, 1), size=rng.integers(0, 3))))
        if axes is not None and rng.random() < 0.2:
            axes = axes[::-1] + [0]
        acc = None if (rng.random() < 0.6 or kind.startswith("norm")) else DTYPES[rng.integers(len(DTYPES))]
        r = xt.Reducer(kind, a, axes, keep_dims=bool(rng.random() < 0.3), dtype=acc)
        o = Oracle({a: np.zeros(shp, np.uint8)})
        try:
            want = ("ok",) + o.reduce_info(r)
        except OracleError as err:
            want = ("err", err.code)
        try:
            got = ("ok",) + xt.reduce_info(r)
        except RuntimeError:
            got = ("err", A.BAD_AXES)
        except TypeError:
            got = ("err", A.TYPE_ERROR)
        except ValueError:
            got = ("err", A.INVALID_ARGUMENT)
        assert got == want, (dt, shp, kind, axes, acc)
        # scans
        axis = None if rng.random() < 0.3 else int(rng.integers(-3, 3))
        sacc = None if rng.random() < 0.7 else DTYPES[rng.integers(len(DTYPES))]
        kindsc = ["cumsum", "cumprod"][rng.integers(2)]
        try:
            want = ("ok",) + o.scan_info(a, axis, kindsc, sacc)
        except OracleError as err:
            want = ("err", err.code)
        d = A.ScanDesc(A.SCAN[kindsc], -1 if sacc is None else xt.dtype_code(sacc),
                       A.AXIS_NONE if axis is None else axis, 0)
        s = xt.serialize(a)
        dtc, ndc = C.c_int32(), C.c_int32()
        shpc = (C.c_int64 * A.MAX_RANK)()
        st = A.lib().xg_scan_info(s.ptr(), C.byref(d), C.byref(dtc), C.byref(ndc), shpc)
        got = ("ok", dtc.value, tuple(shpc[i] for i in range(ndc.value))) if st == 0 else ("err", st)
        assert got == want, (dt, shp, kindsc, axis, sacc)


def test_broadcast_error_and_axes_errors():
    a = _Placeholder(A.FLOAT64, (2, 3))
    b = _Placeholder(A.FLOAT64, (4,))
    with pytest.raises(xt.BroadcastError):
        (a + b).shape
    with pytest.raises(RuntimeError, match="sorted"):
        xt.sum(a, [1, 0]).shape
    with pytest.raises(RuntimeError, match="duplicates"):
        xt.sum(a, [1, 1]).shape
    with pytest.raises(RuntimeError, match="out of bounds"):
        xt.sum(a, [0, 2]).shape


@pytest.mark.parametrize("keys", [
    (slice(None), slice(None), slice(0, 56, 2), slice(None)),
    (slice(None, None, -1), 3, slice(1, None, 3)),
    (Ellipsis, slice(-5, -1)),
    (1, None, slice(None), 2),
    (slice(2, 2), slice(None)),
])
def test_view_strides_match_numpy(keys):
    base = np.zeros((4, 8, 56, 9), np.float32)
    if len([k for k in keys if k is not None and k is not Ellipsis]) < 4 and Ellipsis not in keys:
        pass
    a = _Placeholder(A.FLOAT32, base.shape)
    xk = tuple(xt.newaxis() if k is None else k for k in keys)
    v = xt.view(a, *xk)
    nv = base[keys]
    assert v.shape == nv.shape
    off = (nv.__array_interface__["data"][0] - base.__array_interface__["data"][0]) // 4
    if nv.size:
        assert v.offset == off
        for n, s, ns in zip(v.shape, v.strides, nv.strides):
            if n > 1:
                assert s == ns // 4

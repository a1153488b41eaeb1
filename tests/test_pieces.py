"""How elementwise assignments past 32-bit offsets are cut into pieces (host
planning only: the tensors are descriptors with placeholder addresses that
nothing reads).  Every piece must fit 32-bit offsets and 32-bit unit
indexing; arrays that fit are not cut, and a single index of the first
dimension that is itself too long is refused."""

import ctypes as C

import pytest

from xtensor_amd import _abi as A


def tensor(dtype, shape, data=0x100000):
    t = A.Tensor()
    t.data, t.dtype, t.ndim, t.offset = data, dtype, len(shape), 0
    stride = 1
    for k in reversed(range(len(shape))):
        t.shape[k], t.strides[k] = shape[k], stride
        stride *= shape[k]
    return t


def expr(nodes, leaves, consts=()):
    """nodes: (op name, args, index) in post order."""
    arr = (A.Node * len(nodes))()
    for i, (op, args, index) in enumerate(nodes):
        arr[i].op, arr[i].nargs, arr[i].index = A.OP[op], len(args), index
        for j, a in enumerate(args):
            arr[i].args[j] = a
    lv = (A.Tensor * len(leaves))(*leaves)
    cs = (A.Scalar * max(1, len(consts)))()
    for i, v in enumerate(consts):
        cs[i].dtype, cs[i].v.i = A.INT32, v
    e = A.Expr()
    e.n_nodes, e.n_leaves, e.n_consts, e.flags = len(nodes), len(leaves), len(consts), 0
    e.nodes = C.cast(arr, C.POINTER(A.Node))
    e.leaves = C.cast(lv, C.POINTER(A.Tensor))
    e.consts = C.cast(cs, C.POINTER(A.Scalar))
    e._keep = (arr, lv, cs)
    return e


def pieces(e, out):
    n, compact, units = C.c_int64(), C.c_int32(), C.c_int64()
    st = A.lib().xg_debug_assign_pieces(C.byref(e), C.byref(out), C.byref(n), C.byref(compact), C.byref(units))
    assert st == A.OK, A.lib().xg_last_error_message()
    return n.value, compact.value, units.value


def test_flat_int8_past_2_31_is_three_pieces():
    n = (1 << 31) + (1 << 21) + 3
    x, y = tensor(A.INT8, (n,)), tensor(A.INT8, (n,), data=0x200000)
    e = expr([("LEAF", (), 0), ("CONST", (), 0), ("PLUS", (0, 1), 0)], [x], [3])
    count, compact, units = pieces(e, y)
    assert count == 3 and compact == 1 and units <= 1 << 30


def test_rows_with_broadcast_row_are_one_row_per_piece():
    r, c = 3, (1 << 30) + 5
    x, b = tensor(A.INT16, (r, c)), tensor(A.INT16, (c,), data=0x300000)
    y = tensor(A.INT16, (r, c), data=0x400000)
    e = expr([("LEAF", (), 0), ("LEAF", (), 1), ("MINUS", (0, 1), 0)], [x, b])
    assert pieces(e, y) == (3, 1, c)


def test_fitting_arrays_are_not_cut_and_overlong_rows_are_refused():
    x, y = tensor(A.FLOAT32, (4, 100)), tensor(A.FLOAT32, (4, 100), data=0x200000)
    e = expr([("LEAF", (), 0), ("NEGATE", (0,), 0)], [x])
    assert pieces(e, y)[0] == 0
    n = (1 << 31) + 5
    x, y = tensor(A.INT8, (2, n)), tensor(A.INT8, (2, n), data=0x200000)
    e = expr([("LEAF", (), 0), ("NEGATE", (0,), 0)], [x])
    assert pieces(e, y)[0] == 0

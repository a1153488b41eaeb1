"""How elementwise assignments past 32-bit offsets, and reductions whose
reduced extent passes 32-bit indexing, are cut into pieces (host planning
only: the tensors are descriptors with placeholder addresses that nothing
reads).  Every piece must fit 32-bit offsets and 32-bit unit indexing;
arrays that fit are not cut, and a single index of the first dimension that
is itself too long is refused."""

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


@pytest.mark.parametrize("dtype,shape,count", [
    (A.FLOAT32, (65536, 32768), 2),    # 2^31 contiguous elements: one piece would merge into 2^31
    (A.INT8, (8, 1 << 29), 4),
    (A.FLOAT32, (131072, 32768), 4),
])
def test_power_of_two_rows_are_cut_below_2_31_elements(dtype, shape, count):
    x, y = tensor(dtype, shape), tensor(dtype, shape, data=0x200000)
    e = expr([("LEAF", (), 0), ("NEGATE", (0,), 0)], [x])
    n, compact, units = pieces(e, y)
    assert n == count and compact == 1 and units <= 1 << 30  # elements (one per unit when planned generic)


def red(op, axes=None, acc=-1):
    d = A.ReduceDesc()
    d.op, d.acc_dtype = A.RED[op], acc
    if axes is None:
        d.n_axes = -1
    else:
        d.n_axes = len(axes)
        for i, a in enumerate(axes):
            d.axes[i] = a
    return d


def reduce_pieces(e, r, out):
    n, axis, length, fit = C.c_int64(), C.c_int32(), C.c_int64(), C.c_int32()
    st = A.lib().xg_debug_reduce_pieces(C.byref(e), C.byref(r), C.byref(out), C.byref(n), C.byref(axis),
                                        C.byref(length), C.byref(fit))
    assert st == A.OK, A.lib().xg_last_error_message()
    return n.value, axis.value, length.value, fit.value


def scalar_out(dtype):
    t = tensor(dtype, (), data=0x900000)
    return t


@pytest.mark.parametrize("op,rdt", [("sum", A.INT64), ("amax", A.INT8), ("prod", A.INT64), ("amin", A.INT8),
                                    ("mean", A.FLOAT64), ("norm_l2", A.FLOAT64), ("norm_sq", A.INT64),
                                    ("norm_l1", A.INT64), ("norm_linf", A.INT8), ("norm_l0", A.UINT64)])
@pytest.mark.parametrize("shape,count,axis", [
    (((1 << 31) + 5, 2), 5, 0),   # a long axis that merges with the short inner one: 2^29 rows a piece
    ((2, 1 << 30), 2, 0),         # two dims that merge into 2^31
    ((4, 1 << 29, 4), 8, 1),      # the middle axis is cut, the outer one taken an index at a time
    ((2, (1 << 31) + 5), 6, 1),
])
def test_full_reductions_past_2_31_are_cut(op, rdt, shape, count, axis):
    x = tensor(A.INT8, shape)
    e = expr([("LEAF", (), 0)], [x])
    acc = A.INT64 if op in ("sum", "prod") else -1
    n, cut_axis, length, fit = reduce_pieces(e, red(op, None, acc), scalar_out(rdt))
    assert (n, cut_axis, fit) == (count, axis, 1)
    inner = 1
    for k in range(axis + 1, len(shape)):
        inner *= shape[k]
    assert length * inner <= 1 << 30 < (length + 1) * inner or length == shape[axis]


def test_axis_reductions_past_2_31():
    # sum over the long axis of [2, 2^31 + 5] keeps [2]: three pieces of 2^30
    c = (1 << 31) + 5
    x = tensor(A.INT8, (2, c))
    e = expr([("LEAF", (), 0)], [x])
    assert reduce_pieces(e, red("sum", [1], A.INT64), tensor(A.INT64, (2,), data=0x900000)) == (3, 1, 1 << 30, 1)
    # mean over axis 0 of [2^31 + 5, 3] float32 (outputs contiguous): a column reduction
    x = tensor(A.FLOAT32, ((1 << 31) + 5, 3))
    e = expr([("LEAF", (), 0)], [x])
    assert reduce_pieces(e, red("mean", [0]), tensor(A.FLOAT64, (3,), data=0x900000)) == (3, 0, 1 << 30, 1)


def test_fitting_reductions_are_not_cut():
    x = tensor(A.FLOAT32, (8192, 4096))
    e = expr([("LEAF", (), 0)], [x])
    assert reduce_pieces(e, red("sum"), scalar_out(A.FLOAT32))[0] == 0
    assert reduce_pieces(e, red("sum", [0]), tensor(A.FLOAT32, (4096,), data=0x900000))[0] == 0

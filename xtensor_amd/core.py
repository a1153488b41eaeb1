"""xtensor's lazy-expression API over device-resident arrays.

Mirrors, in Python, the part of xtensor that the evaluator replaces:

* containers: `DeviceArray` is an `xt::xarray` whose storage is HBM
  (containers/xarray.hpp, containers/xcontainer.hpp); views share storage
  (`view`, `transpose`, `broadcast`, numpy-style indexing) with xtensor's
  stride rules: stepped ranges multiply strides, integer indices drop a
  dimension and move the offset (views/xview.hpp:1311-1346, 1536-1572,
  views/xslice.hpp:136-173), transpose reverses shape and strides
  (misc/xmanipulation.hpp:238-259);
* expressions: operators and math functions build a lazy tree, exactly the
  xfunction tree xtensor builds (core/xoperation.hpp:231-545,
  core/xmath.hpp:900-1201); python literals keep C++ literal types (an int
  is `int`, a float is `double`);
* evaluation: `noalias(out).assign(e)` / `out[...] = e` is
  `xt::noalias(out) = e` (core/xnoalias.hpp:160-221); reducers
  (`sum`, `mean`, `amax`, `norm_l2` ...) are lazy like xreducer and are
  evaluated by assignment or `eval`; `cumsum` / `cumprod` are immediate like
  xaccumulator (reducers/xaccumulator.hpp:343-367).

Every evaluation is one call into the C ABI of libxg.so; nothing here
computes element values.  Errors map back to the exception types xtensor
throws.
"""

from __future__ import annotations

import builtins
import ctypes as C
import numbers
import os
from typing import Any, Callable, Iterable, Sequence

import numpy as np

from . import _abi as A


# ----------------------------------------------------------------------------
# dtypes
_NP_OF = {
    A.BOOL: np.bool_, A.INT8: np.int8, A.UINT8: np.uint8, A.INT16: np.int16, A.UINT16: np.uint16,
    A.INT32: np.int32, A.UINT32: np.uint32, A.INT64: np.int64, A.UINT64: np.uint64,
    A.FLOAT32: np.float32, A.FLOAT64: np.float64,
}


def dtype_code(dt: Any) -> int:
    """Our dtype code for a name, numpy dtype or code."""
    if isinstance(dt, (int, np.integer)) and not isinstance(dt, bool):
        return int(dt)
    if isinstance(dt, str):
        return A.DTYPE_NAMES.index(dt)
    nd = np.dtype(dt)
    for code, t in _NP_OF.items():
        if np.dtype(t) == nd:
            return code
    raise TypeError(f"unsupported dtype {dt!r}")


def dtype_name(code: int) -> str:
    return A.DTYPE_NAMES[code]


def numpy_dtype(code: int):
    """numpy dtype used to hold values of `code` on the host (bf16 -> float32)."""
    return np.float32 if code == A.BFLOAT16 else _NP_OF[code]


# ----------------------------------------------------------------------------
# errors (xtensor exception types)
class BroadcastError(RuntimeError):
    """xt::broadcast_error (utils/xexception.hpp:63)."""


class DeviceError(RuntimeError):
    pass


def _check(status: int) -> None:
    if status == A.OK:
        return
    msg = A.lib().xg_last_error_message().decode()
    if status == A.BROADCAST_ERROR:
        raise BroadcastError(msg)
    if status == A.BAD_AXES:
        raise RuntimeError(msg)
    if status == A.TYPE_ERROR:
        raise TypeError(msg)
    if status == A.UNSUPPORTED:
        raise NotImplementedError(msg)
    if status == A.DEVICE_ERROR:
        raise DeviceError(msg)
    if status == A.OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise ValueError(msg)


# ----------------------------------------------------------------------------
# expressions
class Expr:
    """A node of a lazy expression tree (xfunction / xscalar / container)."""

    # ---- arithmetic operators (core/xoperation.hpp:231-545)
    def __add__(self, o): return Func(A.OP["PLUS"], self, o)
    def __radd__(self, o): return Func(A.OP["PLUS"], o, self)
    def __sub__(self, o): return Func(A.OP["MINUS"], self, o)
    def __rsub__(self, o): return Func(A.OP["MINUS"], o, self)
    def __mul__(self, o): return Func(A.OP["MULTIPLIES"], self, o)
    def __rmul__(self, o): return Func(A.OP["MULTIPLIES"], o, self)
    def __truediv__(self, o): return Func(A.OP["DIVIDES"], self, o)
    def __rtruediv__(self, o): return Func(A.OP["DIVIDES"], o, self)
    def __mod__(self, o): return Func(A.OP["MODULUS"], self, o)
    def __rmod__(self, o): return Func(A.OP["MODULUS"], o, self)
    def __and__(self, o): return Func(A.OP["BITWISE_AND"], self, o)
    def __rand__(self, o): return Func(A.OP["BITWISE_AND"], o, self)
    def __or__(self, o): return Func(A.OP["BITWISE_OR"], self, o)
    def __ror__(self, o): return Func(A.OP["BITWISE_OR"], o, self)
    def __xor__(self, o): return Func(A.OP["BITWISE_XOR"], self, o)
    def __rxor__(self, o): return Func(A.OP["BITWISE_XOR"], o, self)
    def __lshift__(self, o): return Func(A.OP["LEFT_SHIFT"], self, o)
    def __rlshift__(self, o): return Func(A.OP["LEFT_SHIFT"], o, self)
    def __rshift__(self, o): return Func(A.OP["RIGHT_SHIFT"], self, o)
    def __rrshift__(self, o): return Func(A.OP["RIGHT_SHIFT"], o, self)
    def __neg__(self): return Func(A.OP["NEGATE"], self)
    def __pos__(self): return Func(A.OP["IDENTITY"], self)
    def __invert__(self): return Func(A.OP["BITWISE_NOT"], self)
    def __lt__(self, o): return Func(A.OP["LESS"], self, o)
    def __le__(self, o): return Func(A.OP["LESS_EQUAL"], self, o)
    def __gt__(self, o): return Func(A.OP["GREATER"], self, o)
    def __ge__(self, o): return Func(A.OP["GREATER_EQUAL"], self, o)
    __hash__ = object.__hash__

    def equal(self, o):
        """Elementwise ==, xt::equal (Python's == stays identity comparison)."""
        return Func(A.OP["EQUAL_TO"], self, o)

    def not_equal(self, o):
        return Func(A.OP["NOT_EQUAL_TO"], self, o)

    # ---- shape / type (xfunction::shape, value_type)
    def _info(self):
        return expr_info(self)

    @property
    def shape(self) -> tuple:
        return self._info()[1]

    @property
    def dtype(self) -> str:
        return dtype_name(self._info()[0])

    @property
    def ndim(self) -> int:
        return len(self.shape)

    @property
    def size(self) -> int:
        return int(np.prod(self.shape, dtype=np.int64))


class Scalar(Expr):
    """xt::xscalar leaf: a value with a C++ type (containers/xscalar.hpp)."""

    def __init__(self, value, dtype=None):
        if dtype is None:
            if isinstance(value, (bool, np.bool_)):
                dtype = A.BOOL
            elif isinstance(value, np.generic):
                dtype = dtype_code(value.dtype)
            elif isinstance(value, numbers.Integral):
                dtype = A.INT32 if -2**31 <= int(value) < 2**31 else A.INT64
            elif isinstance(value, numbers.Real):
                dtype = A.FLOAT64
            else:
                raise TypeError(f"unsupported scalar {value!r}")
        self.dtype_code = dtype_code(dtype)
        self.value = value

    def to_abi(self) -> A.Scalar:
        s = A.Scalar()
        s.dtype = self.dtype_code
        if self.dtype_code in (A.FLOAT32, A.FLOAT64):
            s.v.f = float(self.value)
        elif self.dtype_code in (A.UINT8, A.UINT16, A.UINT32, A.UINT64):
            s.v.u = int(self.value) & 0xFFFFFFFFFFFFFFFF
        else:
            s.v.i = int(self.value)
        return s


def scalar(value, dtype=None) -> Scalar:
    """A typed constant, e.g. scalar(3, 'int64') is `int64_t(3)`."""
    return Scalar(value, dtype)


def _wrap(x) -> Expr:
    if isinstance(x, Expr):
        return x
    return Scalar(x)


class Func(Expr):
    """xfunction node: functor `op` applied to argument expressions."""

    def __init__(self, op: int, *args, index: int = 0):
        self.op = op
        self.args = tuple(_wrap(a) for a in args)
        self.index = index


# ----------------------------------------------------------------------------
# device arrays
class DeviceArray(Expr):
    """A strided array in device memory (an xt::xarray with HBM storage).

    `storage` is a torch byte tensor that owns the memory; shape, strides
    (in elements) and offset describe this view of it.
    """

    def __init__(self, storage, dtype, shape, strides=None, offset=0):
        self.storage = storage
        self.dtype_code = dtype_code(dtype)
        self._shape = tuple(int(s) for s in shape)
        if strides is None:
            strides = _row_major_strides(self._shape)
        self.strides = tuple(int(s) for s in strides)
        self.offset = int(offset)

    # Expr overrides: a container knows its own shape/type
    @property
    def shape(self) -> tuple:
        return self._shape

    @property
    def dtype(self) -> str:
        return dtype_name(self.dtype_code)

    @property
    def itemsize(self) -> int:
        return A.DTYPE_SIZE[self.dtype_code]

    @property
    def data_ptr(self) -> int:
        return int(self.storage.data_ptr())

    @property
    def is_cuda(self) -> bool:
        return bool(getattr(self.storage, "is_cuda", False))

    def to_abi(self) -> A.Tensor:
        t = A.Tensor()
        t.data = self.data_ptr
        t.dtype = self.dtype_code
        t.ndim = len(self._shape)
        t.offset = self.offset
        for k, (n, s) in enumerate(zip(self._shape, self.strides)):
            t.shape[k] = n
            t.strides[k] = s
        return t

    def to_numpy(self) -> np.ndarray:
        """Copy this view to a host numpy array (bf16 widened to float32)."""
        import torch
        n = self.size
        npdt = numpy_dtype(self.dtype_code)
        if n == 0:
            return np.zeros(self._shape, npdt)
        lo = self.offset + builtins.sum((d - 1) * s for d, s in zip(self._shape, self.strides) if s < 0)
        hi = self.offset + builtins.sum((d - 1) * s for d, s in zip(self._shape, self.strides) if s > 0)
        isz = self.itemsize
        raw = self.storage[lo * isz:(hi + 1) * isz].cpu().numpy()
        if self.dtype_code == A.BFLOAT16:
            u16 = raw.view(np.uint16).astype(np.uint32) << 16
            flat = u16.view(np.float32)
        else:
            flat = raw.view(npdt)
        base = self.offset - lo
        view = np.lib.stride_tricks.as_strided(
            flat[base:], shape=self._shape, strides=tuple(s * flat.itemsize for s in self.strides))
        return np.array(view)

    # ---- numpy-style indexing: a[...] is xt::view / xt::strided_view
    def __getitem__(self, key):
        return view(self, *(key if isinstance(key, tuple) else (key,)))

    def __setitem__(self, key, value):
        assign_checked(self[key], value)

    @property
    def T(self):
        return transpose(self)

    def __repr__(self):
        return f"DeviceArray(shape={self._shape}, dtype={self.dtype}, strides={self.strides}, offset={self.offset})"


def _row_major_strides(shape: Sequence[int]) -> tuple:
    st = [1] * len(shape)
    for k in range(len(shape) - 2, -1, -1):
        st[k] = st[k + 1] * max(int(shape[k + 1]), 1)
    return tuple(st)


def _torch_device(device):
    return device or "cuda"


def empty(shape, dtype="float64", device=None) -> DeviceArray:
    """Uninitialised device array (xt::xarray<T>::from_shape)."""
    import torch
    shape = tuple(int(s) for s in (shape if isinstance(shape, (tuple, list)) else (shape,)))
    code = dtype_code(dtype)
    n = int(np.prod(shape, dtype=np.int64)) if shape else 1
    storage = torch.empty(max(n * A.DTYPE_SIZE[code], 1), dtype=torch.uint8, device=_torch_device(device))
    return DeviceArray(storage, code, shape)


def zeros(shape, dtype="float64", device=None) -> DeviceArray:
    a = empty(shape, dtype, device)
    a.storage.zero_()
    return a


def from_numpy(x, dtype=None, device=None) -> DeviceArray:
    """Copy a host array to the device (row-major).  dtype='bfloat16' rounds to bf16."""
    import torch
    x = np.asarray(x)
    code = dtype_code(dtype) if dtype is not None else dtype_code(x.dtype)
    if code == A.BFLOAT16:
        t = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float32)).to(torch.bfloat16)
        raw = t.view(torch.int16).numpy().view(np.uint8).reshape(-1)
    else:
        raw = np.ascontiguousarray(x, dtype=_NP_OF[code]).reshape(-1).view(np.uint8)
    storage = torch.empty(max(raw.size, 1), dtype=torch.uint8, device=_torch_device(device))
    if raw.size:
        storage[:raw.size].copy_(torch.from_numpy(raw.copy()))
    return DeviceArray(storage, code, x.shape)


def asarray(x, dtype=None, device=None) -> DeviceArray:
    return x if isinstance(x, DeviceArray) else from_numpy(x, dtype, device)


# ----------------------------------------------------------------------------
# views (views/xview.hpp, views/xslice.hpp, views/xstrided_view.hpp)
class _All:
    def __repr__(self):
        return "all()"


class _NewAxis:
    def __repr__(self):
        return "newaxis()"


class range_:
    """xt::range(start, stop[, step]); None stands for xt::placeholders::_."""

    def __init__(self, start=None, stop=None, step=None):
        self.start, self.stop, self.step = start, stop, step

    def as_slice(self):
        return slice(self.start, self.stop, self.step)


def all_():
    return _All()


def newaxis():
    return _NewAxis()


def _norm_key(a: DeviceArray, keys) -> list:
    """Expand Ellipsis and pad with all() (xt::view keeps trailing dims)."""
    keys = list(keys)
    n_ell = builtins.sum(1 for k in keys if k is Ellipsis)
    if n_ell > 1:
        raise IndexError("only one Ellipsis allowed")
    consumed = builtins.sum(1 for k in keys if not isinstance(k, _NewAxis) and k is not None and k is not Ellipsis)
    if consumed > a.ndim:
        raise IndexError("too many indices")
    if n_ell:
        i = keys.index(Ellipsis)
        keys[i:i + 1] = [_All()] * (a.ndim - consumed)
    else:
        keys += [_All()] * (a.ndim - consumed)
    return keys


def view(a: DeviceArray, *keys) -> DeviceArray:
    """xt::view / xt::strided_view over a container (no copy).

    Keys: int (drops the dim, views/xslice.hpp integral case), slice or
    range_ (start/stop/step with Python's negative-index rules), all_(),
    newaxis() / None, Ellipsis.
    """
    if not isinstance(a, DeviceArray):
        raise TypeError("views are taken of device arrays")
    keys = _norm_key(a, keys)
    shape, strides = [], []
    offset = a.offset
    dim = 0
    for k in keys:
        if isinstance(k, _NewAxis) or k is None:
            shape.append(1)
            strides.append(0)
            continue
        n, s = a.shape[dim], a.strides[dim]
        if isinstance(k, _All):
            shape.append(n)
            strides.append(s)
        elif isinstance(k, (slice, range_)):
            sl = k.as_slice() if isinstance(k, range_) else k
            start, stop, step = sl.indices(n)
            cnt = len(range(start, stop, step))
            if cnt > 0:
                offset += start * s
            shape.append(cnt)
            strides.append(s * step)
        elif isinstance(k, (numbers.Integral, np.integer)):
            i = int(k)
            if i < 0:
                i += n
            if not 0 <= i < n:
                raise IndexError(f"index {k} out of range for dimension of size {n}")
            offset += i * s
        else:
            raise TypeError(f"unsupported view key {k!r}")
        dim += 1
    return DeviceArray(a.storage, a.dtype_code, shape, strides, offset)


strided_view = view


def transpose(a: DeviceArray, perm: Sequence[int] | None = None) -> DeviceArray:
    """xt::transpose: permuted (default: reversed) shape and strides."""
    if perm is None:
        perm = list(range(a.ndim))[::-1]
    if sorted(perm) != list(range(a.ndim)):
        raise RuntimeError("transpose: permutation is not a permutation of the axes")
    return DeviceArray(a.storage, a.dtype_code, [a.shape[p] for p in perm], [a.strides[p] for p in perm], a.offset)


def broadcast(a, shape: Sequence[int]):
    """xt::broadcast of a container to `shape`: stride-0 view.  A scalar or an
    expression is returned as it is: assignment broadcasts it to the target
    (views/xbroadcast.hpp:157-480 makes the same values available lazily)."""
    if not isinstance(a, DeviceArray):
        return _wrap(a)
    shape = tuple(int(s) for s in shape)
    if len(shape) < a.ndim:
        raise BroadcastError("broadcast: target rank smaller than array rank")
    off = len(shape) - a.ndim
    st = [0] * len(shape)
    for k in range(a.ndim):
        n = a.shape[k]
        if n == shape[off + k]:
            st[off + k] = a.strides[k] if n != 1 else 0
        elif n == 1:
            st[off + k] = 0
        else:
            raise BroadcastError(f"cannot broadcast {a.shape} to {shape}")
    return DeviceArray(a.storage, a.dtype_code, shape, st, a.offset)


def reshape_view(a: DeviceArray, shape: Sequence[int]) -> DeviceArray:
    """xt::reshape_view of a row-major contiguous array."""
    if a.strides != _row_major_strides(a.shape):
        raise RuntimeError("reshape_view needs a row-major contiguous array")
    shape = tuple(int(s) for s in shape)
    if int(np.prod(shape, dtype=np.int64)) != a.size:
        raise RuntimeError("reshape_view: size mismatch")
    return DeviceArray(a.storage, a.dtype_code, shape, None, a.offset)


# ----------------------------------------------------------------------------
# functions (core/xmath.hpp, core/xoperation.hpp)
def _unary(name):
    code = A.OP[name.upper()]

    def f(e):
        return Func(code, e)
    f.__name__ = name
    f.__doc__ = f"xt::{name}(e): std::{name} elementwise"
    return f


for _n in ["abs", "fabs", "exp", "exp2", "expm1", "log", "log2", "log10", "log1p", "sqrt", "cbrt", "sin", "cos",
           "tan", "tanh", "sinh", "cosh", "asin", "acos", "atan", "erf", "floor", "ceil", "trunc", "round"]:
    globals()[_n] = _unary(_n)


def maximum(a, b): return Func(A.OP["MAXIMUM"], a, b)
def minimum(a, b): return Func(A.OP["MINIMUM"], a, b)
def pow(a, b): return Func(A.OP["POW"], a, b)  # noqa: A001
def fmod(a, b): return Func(A.OP["FMOD"], a, b)
def atan2(a, b): return Func(A.OP["ATAN2"], a, b)
def hypot(a, b): return Func(A.OP["HYPOT"], a, b)
def where(c, a, b): return Func(A.OP["WHERE"], c, a, b)
def clip(e, lo, hi): return Func(A.OP["CLIP"], e, lo, hi)
def logical_and(a, b): return Func(A.OP["LOGICAL_AND"], a, b)
def logical_or(a, b): return Func(A.OP["LOGICAL_OR"], a, b)
def logical_not(a): return Func(A.OP["LOGICAL_NOT"], a)
def equal(a, b): return Func(A.OP["EQUAL_TO"], a, b)
def not_equal(a, b): return Func(A.OP["NOT_EQUAL_TO"], a, b)
def less(a, b): return Func(A.OP["LESS"], a, b)
def greater(a, b): return Func(A.OP["GREATER"], a, b)


def cast(e, dtype) -> Func:
    """xt::cast<T>(e): static_cast of every element."""
    return Func(A.OP["CAST"], e, index=dtype_code(dtype))


# ----------------------------------------------------------------------------
# reducers (reducers/xreducer.hpp, core/xmath.hpp, reducers/xnorm.hpp)
class Reducer(Expr):
    """Lazy reduction; evaluated by `eval`, assignment, or use in an expression."""

    def __init__(self, kind, e, axes=None, keep_dims=False, initial=None, dtype=None, ddof=0,
                 evaluation_strategy="lazy"):
        self.kind = kind
        self.ddof = int(ddof)
        if evaluation_strategy not in ("lazy", "immediate"):
            raise ValueError("evaluation_strategy is 'lazy' or 'immediate'")
        # xt::evaluation_strategy::immediate (reducers/xreducer.hpp:944-952):
        # the expression is evaluated first, then reduced
        self.immediate = evaluation_strategy == "immediate"
        self.expr = _wrap(e)
        if axes is not None and isinstance(axes, (numbers.Integral, np.integer)):
            axes = [int(axes)]
        self.axes = None if axes is None else [int(x) for x in axes]
        self.keep_dims = bool(keep_dims)
        self.initial = None if initial is None else _wrap(initial)
        self.acc_dtype = -1 if dtype is None else dtype_code(dtype)

    def to_abi(self) -> A.ReduceDesc:
        d = A.ReduceDesc()
        d.op = A.RED[self.kind]
        d.acc_dtype = self.acc_dtype
        if self.axes is None:
            d.n_axes = -1
        else:
            if len(self.axes) > A.MAX_RANK:
                raise RuntimeError("too many axes")
            d.n_axes = len(self.axes)
            for i, x in enumerate(self.axes):
                d.axes[i] = x
        d.keep_dims = int(self.keep_dims)
        if self.initial is not None:
            d.has_initial = 1
            d.initial = self.initial.to_abi()
        d.ddof = self.ddof
        d.flags = A.REDUCE_IMMEDIATE if self.immediate else 0
        return d

    def _info(self):
        return reduce_info(self)


def _reducer(kind):
    def f(e, axes=None, keep_dims=False, initial=None, dtype=None, evaluation_strategy="lazy"):
        return Reducer(kind, e, axes, keep_dims, initial, dtype, evaluation_strategy=evaluation_strategy)
    f.__name__ = kind
    f.__doc__ = f"xt::{kind}(e, axes, options)"
    return f


for _k in A.RED:
    globals()[_k] = _reducer(_k)


def variance(e, axes=None, ddof=0, keep_dims=False, dtype=None, evaluation_strategy="lazy"):
    """xt::variance<T>(e, axes, ddof) (core/xmath.hpp:2031-2113): mean<T> of
    square(e - mean<T>(e, axes)), divided by (size - ddof) / outputs."""
    return Reducer("variance", e, axes, keep_dims, None, dtype, ddof, evaluation_strategy)


def stddev(e, axes=None, keep_dims=False, dtype=None, evaluation_strategy="lazy"):
    """xt::stddev<T>(e, axes) = sqrt(variance<T>(e, axes)) (core/xmath.hpp:2046, 2143)."""
    return Reducer("stddev", e, axes, keep_dims, None, dtype, evaluation_strategy=evaluation_strategy)


def minmax(e):
    """xt::minmax(e) (core/xmath.hpp:2195-2230): [min, max] of every element,
    folded with std::min / std::max from {max(), lowest()} of the value type."""
    return Reducer("minmax", e)


# ----------------------------------------------------------------------------
# serialisation to the C ABI
class Serialized:
    """An xg_expr with the buffers it points into kept alive."""

    def __init__(self, nodes, leaves, consts, keep):
        self.nodes = (A.Node * max(len(nodes), 1))(*nodes)
        self.leaves = (A.Tensor * max(len(leaves), 1))(*leaves)
        self.consts = (A.Scalar * max(len(consts), 1))(*consts)
        self.keep = keep
        self.expr = A.Expr(len(nodes), len(leaves), len(consts), 0, self.nodes, self.leaves, self.consts)

    def ptr(self):
        return C.byref(self.expr)


def serialize(root: Expr, leaf_fn: Callable[[DeviceArray], A.Tensor] | None = None,
              reducer_fn: Callable[[Reducer], Any] | None = None) -> Serialized:
    """Flatten a tree into xg_expr nodes (arguments before users).

    The same DeviceArray object used twice becomes one leaf, the same
    sub-expression object used twice one node.  Reducers inside a tree are
    evaluated first by `reducer_fn` (a lazy reducer yields the same value at
    every element it is read from).
    """
    nodes: list = []
    leaves: list = []
    consts: list = []
    keep: list = []
    leaf_of: dict = {}
    memo: dict = {}
    leaf_fn = leaf_fn or (lambda a: a.to_abi())

    def add(op, args=(), index=0):
        n = A.Node()
        n.op = op
        n.nargs = len(args)
        for i, x in enumerate(args):
            n.args[i] = x
        n.index = index
        nodes.append(n)
        if len(nodes) > A.MAX_NODES:
            raise NotImplementedError("expression has too many nodes")
        return len(nodes) - 1

    def leaf_slot(arr):
        k = id(arr)
        if k not in leaf_of:
            leaf_of[k] = len(leaves)
            leaves.append(leaf_fn(arr))
            keep.append(arr)
        return leaf_of[k]

    def leaf(arr):
        return add(A.OP["LEAF"], (), leaf_slot(arr))

    def visit(e):
        k = id(e)
        if k in memo:
            return memo[k]
        if isinstance(e, DeviceArray):
            r = leaf(e)
        elif isinstance(e, IndexView):  # a gather of the source at the offsets' values
            off = visit(e.offsets)
            r = add(A.OP["GATHER"], (off,), leaf_slot(e.source))
        elif isinstance(e, Reducer):
            if reducer_fn is None:
                raise RuntimeError("reducer inside an expression needs an evaluation context")
            tmp = reducer_fn(e)
            keep.append(tmp)
            r = leaf(tmp)
        elif isinstance(e, Scalar):
            consts.append(e.to_abi())
            r = add(A.OP["CONST"], (), len(consts) - 1)
        elif isinstance(e, Func):
            args = [visit(x) for x in e.args]
            r = add(e.op, args, e.index)
        else:
            raise TypeError(f"not an expression: {e!r}")
        memo[k] = r
        keep.append(e)
        return r

    visit(root)
    return Serialized(nodes, leaves, consts, keep)


def _shape_of(ndim, shp):
    return tuple(int(shp[i]) for i in range(ndim))


def expr_info(e: Expr):
    """(dtype code, broadcast shape) of an expression, from the library."""
    s = serialize(e, reducer_fn=lambda r: _placeholder_for(r))
    dt, nd = C.c_int32(), C.c_int32()
    shp = (C.c_int64 * A.MAX_RANK)()
    _check(A.lib().xg_expr_info(s.ptr(), C.byref(dt), C.byref(nd), shp))
    return dt.value, _shape_of(nd.value, shp)


class _Placeholder(DeviceArray):
    """Shape/type stand-in for an unevaluated reducer (no storage)."""

    class _NoStorage:
        def data_ptr(self):
            return 0

    def __init__(self, dtype, shape):
        super().__init__(self._NoStorage(), dtype, shape)


def _placeholder_for(r: Reducer):
    dt, shp = reduce_info(r)
    return _Placeholder(dt, shp)


def reduce_info(r: Reducer):
    s = serialize(r.expr, reducer_fn=lambda x: _placeholder_for(x))
    d = r.to_abi()
    dt, nd = C.c_int32(), C.c_int32()
    shp = (C.c_int64 * A.MAX_RANK)()
    _check(A.lib().xg_reduce_info(s.ptr(), C.byref(d), C.byref(dt), C.byref(nd), shp))
    return dt.value, _shape_of(nd.value, shp)


# ----------------------------------------------------------------------------
# evaluation
def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _require_device(arrs: Iterable[DeviceArray]):
    for a in arrs:
        if isinstance(a, DeviceArray) and not isinstance(a, _Placeholder) and not a.is_cuda:
            raise RuntimeError("evaluation needs device (cuda) arrays")


def _leaves(e) -> list:
    out = []

    def walk(x):
        if isinstance(x, DeviceArray):
            out.append(x)
        elif isinstance(x, Func):
            for y in x.args:
                walk(y)
        elif isinstance(x, Reducer):
            walk(x.expr)
        elif isinstance(x, IndexView):
            out.append(x.source)
            walk(x.offsets)
    walk(e)
    return out


def _eval_reducer(r: Reducer) -> DeviceArray:
    dt, shp = reduce_info(r)
    out = empty(shp, dt)
    _reduce_into(r, out)
    return out


def _reduce_into(r: Reducer, out: DeviceArray) -> None:
    _require_device(_leaves(r.expr) + [out])
    s = serialize(r.expr, reducer_fn=_eval_reducer)
    d = r.to_abi()
    t = out.to_abi()
    _check(A.lib().xg_reduce(s.ptr(), C.byref(d), C.byref(t), _stream()))


def assign(out, e):
    """xt::noalias(out) = e (or assignment into a view).  Returns out."""
    e = _wrap(e)
    if isinstance(out, IndexView):
        return out.assign(e)
    if isinstance(e, Reducer):
        _reduce_into(e, out)
        return out
    _require_device(_leaves(e) + [out])
    s = serialize(e, reducer_fn=_eval_reducer)
    t = out.to_abi()
    _check(A.lib().xg_assign(s.ptr(), C.byref(t), _stream()))
    return out


def assign_reduce(out: DeviceArray, e, rows=None, cols=None) -> DeviceArray:
    """noalias(out) = e for a 2-D row-major `out`, and in the same pass over
    the stored values the reduction of every row and / or every column:
    rows=("sum", s) puts sum(out, {1}) into s, cols=("amax", m) puts
    amax(out, {0}) into m (a third item, a dict, passes keep_dims= /
    initial= / dtype=).
    The results are those of the three statements
        out = e; s = sum(out, {1}); m = amax(out, {0})
    (core/xassign.hpp:208-226, reducers/xreducer.hpp:1778-1828), which the
    library runs instead where the single pass does not apply.  Returns out."""
    e = _wrap(e)
    s, args, _ = _assign_reduce_args(out, e, rows, cols)
    _check(A.lib().xg_assign_reduce2d(s.ptr(), *args, _stream()))
    return out


def _assign_reduce_args(out, e, rows, cols):
    """The serialized expression and the C arguments (out, row desc, row out,
    col desc, col out) of xg_assign_reduce2d / xg_plan_assign_reduce2d, and
    the Python objects they point into."""
    def desc(spec, axis):
        if spec is None:
            return None, None
        kind, dst = spec[0], spec[1]
        kw = dict(spec[2]) if len(spec) > 2 else {}
        r = Reducer(kind, out, axes=[axis], keep_dims=kw.get("keep_dims", False), initial=kw.get("initial"),
                    dtype=kw.get("dtype"))
        if not isinstance(dst, DeviceArray):
            raise TypeError("a reduction's output must be a DeviceArray")
        return r.to_abi(), dst

    rd, ro = desc(rows, 1)
    cd, co = desc(cols, 0)
    _require_device(_leaves(e) + [out] + [x for x in (ro, co) if x is not None])
    s = serialize(e, reducer_fn=_eval_reducer)
    t = out.to_abi()
    rt = ro.to_abi() if ro is not None else None
    ct = co.to_abi() if co is not None else None
    ref = lambda x: C.byref(x) if x is not None else None  # noqa: E731
    keep = [t, rd, rt, cd, ct, ro, co]
    return s, (ref(t), ref(rd), ref(rt), ref(cd), ref(ct)), keep


def _same_view(a: DeviceArray, b: DeviceArray) -> bool:
    return (a.storage is b.storage and a.offset == b.offset and a.shape == b.shape
            and all(s1 == s2 for n, s1, s2 in zip(a.shape, a.strides, b.strides) if n > 1))


def assign_checked(out: DeviceArray, e) -> DeviceArray:
    """`out = e` (xsemantic operator=, core/xsemantic.hpp:620-640): when an
    operand shares memory with `out` other than element for element, the
    expression is first evaluated into a temporary."""
    e = _wrap(e)
    for leaf in _leaves(e):
        if leaf.storage is out.storage and not _same_view(leaf, out):
            e = eval(e)
            break
    return assign(out, e)


class noalias:
    """xt::noalias(out): assignment without a temporary (core/xnoalias.hpp)."""

    def __init__(self, target: DeviceArray):
        self.target = target

    def assign(self, e):
        return assign(self.target, e)

    __lshift__ = assign

    def plus_assign(self, e): return assign(self.target, self.target + e)
    def minus_assign(self, e): return assign(self.target, self.target - e)
    def multiplies_assign(self, e): return assign(self.target, self.target * e)
    def divides_assign(self, e): return assign(self.target, self.target / e)
    def modulus_assign(self, e): return assign(self.target, self.target % e)
    def bit_and_assign(self, e): return assign(self.target, self.target & e)
    def bit_or_assign(self, e): return assign(self.target, self.target | e)
    def bit_xor_assign(self, e): return assign(self.target, self.target ^ e)


def eval(e) -> DeviceArray:  # noqa: A001
    """xt::eval: materialise an expression into a new row-major array."""
    e = _wrap(e)
    if isinstance(e, DeviceArray) and e.offset == 0 and e.strides == _row_major_strides(e.shape):
        return e  # already a row-major container
    if isinstance(e, Reducer):
        return _eval_reducer(e)
    dt, shp = expr_info(e)
    out = empty(shp, dt)
    return assign(out, e)


# ----------------------------------------------------------------------------
# index views, filters, filtrations, masked views (views/xindex_view.hpp,
# views/xmasked_view.hpp)
class IndexView(Expr):
    """xt::index_view(e, indices) (views/xindex_view.hpp:786-797): the 1-d
    sequence of e's elements at `indices` — multi-indices of shape
    [n, e.ndim] (or positions [n] for a 1-d e), on the host or the device.
    Read as an operand it is a gather (one fused launch with the rest of the
    expression); assigned into, it scatters.  Indices outside e read as 0 and
    are skipped on assignment (xtensor leaves them undefined)."""

    def __init__(self, source, indices):
        if not isinstance(source, DeviceArray):  # a function: index its evaluated values (read-only)
            source = eval(_wrap(source))
        self.source = source
        nd = len(source.shape)
        if isinstance(indices, DeviceArray):
            idx = indices
            shape = tuple(idx.shape)
        else:
            h = np.asarray(indices, dtype=np.int64)
            if h.ndim == 1 and nd != 1:
                h = h.reshape(-1, nd) if h.size % max(nd, 1) == 0 and h.size else h.reshape(0, nd)
            idx = from_numpy(h)
            shape = h.shape
        self.index = idx  # the index array as given (or uploaded)
        if len(shape) == 1:
            if nd != 1:
                raise ValueError("positions need a 1-d source; give [n, ndim] multi-indices")
            shape = (shape[0], 1)
            idx = DeviceArray(idx.storage, idx.dtype, shape, (idx.strides[0], 0), idx.offset)
        if shape[1] != nd:
            raise ValueError(f"indices have {shape[1]} components for a {nd}-d source")
        self.n = int(shape[0])
        # offset of element i: sum over k of indices[i, k] * stride_k
        off = None
        for k in range(nd):
            col = DeviceArray(idx.storage, idx.dtype, (self.n,), (idx.strides[0],), idx.offset + k * idx.strides[1])
            col.parent = self.index  # a view of the index array (lets a checker map it to host data)
            term = col * scalar(int(source.strides[k]), "int64")
            off = term if off is None else off + term
        self.offsets = off if off is not None else scalar(0, "int64")
        if nd == 0:
            raise ValueError("index_view of a 0-d array")

    def fill(self, value):
        """xindex_view::fill: every indexed element = value."""
        return self.assign(value)

    def assign(self, e):
        """index_view(e, indices) = values: a scatter of the evaluated values."""
        e = _wrap(e)
        _require_device(_leaves(e) + _leaves(self.offsets) + [self.source])
        sv = serialize(e, reducer_fn=_eval_reducer)
        so = serialize(self.offsets)
        t = self.source.to_abi()
        _check(A.lib().xg_assign_indexed(sv.ptr(), so.ptr(), C.byref(t), _stream()))
        return self


def index_view(e, indices) -> IndexView:
    return IndexView(e, indices)


def functor_view(e, f):
    """xt::xfunctor_view<F, E> (views/xfunctor_view.hpp:400-470): e's elements
    seen through the functor f, applied lazily when the view is read; f maps
    an expression to an expression (f = lambda x: xt.abs(x), ...), so the
    view is one more node of the fused expression.  xfunctor_view is an
    lvalue only for functors that return a reference into the element (the
    real / imaginary parts of complex values, views/xcomplex.hpp); with the
    real element types here, real(e) is that view and is assignable."""
    return f(_wrap(e))


def real(e):
    """xt::real(e) of a real-valued expression: e itself (an lvalue view of a
    container, as xt::real's functor view is)."""
    return e if isinstance(e, DeviceArray) else _wrap(e)


def imag(e):
    """xt::imag(e) of a real-valued expression: zeros of e's shape and value
    type (read-only), also where e is NaN."""
    e = _wrap(e)
    dt, _ = expr_info(e)
    return where(scalar(True, "bool"), scalar(0, dtype_name(dt)), e)


def count_nonzero(e) -> int:
    """xt::count_nonzero(e) (core/xmath.hpp:2495-2505): the number of
    elements that convert to true (a device reduction; the count is read
    back)."""
    return int(eval(sum(cast(cast(_wrap(e), "bool"), "int64"))).to_numpy())


def argwhere(cond, layout="row_major") -> DeviceArray:
    """xt::argwhere<L> (core/xoperation.hpp:893-930): int64 [count, ndim]
    multi-indices of the true elements in row-major (or column-major)
    traversal order, listed on the device (xg_argwhere: a scan of the flags
    and one compaction kernel); only the count is read back, to size the
    result."""
    c = _wrap(cond)
    if layout not in ("row_major", "column_major"):
        raise ValueError("layout is 'row_major' or 'column_major'")
    _, shp = expr_info(c)
    n = count_nonzero(c)
    out = empty((n, len(shp)), "int64")
    if n:
        _require_device(_leaves(c) + [out])
        s = serialize(c, reducer_fn=_eval_reducer)
        t = out.to_abi()
        _check(A.lib().xg_argwhere(s.ptr(), 1 if layout == "column_major" else 0, C.byref(t), _stream()))
    return out


def filter(e, cond, layout="row_major") -> IndexView:  # noqa: A001
    """xt::filter<L>(e, cond) = index_view(e, argwhere<L>(cond)) (views/xindex_view.hpp:820-826)."""
    return IndexView(e, argwhere(cond, layout))


class Filtration:
    """xt::filtration(e, cond) (views/xindex_view.hpp:845-850): scalar (and
    computed) assignment to the elements of e where cond holds — on the
    device e = where(cond, value, e), one fused launch, e read and written
    element by element."""

    def __init__(self, e, cond):
        if not isinstance(e, DeviceArray):
            raise TypeError("filtration needs a device array or a strided view of one")
        self.e = e
        self.cond = _wrap(cond)

    def assign(self, v):
        return assign(self.e, where(self.cond, _wrap(v), self.e))

    def __iadd__(self, v):
        assign(self.e, where(self.cond, self.e + _wrap(v), self.e))
        return self

    def __isub__(self, v):
        assign(self.e, where(self.cond, self.e - _wrap(v), self.e))
        return self

    def __imul__(self, v):
        assign(self.e, where(self.cond, self.e * _wrap(v), self.e))
        return self

    def __itruediv__(self, v):
        assign(self.e, where(self.cond, self.e / _wrap(v), self.e))
        return self


def filtration(e, cond) -> Filtration:
    return Filtration(e, cond)


def masked_view(e, mask) -> Filtration:
    """xt::masked_view(e, mask) (views/xmasked_view.hpp) for assignment:
    `masked_view(a, m).assign(v)` sets a's elements where the (broadcast)
    mask is true; reading masked values is not offered on the device."""
    return Filtration(e, mask)


# ----------------------------------------------------------------------------
# .npy files (io/xnpy.hpp), straight to and from device memory
def npy_info(path):
    """(dtype name, shape, fortran_order) of an .npy file's header."""
    dt, nd, fo = C.c_int32(), C.c_int32(), C.c_int32()
    shp = (C.c_int64 * A.MAX_RANK)()
    _check(A.lib().xg_npy_info(os.fsencode(path), C.byref(dt), C.byref(nd), shp, C.byref(fo)))
    return dtype_name(dt.value), tuple(shp[k] for k in range(nd.value)), bool(fo.value)


def load_npy(path, dtype=None) -> DeviceArray:
    """xt::load_npy<T>(path) (io/xnpy.hpp:775-800): the file's data in a new
    device array; a Fortran-order file gives a column-major array.  `dtype`,
    if given, must be the file's ("Cast error: formats not matching")."""
    dt, shape, fortran = npy_info(path)
    if dtype is not None and dtype_code(dtype) != dtype_code(dt):
        raise RuntimeError(f"Cast error: formats not matching: file holds {dt}, asked for {dtype_name(dtype_code(dtype))}")
    a = empty(shape, dt)
    if fortran:
        st, s = [], 1
        for n in shape:
            st.append(s)
            s *= max(int(n), 1)
        a = DeviceArray(a.storage, a.dtype, shape, tuple(st), 0)
    t = a.to_abi()
    _check(A.lib().xg_npy_load(os.fsencode(path), C.byref(t), _stream()))
    return a


def dump_npy(path, e) -> None:
    """xt::dump_npy(path, e) (io/xnpy.hpp:709-755): evaluates e and writes a
    C-order .npy file with the reference's header layout."""
    a = e if isinstance(e, DeviceArray) else eval(e)
    _require_device([a])
    t = a.to_abi()
    _check(A.lib().xg_npy_dump(os.fsencode(path), C.byref(t), _stream()))


# ----------------------------------------------------------------------------
# accumulators (reducers/xaccumulator.hpp) — immediate, like xtensor's
class Scan:
    """A scan to prepare with `plan(out, Scan(...))`: cumsum / cumprod of e
    along `axis` (a plan needs an axis; the flattened form runs through
    cumsum(e) directly)."""

    def __init__(self, kind, e, axis, dtype=None):
        self.kind = kind
        self.expr = _wrap(e)
        self.axis = int(axis)
        self.dtype = dtype

    def to_abi(self) -> A.ScanDesc:
        d = A.ScanDesc()
        d.op = A.SCAN[self.kind]
        d.acc_dtype = -1 if self.dtype is None else dtype_code(self.dtype)
        d.axis = self.axis
        return d


def _scan(kind, e, axis=None, dtype=None) -> DeviceArray:
    e = _wrap(e)
    d = A.ScanDesc()
    d.op = A.SCAN[kind]
    d.acc_dtype = -1 if dtype is None else dtype_code(dtype)
    d.axis = A.AXIS_NONE if axis is None else int(axis)
    s = serialize(e, reducer_fn=_eval_reducer)
    dt, nd = C.c_int32(), C.c_int32()
    shp = (C.c_int64 * A.MAX_RANK)()
    _check(A.lib().xg_scan_info(s.ptr(), C.byref(d), C.byref(dt), C.byref(nd), shp))
    out = empty(_shape_of(nd.value, shp), dt.value)
    _require_device(_leaves(e) + [out])
    t = out.to_abi()
    _check(A.lib().xg_scan(s.ptr(), C.byref(d), C.byref(t), _stream()))
    return out


def cumsum(e, axis=None, dtype=None) -> DeviceArray:
    """xt::cumsum(e[, axis]) (core/xmath.hpp:2247-2266)."""
    return _scan("cumsum", e, axis, dtype)


def cumprod(e, axis=None, dtype=None) -> DeviceArray:
    """xt::cumprod(e[, axis])."""
    return _scan("cumprod", e, axis, dtype)


class Plan:
    """A prepared evaluation (xg_plan_*): planned once, executed many times.

    `plan(out, e)` prepares `noalias(out) = e` (or a reduction into out);
    `run()` enqueues it on the current stream with ~no host work.  The plan
    keeps its operands alive and reads whatever they hold at run time.
    """

    def __init__(self, out: DeviceArray, e, rows=None, cols=None):
        e = e if isinstance(e, Scan) else _wrap(e)
        self._keep = [out, e]
        h = C.c_void_p()
        L = A.lib()
        if rows is not None or cols is not None:  # out = e with its row / column reductions in one pass
            s, args, keep = _assign_reduce_args(out, e, rows, cols)
            self._keep += keep
            _check(L.xg_plan_assign_reduce2d(s.ptr(), *args, C.byref(h)))
        elif isinstance(e, Scan):
            _require_device(_leaves(e.expr) + [out])
            s = serialize(e.expr, reducer_fn=_eval_reducer)
            d = e.to_abi()
            t = out.to_abi()
            _check(L.xg_plan_scan(s.ptr(), C.byref(d), C.byref(t), C.byref(h)))
        elif isinstance(e, Reducer):
            _require_device(_leaves(e.expr) + [out])
            s = serialize(e.expr, reducer_fn=_eval_reducer)
            d = e.to_abi()
            t = out.to_abi()
            _check(L.xg_plan_reduce(s.ptr(), C.byref(d), C.byref(t), C.byref(h)))
        else:
            _require_device(_leaves(e) + [out])
            s = serialize(e, reducer_fn=_eval_reducer)
            t = out.to_abi()
            _check(L.xg_plan_assign(s.ptr(), C.byref(t), C.byref(h)))
        self._keep.append(s)
        self.handle = h
        self._exec = L.xg_plan_execute

    def run(self, stream=None):
        st = _stream() if stream is None else C.c_void_p(stream)
        r = self._exec(self.handle, None, st)
        if r:
            _check(r)

    def __del__(self):
        try:
            if self.handle:
                A.lib().xg_plan_destroy(self.handle)
        except Exception:
            pass


def plan(out: DeviceArray, e, rows=None, cols=None) -> Plan:
    """Prepare noalias(out) = e (or a reduction / scan into out); with rows= /
    cols= as for assign_reduce, out = e with those reductions in one pass."""
    return Plan(out, e, rows, cols)


def execute_many(plans: Sequence[Plan], stream=None) -> None:
    """Enqueue independent plans; elementwise plans sharing a compiled program
    and rank run as one fused launch (xg_plan_execute_many)."""
    arr = (C.c_void_p * max(1, len(plans)))(*[p.handle.value for p in plans])
    st = _stream() if stream is None else C.c_void_p(stream)
    _check(A.lib().xg_plan_execute_many(arr, len(plans), st))


def last_kernel() -> str:
    """Kernel family the last call on this thread ran (diagnostic)."""
    f = getattr(A.lib(), "xg_last_kernel", None)
    return f().decode() if f else ""

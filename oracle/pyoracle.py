"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke test and the CPU-baseline
leg of bench.py as the checker.  The product (xtensor_amd) never imports it.

The oracle consumes the same expression trees the front end builds; each
device leaf is replaced by the host numpy array the caller maps it to, so the
oracle evaluates on host data with its own broadcasting, typing and
evaluation order (see oracle.cpp).
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from xtensor_amd import _abi as A
from xtensor_amd import core

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle.so")
_lib = None

_NP = {A.BOOL: np.bool_, A.INT8: np.int8, A.UINT8: np.uint8, A.INT16: np.int16, A.UINT16: np.uint16,
       A.INT32: np.int32, A.UINT32: np.uint32, A.INT64: np.int64, A.UINT64: np.uint64,
       A.FLOAT32: np.float32, A.FLOAT64: np.float64, A.BFLOAT16: np.uint16}


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run `make -C oracle`")
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        L.xo_last_error.restype = C.c_char_p
        for name, extra in [("xo_expr_info", [P(C.c_int32), P(C.c_int32), P(C.c_int64)]),
                            ("xo_assign", [P(A.Tensor)]),
                            ("xo_reduce_info", [P(A.ReduceDesc), P(C.c_int32), P(C.c_int32), P(C.c_int64)]),
                            ("xo_reduce", [P(A.ReduceDesc), P(A.Tensor)]),
                            ("xo_reduce_hp", [P(A.ReduceDesc), P(A.Tensor)]),
                            ("xo_scan_info", [P(A.ScanDesc), P(C.c_int32), P(C.c_int32), P(C.c_int64)]),
                            ("xo_scan", [P(A.ScanDesc), P(A.Tensor)]),
                            ("xo_scan_hp", [P(A.ScanDesc), P(A.Tensor)])]:
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = [P(A.Expr)] + extra
        _lib = L
    return _lib


class OracleError(Exception):
    def __init__(self, code, msg):
        super().__init__(f"oracle status {code}: {msg}")
        self.code = code


def _check(st):
    if st != 0:
        raise OracleError(st, lib().xo_last_error().decode())


def host_tensor(arr: np.ndarray, dtype: int) -> A.Tensor:
    """xg_tensor over a host numpy array (strides converted to elements)."""
    t = A.Tensor()
    t.data = arr.ctypes.data
    t.dtype = dtype
    t.ndim = arr.ndim
    t.offset = 0
    isz = arr.itemsize
    for k in range(arr.ndim):
        if arr.strides[k] % isz:
            raise ValueError("strides must be whole elements")
        t.shape[k] = arr.shape[k]
        t.strides[k] = arr.strides[k] // isz
    return t


class Oracle:
    """Evaluates front-end expressions on host arrays.

    `leaves` maps each DeviceArray (by object) to the numpy array holding
    its values on the host (bf16 leaves: uint16 bit patterns).
    """

    def __init__(self, leaves: dict):
        self.leaves = {id(k): v for k, v in leaves.items()}
        self._keep = []

    def _leaf(self, a):
        if isinstance(a, tuple):  # (numpy array, dtype code) of an evaluated reducer
            arr, code = a
            self._keep.append(arr)
            return host_tensor(arr, code)
        arr = self.leaves.get(id(a))
        if arr is None and id(getattr(a, "parent", None)) in self.leaves:
            # a view the front end derived from a registered dense array: the
            # same element offsets over that array's host copy
            base = np.ascontiguousarray(self.leaves[id(a.parent)]).reshape(-1)
            start = a.offset - a.parent.offset
            arr = np.lib.stride_tricks.as_strided(base[start:], shape=a.shape,
                                                  strides=[s * base.itemsize for s in a.strides])
        if arr is None:
            raise KeyError(f"no host data registered for {a!r}")
        self._keep.append(arr)
        return host_tensor(arr, a.dtype_code)

    def _reducer(self, r):
        arr = self.reduce(r)
        dt, _ = self.reduce_info(r)
        return (arr, dt)

    def serialize(self, e):
        return core.serialize(core._wrap(e), leaf_fn=self._leaf, reducer_fn=self._reducer)

    def info(self, e):
        s = self.serialize(e)
        dt, nd = C.c_int32(), C.c_int32()
        shp = (C.c_int64 * A.MAX_RANK)()
        _check(lib().xo_expr_info(s.ptr(), C.byref(dt), C.byref(nd), shp))
        return dt.value, tuple(shp[i] for i in range(nd.value))

    def assign(self, e, out: np.ndarray | None = None, dtype=None, out_code: int | None = None) -> np.ndarray:
        """Evaluate e into `out` (a numpy array / view) or a new row-major array."""
        s = self.serialize(e)
        if out is None:
            dt, shp = self.info(e)
            code = core.dtype_code(dtype) if dtype is not None else dt
            out = np.zeros(shp, _NP[code])
            out_code = code
        elif out_code is None:
            out_code = core.dtype_code(out.dtype)
        t = host_tensor(out, out_code)
        _check(lib().xo_assign(s.ptr(), C.byref(t)))
        return out

    def reduce_info(self, r: core.Reducer):
        s = self.serialize(r.expr)
        d = r.to_abi()
        dt, nd = C.c_int32(), C.c_int32()
        shp = (C.c_int64 * A.MAX_RANK)()
        _check(lib().xo_reduce_info(s.ptr(), C.byref(d), C.byref(dt), C.byref(nd), shp))
        return dt.value, tuple(shp[i] for i in range(nd.value))

    def reduce(self, r: core.Reducer, high_precision: bool = False) -> np.ndarray:
        """xtensor-order reduction, or (high_precision) the long-double reference as float64."""
        s = self.serialize(r.expr)
        d = r.to_abi()
        dt, shp = self.reduce_info(r)
        code = A.FLOAT64 if high_precision else dt
        out = np.zeros(shp, _NP[code])
        t = host_tensor(out, code)
        f = lib().xo_reduce_hp if high_precision else lib().xo_reduce
        _check(f(s.ptr(), C.byref(d), C.byref(t)))
        return out

    def scan_info(self, e, axis=None, kind="cumsum", dtype=None):
        s = self.serialize(e)
        d = self._scan_desc(axis, kind, dtype)
        dt, nd = C.c_int32(), C.c_int32()
        shp = (C.c_int64 * A.MAX_RANK)()
        _check(lib().xo_scan_info(s.ptr(), C.byref(d), C.byref(dt), C.byref(nd), shp))
        return dt.value, tuple(shp[i] for i in range(nd.value))

    @staticmethod
    def _scan_desc(axis, kind, dtype):
        d = A.ScanDesc()
        d.op = A.SCAN[kind]
        d.acc_dtype = -1 if dtype is None else core.dtype_code(dtype)
        d.axis = A.AXIS_NONE if axis is None else int(axis)
        return d

    def scan(self, e, axis=None, kind="cumsum", dtype=None, high_precision=False) -> np.ndarray:
        s = self.serialize(e)
        d = self._scan_desc(axis, kind, dtype)
        dt, shp = self.scan_info(e, axis, kind, dtype)
        code = A.FLOAT64 if high_precision else dt
        out = np.zeros(shp, _NP[code])
        t = host_tensor(out, code)
        f = lib().xo_scan_hp if high_precision else lib().xo_scan
        _check(f(s.ptr(), C.byref(d), C.byref(t)))
        return out

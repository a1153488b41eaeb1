"""Batches of uneven instances evaluated by one persistent launch.

`Batch("softmax", xs, outs)` prepares, for every pair (x, out) of 2-D
device arrays, the row-wise template

    softmax: out = exp(x - amax(x, {1})) / sum(exp(x - amax(x, {1})), {1})
    sum:     out = sum(x, {1})
    amax:    out = amax(x, {1})

(xtensor's row softmax of test/test_xmath.cpp:985-1027 and the reducers of
core/xmath.hpp:1803-1823), and evaluates all of them in one work-queue
kernel (`launch()`) or one HIP graph replay (`execute()`).  `autotune()`
times the candidate launch shapes and caches the choice for the batch's
shape class.  Inputs need unit column stride; rows may be padded.
"""

from __future__ import annotations

import ctypes as C
from typing import Sequence

from . import _abi as A
from .core import DeviceArray, _check, _stream

OPS = {"softmax": 0, "sum": 1, "amax": 2}


def _row_stride(a: DeviceArray, what: str) -> int:
    if a.ndim != 2:
        raise ValueError(f"{what} must be 2-D, got shape {a.shape}")
    rows, cols = a.shape
    if cols > 1 and a.strides[1] != 1:
        raise ValueError(f"{what} must have unit column stride")
    return a.strides[0] if rows > 1 else max(cols, 1)


class Batch:
    def __init__(self, op: str, inputs: Sequence[DeviceArray], outputs: Sequence[DeviceArray]):
        if op not in OPS:
            raise ValueError(f"unknown batch op {op!r}")
        if len(inputs) != len(outputs):
            raise ValueError("inputs and outputs differ in length")
        self.op = op
        dts = {x.dtype_code for x in inputs} | {o.dtype_code for o in outputs}
        if len(dts) > 1:
            raise TypeError("a batch holds one dtype")
        self.dtype_code = dts.pop() if dts else A.FLOAT32
        items = (A.BatchItem * max(1, len(inputs)))()
        for k, (x, o) in enumerate(zip(inputs, outputs)):
            rows, cols = x.shape if x.ndim == 2 else (None, None)
            irs = _row_stride(x, f"input {k}")
            if op == "softmax":
                if tuple(o.shape) != tuple(x.shape):
                    raise ValueError(f"output {k} shape {o.shape} != input shape {x.shape}")
                ors = _row_stride(o, f"output {k}")
            else:
                if tuple(o.shape) != (rows,):
                    raise ValueError(f"output {k} must have shape ({rows},)")
                ors = o.strides[0] if rows > 1 else 1
            it = items[k]
            it.in_ = x.data_ptr + x.offset * x.itemsize
            it.out = o.data_ptr + o.offset * o.itemsize
            it.rows, it.cols = rows, cols
            it.in_row_stride, it.out_row_stride = irs, ors
        self._keep = (list(inputs), list(outputs))
        h = C.c_void_p()
        _check(A.lib().xg_batch_create(OPS[op], self.dtype_code, items, len(inputs), C.byref(h)))
        self.handle = h

    def autotune(self, stream=None) -> dict:
        st = _stream() if stream is None else C.c_void_p(stream)
        _check(A.lib().xg_batch_autotune(self.handle, st))
        return self.tuning

    @property
    def tuning(self) -> dict:
        t = A.BatchTuning()
        _check(A.lib().xg_batch_get_tuning(self.handle, C.byref(t)))
        return {"block": t.block, "blocks_per_cu": t.blocks_per_cu, "task_elems": t.task_elems,
                "dyn_percent": t.dyn_percent}

    @tuning.setter
    def tuning(self, d: dict) -> None:
        cur = self.tuning
        cur.update(d)
        t = A.BatchTuning(cur["block"], cur["blocks_per_cu"], cur["task_elems"], cur["dyn_percent"])
        _check(A.lib().xg_batch_set_tuning(self.handle, C.byref(t)))

    def info(self) -> dict:
        nt, ne = C.c_int64(), C.c_int64()
        _check(A.lib().xg_batch_info(self.handle, C.byref(nt), C.byref(ne)))
        return {"tasks": nt.value, "elements": ne.value}

    def launch(self, stream=None) -> None:
        st = _stream() if stream is None else C.c_void_p(stream)
        _check(A.lib().xg_batch_launch(self.handle, st))

    def execute(self, stream=None) -> None:
        st = _stream() if stream is None else C.c_void_p(stream)
        _check(A.lib().xg_batch_execute(self.handle, st))

    def __del__(self):
        try:
            if self.handle:
                A.lib().xg_batch_destroy(self.handle)
        except Exception:
            pass

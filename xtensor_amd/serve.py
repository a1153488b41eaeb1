"""Request server: continuous batching of device-resident requests.

`Server(budget_bytes=64 << 30, streams=4, max_batch=64)` starts the library's
request loop (xg_server_* in include/xg.h).  `submit(...)` queues a request
whose operands are DeviceArrays; the loop admits requests in arrival order
within the HBM budget, fuses admitted requests of one template / dtype /
shape class into one launch and keeps up to `streams` launches in flight.
`poll()` returns completions with their arrival, admission, launch and
completion times (the server's monotonic clock, `now_ns()`).

Templates (each is the xtensor expression of the same name):
    affine_clip  out = clip((x - mean) / sqrt(var + 1e-5) * gamma + beta, -3, 3)
                 x, out [N, C, H, W]; mean, var, gamma, beta [C]
    row_sum      out = sum(x, {1}), x [rows, cols]
    row_softmax  out = exp(x - amax(x, {1})) / sum(exp(x - amax(x, {1})), {1})
    bitmix       out = (x * 3 + (x >> 2)) & 0xFFFF, x int32 [n]
    expr         out = e for any expression e (`submit_expr(out, e)`); requests
                 of one compiled program and rank are fused
    reduce       out = a reducer such as sum(x, [0]) (`submit_expr(out, r)`),
                 any axes; each runs alone on a server stream
"""

from __future__ import annotations

import ctypes as C
from typing import Sequence

from . import _abi as A
from .core import DeviceArray, _check

TEMPLATES = {"affine_clip": 0, "row_sum": 1, "row_softmax": 2, "bitmix": 3, "expr": 4, "reduce": 5}


def now_ns() -> int:
    return int(A.lib().xg_server_now_ns())


def _ptr(a: DeviceArray) -> int:
    return a.data_ptr + a.offset * a.itemsize


def _contig(a: DeviceArray, what: str):
    exp = 1
    for n, s in zip(reversed(a.shape), reversed(a.strides)):
        if n > 1 and s != exp:
            raise ValueError(f"{what} must be contiguous row-major")
        exp *= n


class Server:
    def __init__(self, budget_bytes: int = 64 << 30, streams: int = 4, max_batch: int = 64,
                 queue_capacity: int = 1 << 16, plain: bool = False):
        cfg = A.ServerConfig(budget_bytes, streams, max_batch, queue_capacity, int(plain))
        h = C.c_void_p()
        _check(A.lib().xg_server_create(C.byref(cfg), C.byref(h)))
        self.handle = h
        self._keep = {}

    def submit(self, template: str, x: DeviceArray, out: DeviceArray, params: Sequence[DeviceArray] = (),
               tag: int = 0, arrival_ns: int = 0) -> int:
        """Queue a request; returns the library status (0 = queued, 8 = queue full)."""
        return self.submit_prepared(self.prepare(template, x, out, params, tag), arrival_ns)

    def submit_expr(self, out: DeviceArray, e, tag: int = 0, arrival_ns: int = 0) -> int:
        """Queue `noalias(out) = e` for any expression (template "expr"), or a
        reduction into out when e is a reducer such as xt.sum(x, [0])
        (template "reduce").  The descriptors are copied at submission; the
        arrays stay referenced until the request completes."""
        from .core import Reducer, _leaves, _require_device, _wrap, _eval_reducer, serialize
        e = _wrap(e)
        red = isinstance(e, Reducer)
        body = e.expr if red else e
        _require_device(_leaves(body) + [out])
        s = serialize(body, reducer_fn=_eval_reducer)
        t = out.to_abi()
        r = A.Request()
        r.tpl = TEMPLATES["reduce" if red else "expr"]
        r.dtype = out.dtype_code
        r.expr = C.addressof(s.expr)
        r.out_tensor = C.addressof(t)
        d = e.to_abi() if red else None
        if red:
            r.reduce = C.addressof(d)
        r.tag = tag
        r.arrival_ns = arrival_ns
        st = A.lib().xg_server_submit(self.handle, C.byref(r))
        if st == A.BUSY:
            return st
        _check(st)
        self._keep[tag] = (out, e, _leaves(body))
        return st

    def submit_prepared(self, prepared, arrival_ns: int = 0) -> int:
        """Queue a request built by prepare() (its checks already done), with
        its arrival time; returns the library status as submit() does."""
        r, keep = prepared
        r.arrival_ns = arrival_ns
        st = A.lib().xg_server_submit(self.handle, C.byref(r))
        if st == A.BUSY:
            return st
        _check(st)
        self._keep[r.tag] = keep
        return st

    @staticmethod
    def prepare(template: str, x: DeviceArray, out: DeviceArray, params: Sequence[DeviceArray] = (),
                tag: int = 0):
        """Check a request and build its descriptor once, ahead of submission."""
        r = A.Request()
        r.tpl = TEMPLATES[template]
        r.dtype = x.dtype_code
        _contig(x, "input")
        _contig(out, "output")
        if template == "affine_clip":
            if x.ndim != 4 or len(params) != 4:
                raise ValueError("affine_clip takes x [N, C, H, W] and four [C] parameters")
            for k, p in enumerate(params):
                if tuple(p.shape) != (x.shape[1],):
                    raise ValueError("affine_clip parameters must have shape [C]")
                _contig(p, "parameter")
                r.params[k] = _ptr(p)
        elif template in ("row_sum", "row_softmax"):
            if x.ndim != 2:
                raise ValueError(f"{template} takes a 2-D input")
        elif x.ndim != 1:
            raise ValueError("bitmix takes a 1-D input")
        for k, n in enumerate(x.shape):
            r.shape[k] = n
        r.in_ = _ptr(x)
        r.out = _ptr(out)
        r.tag = tag
        return r, (x, out, tuple(params))

    def poll(self, cap: int = 4096) -> list:
        buf = (A.Completion * cap)()
        n = C.c_int32()
        _check(A.lib().xg_server_poll(self.handle, buf, cap, C.byref(n)))
        out = []
        for i in range(n.value):
            c = buf[i]
            self._keep.pop(c.tag, None)
            out.append({"tag": c.tag, "status": c.status, "batch_size": c.batch_size, "arrival_ns": c.arrival_ns,
                        "admit_ns": c.admit_ns, "launch_ns": c.launch_ns, "done_ns": c.done_ns})
        return out

    def hold(self, on: bool = True) -> None:
        """Keep admitting but launch nothing (True), or resume (False)."""
        _check(A.lib().xg_server_hold(self.handle, int(on)))

    def drain(self, timeout_ms: int = -1) -> None:
        _check(A.lib().xg_server_drain(self.handle, timeout_ms))

    def stats(self) -> dict:
        s = A.ServerStats()
        _check(A.lib().xg_server_stats_get(self.handle, C.byref(s)))
        return {f: getattr(s, f) for f, _ in A.ServerStats._fields_}

    def close(self) -> None:
        if self.handle:
            A.lib().xg_server_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

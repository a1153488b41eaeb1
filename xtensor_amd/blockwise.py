"""Blockwise reducers (reducers/xblockwise_reducer.hpp,
reducers/xblockwise_reducer_functors.hpp): reduce an array chunk by chunk
and merge the chunk results, so an input larger than device memory — a host
array or a numpy memmap — streams through the GPU one chunk at a time.

`xt.blockwise.sum(e, chunk_shape, axes)` etc.: every chunk is reduced on the
device with kept dims (host chunks are uploaded one ahead of the
computation, on a stream of their own, through pinned buffers), and its result merged into the slice of the output
its kept coordinates cover — `+=` for sums, `*=` for products, max / min for
extrema (the merges of xblockwise_reducer_functors.hpp).  Float sums merge
in float64 and round once at the end; integer merges wrap exactly like the
single-pass reducer.  mean divides the merged sum by the reduced count;
norm_l2 is the square root of the merged norm_sq; variance / stddev run two
blockwise passes (the mean, then the mean of squared deviations), as
xt::variance itself does.  The result has the single-pass reducer's value
type and shape.
"""

from __future__ import annotations

import itertools

import numpy as np

from . import core as _c
from .chunked import ChunkedArray

_MERGE = {"sum": "add", "prod": "mul", "amax": "max", "amin": "min", "mean": "add", "norm_l0": "add",
          "norm_l1": "add", "norm_sq": "add", "norm_l2": "add", "norm_linf": "max"}


def _norm_axes(axes, nd):
    if axes is None:
        return list(range(nd))
    if isinstance(axes, (int, np.integer)):
        axes = [int(axes)]
    ax = [a + nd if a < 0 else a for a in axes]
    if sorted(ax) != ax or len(set(ax)) != len(ax) or (ax and (ax[0] < 0 or ax[-1] >= nd)):
        raise RuntimeError("bad reduction axes")
    return ax


class _HostUpload:
    """Host chunks onto the device one chunk ahead of the computation: two
    pinned staging buffers and two device buffers, the copies on their own
    stream.  Chunk k+1 is gathered into pinned memory and its upload enqueued
    before chunk k is handed to the caller, whose work on the current stream
    waits only for chunk k's upload; a buffer is reused once the work that
    read it (recorded after the caller's turn) has finished."""

    def __init__(self, get, code, chunk_shape):
        import torch
        self.get = get  # slices -> the host chunk
        self.code = code
        itemsize = _c.A.DTYPE_SIZE[code]
        nbytes = max(1, int(np.prod(chunk_shape, dtype=np.int64)) * itemsize)
        self.pinned = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
        self.dev = [torch.empty(nbytes, dtype=torch.uint8, device="cuda") for _ in range(2)]
        self.copy = torch.cuda.Stream()
        self.ready = [torch.cuda.Event() for _ in range(2)]  # upload into slot done
        self.freed = [torch.cuda.Event() for _ in range(2)]  # work reading slot done
        self.used = [False, False]

    def _upload(self, k, sl):
        import torch
        slot = k % 2
        if self.used[slot]:
            self.ready[slot].synchronize()  # the pinned buffer's previous upload has left it
        arr = np.ascontiguousarray(self.get(sl))
        n = arr.nbytes
        np.copyto(self.pinned[slot][:n].numpy(), arr.reshape(-1).view(np.uint8))
        with torch.cuda.stream(self.copy):
            if self.used[slot]:
                self.copy.wait_event(self.freed[slot])
            self.dev[slot][:n].copy_(self.pinned[slot][:n], non_blocking=True)
            self.ready[slot].record(self.copy)
        self.used[slot] = True
        return _c.DeviceArray(self.dev[slot], self.code, arr.shape)

    def chunks(self, slices):
        import torch
        slices = list(slices)
        if not slices:
            return
        compute = torch.cuda.current_stream()
        nxt = self._upload(0, slices[0])
        for k, sl in enumerate(slices):
            cur = nxt
            if k + 1 < len(slices):
                nxt = self._upload(k + 1, slices[k + 1])
            compute.wait_event(self.ready[k % 2])
            yield sl, cur
            self.freed[k % 2].record(compute)  # the caller's work on this chunk is enqueued


def _source(e, chunk_shape):
    """(shape, dtype code, iterator factory) of a host array or host chunked
    array (chunks uploaded one ahead of the computation), a device array
    (chunks are views) or a device chunked array (its chunks)."""
    if isinstance(e, ChunkedArray):
        if tuple(chunk_shape) != e.chunk_shape:
            raise ValueError("a chunked array is reduced with its own chunk shape")

        def grid_index(sl):
            return tuple(s.start // c for s, c in zip(sl, e.chunk_shape))
        if e.storage == "device":
            def it(slices):
                for sl in slices:
                    yield sl, e.chunk(*grid_index(sl))
            return e.shape, e.dtype_code, it
        return e.shape, e.dtype_code, lambda slices: _HostUpload(
            lambda sl: e.chunk(*grid_index(sl)), e.dtype_code, e.chunk_shape).chunks(slices)
    if isinstance(e, _c.DeviceArray):
        def it(slices):
            for sl in slices:
                yield sl, _c.view(e, *[_c.range_(s.start, s.stop) for s in sl])
        return tuple(e.shape), e.dtype_code, it
    h = np.asarray(e) if not isinstance(e, np.memmap) else e
    code = _c.dtype_code(h.dtype)
    return tuple(h.shape), code, lambda slices: _HostUpload(lambda sl: h[sl], code, tuple(chunk_shape)).chunks(slices)


def _chunks(shape, chunk_shape):
    grid = [range(0, n, c) for n, c in zip(shape, chunk_shape)]
    for start in itertools.product(*grid):
        yield tuple(slice(s, min(s + c, n)) for s, c, n in zip(start, chunk_shape, shape))


def _whole(e, shape):
    """The whole source as one device array (for empty sources)."""
    if isinstance(e, _c.DeviceArray):
        return e
    if isinstance(e, ChunkedArray):
        return _c.from_numpy(e.to_numpy())
    return _c.from_numpy(np.ascontiguousarray(np.asarray(e)))


def _reduce(kind, e, chunk_shape, axes=None, keep_dims=False, dtype=None, ddof=0):
    if chunk_shape is None and isinstance(e, ChunkedArray):
        chunk_shape = e.chunk_shape
    nd = len(e.shape) if hasattr(e, "shape") else np.ndim(e)
    if chunk_shape is None or len(chunk_shape) != nd or any(int(c) <= 0 for c in chunk_shape):
        raise ValueError("chunk_shape needs one positive extent per dimension")
    chunk_shape = tuple(int(c) for c in chunk_shape)
    shape, code, chunks_of = _source(e, chunk_shape)
    ax = _norm_axes(axes, nd)
    ph = _c._Placeholder(code, shape)
    if kind in ("variance", "stddev"):
        return _moment(kind, e, chunk_shape, ax, keep_dims, dtype, ddof, ph)
    rdt, rshape = _c.reduce_info(_c.Reducer(kind, ph, ax, keep_dims, None, dtype))
    # the merged accumulator: the value type of the partial reduction (the
    # sum for mean, norm_sq for norm_l2), float64 for float sums
    pkind = {"mean": "sum", "norm_l2": "norm_sq"}.get(kind, kind)
    pdt, _ = _c.reduce_info(_c.Reducer(pkind, ph, ax, True, None, dtype))
    acc_dt = pdt
    if _MERGE[kind] == "add" and _c.dtype_name(pdt) in ("float32", "float64"):
        acc_dt = _c.dtype_code("float64")
    kshape = tuple(1 if k in ax else n for k, n in enumerate(shape))  # kept-dims result shape
    acc = _c.empty(kshape, acc_dt)
    seen = np.zeros([(-(-n // c)) if k not in ax else 1 for k, (n, c) in enumerate(zip(shape, chunk_shape))], bool)
    for sl, chunk in chunks_of(_chunks(shape, chunk_shape)):
        part = _c.eval(_c.Reducer(pkind, chunk, ax, True, None, dtype))
        osl = tuple(slice(0, 1) if k in ax else s for k, s in enumerate(sl))
        dst = _c.view(acc, *[_c.range_(s.start, s.stop) for s in osl])
        gi = tuple(0 if k in ax else s.start // chunk_shape[k] for k, s in enumerate(sl))
        op = _MERGE[kind]
        if not seen[gi]:
            _c.assign(dst, part)
            seen[gi] = True
        elif op == "add":
            _c.assign(dst, dst + part)
        elif op == "mul":
            _c.assign(dst, dst * part)
        elif op == "max":
            _c.assign(dst, _c.maximum(dst, part))
        else:
            _c.assign(dst, _c.minimum(dst, part))
    if 0 in shape:  # no chunks: the single-pass reducer gives the identity / NaN mean
        return _c.eval(_c.Reducer(kind, _whole(e, shape), ax, keep_dims, None, dtype))
    if kind == "mean":
        # mean_division: the sum divided by T-or-double(size / outputs)
        n_red = int(np.prod([shape[k] for k in ax], dtype=np.int64)) if ax else 1
        res = _c.cast(acc, pdt) / _c.scalar(n_red, "float64" if dtype is None else dtype)
    elif kind == "norm_l2":
        res = _c.sqrt(_c.cast(acc, pdt))
    else:
        res = acc
    val = _c.eval(_c.cast(res, rdt))  # kept-dims shape, row-major
    return val if keep_dims else _c.reshape_view(val, rshape)


def _moment(kind, e, chunk_shape, ax, keep_dims, dtype, ddof, ph):
    if dtype is not None and _c.dtype_name(_c.dtype_code(dtype)) not in ("float32", "float64"):
        raise NotImplementedError("blockwise variance / stddev with an integral accumulator type")
    shape, code, chunks_of = _source(e, chunk_shape)
    rdt, rshape = _c.reduce_info(_c.Reducer(kind, ph, ax, keep_dims, None, dtype, ddof))
    m = _reduce("mean", e, chunk_shape, ax, True, dtype)          # kept dims
    # sum of squared deviations, merged in float64
    kshape = tuple(1 if k in ax else n for k, n in enumerate(shape))
    acc = _c.zeros(kshape, "float64")
    for sl, chunk in chunks_of(_chunks(shape, chunk_shape)):
        osl = [_c.range_(0, 1) if k in ax else _c.range_(s.start, s.stop) for k, s in enumerate(sl)]
        mv = _c.view(m, *osl)
        d = chunk - mv
        part = _c.eval(_c.Reducer("sum", d * d, ax, True, None, None))
        dst = _c.view(acc, *osl)
        _c.assign(dst, dst + part)
    size = int(np.prod(shape, dtype=np.int64))
    n_out = int(np.prod([shape[k] for k in range(len(shape)) if k not in ax], dtype=np.int64))
    q = (size - ddof) // n_out if n_out else 0  # detail::mean's divisor
    v = acc / _c.scalar(q, "float64")
    if kind == "stddev":
        v = _c.sqrt(v)
    val = _c.eval(_c.cast(v, rdt))
    return val if keep_dims else _c.reshape_view(val, rshape)


def _make(kind):
    def f(e, chunk_shape, axes=None, keep_dims=False, dtype=None):
        return _reduce(kind, e, chunk_shape, axes, keep_dims, dtype)
    f.__name__ = kind
    f.__doc__ = f"xt::blockwise::{kind}(e, chunk_shape, axes): chunk-by-chunk {kind} (reducers/xblockwise_reducer.hpp:284-330)"
    return f


sum = _make("sum")  # noqa: A001
prod = _make("prod")
amax = _make("amax")
amin = _make("amin")
mean = _make("mean")
norm_l0 = _make("norm_l0")
norm_l1 = _make("norm_l1")
norm_sq = _make("norm_sq")
norm_l2 = _make("norm_l2")
norm_linf = _make("norm_linf")


def variance(e, chunk_shape, axes=None, ddof=0, keep_dims=False, dtype=None):
    return _reduce("variance", e, chunk_shape, axes, keep_dims, dtype, ddof)


def stddev(e, chunk_shape, axes=None, keep_dims=False, dtype=None):
    return _reduce("stddev", e, chunk_shape, axes, keep_dims, dtype)

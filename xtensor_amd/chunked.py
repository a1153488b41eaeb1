"""Chunked arrays (chunk/xchunked_array.hpp, chunk/xchunked_assign.hpp).

`xt.chunked_array(shape, chunk_shape, dtype)` is an array stored as a grid of
chunks, each its own allocation: on the device (storage="device") or in host
memory (storage="host": arrays larger than the device, streamed through it).
Chunks at the upper edges are clipped to the array's shape.

Assignment is xchunked_assigner's loop (chunk/xchunked_assign.hpp): every
chunk receives the expression restricted to that chunk's region,
`noalias(chunk) = strided_view(e, chunk slices)`.  Device chunks are
assigned by one fused launch for all chunks that share a compiled program
(prepared plans run through xg_plan_execute_many); host chunks are computed
on the device one at a time into two device buffers and copied back through
pinned memory while the next chunk is computed.  A chunked array may appear
in the expression it is assigned from (or in any expression assigned into a
chunked array with the same chunk grid): each chunk then reads the matching
chunk.  `xt.blockwise.*` reduces chunked arrays chunk by chunk.
"""

from __future__ import annotations

import itertools

import numpy as np

from . import core as _c


class ChunkedArray(_c.Expr):
    """xt::xchunked_array: shape, chunk_shape, a grid of chunks.  It is an
    expression operand (arithmetic builds expressions over it) that the
    chunked assignment and the blockwise reducers evaluate chunk by chunk."""

    def __init__(self, shape, chunk_shape, dtype="float64", storage="device"):
        shape = tuple(int(n) for n in shape)
        chunk_shape = tuple(int(c) for c in chunk_shape)
        if len(chunk_shape) != len(shape) or any(c <= 0 for c in chunk_shape):
            raise ValueError("chunk_shape needs one positive extent per dimension")
        if storage not in ("device", "host"):
            raise ValueError("storage is 'device' or 'host'")
        self._shape = shape
        self.chunk_shape = chunk_shape
        self.dtype_code = _c.dtype_code(dtype)
        self.storage = storage
        self.grid_shape = tuple(-(-n // c) for n, c in zip(shape, chunk_shape))
        self._chunks = {}
        for gi in itertools.product(*[range(g) for g in self.grid_shape]):
            cs = tuple(min(c, n - i * c) for i, c, n in zip(gi, chunk_shape, shape))
            if storage == "device":
                self._chunks[gi] = _c.empty(cs, self.dtype_code)
            else:
                self._chunks[gi] = np.empty(cs, _c.numpy_dtype(self.dtype_code))

    @property
    def shape(self) -> tuple:
        return self._shape

    @property
    def dtype(self) -> str:
        return _c.dtype_name(self.dtype_code)

    @property
    def ndim(self) -> int:
        return len(self._shape)

    @property
    def size(self) -> int:
        return int(np.prod(self._shape, dtype=np.int64))

    def slices(self, gi):
        """The region of chunk `gi` in the array, as slices."""
        return tuple(slice(i * c, min((i + 1) * c, n)) for i, c, n in zip(gi, self.chunk_shape, self.shape))

    def chunk(self, *gi):
        return self._chunks[tuple(int(i) for i in gi)]

    def chunks(self):
        """(grid index, slices, chunk) in row-major grid order."""
        for gi in itertools.product(*[range(g) for g in self.grid_shape]):
            yield gi, self.slices(gi), self._chunks[gi]

    def __getitem__(self, idx):
        """Element read: the chunk holding `idx`, at its in-chunk position."""
        idx = tuple(int(i) for i in (idx if isinstance(idx, tuple) else (idx,)))
        if len(idx) != len(self.shape):
            raise IndexError("one index per dimension")
        idx = tuple(i + n if i < 0 else i for i, n in zip(idx, self.shape))
        gi = tuple(i // c for i, c in zip(idx, self.chunk_shape))
        local = tuple(i - g * c for i, g, c in zip(idx, gi, self.chunk_shape))
        ch = self._chunks[gi]
        if self.storage == "host":
            return ch[local]
        return _c.view(ch, *[_c.range_(i, i + 1) for i in local]).to_numpy().reshape(())[()]

    def __setitem__(self, idx, value):
        """Element write (xchunked_array::operator[] / operator() as lvalues)."""
        idx = tuple(int(i) for i in (idx if isinstance(idx, tuple) else (idx,)))
        if len(idx) != len(self.shape):
            raise IndexError("one index per dimension")
        idx = tuple(i + n if i < 0 else i for i, n in zip(idx, self.shape))
        gi = tuple(i // c for i, c in zip(idx, self.chunk_shape))
        local = tuple(i - g * c for i, g, c in zip(idx, gi, self.chunk_shape))
        ch = self._chunks[gi]
        if self.storage == "host":
            ch[local] = value
        else:
            _c.assign(_c.view(ch, *[_c.range_(i, i + 1) for i in local]), _c.scalar(value, self.dtype))

    def __iter__(self):
        """Elements in row-major order (the container iterator)."""
        return iter(self.to_numpy().reshape(-1).tolist())

    # computed assignments (xsemantic_base::operator+= and friends)
    def __iadd__(self, e): return self.assign(self + e)
    def __isub__(self, e): return self.assign(self - e)
    def __imul__(self, e): return self.assign(self * e)
    def __itruediv__(self, e): return self.assign(self / e)

    def to_numpy(self) -> np.ndarray:
        out = np.empty(self.shape, _c.numpy_dtype(self.dtype_code))
        for _, sl, ch in self.chunks():
            out[sl] = ch if self.storage == "host" else ch.to_numpy()
        return out

    def fill(self, value):
        return self.assign(_c.scalar(value, self.dtype))

    def assign(self, e):
        """noalias(self) = e, chunk by chunk (xchunked_semantic::assign_xexpression)."""
        e = e if isinstance(e, ChunkedArray) else _c._wrap(e)
        eshape = e.shape if isinstance(e, ChunkedArray) else _c.expr_info(_shape_expr(e))[1]
        if not _broadcastable(eshape, self.shape):
            raise RuntimeError(f"broadcast error: cannot assign {eshape} to {self.shape}")
        if self.storage == "device":
            plans = [_c.plan(ch, _slice_expr(e, sl, self.chunk_shape, self.shape)) for _, sl, ch in self.chunks()]
            _c.execute_many(plans)
            return self
        _assign_host(self, e)
        return self


def _broadcastable(src, dst):
    if len(src) > len(dst):
        return False
    return all(s == 1 or s == d for s, d in zip(src[::-1], dst[::-1]))


def _slice_expr(e, sl, chunk_shape, out_shape):
    """e restricted to the region `sl` of the output (None: the whole, with
    chunked arrays read whole): leaves sliced along the dims they span
    (broadcast dims of extent 1 stay whole), reducers and index views
    evaluated first, chunked arrays replaced by the chunk covering `sl`."""
    R = len(out_shape)

    def leaf_view(a):
        if sl is None:
            return a
        r = len(a.shape)
        keys = []
        for j, n in enumerate(a.shape):
            k = R - r + j
            keys.append(_c.range_(0, 1) if n == 1 else _c.range_(sl[k].start, sl[k].stop))
        return _c.view(a, *keys)

    def visit(x):
        if isinstance(x, ChunkedArray):
            if sl is None:
                return _c.from_numpy(x.to_numpy()) if x.storage == "host" else _gather_device(x)
            r = len(x.shape)
            sub = sl[R - r:]
            for k, (s, n) in enumerate(zip(sub, x.shape)):
                if n == 1:
                    continue
                c = x.chunk_shape[k]
                if s.start % c or (s.stop - s.start) > c or (s.stop != n and s.stop - s.start != c):
                    raise NotImplementedError("a chunked operand needs the destination's chunk grid")
            gi = tuple(0 if n == 1 else s.start // c for s, n, c in zip(sub, x.shape, x.chunk_shape))
            ch = x.chunk(*gi)
            return ch if x.storage == "device" else _c.from_numpy(ch)
        if isinstance(x, _c.DeviceArray):
            return leaf_view(x)
        if isinstance(x, _c.Scalar):
            return x
        if isinstance(x, _c.Reducer):
            return leaf_view(_c._eval_reducer(x))
        if isinstance(x, _c.IndexView):
            return leaf_view(_c.eval(x))
        if isinstance(x, _c.Func):
            return _c.Func(x.op, *[visit(a) for a in x.args], index=x.index)
        raise TypeError(f"not an expression: {x!r}")

    return visit(e)


def _shape_expr(e):
    """e with its chunked operands replaced by shape / type stand-ins."""
    if isinstance(e, ChunkedArray):
        return _c._Placeholder(e.dtype_code, e.shape)
    if isinstance(e, _c.Func):
        return _c.Func(e.op, *[_shape_expr(a) for a in e.args], index=e.index)
    return e


def _gather_device(x: ChunkedArray):
    """A chunked device array copied into one contiguous device array."""
    out = _c.empty(x.shape, x.dtype_code)
    plans = [_c.plan(_c.view(out, *[_c.range_(s.start, s.stop) for s in sl]), ch) for _, sl, ch in x.chunks()]
    _c.execute_many(plans)
    return out


def _assign_host(dst: ChunkedArray, e):
    """Host chunks: each computed on the device into one of two buffers, and
    copied back through pinned memory while the next one is computed."""
    import torch
    nbytes = max(1, int(np.prod(dst.chunk_shape, dtype=np.int64)) * _c.A.DTYPE_SIZE[dst.dtype_code])
    bufs = [torch.empty(nbytes, dtype=torch.uint8, device="cuda") for _ in range(2)]
    pinned = [torch.empty(nbytes, dtype=torch.uint8, pin_memory=True) for _ in range(2)]
    back = torch.cuda.Stream()
    done = [torch.cuda.Event() for _ in range(2)]     # copy back of slot finished
    computed = [torch.cuda.Event() for _ in range(2)]  # chunk in slot computed
    pending = [None, None]  # (host chunk, bytes) waiting in the slot's pinned buffer
    compute = torch.cuda.current_stream()

    def drain(slot):
        if pending[slot] is not None:
            done[slot].synchronize()
            ch, n = pending[slot]
            ch.reshape(-1).view(np.uint8)[:] = pinned[slot][:n].numpy()
            pending[slot] = None

    for k, (_, sl, ch) in enumerate(dst.chunks()):
        slot = k % 2
        drain(slot)  # the slot's previous chunk is home; its buffers are free
        compute.wait_event(done[slot])
        d = _c.DeviceArray(bufs[slot], dst.dtype_code, ch.shape)
        _c.assign(d, _slice_expr(e, sl, dst.chunk_shape, dst.shape))
        computed[slot].record(compute)
        n = ch.size * ch.itemsize
        with torch.cuda.stream(back):
            back.wait_event(computed[slot])
            pinned[slot][:n].copy_(bufs[slot][:n], non_blocking=True)
            done[slot].record(back)
        pending[slot] = (ch, n)
    for slot in range(2):
        drain(slot)


def is_chunked(e) -> bool:
    """xt::is_chunked (chunk/xchunked_array.hpp:205-209)."""
    return isinstance(e, ChunkedArray)


def chunked_array(shape_or_expr, chunk_shape=None, dtype=None, storage="device") -> ChunkedArray:
    """xt::chunked_array<T>(shape, chunk_shape) (chunk/xchunked_array.hpp:228-
    257), or chunked_array(e[, chunk_shape]): a chunked array holding the
    values of an expression, a device array or a host array (without a chunk
    shape: e's own if it is chunked, else one chunk of e's shape,
    chunk/xchunked_array.hpp:274-276)."""
    if isinstance(shape_or_expr, (tuple, list)) and all(isinstance(n, (int, np.integer)) for n in shape_or_expr):
        if chunk_shape is None:
            raise ValueError("chunked_array(shape, chunk_shape) needs a chunk shape")
        return ChunkedArray(shape_or_expr, chunk_shape, dtype or "float64", storage)
    src = shape_or_expr
    if chunk_shape is None:
        chunk_shape = src.chunk_shape if isinstance(src, ChunkedArray) else tuple(np.shape(src) if isinstance(
            src, np.ndarray) else _c.expr_info(_shape_expr(_c._wrap(src)))[1])
    if isinstance(src, np.ndarray):
        a = ChunkedArray(src.shape, chunk_shape, dtype or _c.dtype_name(_c.dtype_code(src.dtype)), storage)
        for _, sl, ch in a.chunks():
            if storage == "host":
                ch[...] = src[sl]
            else:
                _c.assign(ch, _c.from_numpy(np.ascontiguousarray(src[sl])))
        return a
    e = src if isinstance(src, ChunkedArray) else _c._wrap(src)
    if isinstance(e, ChunkedArray):
        shape, code = e.shape, e.dtype_code
    else:
        code, shape = _c.expr_info(_shape_expr(e))
    return ChunkedArray(shape, chunk_shape, dtype or _c.dtype_name(code), storage).assign(e)

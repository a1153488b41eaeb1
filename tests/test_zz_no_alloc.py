"""xg_reduce_noalloc: the allocation-free variant of xg_reduce.

Each case calls xg_reduce and then xg_reduce_noalloc through the C ABI on
the same descriptors; the two must agree bit for bit (same kernels, same
merge order), and the result is checked against float64 / exact integer
arithmetic to the bar of the reduction tests (float32 sums within 2e-7 of
float64, integers exact).  Every case runs twice in a row, the second time
with a smaller input, since all calls share one scratch array per device.
"""

import ctypes as C

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd import _abi as A
from xtensor_amd.core import reduce_info, serialize

pytestmark = pytest.mark.gpu


def _sync():
    import torch
    torch.cuda.synchronize()


def _stream():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _call(fn, r: xt.Reducer, out):
    s = serialize(r.expr)
    d = r.to_abi()
    t = out.to_abi()
    return fn(s.ptr(), C.byref(d), C.byref(t), _stream())


def _both(r: xt.Reducer):
    """(xg_reduce result, xg_reduce_noalloc result) as numpy arrays."""
    L = A.lib()
    dt, shp = reduce_info(r)
    ref = xt.empty(shp, dt)
    got = xt.empty(shp, dt)
    assert _call(L.xg_reduce, r, ref) == A.OK, A.lib().xg_last_error_message()
    # start from a pattern the reduction must overwrite
    xt.noalias(got).assign(xt.cast(xt.from_numpy(np.full(shp, 7, np.int64)), A.DTYPE_NAMES[dt]))
    _sync()
    assert _call(L.xg_reduce_noalloc, r, got) == A.OK, A.lib().xg_last_error_message()
    _sync()
    return ref.to_numpy(), got.to_numpy()


def _rel(got, ref):
    ref = np.asarray(ref, np.float64)
    return float(np.max(np.abs(np.asarray(got, np.float64) - ref) / np.maximum(np.abs(ref), 1e-300)))


def _cases(rows, cols, seed):
    rng = np.random.default_rng(seed)
    x = rng.normal(1.0, 1.0, (rows, cols)).astype(np.float32)
    i = rng.integers(-1000, 1001, (rows, cols)).astype(np.int32)
    X, I = xt.from_numpy(x), xt.from_numpy(i)
    x64 = x.astype(np.float64)
    return [
        ("sum axis 0 f32", xt.sum(X, [0]), x64.sum(0), 2e-7),
        ("sum axis 1 f32", xt.sum(X, [1]), x64.sum(1), 2e-7),
        ("mean axis 0 f32", xt.mean(X, [0]), x64.mean(0), 2e-7),  # a float32 sum divided in double
        ("amax all f32", xt.amax(X), x.max(), 0.0),
        # the squares are float32 products (a float's scalar norm_sq), summed in double
        ("norm_l2 all f32", xt.norm_l2(X), np.sqrt((x64 * x64).sum()), 4e-7),
        ("variance axis 1 f32", xt.variance(X, [1]), x64.var(1), 1e-10),
        ("sum<int64> axis 0 i32", xt.sum(I, [0], dtype="int64"), i.astype(np.int64).sum(0), 0.0),
        # the evaluated float32 temporary X * X + X is exact enough that its
        # float64 sum bounds the float32 result like a plain sum
        ("immediate sum axis 0 of x*x+x", xt.sum(X * X + X, [0], evaluation_strategy="immediate"),
         (x * x + x).astype(np.float64).sum(0), 2e-7),
    ]


@pytest.mark.parametrize("idx", range(8))
def test_noalloc_matches_xg_reduce_twice(idx):
    # first a large input (split partials and a 12 MiB temporary in the
    # scratch), then a smaller one
    for rows, cols, seed in ((2048, 1536, 11), (1000, 300, 12)):
        name, r, want, tol = _cases(rows, cols, seed)[idx]
        ref, got = _both(r)
        assert got.dtype == ref.dtype and got.shape == ref.shape, name
        assert got.tobytes() == ref.tobytes(), f"{name} [{rows}x{cols}]: noalloc differs from xg_reduce"
        if tol == 0.0:
            assert np.array_equal(got, want), name
        else:
            assert _rel(got, want) <= tol, f"{name}: rel err {_rel(got, want)}"


def test_noalloc_allocates_nothing():
    import torch
    rng = np.random.default_rng(5)
    X = xt.from_numpy(rng.normal(0.0, 1.0, (2048, 1024)).astype(np.float32))
    out = xt.empty((1024,), "float32")
    r = xt.sum(X, [0])
    L = A.lib()
    assert _call(L.xg_reduce_noalloc, r, out) == A.OK  # set-up: the scratch address is resolved
    _sync()
    s = torch.cuda.Stream()  # a stream xg_reduce has no workspace for yet
    with torch.cuda.stream(s):
        torch.zeros(1, device="cuda").add_(1)  # the runtime's own set-up of the stream
    _sync()
    free0 = torch.cuda.mem_get_info()[0]
    with torch.cuda.stream(s):
        for _ in range(3):
            assert _call(L.xg_reduce_noalloc, r, out) == A.OK
    _sync()
    # xg_reduce would have allocated at least 1 MiB of workspace for this stream
    assert torch.cuda.mem_get_info()[0] == free0


def test_noalloc_rejects_beyond_its_scratch():
    L = A.lib()
    cap = L.xg_reduce_noalloc_max_bytes()
    assert cap == 16 << 20
    n = cap // 4 + 1024  # an immediate reduction's temporary of n float32 values
    X = xt.from_numpy(np.ones(n, np.float32))
    out = xt.from_numpy(np.array(-1.0, np.float32))
    r = xt.sum(X * 2.0, evaluation_strategy="immediate")
    assert _call(L.xg_reduce_noalloc, r, out) == A.OUT_OF_MEMORY
    assert b"scratch" in L.xg_last_error_message()
    _sync()
    assert out.to_numpy()[()] == -1.0  # nothing was enqueued
    # the same reduction, lazily, fits (no temporary of the expression)
    r2 = xt.sum(X * 2.0)
    assert _call(L.xg_reduce_noalloc, r2, out) == A.OK
    _sync()
    assert out.to_numpy()[()] == 2.0 * n

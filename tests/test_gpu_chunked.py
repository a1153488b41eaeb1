"""Chunked arrays (chunk/xchunked_array.hpp, chunk/xchunked_assign.hpp) on
the device and in host memory.

Each chunk is assigned the expression restricted to its region (the
per-chunk noalias of xchunked_assigner), so every element is computed by the
same program as a single-pass assignment: results are compared bit for bit
with xg_assign over the whole array, and with the reference's known
answers (test/test_xchunked_array.cpp)."""

import numpy as np
import pytest

import xtensor_amd as xt

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("storage", ["device", "host"])
def test_kat_indexed_access_and_assign(storage):
    """test_xchunked_array.cpp:20-82, 119-128."""
    a = xt.chunked_array((10, 10, 10), (2, 3, 4), storage=storage)
    a[3, 9, 8] = 1.0
    assert a[3, 9, 8] == 1.0
    a[3, 9, 8] = 2.0
    assert a[3, 9, 8] == 2.0
    a.fill(3.0)
    assert all(v == 3.0 for v in a)
    a1 = xt.chunked_array((2, 2, 2), (2, 3, 4), storage=storage)
    a1.assign(xt.broadcast(xt.scalar(3.0), (2, 2, 2)))
    assert all(v == 3.0 for v in a1)
    a2 = xt.chunked_array((32, 10, 10), (2, 3, 4), storage=storage)
    a2.assign(xt.broadcast(xt.scalar(3.0), (32, 10, 10)))
    a2 += a2
    assert all(v == 6.0 for v in a2)
    a2 += 2.0
    assert all(v == 8.0 for v in a2)
    a3 = np.arange(1.0, 10.0).reshape(3, 3)
    a4 = xt.chunked_array(a3, (2, 2), storage=storage)
    assert xt.is_chunked(a4) and not xt.is_chunked(xt.from_numpy(a3))
    assert list(a4) == list(range(1, 10))
    a6 = xt.chunked_array(a3, storage=storage)
    assert a6.chunk_shape == (3, 3)
    b = np.arange(1000.0).reshape(10, 10, 10)
    n = xt.chunked_array((10, 10, 10), (2, 2, 2), storage=storage)
    n.assign(xt.from_numpy(b))
    assert np.array_equal(n.to_numpy(), b)
    # chunk order: row-major over the grid; the first element of chunk k
    firsts = [c[(0, 0, 0)] if storage == "host" else c.to_numpy()[0, 0, 0] for _, _, c in n.chunks()]
    assert firsts[2] == b[0, 0, 4] and firsts[25] == b[2, 0, 0]


@pytest.mark.parametrize("storage", ["device", "host"])
def test_chunked_assign_matches_single_pass(storage):
    """c = a*b + exp(-a*a) with b broadcast, float64 [1000, 700] in chunks of
    [256, 300] (partial chunks on both edges): bit-identical to xg_assign of
    the whole array; then a chunked operand in the expression (the same chunk
    grid) reads its own chunks."""
    rng = np.random.default_rng(61)
    a_h = rng.uniform(-1, 1, (1000, 700))
    b_h = rng.uniform(-1, 1, 700)
    a, b = xt.from_numpy(a_h), xt.from_numpy(b_h)
    whole = xt.eval(a * b + xt.exp(-a * a)).to_numpy()
    c = xt.chunked_array((1000, 700), (256, 300), "float64", storage=storage)
    c.assign(a * b + xt.exp(-a * a))
    assert np.array_equal(c.to_numpy(), whole)
    d = xt.chunked_array((1000, 700), (256, 300), "float64", storage=storage)
    d.assign(c * 2.0 + b)
    assert np.array_equal(d.to_numpy(), whole * 2.0 + b_h)
    e = xt.chunked_array((1000, 700), (100, 100), "float64", storage=storage)
    with pytest.raises(NotImplementedError):
        e.assign(c + 1.0)
    with pytest.raises(RuntimeError):
        e.assign(xt.from_numpy(np.zeros((3, 700))))


@pytest.mark.parametrize("storage", ["device", "host"])
def test_blockwise_over_chunked(storage):
    """xt.blockwise reducers over a chunked array's own chunks: exact int64
    sums and float extrema against numpy."""
    rng = np.random.default_rng(62)
    h = rng.integers(-1000, 1000, (700, 500)).astype(np.int32)
    c = xt.chunked_array(h, (128, 96), storage=storage)
    got = xt.blockwise.sum(c, None, [0], dtype="int64").to_numpy()
    assert np.array_equal(got, h.astype(np.int64).sum(0))
    f = rng.normal(0, 1, (700, 500))
    cf = xt.chunked_array(f, (200, 200), storage=storage)
    assert xt.blockwise.amax(cf, None).to_numpy()[()] == f.max()
    np.testing.assert_allclose(xt.blockwise.sum(cf, None, [1]).to_numpy(), f.sum(1), rtol=1e-12)

"""Index views, filters, filtrations and masked views on the device
(views/xindex_view.hpp, views/xmasked_view.hpp), against the reference's
tests (test/test_xindex_view.cpp, arrays via tests/golden/kat_arrays.json),
numpy fancy indexing and the oracle's gather.

Reads are gathers fused into the expression's launch; assignments into an
index view scatter evaluated values; filtrations and masked views assign
where(cond, value, e).  All bit-exact (copies and IEEE + - *).
"""

import json
import os

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd import _abi as A

from helpers import Dual, oracle_for, same_bits
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_arrays.json")))


def kat(test, name, dtype=np.float64):
    r = KATS[f"test_xindex_view.cpp:xindex_view.{test}:{name}"]
    return np.array(r["values"], dtype=dtype).reshape(r["shape"])


def sync():
    import torch
    torch.cuda.synchronize()


def host(a):
    sync()
    return a.to_numpy()


IDX = [[1, 1], [1, 2], [2, 2]]


def test_kat_indices():
    """test_xindex_view.cpp:24-57."""
    h = np.random.default_rng(0).random((3, 3))
    e = xt.from_numpy(h)
    v = xt.index_view(e, IDX)
    assert v.shape == (3,)
    assert np.array_equal(host(xt.eval(v)), [h[1, 1], h[1, 2], h[2, 2]])
    xt.noalias(v).plus_assign(3)                        # v += 3
    assert host(e)[1, 1] == h[1, 1] + 3
    t = xt.eval(v + 3)
    assert host(t)[0] == h[1, 1] + 6
    v.assign(xt.broadcast(123, (3,)))                   # v = broadcast(123, v.shape())
    assert host(e)[1, 1] == host(e)[1, 2] == host(e)[2, 2] == 123
    v.assign(xt.from_numpy(kat("indices", "as")))       # v = {3, 3, 3}
    assert np.array_equal(host(xt.eval(v)), [3, 3, 3]) and host(e)[2, 2] == 3
    assert host(e)[0, 0] == h[0, 0]


def test_kat_indices_on_function():
    """test_xindex_view.cpp:112-129: index_view of an expression."""
    h = np.random.default_rng(1).random((3, 3))
    e = xt.from_numpy(h)
    v = xt.index_view(e * 3 - 120, IDX)
    fn = h * 3 - 120
    assert np.array_equal(host(xt.eval(v)), [fn[1, 1], fn[1, 2], fn[2, 2]])


def test_kat_boolean_and_fill():
    """test_xindex_view.cpp:68-103."""
    e = xt.from_numpy(kat("boolean", "e"))
    v = xt.filter(e, e > 0)
    assert host(xt.eval(v))[0] == 1
    xt.noalias(v).plus_assign(2)
    assert host(e)[1, 1] == 3 and host(xt.eval(v))[1] == 3
    xt.noalias(v).plus_assign(xt.from_numpy(np.array([1.0, 2.0, 3.0])))
    assert host(e)[1, 1] == 5 and host(e)[2, 2] == 6
    h2 = np.random.default_rng(2).random((3, 3, 3, 3))
    e2 = xt.from_numpy(h2)
    v2 = xt.filter(e2, e2 > 0.5)
    xt.noalias(v2).multiplies_assign(0)
    assert not (host(e2) > 0.5).any() and np.array_equal(host(e2), np.where(h2 > 0.5, 0.0, h2))
    f = xt.from_numpy(kat("fill", "e"))
    xt.filter(f, f < 1).fill(2)
    assert np.array_equal(host(f), kat("fill", "res"))


def test_kat_assign_scalar_filtration_filter():
    """test_xindex_view.cpp:140-165, 183-191."""
    a = xt.from_numpy(kat("assign_scalar", "a"))
    v = xt.filter(a, a >= 5)
    v.assign(100)
    assert np.array_equal(host(xt.eval(v)), [100, 100, 100])
    b = xt.from_numpy(kat("filtration", "a"))
    f = xt.filtration(b, b >= 5)
    f += 2
    assert np.array_equal(host(b), kat("filtration", "expected"))
    a = xt.from_numpy(kat("filter", "a"))
    bb = xt.from_numpy(kat("filter", "b"))
    xt.noalias(xt.filter(a, a > 3)).plus_assign(xt.filter(bb, bb > 3))
    assert np.array_equal(host(a), kat("filter", "expected"))
    c = xt.from_numpy(kat("const_adapt_filter", "b"))
    ad = xt.from_numpy(np.array([1.0, 2, 3, 4, 5, 6]).reshape(3, 2))
    xt.noalias(xt.filter(c, c > 3)).plus_assign(xt.filter(ad, ad < 4))
    assert np.array_equal(host(c), kat("const_adapt_filter", "expected"))


def test_kat_filter_column_major():
    """test_xindex_view.cpp:167-181."""
    a = xt.from_numpy(kat("filter_column_major", "a", np.int32))
    cond = xt.from_numpy(np.array([[[1, 1], [0, 0]], [[1, 1], [0, 0]]], bool))
    rc = host(xt.eval(xt.filter(a, cond, layout="column_major")))
    rr = host(xt.eval(xt.filter(a, cond)))
    assert np.array_equal(rc, kat("filter_column_major", "expc", np.int32))
    assert np.array_equal(rr, kat("filter_column_major", "expr", np.int32))


@pytest.mark.parametrize("shape,n,dtype", [((300, 500), 100000, np.float32), ((64, 3, 77), 5000, np.int32),
                                           ((1 << 20,), 1 << 20, np.float64)])
def test_gather_matches_numpy_and_oracle(shape, n, dtype):
    rng = np.random.default_rng(3)
    h = (rng.normal(size=shape) * 100).astype(dtype)
    idx = np.stack([rng.integers(0, s, n) for s in shape], axis=1)
    src = Dual.of(h)
    v = xt.index_view(src.dev, idx)
    got = host(xt.eval(v * 2 + 1))
    ref = h[tuple(idx.T)] * 2 + 1
    assert np.array_equal(got, ref.astype(got.dtype))
    o = Oracle({src.dev: h, v.index: idx.astype(np.int64)})
    assert same_bits(got, o.assign(v * 2 + 1))
    # a reduction over gathered values
    s = host(xt.eval(xt.sum(xt.cast(v, "float64"))))
    np.testing.assert_allclose(s, h[tuple(idx.T)].astype(np.float64).sum(), rtol=1e-12)


def test_gather_of_strided_view_and_out_of_range():
    rng = np.random.default_rng(4)
    h = rng.normal(size=(40, 50)).astype(np.float32)
    a = xt.from_numpy(h)
    sv = xt.view(a, xt.range_(39, None, -2), xt.range_(1, 50, 3))   # reversed, stepped
    hv = h[39::-2, 1:50:3]
    idx = np.stack([rng.integers(0, hv.shape[0], 1000), rng.integers(0, hv.shape[1], 1000)], axis=1)
    got = host(xt.eval(xt.index_view(sv, idx)))
    assert np.array_equal(got, hv[tuple(idx.T)])
    bad = np.array([[0, 0], [1000, 0], [-5, 2]])      # outside the view's extent read as 0
    got = host(xt.eval(xt.index_view(a, bad)))
    assert got[0] == h[0, 0] and got[1] == 0 and got[2] == 0


def test_scatter_large_and_skip_out_of_range():
    rng = np.random.default_rng(5)
    n = 1 << 20
    h = np.zeros((1024, 2048), np.float64)
    a = xt.from_numpy(h)
    flat = rng.permutation(h.size)[:n]
    idx = np.stack(np.unravel_index(flat, h.shape), axis=1)
    vals = rng.normal(size=n)
    xt.index_view(a, idx).assign(xt.from_numpy(vals))
    exp = h.copy()
    exp[tuple(idx.T)] = vals
    assert np.array_equal(host(a), exp)
    xt.index_view(a, np.array([[5000, 0], [0, 0]])).assign(xt.from_numpy(np.array([7.0, 9.0])))
    r = host(a)
    assert r[0, 0] == 9.0 and np.array_equal(r[1:], exp[1:])


def test_masked_view_assignment():
    """views/xmasked_view.hpp: assignment through a (broadcast) mask."""
    h = np.arange(24, dtype=np.float32).reshape(4, 6)
    a = xt.from_numpy(h)
    mask = xt.from_numpy(np.array([True, False, True, False, False, True]))
    xt.masked_view(a, mask).assign(-1)
    exp = h.copy()
    exp[:, [0, 2, 5]] = -1
    assert np.array_equal(host(a), exp)
    m = xt.masked_view(a, a > 10)
    m *= 2
    exp2 = np.where(exp > 10, exp * 2, exp)
    assert np.array_equal(host(a), exp2)


def test_gather_with_runtime_compiled_kernel():
    rng = np.random.default_rng(6)
    h = rng.normal(size=(2000, 3000))
    a = xt.from_numpy(h)
    idx = np.stack([rng.integers(0, 2000, 1 << 21), rng.integers(0, 3000, 1 << 21)], axis=1)
    A.lib().xg_debug_jit(2)
    try:
        out = xt.eval(xt.exp(xt.index_view(a, idx)) - 1.0)
        kern = xt.last_kernel()
    finally:
        A.lib().xg_debug_jit(1)
    A.lib().xg_debug_force_dynamic(1)
    try:
        out2 = xt.eval(xt.exp(xt.index_view(a, idx)) - 1.0)
        assert xt.last_kernel() == "dynamic"
    finally:
        A.lib().xg_debug_force_dynamic(0)
    assert kern.startswith("jit:")
    assert same_bits(host(out), host(out2))


# ---------------------------------------------------------------- argwhere on the device
def test_kat_argwhere():
    """test_xoperation.cpp:562-595: argwhere of a 1-d int array, a 2-d int
    array, equal(b, 0), and an all-true [3, 3, 3] bool array."""
    a = xt.from_numpy(np.array([1, 0, 3], np.int32))
    assert np.array_equal(host(xt.argwhere(a)), [[0], [2]])
    assert xt.last_kernel() == "argwhere"
    b = xt.from_numpy(np.array([[0, 2, 1], [2, 1, 0]], np.int32))
    assert np.array_equal(host(xt.argwhere(b)), [[0, 1], [0, 2], [1, 0], [1, 1]])
    assert np.array_equal(host(xt.argwhere(xt.equal(b, 0))), [[0, 0], [1, 2]])
    d = xt.from_numpy(np.ones((3, 3, 3), bool))
    dn = host(xt.argwhere(d))
    assert dn.shape == (27, 3) and np.array_equal(dn[-1], [2, 2, 2])
    assert xt.count_nonzero(d) == 27


@pytest.mark.parametrize("layout", ["row_major", "column_major"])
def test_argwhere_against_numpy(layout):
    """Random conditions against numpy's argwhere (column-major order: the
    argwhere of the transpose with the index components reversed): a float
    comparison, an integer array read as truth values, a 3-d comparison with
    a broadcast operand, an empty result, and a [8192, 4096] float32 case."""
    rng = np.random.default_rng(51)

    def ref(c):
        c = np.asarray(c, bool)
        if layout == "column_major":
            return np.argwhere(c.T)[:, ::-1]
        return np.argwhere(c)

    x = rng.uniform(0, 1, (1000, 300)).astype(np.float32)
    X = xt.from_numpy(x)
    assert np.array_equal(host(xt.argwhere(X > 0.5, layout)), ref(x > 0.5))
    k = (rng.integers(0, 3, (77, 33))).astype(np.int64)
    assert np.array_equal(host(xt.argwhere(xt.from_numpy(k), layout)), ref(k != 0))
    a3 = rng.normal(0, 1, (20, 30, 40))
    b1 = rng.normal(0, 1, (40,))
    A3, B1 = xt.from_numpy(a3), xt.from_numpy(b1)
    assert np.array_equal(host(xt.argwhere(A3 > B1, layout)), ref(a3 > b1))
    got = xt.argwhere(A3 > 100.0, layout)
    assert got.shape == (0, 3)
    big = rng.normal(1.0, 1.0, (8192, 4096)).astype(np.float32)
    BIG = xt.from_numpy(big)
    assert np.array_equal(host(xt.argwhere(BIG > 1.5, layout)), ref(big > 1.5))


def test_filter_reads_device_indices():
    """filter(e, cond) = index_view(e, argwhere(cond)) with the indices
    listed on the device: values and their order match numpy's boolean
    selection (row-major), also for a large array."""
    rng = np.random.default_rng(52)
    e = rng.normal(0, 1, (4096, 1000)).astype(np.float32)
    E = xt.from_numpy(e)
    got = host(xt.eval(xt.filter(E, E > 1.0)))
    assert np.array_equal(got, e[e > 1.0])


def test_functor_view_real_imag():
    """xfunctor_view (views/xfunctor_view.hpp:400-470) as a node of the fused
    expression; real(e) of real data is e itself and assignable; imag(e) is
    zeros of e's type, NaN included."""
    rng = np.random.default_rng(53)
    x = rng.normal(0, 1, (300, 200))
    x[0, 0] = np.nan
    X = xt.from_numpy(x)
    fv = xt.eval(xt.functor_view(X, lambda v: xt.abs(v))).to_numpy()
    assert np.array_equal(fv, np.abs(x), equal_nan=True)
    assert np.array_equal(xt.eval(xt.real(X)).to_numpy(), x, equal_nan=True)
    im = xt.eval(xt.imag(X)).to_numpy()
    assert im.dtype == np.float64 and np.array_equal(im, np.zeros_like(x))
    xt.noalias(xt.real(X)).assign(X * 2.0)
    assert np.array_equal(X.to_numpy(), x * 2.0, equal_nan=True)


@pytest.mark.parametrize("w,dtype,n", [(56, "float32", 64), (60, "float32", 64), (28, "int16", 160),
                                       (28, "float64", 64)])
def test_strided_rows_with_granule_fill_reads(w, dtype, n):
    """y[:, :, 0:56:2, :] = x[:, :, 0:56:2, :] + 1 on outputs past 64 MiB whose
    rows end (and for w = 60 start) inside 64-byte granules: the kernels read
    the rest of those granules before storing.  The written rows are exact,
    every byte of the rows left out keeps its (random, NaN-holding) pattern,
    and the result is bit-identical with the reads switched off."""
    rng = np.random.default_rng(61)
    shape = (n, 256, 56, w)
    if dtype == "int16":
        x = rng.integers(-30000, 30000, shape, dtype=np.int16)
    else:
        x = rng.normal(0, 1, shape).astype(dtype)
    pattern = rng.integers(0, 256, int(np.prod(shape)) * x.itemsize, dtype=np.uint8).view(x.dtype).reshape(shape)
    X = xt.from_numpy(x)
    outs = []
    try:
        for fill in (1, 0):
            A.lib().xg_debug_fill_granules(fill)
            y = xt.from_numpy(pattern)
            view = (xt.all_(), xt.all_(), xt.range_(0, 56, 2), xt.all_())
            xt.noalias(xt.view(y, *view)).assign(xt.cast(xt.view(X, *view) + 1, dtype))
            sync()
            outs.append(y.to_numpy())
    finally:
        A.lib().xg_debug_fill_granules(1)
    got = outs[0]
    assert same_bits(got, outs[1])
    assert same_bits(got[:, :, 1::2, :], pattern[:, :, 1::2, :])
    assert same_bits(got[:, :, 0::2, :], (x[:, :, 0::2, :] + 1).astype(dtype))

"""GPU parity of axis reductions (xg_reduce) against the oracle.

Bar: integer reductions, amax / amin and norm_linf bit-exact (exact
arithmetic, order independent); float sums / means / norms within a stated
relative error of the long-double reference (the GPU tree order differs
from xtensor's sequential fold by design, and is more accurate):
  * float64 data: rel <= 1e-14
  * float32 / bf16 data (float32 accumulation with compensation): rel <= 2e-7
    on the float32 result (1 ulp is 6e-8 relative), 1e-12 for results kept
    in double before rounding (mean, norms).
"""

import math

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd import _abi as A

from helpers import Dual, oracle_for, same_bits, to_bf16_bits

pytestmark = pytest.mark.gpu


def sync():
    import torch
    torch.cuda.synchronize()


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    den = np.maximum(np.abs(ref), 1e-300)
    return float(np.max(np.abs(got - ref) / den)) if ref.size else 0.0


def run(r: xt.Reducer):
    out = xt.eval(r)
    sync()
    return out.to_numpy(), xt.last_kernel()


# ---------------------------------------------------------------- reference KATs on the device
def test_kat_features_sum():
    """test_xreducer.cpp:222-240, 355-362, 475-481."""
    a = np.ones((3, 2, 4, 6, 5))
    a[1, :, 1, :, 1] = 2
    d = xt.from_numpy(a)
    r, _ = run(xt.sum(d, [1, 3]))
    exp = np.full((3, 4, 5), 12.0)
    exp[1, 1, 1] = 24
    assert np.array_equal(r, exp)
    assert run(xt.sum(d))[0][()] == 732


def test_kat_small_sums_and_means():
    """test_xreducer.cpp:390-397, 442-473, 542-569; test_xmath.cpp:857-870."""
    m = xt.from_numpy(np.array([[1.0, 2.0], [3.0, 4.0]]))
    assert np.array_equal(run(xt.sum(m, [0]))[0], [4.0, 6.0])
    u = xt.from_numpy(np.ones((2, 4)))
    assert np.array_equal(run(xt.sum(u, [-1]))[0], [4.0, 4.0])
    c = xt.from_numpy(np.ones(1000, np.uint8))
    r, _ = run(xt.sum(c))
    assert r.dtype == np.int32 and r[()] == 1000
    x = xt.from_numpy(np.array([[-1.0, 0.0], [1.0, 0.0]]))
    assert run(xt.mean(x))[0][()] == 0.0
    assert np.array_equal(run(xt.mean(x, [1]))[0], [-0.5, 0.5])
    c2 = xt.from_numpy(np.array([1, 2], np.uint8))
    assert run(xt.mean(c2))[0][()] == 1.5
    v = xt.from_numpy(np.array([[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]]))
    assert np.array_equal(run(xt.mean(v, [0]))[0], [1.5, 1.5, 1.5])
    assert run(xt.mean(v, [0, 1]))[0][()] == 1.5


def test_kat_keep_dims_initial_empty():
    """test_xreducer.cpp:964-1079."""
    a_h = np.arange(625, dtype=np.float64).reshape(5, 5, 5, 5)
    a = xt.from_numpy(a_h)
    r, _ = run(xt.sum(a, [0, 1], keep_dims=True))
    assert r.shape == (1, 1, 5, 5) and np.array_equal(r[0, 0], a_h.sum(axis=(0, 1)))
    base, _ = run(xt.sum(a, [0, 2], keep_dims=True))
    assert np.array_equal(run(xt.sum(a, [0, 2], keep_dims=True, initial=5))[0], base + 5)
    assert np.all(run(xt.prod(a, [1, 2], initial=0))[0] == 0)
    x = xt.from_numpy(np.zeros((0, 1), np.int32))
    r0, _ = run(xt.sum(x, [0], keep_dims=True))
    assert r0.shape == (1, 1) and r0[0, 0] == 0
    assert run(xt.sum(x, [1], keep_dims=True))[0].shape == (0, 1)
    e = xt.from_numpy(np.ones((1, 2, 0, 1)))
    assert math.isnan(run(xt.mean(e))[0][()])
    r2, _ = run(xt.mean(e, [2]))
    assert r2.shape == (1, 2, 1) and np.all(np.isnan(r2))
    ones = xt.from_numpy(np.ones((1, 1, 2, 4), np.int32))
    assert np.array_equal(run(xt.sum(ones, [1, 2], keep_dims=True))[0], [[[[2, 2, 2, 2]]]])
    b = xt.from_numpy(np.array([[1, 2, 3], [4, 5, 6]], np.int32))
    assert np.array_equal(run(xt.sum(b, []))[0], [[1, 2, 3], [4, 5, 6]])


def test_kat_norms_and_extrema():
    """test_xnorm.cpp:170-187, test_xmath.cpp:147-168."""
    a = xt.from_numpy(np.array([[-1.0, 2.0], [-3.0, 4.0]]))
    assert run(xt.norm_l0(a))[0][()] == 4
    assert run(xt.norm_l1(a))[0][()] == 10.0
    assert run(xt.norm_sq(a))[0][()] == 30.0
    assert run(xt.norm_linf(a))[0][()] == 4.0
    for vals, mn, mx in [([-10.0], -10.0, -10.0), ([-10.0, -20.0], -20.0, -10.0), ([-10.0, 20.0], -10.0, 20.0)]:
        d = xt.from_numpy(np.array(vals))
        assert run(xt.amin(d))[0][()] == mn and run(xt.amax(d))[0][()] == mx


@pytest.mark.parametrize("axes", [[1, 0], [0, 2], [0, 0]])
def test_bad_axes_raise(axes):
    a = xt.from_numpy(np.ones((3, 2), np.int32))
    with pytest.raises(RuntimeError):
        xt.eval(xt.sum(a, axes))


# ---------------------------------------------------------------- fused expression + reductions (f64)
def test_fused_row_sum_and_column_amax_f64():
    """s = sum(c, {1}), m = amax(c, {0}) for c = a*b + exp(-a*a) [2048, 1024] f64."""
    rng = np.random.default_rng(0)
    a_h = rng.uniform(-1, 1, (2048, 1024))
    b_h = rng.uniform(-1, 1, (1024,))
    a, b = xt.from_numpy(a_h), xt.from_numpy(b_h)
    c = xt.eval(a * b + xt.exp(-a * a))
    c_h = c.to_numpy()
    C_ = Dual(c, c_h)
    o = oracle_for(C_)
    s, ks = run(xt.sum(c, [1]))
    m, km = run(xt.amax(c, [0]))
    assert rel_err(s, o.reduce(xt.sum(c, [1]))) <= 1e-14
    assert rel_err(s, o.reduce(xt.sum(c, [1]), high_precision=True)) <= 1e-14
    assert np.array_equal(m, o.reduce(xt.amax(c, [0])))
    # fused: the expression evaluated inside the reduction
    sf, _ = run(xt.sum(a * b + xt.exp(-a * a), [1]))
    assert rel_err(sf, s) <= 1e-14
    print("kernels", ks, km)


# ---------------------------------------------------------------- float32 / bf16 / int32 matrix reductions
def matrix(rows, cols, seed=2):
    rng = np.random.default_rng(seed)
    return rng.normal(1, 1, (rows, cols)).astype(np.float32)


@pytest.mark.parametrize("rows,cols", [(512, 4096), (8192, 4096), (3, 1000), (1000, 3), (1, 1)])
def test_f32_sum_mean_amax_norm(rows, cols):
    x = matrix(rows, cols)
    X = Dual.of(x)
    o = oracle_for(X)
    errs = {}
    for ax in ([0], [1]):
        for kind in ("sum", "mean"):
            r = xt.Reducer(kind, X.dev, ax)
            got, k = run(r)
            ref = o.reduce(r, high_precision=True)
            tol = 2e-7 if kind == "sum" else 2e-7
            errs[(kind, ax[0], k)] = rel_err(got, ref)
            assert errs[(kind, ax[0], k)] <= tol, (kind, ax, k)
    got, _ = run(xt.amax(X.dev))
    assert got[()] == x.max()
    got, _ = run(xt.amax(X.dev, [0]))
    assert np.array_equal(got, x.max(axis=0))
    r = xt.norm_l2(X.dev)
    got, k = run(r)
    errs[("norm_l2", k)] = rel_err(got, o.reduce(r, high_precision=True))
    assert errs[("norm_l2", k)] <= 1e-12 or errs[("norm_l2", k)] <= 2e-7 * 2
    print(errs)


def test_f32_transposed_view():
    x = matrix(1024, 2048, seed=5)
    X = Dual.of(x)
    XT = X.transpose()
    o = oracle_for(XT)
    for ax in ([0], [1]):
        r = xt.sum(XT.dev, ax)
        got, k = run(r)
        assert rel_err(got, o.reduce(r, high_precision=True)) <= 2e-7, (ax, k)


def test_bf16_reductions():
    x = matrix(512, 4096, seed=7)
    X = Dual.of(x, "bfloat16")
    o = oracle_for(X)
    for r in (xt.sum(X.dev, [0]), xt.sum(X.dev, [1]), xt.mean(X.dev, [1]), xt.norm_l2(X.dev)):
        got, k = run(r)
        assert rel_err(got, o.reduce(r, high_precision=True)) <= 2e-7, k
    got, _ = run(xt.amax(X.dev))
    assert got[()] == o.reduce(xt.amax(X.dev))[()]


@pytest.mark.parametrize("shape", [(8192, 4096), (257, 3001)])
def test_int32_sum_into_int64_exact(shape):
    rng = np.random.default_rng(2)
    x = rng.integers(-1000, 1001, shape, dtype=np.int32)
    X = Dual.of(x)
    for ax in ([0], [1]):
        got, k = run(xt.sum(X.dev, ax, dtype="int64"))
        assert got.dtype == np.int64
        assert np.array_equal(got, x.astype(np.int64).sum(axis=ax[0])), k


@pytest.mark.parametrize("rows,cols", [(8191, 4096), (1001, 4100), (129, 2050), (7, 4097), (3000, 1023),
                                       (64, 100003), (2, 1 << 20)])
def test_fast_kernel_tile_and_split_edges(rows, cols):
    """Compiled fast kernels around their edges: column tiles of 128 or 64
    lanes, a tail chunk (cols not a multiple of the vector width), splits whose
    last row or chunk range is short (rows past the end are clamped and
    dropped on the whole-tile path).  Integer sums into int64 and maxima are
    exact; float64 sums within 1e-14 of the long-double oracle."""
    rng = np.random.default_rng(rows * 7 + cols)
    xi = rng.integers(-1000, 1001, (rows, cols), dtype=np.int32)
    XI = Dual.of(xi)
    xf = rng.normal(0.5, 1, (rows, cols))
    XF = Dual.of(xf)
    xs = xf.astype(np.float32)
    XS = Dual.of(xs)
    for ax in ([0], [1], None):
        npax = None if ax is None else ax[0]
        got, k = run(xt.sum(XI.dev, ax, dtype="int64"))
        assert np.array_equal(got, xi.astype(np.int64).sum(axis=npax)), (ax, k)
        got, k = run(xt.amax(XI.dev, ax))
        assert np.array_equal(got, xi.max(axis=npax)), (ax, k)
        got, k = run(xt.amax(XS.dev, ax))
        assert np.array_equal(got, xs.max(axis=npax)), (ax, k)
        got, k = run(xt.sum(XF.dev, ax))
        ref = oracle_for(XF).reduce(xt.sum(XF.dev, ax), high_precision=True)
        assert rel_err(got, ref) <= 1e-14, (ax, k)


def test_column_split_merge_repeats_and_streams():
    """Split column sums (partials in the per-stream workspace, merged by a
    second kernel): repeated launches, two shapes interleaved, and launches on
    two streams at once all give the exact int64 sums."""
    import torch
    rng = np.random.default_rng(5)
    xs = [rng.integers(-1000, 1001, s, dtype=np.int32) for s in ((8192, 4096), (3001, 1536))]
    X = [Dual.of(x) for x in xs]
    refs = [x.astype(np.int64).sum(axis=0) for x in xs]
    outs = [xt.empty((x.shape[1],), "int64") for x in xs]
    plans = [xt.plan(o, xt.sum(d.dev, [0], dtype="int64")) for o, d in zip(outs, X)]
    for rep in range(6):
        plans[rep % 2].run()
        sync()
        assert np.array_equal(outs[rep % 2].to_numpy(), refs[rep % 2]), rep
    streams = [torch.cuda.Stream() for _ in range(2)]
    for rep in range(3):
        for k, st in enumerate(streams):
            plans[k].run(stream=st.cuda_stream)
        torch.cuda.synchronize()
        for k in range(2):
            assert np.array_equal(outs[k].to_numpy(), refs[k]), (rep, k)


def test_int32_sum_wraps_like_cpp():
    """sum(int32) has value type int: the total wraps modulo 2^32 exactly as xtensor's does."""
    x = np.full((4, 1 << 16), 1 << 20, np.int32)
    X = Dual.of(x)
    got, _ = run(xt.sum(X.dev, [1]))
    assert np.array_equal(got, oracle_for(X).reduce(xt.sum(X.dev, [1])))


# ---------------------------------------------------------------- fuzz over shapes / axes / kinds / dtypes
KINDS = ["sum", "prod", "amax", "amin", "mean", "norm_sq", "norm_l1", "norm_linf", "norm_l0", "norm_l2"]


@pytest.mark.parametrize("seed", range(10))
def test_random_reductions(seed):
    rng = np.random.default_rng(500 + seed)
    for _ in range(12):
        nd = int(rng.integers(1, 5))
        shape = tuple(int(rng.integers(1, 9)) for _ in range(nd))
        dt = ["int32", "int64", "uint8", "float32", "float64", "int16"][rng.integers(6)]
        kind = KINDS[rng.integers(len(KINDS))]
        if dt.startswith("float"):
            v = rng.uniform(0.5, 2.0, shape).astype(dt)
        else:
            v = rng.integers(-3 if dt != "uint8" else 0, 4, shape).astype(dt)
        X = Dual.of(v)
        # views: maybe transpose or reverse
        if nd >= 2 and rng.random() < 0.3:
            X = X.transpose()
        axes = sorted(set(int(a) for a in rng.integers(0, nd, size=rng.integers(0, nd + 1))))
        if rng.random() < 0.2:
            axes = None
        r = xt.Reducer(kind, X.dev, axes, keep_dims=bool(rng.random() < 0.3))
        got, k = run(r)
        o = oracle_for(X)
        ref = o.reduce(r)
        assert got.shape == ref.shape and got.dtype == ref.dtype
        if ref.dtype.kind in "iub" or kind in ("amax", "amin", "norm_linf"):
            assert np.array_equal(got, ref), (kind, dt, shape, axes, k)
        else:
            hp = o.reduce(r, high_precision=True)
            tol = 1e-13 if ref.dtype == np.float64 and dt == "float64" else 3e-7
            if kind == "prod":
                tol = 1e-5 if dt == "float32" else 1e-12
            assert rel_err(got, hp) <= tol, (kind, dt, shape, axes, k, rel_err(got, hp))


# ---------------------------------------------------------------- variance, stddev, minmax, mean<T>
# variance / stddev: two passes (the inner mean, then the mean of squared
# deviations), both evaluated in the types xtensor gives them (double for
# float32 data); bar rel <= 1e-12 against the long-double reference.
# minmax: exact (min / max are exact; test data has no signed-zero ties).
_KATS = None


def xr_kat(name):
    import json
    import os
    global _KATS
    if _KATS is None:
        _KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_arrays.json")))
    r = _KATS[f"test_extended_xmath_reducers.cpp:xtest_extended_xmath.var_std:{name}"]
    return np.array(r["values"], np.float64).reshape(r["shape"])


def test_variance_stddev_kats_on_device():
    """test_extended_xmath_reducers.cpp:1824-2833 (numpy goldens, 16 digits)."""
    a = xt.from_numpy(xr_kat("py_a"))
    tol = dict(rtol=1e-13, atol=0)
    np.testing.assert_allclose(run(xt.variance(a))[0][()], xr_kat("py_vr_all"), **tol)
    np.testing.assert_allclose(run(xt.stddev(a))[0][()], xr_kat("py_st_all"), **tol)
    np.testing.assert_allclose(run(xt.variance(a, [0, 1, 2, 3], ddof=1))[0][()], xr_kat("py_vr_all_ddof"), **tol)
    np.testing.assert_allclose(run(xt.variance(a, [0, 2]))[0], xr_kat("py_vr"), **tol)
    np.testing.assert_allclose(run(xt.variance(a, [0, 2], ddof=1))[0], xr_kat("py_vr_ddof"), **tol)
    np.testing.assert_allclose(run(xt.stddev(a, [0, 2]))[0], xr_kat("py_st"), **tol)


@pytest.mark.parametrize("dtype,shape,axes,ddof", [
    ("float32", (512, 3000), [0], 0), ("float32", (512, 3000), [1], 1), ("float32", (64, 33, 70), [0, 2], 0),
    ("float64", (300, 777), None, 1), ("int32", (200, 999), [1], 0), ("float32", (2048, 4096), [1], 0),
    ("float32", (4096, 2048), [0], 3), ("float32", (7, 1, 5), [1], 0),
])
def test_variance_stddev_match_oracle(dtype, shape, axes, ddof):
    rng = np.random.default_rng(11)
    if dtype == "int32":
        h = rng.integers(-1000, 1000, shape, dtype=np.int32)
    else:
        h = rng.normal(2.0, 1.5, shape).astype(dtype)
    a = Dual.of(h)
    o = oracle_for(a)
    for rf in (lambda e: xt.variance(e, axes, ddof=ddof), lambda e: xt.stddev(e, axes)):
        got = xt.eval(rf(a.dev))
        sync()
        ref = o.reduce(rf(a.dev), high_precision=True)
        dt, shp = o.reduce_info(rf(a.dev))
        assert got.dtype == xt.dtype_name(dt) and tuple(got.shape) == tuple(shp)
        assert rel_err(got.to_numpy(), ref) <= 1e-12


def test_variance_keep_dims_and_bad_ddof():
    h = np.random.default_rng(3).normal(size=(6, 10))
    a = xt.from_numpy(h)
    r, _ = run(xt.variance(a, [1], keep_dims=True))
    assert r.shape == (6, 1)
    np.testing.assert_allclose(r[:, 0], h.var(axis=1), rtol=1e-13)
    with pytest.raises(Exception):
        run(xt.variance(a, [1], ddof=61))
    with pytest.raises(Exception):
        xt.eval(xt.Reducer("sum", a, [1], ddof=1))


@pytest.mark.parametrize("dtype,shape", [("float32", (1000, 1000)), ("float64", (333, 77)), ("int32", (4096, 513)),
                                         ("uint8", (100, 37)), ("float32", (8192, 4096)), ("int16", (3,))])
def test_minmax_matches_oracle(dtype, shape):
    rng = np.random.default_rng(5)
    if dtype.startswith("float"):
        h = rng.normal(0, 10, shape).astype(dtype)
        h[np.abs(h) < 1e-3] = 1.0  # no signed-zero ties
        if h.size > 10:
            h.flat[rng.integers(0, h.size, 7)] = np.nan  # std::min / std::max skip NaN
    else:
        info = np.iinfo(dtype)
        h = rng.integers(info.min, info.max, shape, dtype=dtype, endpoint=True)
    a = Dual.of(h)
    got, _ = run(xt.minmax(a.dev))
    ref = oracle_for(a).reduce(xt.minmax(a.dev))
    assert got.dtype == np.dtype(dtype) and got.shape == (2,)
    assert np.array_equal(got, ref)
    if dtype.startswith("float"):
        assert got[0] == np.nanmin(h) and got[1] == np.nanmax(h)


def test_minmax_empty_and_expression():
    e = xt.from_numpy(np.zeros((0, 4), np.float32))
    got, _ = run(xt.minmax(e))
    assert got[0] == np.finfo(np.float32).max and got[1] == -np.finfo(np.float32).max
    h = np.random.default_rng(9).normal(size=(300, 400))
    a = Dual.of(h)
    got, _ = run(xt.minmax(xt.exp(a.dev) - 2.0 * a.dev))
    ref = oracle_for(a).reduce(xt.minmax(xt.exp(a.dev) - 2.0 * a.dev))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("shape,axes,acc", [((300, 1001), [1], "int64"), ((64, 64), None, "int32"),
                                            ((1000, 37), [0], "int64"), ((5, 7, 9), [0, 2], "int32")])
def test_mean_integer_accumulator_divides_in_integers(shape, axes, acc):
    """core/xmath.hpp:1827-1873: mean<T> with integral T divides by T(n) in integers."""
    h = np.random.default_rng(2).integers(-5000, 5000, shape, dtype=np.int32)
    a = Dual.of(h)
    r = xt.mean(a.dev, axes, dtype=acc)
    got, _ = run(r)
    ref = oracle_for(a).reduce(r)
    assert got.dtype == ref.dtype == np.dtype(acc)
    assert np.array_equal(got, ref)


def test_mean_float_accumulator():
    h = np.random.default_rng(4).normal(1, 1, (777, 1234)).astype(np.float32)
    a = Dual.of(h)
    r = xt.mean(a.dev, [1], dtype="float32")
    got, _ = run(r)
    assert got.dtype == np.float32
    ref = oracle_for(a).reduce(r, high_precision=True)
    assert rel_err(got, ref) <= 2e-7


# ---------------------------------------------------------------- evaluation_strategy::immediate
IMMEDIATE_AXES = {(3, 3, 3): [None, [1], [0, 2], [1, 2]],
                  (4, 3, 6, 2, 7): [None, [1], [0, 2], [1, 2], [1, 3, 4], [0, 1, 4], [0, 1, 3], [0, 2, 3],
                                    [1, 2, 3]]}


@pytest.mark.parametrize("shape", list(IMMEDIATE_AXES))
def test_immediate_matches_lazy_kat(shape):
    """test_xreducer.cpp:704-771 on the device: immediate == lazy == the
    oracle's restatement of reduce_immediate, on arange data (exact)."""
    h = np.arange(np.prod(shape), dtype=np.float64).reshape(shape)
    a = Dual.of(h)
    o = oracle_for(a)
    for axes in IMMEDIATE_AXES[shape]:
        lazy, _ = run(xt.sum(a.dev, axes))
        imm, _ = run(xt.sum(a.dev, axes, evaluation_strategy="immediate"))
        ref = o.reduce(xt.sum(a.dev, axes, evaluation_strategy="immediate"))
        assert np.array_equal(lazy, imm) and np.array_equal(imm, ref), axes
        # an expression is evaluated first, then reduced
        imm2, _ = run(xt.sum(a.dev * 2.0 - 1.0, axes, evaluation_strategy="immediate"))
        assert np.array_equal(imm2, o.reduce(xt.sum(a.dev * 2.0 - 1.0, axes, evaluation_strategy="immediate"))), axes


def test_immediate_column_major_and_float_results():
    """Column-major shorts (test_xreducer.cpp:765-770) exactly, and float32
    data through an expression within the float bar of the long-double
    reference, for lazy and immediate alike."""
    rng = np.random.default_rng(51)
    h = rng.integers(-1000, 1000, (3, 2, 1, 5, 1)).astype(np.int16)  # the transpose of [1,5,1,2,3]
    base = Dual.of(h)
    ct = base.transpose()  # shape [1, 5, 1, 2, 3], column-major strides
    o = oracle_for(base, ct)
    for axes in ([0, 1, 2], [2, 3], [1, 3]):
        imm, _ = run(xt.sum(ct.dev, axes, evaluation_strategy="immediate"))
        lazy, _ = run(xt.sum(ct.dev, axes))
        ref = o.reduce(xt.sum(ct.dev, axes, evaluation_strategy="immediate"))
        assert np.array_equal(imm, lazy) and np.array_equal(imm, ref), axes
    x = rng.normal(1.0, 1.0, (64, 33, 129)).astype(np.float32)
    X = Dual.of(x)
    o = oracle_for(X)
    for axes in (None, [0], [2], [0, 2], [1, 2]):
        e = xt.sum(X.dev * X.dev + 1.0, axes, evaluation_strategy="immediate")
        got, _ = run(e)
        hp = o.reduce(e, high_precision=True)
        assert rel_err(got, hp) <= 1e-12, axes  # double sums of float32 values
        got, _ = run(xt.mean(X.dev, axes, evaluation_strategy="immediate"))
        # mean<void>(float32): a float32 sum (compensated), then divided in double
        assert rel_err(got, o.reduce(xt.mean(X.dev, axes), high_precision=True)) <= 2e-7, axes
    # variance: both passes evaluate first
    got, _ = run(xt.variance(X.dev, [1], evaluation_strategy="immediate"))
    lazy, _ = run(xt.variance(X.dev, [1]))
    assert rel_err(got, lazy) <= 1e-12


def test_prepared_multi_pass_and_immediate_reductions():
    """variance / stddev / minmax and evaluation_strategy::immediate as
    prepared plans: kept as descriptors, run as xg_reduce each execution,
    bit-identical to the direct calls, also after the inputs change."""
    import torch
    rng = np.random.default_rng(71)
    x = xt.from_numpy(rng.normal(0, 1, (300, 257)).astype(np.float32))
    cases = [xt.variance(x, [1], ddof=1), xt.stddev(x, [0]), xt.minmax(x),
             xt.sum(x * 2.0, [1], evaluation_strategy="immediate")]
    for r in cases:
        first = xt.eval(r)
        out = xt.empty(first.shape, first.dtype)
        pl = xt.plan(out, r)
        for rep in range(2):
            x.storage.view(torch.float32).mul_(1.5)
            pl.run()
            ref = xt.eval(r)
            torch.cuda.synchronize()
            assert same_bits(out.to_numpy(), ref.to_numpy()), (rep, r.kind)

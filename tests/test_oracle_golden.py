"""Pins the CPU oracle against the reference's own known-answer tests.

Every case cites the reference test it replays (file:line).  Arrays come from
tests/golden/kat_arrays.json (extracted from those tests by
tests/golden/extract_kats.py); expected values the tests state inline are
restated here with their citation.  Cases that no reference test fixes are
pinned by closed forms (exact integer / IEEE arithmetic against numpy, which
computes the same correctly rounded operations) further below.
"""

import json
import math
import os

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd import _abi as A

from helpers import Dual, oracle_for
from oracle.pyoracle import Oracle, OracleError

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_arrays.json")))
NP_OF_CTYPE = {"long": np.int64, "int": np.int32, "double": np.float64, "short": np.int16,
               "float": np.float32, "uint8_t": np.uint8, "size_t": np.uint64}


def kat(file, test, name):
    r = KATS[f"{file}:{test}:{name}"]
    return np.array(r["values"], dtype=NP_OF_CTYPE[r["ctype"]]).reshape(r["shape"])


def host(x):
    return Dual.of(np.asarray(x), device=False)


# ---------------------------------------------------------------- broadcasting
BV = "test_extended_broadcast_view.cpp"


def test_broadcasting_kats():
    """test_extended_broadcast_view.cpp:114-500 (numpy goldens)."""
    T = "xtest_extended.broadcasting"
    a1, a2 = host(kat(BV, T, "py_a1")), host(kat(BV, T, "py_a2"))
    b1, b2, b3 = host(kat(BV, T, "py_b1")), host(kat(BV, T, "py_b2")), host(kat(BV, T, "py_b3"))
    o = oracle_for(a1, a2, b1, b2, b3)
    r = o.assign(a1.dev + a2.dev)
    assert r.dtype == np.int64 and np.array_equal(r, kat(BV, T, "py_a1_a2"))
    assert np.array_equal(o.assign(a1.dev + b1.dev), kat(BV, T, "py_b1_a1"))
    assert np.array_equal(o.assign(a1.dev + b2.dev), kat(BV, T, "py_a1_b2"))
    got = o.assign(a2.dev + b3.dev)
    # the golden values were printed with 16 significant digits
    np.testing.assert_allclose(got, kat(BV, T, "py_a2_b3"), rtol=2e-16 * 8, atol=0)


def test_negative_slices_twod():
    """test_extended_broadcast_view.cpp:31-72: reversed / stepped strided views."""
    T = "xtest_extended.negative_slices_twod"
    a = kat(BV, T, "py_a")
    for sl, name in [((slice(None, -2), slice(None, None, -1)), "py_av0"),
                     ((slice(None, None, -1), slice(None, None, -1)), "py_av1"),
                     ((slice(1, -3), slice(-3, 2, -1)), "py_av2"),
                     ((slice(-1, -4, -1), slice(-3, 1, -2)), "py_av3"),
                     ((slice(-5, -2), slice(-3, 10)), "py_av5")]:
        v = host(a[sl])
        got = oracle_for(v).assign(v.dev)
        assert np.array_equal(got, kat(BV, T, name)), name


def test_negative_slices_math():
    """test_extended_broadcast_view.cpp:82-112: `view += view` (xsemantic
    operator+= goes through operator=, which uses a temporary when the
    operands overlap, core/xsemantic.hpp:408-411, 620-640)."""
    T = "xtest_extended.negative_slices_math"
    a = kat(BV, T, "py_a0").copy()
    lhs, rhs = host(a[0:-2]), host(a[:3:-1])
    tmp = oracle_for(lhs, rhs).assign(lhs.dev + rhs.dev)
    a[0:-2] = tmp
    assert np.array_equal(a, kat(BV, T, "py_a"))
    lhs, rhs = host(a[::-2]), host(a[::2])
    tmp = oracle_for(lhs, rhs).assign(lhs.dev + rhs.dev)
    a[::-2] = tmp
    assert np.array_equal(a, kat(BV, T, "py_at"))


def test_broadcast_into_view():
    """test_extended_broadcast_view.cpp:811-953: assignment into views with broadcasting."""
    T = "xtest_extended.broadcast_into_view"
    a = kat(BV, T, "py_a0").copy()
    b1 = host(kat(BV, T, "py_b1"))
    Oracle({b1.dev: b1.host}).assign(b1.dev, out=a[:, :, 2])
    assert np.array_equal(a, kat(BV, T, "py_a2"))
    ar2 = host(kat(BV, T, "py_ar2"))
    Oracle({ar2.dev: ar2.host}).assign(ar2.dev, out=a[:, 0, np.newaxis, :])
    assert np.array_equal(a, kat(BV, T, "py_a3"))
    r7 = host(np.arange(7, dtype=np.int64))
    Oracle({r7.dev: r7.host}).assign(r7.dev, out=a[:, 2, :])
    assert np.array_equal(a, kat(BV, T, "py_a4"))
    for sl, name in [((1, slice(None), slice(None)), "py_ax1"),
                     ((slice(0, 2), slice(0, 2), slice(None)), "py_ax2"),
                     ((slice(None), slice(0, 2), slice(None)), "py_ax4"),
                     ((1, slice(0, 2, 2), slice(None)), "py_ax3")]:
        a = kat(BV, T, "py_a0").copy()
        Oracle({}).assign(xt.scalar(-1), out=a[sl])
        assert np.array_equal(a, kat(BV, T, name)), name


def test_broadcast_funcs_view():
    """test_extended_broadcast_view.cpp:971-1031."""
    T = "xtest_extended.broadcast_funcs_view"
    a0 = host(kat(BV, T, "py_a0"))
    a1 = host(kat(BV, T, "py_a1"))
    a4 = host(kat(BV, T, "py_a4").reshape(3, 1))
    a = kat(BV, T, "py_a0").copy()
    oracle_for(a0, a4).assign(a0.dev * 2 + a4.dev, out=a[0:2])
    assert np.array_equal(a, kat(BV, T, "py_a0_cp1"))
    # view(a, 0, all, all) *= a1 * 7 + a4  -> a[0] = a[0] * (a1 * 7 + a4)
    cur = host(a[0])
    tmp = oracle_for(cur, a1, a4).assign(cur.dev * (a1.dev * 7 + a4.dev))
    a[0] = tmp
    assert np.array_equal(a, kat(BV, T, "py_a0_cp2"))
    x0 = host(kat(BV, T, "py_x0")[1])
    oracle_for(x0).assign(x0.dev, out=a[0, :, ::2])
    assert np.array_equal(a, kat(BV, T, "py_a0_cp3"))
    full = oracle_for(a0).assign(a0.dev * a0.dev + 3)
    assert np.array_equal(full[0], kat(BV, T, "py_x1"))


# ---------------------------------------------------------------- accumulators
AC = "test_xaccumulator.cpp"


def test_cumsum_one_d_promotion():
    """test_xaccumulator.cpp:23-34: cumsum(short) is int, {1,3,6,10}."""
    a = host(np.array([1, 2, 3, 4], np.int16))
    o = oracle_for(a)
    for axis in (None, 0):
        dt, _ = o.scan_info(a.dev, axis)
        assert dt == A.INT32
        assert np.array_equal(o.scan(a.dev, axis), kat(AC, "xaccumulator.one_d", "expected"))


def test_cumsum_four_d_all_axes():
    """test_xaccumulator.cpp:47-152: every axis, negative axis, flattened."""
    T = "xaccumulator.four_d"
    a = host(kat(AC, T, "arg_0"))
    o = oracle_for(a)
    assert np.array_equal(o.scan(a.dev), kat(AC, T, "expected"))
    for ax in range(4):
        assert np.array_equal(o.scan(a.dev, ax), kat(AC, T, f"expected_{ax}")), ax
    assert np.array_equal(o.scan(a.dev, -1), kat(AC, T, "expected_3"))


def test_cumsum_misc():
    """test_xaccumulator.cpp:35-45 (dim one), 154-162 (xtensor), 196-210 (empty)."""
    arr = host(kat(AC, "xaccumulator.dim_one", "arr"))
    assert np.array_equal(oracle_for(arr).scan(arr.dev, 0), arr.host)
    arrt = arr.transpose()
    assert np.array_equal(oracle_for(arrt).scan(arrt.dev, 1), arrt.host)
    x = host(kat(AC, "xaccumulator.xtensor", "arr"))
    assert np.array_equal(oracle_for(x).scan(x.dev, 0), kat(AC, "xaccumulator.xtensor", "expected"))
    e = host(np.ones((3, 4, 0)))
    o = oracle_for(e)
    assert o.scan(e.dev).shape == (0,)
    assert o.scan(e.dev, 2).shape == (3, 4, 0)
    assert o.scan(e.dev, 1).shape == (3, 4, 0)


def test_cumprod():
    """test_xaccumulator.cpp:164-179."""
    T = "xaccumulator.cumprod"
    a = host(kat(AC, T, "arg_0"))
    o = oracle_for(a)
    assert np.array_equal(o.scan(a.dev, kind="cumprod"), kat(AC, T, "expected"))
    assert np.array_equal(o.scan(a.dev, 0, kind="cumprod"), kat(AC, T, "expected_0"))
    assert np.array_equal(o.scan(a.dev, 1, kind="cumprod"), kat(AC, T, "expected_1"))


def test_scan_bad_axis():
    """reducers/xaccumulator.hpp:231-234: axis >= dimension throws."""
    a = host(np.ones((2, 3)))
    with pytest.raises(OracleError) as ei:
        oracle_for(a).scan(a.dev, 2)
    assert ei.value.code == A.BAD_AXES


# ---------------------------------------------------------------- reducers
def features_array():
    """test_xreducer.cpp:75-88: ones({3,2,4,6,5}) with m_a(1,i,1,j,1) = 2."""
    a = np.ones((3, 2, 4, 6, 5))
    a[1, :, 1, :, 1] = 2
    return a


def test_reducer_features():
    """test_xreducer.cpp:222-240, 277-286, 355-362, 475-481: sum over {1,3}
    is 12 everywhere but 24 at (1,1,1); shape {3,4,5}; sum of all is 732."""
    a = host(features_array())
    o = oracle_for(a)
    r = o.reduce(xt.sum(a.dev, [1, 3]))
    assert r.shape == (3, 4, 5)
    exp = np.full((3, 4, 5), 12.0)
    exp[1, 1, 1] = 24
    assert np.array_equal(r, exp)
    assert o.reduce(xt.sum(a.dev))[()] == 732


def test_reducer_sum_small():
    """test_xreducer.cpp:390-397, 442-473."""
    m = host(np.array([[1.0, 2.0], [3.0, 4.0]]))
    assert np.array_equal(oracle_for(m).reduce(xt.sum(m.dev, [0])), [4.0, 6.0])
    u = host(np.ones((2, 4)))
    o = oracle_for(u)
    assert np.array_equal(o.reduce(xt.sum(u.dev, [0])), 2 * np.ones(4))
    assert np.array_equal(o.reduce(xt.sum(u.dev, [1])), 4 * np.ones(2))
    assert np.array_equal(o.reduce(xt.sum(u.dev, [-1])), 4 * np.ones(2))
    c = host(np.ones(1000, np.uint8))  # sum(uint8) is int: no overflow
    oc = oracle_for(c)
    assert oc.reduce_info(xt.sum(c.dev))[0] == A.INT32
    assert oc.reduce(xt.sum(c.dev))[()] == 1000


def test_reducer_mean():
    """test_xreducer.cpp:542-569 and test_xmath.cpp:857-870."""
    x = host(np.array([[-1.0, 0.0], [1.0, 0.0]]))
    o = oracle_for(x)
    assert o.reduce(xt.mean(x.dev))[()] == 0.0
    assert np.array_equal(o.reduce(xt.mean(x.dev, [0])), [0.0, 0.0])
    assert np.array_equal(o.reduce(xt.mean(x.dev, [1])), [-0.5, 0.5])
    c = host(np.array([1, 2], np.uint8))
    assert oracle_for(c).reduce(xt.mean(c.dev))[()] == 1.5
    v = host(np.array([[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]]))
    ov = oracle_for(v)
    assert np.array_equal(ov.reduce(xt.mean(v.dev, [0])), [1.5, 1.5, 1.5])
    assert np.array_equal(ov.reduce(xt.mean(v.dev, [1])), [1.0, 2.0])
    assert ov.reduce(xt.mean(v.dev))[()] == 9.0 / 6.0
    assert ov.reduce(xt.mean(v.dev, [0, 1]))[()] == 9.0 / 6.0


def test_reducer_keep_dims_and_initial():
    """test_xreducer.cpp:964-1018."""
    a = host(np.arange(5 * 5 * 5 * 5, dtype=np.float64).reshape(5, 5, 5, 5))
    o = oracle_for(a)
    r = o.reduce(xt.sum(a.dev, [0, 1], keep_dims=True))
    assert r.shape == (1, 1, 5, 5)
    assert np.array_equal(r[0, 0], a.host.sum(axis=(0, 1)))
    assert o.reduce(xt.sum(a.dev, keep_dims=True)).shape == (1, 1, 1, 1)
    assert o.reduce(xt.sum(a.dev)).shape == ()
    base = o.reduce(xt.sum(a.dev, [0, 2], keep_dims=True))
    assert np.array_equal(o.reduce(xt.sum(a.dev, [0, 2], keep_dims=True, initial=5)), base + 5)
    assert np.all(o.reduce(xt.prod(a.dev, [1, 2], keep_dims=True, initial=0)) == 0.0)
    assert np.all(o.reduce(xt.prod(a.dev, [1, 2], initial=0)) == 0.0)


def test_reducer_ones_first_and_empty_axes():
    """test_xreducer.cpp:1020-1040."""
    a = host(np.ones((1, 1, 2, 4), np.int32))
    o = oracle_for(a)
    assert np.array_equal(o.reduce(xt.sum(a.dev, [1], keep_dims=True)), a.host)
    assert np.array_equal(o.reduce(xt.sum(a.dev, [1, 2], keep_dims=True)), [[[[2, 2, 2, 2]]]])
    b = host(np.array([[1, 2, 3], [4, 5, 6]], np.int32))
    assert np.array_equal(oracle_for(b).reduce(xt.sum(b.dev, [])), b.host)


def test_reducer_zero_shape_and_empty_mean():
    """test_xreducer.cpp:1042-1079."""
    x = host(np.zeros((0, 1), np.int32))
    o = oracle_for(x)
    r0 = o.reduce(xt.sum(x.dev, [0], keep_dims=True))
    assert r0.shape == (1, 1) and r0[0, 0] == 0
    assert o.reduce(xt.sum(x.dev, [1], keep_dims=True)).shape == (0, 1)
    r2 = o.reduce(xt.sum(x.dev, keep_dims=True))
    assert r2.shape == (1, 1) and r2[0, 0] == 0
    a = host(np.ones((1, 2, 0, 1)))
    oa = oracle_for(a)
    assert math.isnan(oa.reduce(xt.mean(a.dev))[()])
    assert oa.reduce(xt.mean(a.dev, [1])).shape == (1, 0, 1)
    r = oa.reduce(xt.mean(a.dev, [2]))
    assert r.shape == (1, 2, 1) and np.all(np.isnan(r))


@pytest.mark.parametrize("axes", [[1, 0], [0, 2], [0, 0], [1, 1]])
def test_reducer_bad_axes(axes):
    """test_xreducer.cpp:213-220, 1081-1088: unsorted / duplicate / out of range axes throw."""
    a = host(np.ones((3, 2), np.int32))
    with pytest.raises(OracleError) as ei:
        oracle_for(a).reduce(xt.sum(a.dev, axes))
    assert ei.value.code == A.BAD_AXES


def test_norms_matrix():
    """test_xnorm.cpp:170-187."""
    a = host(np.array([[-1.0, 2.0], [-3.0, 4.0]]))
    o = oracle_for(a)
    assert o.reduce(xt.norm_l0(a.dev))[()] == 4
    assert o.reduce_info(xt.norm_l0(a.dev))[0] == A.UINT64
    assert o.reduce(xt.norm_l1(a.dev))[()] == 10.0
    assert o.reduce(xt.norm_sq(a.dev))[()] == 30.0
    assert o.reduce(xt.norm_linf(a.dev))[()] == 4.0
    assert o.reduce(xt.norm_l2(a.dev))[()] == math.sqrt(30.0)


def test_amin_amax():
    """test_xmath.cpp:147-168."""
    for vals, mn, mx in [([-10.0], -10.0, -10.0), ([-10.0, -20.0], -20.0, -10.0), ([-10.0, 20.0], -10.0, 20.0),
                         ([10.0, 20.0], 10.0, 20.0), ([10.0], 10.0, 10.0)]:
        a = host(np.array(vals))
        o = oracle_for(a)
        assert o.reduce(xt.amin(a.dev))[()] == mn
        assert o.reduce(xt.amax(a.dev))[()] == mx


def test_clip():
    """test_xmath.cpp:210-232: clip, and lo > hi gives hi."""
    a = host(kat("test_xmath.cpp", "xmath.clip", "a"))
    assert np.array_equal(oracle_for(a).assign(xt.clip(a.dev, 2.0, 4.0)), kat("test_xmath.cpp", "xmath.clip", "res"))
    b = host(kat("test_xmath.cpp", "xmath.clip_amin_greater_than_amax", "arr"))
    r = oracle_for(b).assign(xt.clip(b.dev, 8, 1))
    assert r.dtype == np.int32
    assert np.array_equal(r, kat("test_xmath.cpp", "xmath.clip_amin_greater_than_amax", "expected"))


def test_softmax_rows_sum_to_one():
    """test_xmath.cpp:985-1027 (keep_dims softmax pattern): exp(log-softmax) rows sum to 1."""
    x = host(np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]))
    o = oracle_for(x)
    mx = xt.amax(x.dev, [1], keep_dims=True)
    sh = x.dev - mx
    ls = sh - xt.log(xt.sum(xt.exp(sh), [1], keep_dims=True))
    res = o.assign(ls)
    rs = Dual.of(res, device=False)
    sums = oracle_for(rs).reduce(xt.sum(xt.exp(rs.dev), [1]))
    np.testing.assert_allclose(sums, [1.0, 1.0], rtol=1e-12)


# ---------------------------------------------------------------- result types
# test_xmath_result_type.cpp:115-253 (one CHECK_RESULT_TYPE per row)
RESULT_TYPES = [
    ("uint8", lambda a: a + a, "int32"), ("uint8", lambda a: 2 * a, "int32"), ("uint8", lambda a: 2.0 * a, "float64"),
    ("uint8", lambda a: xt.sqrt(a), "float64"), ("uint8", lambda a: xt.abs(a), "uint8"),
    ("uint8", lambda a: xt.sum(a), "int32"), ("uint8", lambda a: xt.mean(a), "float64"),
    ("int16", lambda a: a + a, "int32"), ("int16", lambda a: 2 * a, "int32"), ("int16", lambda a: 2.0 * a, "float64"),
    ("int16", lambda a: xt.sqrt(a), "float64"), ("int16", lambda a: xt.abs(a), "int32"),
    ("int16", lambda a: xt.sum(a), "int32"), ("int16", lambda a: xt.mean(a), "float64"),
    ("uint16", lambda a: a + a, "int32"), ("uint16", lambda a: xt.scalar(2, "uint32") * a, "uint32"),
    ("uint16", lambda a: xt.abs(a), "uint16"), ("uint16", lambda a: xt.sum(a), "int32"),
    ("int32", lambda a: a + a, "int32"), ("int32", lambda a: 2.0 * a, "float64"), ("int32", lambda a: xt.sqrt(a), "float64"),
    ("int32", lambda a: xt.abs(a), "int32"), ("int32", lambda a: xt.sum(a), "int32"), ("int32", lambda a: xt.mean(a), "float64"),
    ("uint32", lambda a: a + a, "uint32"), ("uint32", lambda a: xt.scalar(2, "uint32") * a, "uint32"),
    ("uint32", lambda a: xt.abs(a), "uint32"), ("uint32", lambda a: xt.sum(a), "uint32"),
    ("int64", lambda a: a + a, "int64"), ("int64", lambda a: 2 * a, "int64"), ("int64", lambda a: xt.sqrt(a), "float64"),
    ("int64", lambda a: xt.abs(a), "int64"), ("int64", lambda a: xt.sum(a), "int64"),
    ("uint64", lambda a: a + a, "uint64"), ("uint64", lambda a: xt.abs(a), "uint64"), ("uint64", lambda a: xt.sum(a), "uint64"),
    ("float32", lambda a: a + a, "float32"), ("float32", lambda a: xt.scalar(2.0, "float32") * a, "float32"),
    ("float32", lambda a: 2.0 * a, "float64"), ("float32", lambda a: xt.sqrt(a), "float32"),
    ("float32", lambda a: xt.abs(a), "float32"), ("float32", lambda a: xt.sum(a), "float32"),
    ("float32", lambda a: xt.mean(a), "float64"),
    ("float64", lambda a: a + a, "float64"), ("float64", lambda a: xt.sqrt(a), "float64"),
    ("float64", lambda a: xt.sum(a), "float64"), ("float64", lambda a: xt.mean(a), "float64"),
    # templated accumulator: sum<T>(a) has type decltype(T() + value_type()) (core/xmath.hpp:1790-1803)
    ("int32", lambda a: xt.sum(a, dtype="int64"), "int64"), ("float32", lambda a: xt.sum(a, dtype="float64"), "float64"),
    ("int8", lambda a: xt.sum(a, dtype="int8"), "int32"), ("int32", lambda a: xt.prod(a, dtype="float32"), "float32"),
    # minmax(auchar) is std::array<unsigned char, 2> (test_xmath_result_type.cpp:129)
    ("uint8", lambda a: xt.minmax(a), "uint8"), ("float32", lambda a: xt.minmax(a), "float32"),
    # mean<T>, variance<T>, stddev<T> (CHECK_TEMPLATED_RESULT_TYPE_FOR_ALL,
    # test_xmath_result_type.cpp:72-114): the division is in T's type
    # (mean_division's value_type = T), sqrt of an integer is double
    ("uint8", lambda a: xt.mean(a, dtype="uint8"), "int32"), ("int32", lambda a: xt.mean(a, dtype="int64"), "int64"),
    ("float32", lambda a: xt.mean(a, dtype="float32"), "float32"), ("uint8", lambda a: xt.mean(a, dtype="float32"), "float32"),
    ("uint8", lambda a: xt.variance(a), "float64"), ("float32", lambda a: xt.variance(a), "float64"),
    ("uint8", lambda a: xt.variance(a, dtype="uint8"), "int32"), ("uint8", lambda a: xt.stddev(a, dtype="uint8"), "float64"),
    ("float32", lambda a: xt.stddev(a, dtype="uint8"), "float32"), ("float32", lambda a: xt.variance(a, dtype="float32"), "float32"),
    ("float64", lambda a: xt.stddev(a), "float64"),
]


XR = "test_extended_xmath_reducers.cpp"


def test_variance_stddev_kats():
    """test_extended_xmath_reducers.cpp:1824-2833 (numpy goldens printed with
    16 significant digits): variance / stddev over all axes, over {0, 2}, and
    with ddof = 1."""
    T = "xtest_extended_xmath.var_std"
    a = host(kat(XR, T, "py_a"))
    o = oracle_for(a)
    tol = dict(rtol=1e-14, atol=0)
    np.testing.assert_allclose(o.reduce(xt.variance(a.dev))[()], kat(XR, T, "py_vr_all"), **tol)
    np.testing.assert_allclose(o.reduce(xt.stddev(a.dev))[()], kat(XR, T, "py_st_all"), **tol)
    np.testing.assert_allclose(o.reduce(xt.variance(a.dev, [0, 1, 2, 3], ddof=1))[()], kat(XR, T, "py_vr_all_ddof"),
                               **tol)
    np.testing.assert_allclose(o.reduce(xt.variance(a.dev, [0, 2])), kat(XR, T, "py_vr"), **tol)
    np.testing.assert_allclose(o.reduce(xt.variance(a.dev, [0, 2], ddof=1)), kat(XR, T, "py_vr_ddof"), **tol)
    np.testing.assert_allclose(o.reduce(xt.stddev(a.dev, [0, 2])), kat(XR, T, "py_st"), **tol)
    # the divisor is (size - ddof) / outputs in size_t arithmetic (detail::mean,
    # core/xmath.hpp:1836-1850): 840 elements, 35 outputs, ddof 2 -> 838 / 35 = 23
    # (numpy's n - ddof would be 22)
    v0 = o.reduce(xt.variance(a.dev, [0, 2]))
    np.testing.assert_allclose(o.reduce(xt.variance(a.dev, [0, 2], ddof=2)), v0 * 24 / 23, rtol=1e-14)


def test_minmax():
    """core/xmath.hpp:2195-2230 restated: std::min / std::max folds from
    {max(), lowest()} of the value type; a NaN never replaces the running
    value (std::min(r, v) = v < r ? v : r)."""
    f = host(np.array([[3.0, np.nan], [-1.0, 2.0]]))
    assert np.array_equal(oracle_for(f).reduce(xt.minmax(f.dev)), [-1.0, 3.0])
    i = host(np.array([7, -4, 9, 0], np.int32))
    assert np.array_equal(oracle_for(i).reduce(xt.minmax(i.dev)), [-4, 9])
    e = host(np.zeros((0, 3), np.uint8))
    assert np.array_equal(oracle_for(e).reduce(xt.minmax(e.dev)), [255, 0])
    n = host(np.array([np.nan, np.nan], np.float32))
    r = oracle_for(n).reduce(xt.minmax(n.dev))
    assert r[0] == np.finfo(np.float32).max and r[1] == -np.finfo(np.float32).max


def test_mean_with_accumulator_type():
    """core/xmath.hpp:1827-1873: mean<T> divides by static_cast<T>(n), so an
    integral T divides in integers (truncating) and float T in float."""
    a = host(np.array([[1, 2], [3, 4]], np.int32))
    o = oracle_for(a)
    assert o.reduce(xt.mean(a.dev, dtype="int64"))[()] == 2          # 10 / 4
    assert np.array_equal(o.reduce(xt.mean(a.dev, [1], dtype="int32")), [1, 3])  # 3 / 2, 7 / 2
    assert np.array_equal(o.reduce(xt.mean(a.dev, [1])), [1.5, 3.5])
    f = host(np.array([1.0, 2.0, 4.0], np.float32))
    r = oracle_for(f).reduce(xt.mean(f.dev, dtype="float32"))
    assert r.dtype == np.float32 and r[()] == np.float32(7.0) / np.float32(3.0)


def test_gather_index_view_kats():
    """test_xindex_view.cpp:24-38, 167-181: index_view(e, {{1,1},{1,2},{2,2}})
    reads e(1,1), e(1,2), e(2,2); filter(a, cond) in row- and column-major
    order; positions outside the source read as 0 (the oracle's defined value
    where xtensor's is undefined)."""
    h = np.random.default_rng(0).random((3, 3))
    e = host(h)
    ix = host(np.array([[1, 1], [1, 2], [2, 2], [3, 0]], np.int64))
    v = xt.index_view(e.dev, ix.dev)
    got = oracle_for(e, ix).assign(v)
    assert np.array_equal(got, [h[1, 1], h[1, 2], h[2, 2], 0.0])
    a_h = np.array([[[1, 3], [2, 4]], [[5, 7], [6, 8]]], np.int32)
    cond = np.array([[[1, 1], [0, 0]], [[1, 1], [0, 0]]], bool)
    a = host(a_h)
    for order, exp in (("C", [1, 3, 5, 7]), ("F", [1, 5, 3, 7])):
        idx = np.argwhere(cond) if order == "C" else np.argwhere(cond.T)[:, ::-1].copy()
        ixd = host(idx.astype(np.int64))
        assert np.array_equal(oracle_for(a, ixd).assign(xt.index_view(a.dev, ixd.dev)), exp)
    # a gather through a transposed (strided) source and inside arithmetic
    t = host(h.T.copy())
    tv = Dual(xt.transpose(e.dev), h.T)
    ixt = host(np.array([[0, 2], [2, 1]], np.int64))
    r = oracle_for(tv, ixt).assign(xt.index_view(tv.dev, ixt.dev) * 2.0)
    assert np.array_equal(r, [h.T[0, 2] * 2, h.T[2, 1] * 2])


@pytest.mark.parametrize("case", range(len(RESULT_TYPES)))
def test_result_types(case):
    dt, fn, expect = RESULT_TYPES[case]
    a = host(np.zeros((3, 2), dt))
    o = oracle_for(a)
    e = fn(a.dev)
    got = o.reduce_info(e)[0] if isinstance(e, xt.Reducer) else o.info(e)[0]
    assert xt.dtype_name(got) == expect


# ---------------------------------------------------------------- closed forms
# For IEEE + - * / and sqrt (correctly rounded) and integer arithmetic, numpy
# computes the same values; a dropped term, swapped operand or wrong
# promotion fails these.
def test_elementwise_closed_forms():
    rng = np.random.default_rng(11)
    a = rng.uniform(-4, 4, (37, 19))
    b = rng.uniform(0.5, 3, (19,))
    A_, B_ = host(a), host(b)
    o = oracle_for(A_, B_)
    assert np.array_equal(o.assign(A_.dev * B_.dev + A_.dev / B_.dev - B_.dev), a * b + a / b - b)
    assert np.array_equal(o.assign(xt.sqrt(B_.dev) * A_.dev), np.sqrt(b) * a)
    i = rng.integers(-2**20, 2**20, (64,), dtype=np.int32)
    I = host(i)
    got = oracle_for(I).assign((I.dev * 3 + (I.dev >> 2)) & 0xFFFF)
    assert np.array_equal(got, (i * 3 + (i >> 2)) & 0xFFFF)
    f = rng.normal(size=(5, 6)).astype(np.float32)
    F = host(f)
    F0 = F.view(xt.range_(0, 1))
    # float - float stays float; the double divide promotes, then narrows on store
    got = oracle_for(F, F0).assign(xt.clip((F.dev - F0.dev) / xt.sqrt(F.dev * F.dev + 1e-5), -3, 3), dtype="float32")
    ref = np.clip((f - f[0:1]).astype(np.float64) / np.sqrt((f * f).astype(np.float64) + 1e-5), -3, 3).astype(np.float32)
    assert np.array_equal(got, ref)


def test_sequential_reduction_order():
    """The oracle folds left to right in the value type: compare to an explicit loop."""
    rng = np.random.default_rng(5)
    x = rng.normal(size=(7, 1000)).astype(np.float32)
    X = host(x)
    got = oracle_for(X).reduce(xt.sum(X.dev, [1]))
    for r in range(7):
        acc = np.float32(0)
        for v in x[r]:
            acc = np.float32(acc + v)
        assert got[r] == acc
    gotc = oracle_for(X).reduce(xt.sum(X.dev, [0, 1]))
    acc = np.float32(0)
    for r in range(7):  # aggregate_impl: inner axis folded first, rows merged in order
        racc = np.float32(0)
        for v in x[r]:
            racc = np.float32(racc + v)
        acc = np.float32(acc + racc)
    assert gotc[()] == acc


def test_high_precision_reference_is_exact_on_integers():
    x = np.arange(1, 10001, dtype=np.float64)
    X = host(x)
    assert oracle_for(X).reduce(xt.sum(X.dev), high_precision=True)[()] == 10000 * 10001 / 2

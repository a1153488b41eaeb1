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
from xtensor_amd import core

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


# ---------------------------------------------------------------- prod
def test_prod_kat_no_overflow():
    """test_xreducer.cpp:507-512: prod<long long>(2 * ones<uint8_t>({34})) == 1 << 34."""
    c = host(2 * np.ones(34, np.uint8))
    o = oracle_for(c)
    r = xt.prod(c.dev, None, dtype="int64")
    assert xt.dtype_name(o.reduce_info(r)[0]) == "int64"
    assert int(o.reduce(r)[()]) == 1 << 34


def test_prod_closed_forms():
    """prod over small integers wraps like C++ unsigned arithmetic in the
    value type (numpy's integer product, exact); float prod is the
    sequential left fold in the value type."""
    rng = np.random.default_rng(21)
    i = rng.integers(-5, 6, (6, 40), dtype=np.int32)
    I = host(i)
    o = oracle_for(I)
    assert np.array_equal(o.reduce(xt.prod(I.dev, [1])), np.prod(i, axis=1, dtype=np.int32))
    assert np.array_equal(o.reduce(xt.prod(I.dev, [0], dtype="int64")), np.prod(i.astype(np.int64), axis=0))
    f = rng.uniform(0.5, 1.5, (5, 33)).astype(np.float32)
    F = host(f)
    got = oracle_for(F).reduce(xt.prod(F.dev, [1]))
    for r in range(5):
        acc = np.float32(1)
        for v in f[r]:
            acc = np.float32(acc * v)
        assert got[r] == acc
    # initial is merged once, after the fold: prod(x, initial=3) = 3 * prod(x)
    small = host(np.array([2, 3, 4], np.int64))
    assert int(oracle_for(small).reduce(xt.prod(small.dev, None, initial=3))[()]) == 72


# ---------------------------------------------------------------- every opcode
# C++ value types of the operands (xtensor containers hold these), the
# integral promotions and the usual arithmetic conversions ([conv.prom],
# [expr.arith.conv]) that decide each operator's result type, restated
# from the C++ standard for the LP64 types of this platform.
_INT_RANK = {np.int8: 1, np.uint8: 1, np.int16: 2, np.uint16: 2, np.int32: 3, np.uint32: 3,
             np.int64: 4, np.uint64: 4}


def _promote(t):
    t = np.dtype(t).type
    if t in (np.bool_, np.int8, np.uint8, np.int16, np.uint16):
        return np.int32
    return t


def _usual(a, b):
    a, b = np.dtype(a).type, np.dtype(b).type
    if np.float64 in (a, b):
        return np.float64
    if np.float32 in (a, b):
        return np.float32
    a, b = _promote(a), _promote(b)
    if a == b:
        return a
    sa, sb = np.issubdtype(a, np.signedinteger), np.issubdtype(b, np.signedinteger)
    if sa == sb:
        return a if _INT_RANK[a] >= _INT_RANK[b] else b
    u, s = (b, a) if sa else (a, b)
    return u if _INT_RANK[u] >= _INT_RANK[s] else s


def _operands(dt, rng, n=64, nonzero=False, shift=False):
    t = np.dtype(dt).type
    if t == np.bool_:
        v = rng.integers(0, 2, n).astype(np.bool_)
    elif np.issubdtype(t, np.floating):
        v = rng.uniform(-20, 20, n).astype(t)
        if nonzero:
            v[v == 0] = 1
    else:
        lo = 0 if shift else (max(np.iinfo(t).min, -100) if np.issubdtype(t, np.signedinteger) else 0)
        hi = 16 if shift else min(np.iinfo(t).max, 100)
        v = rng.integers(lo, hi, n).astype(t)
        if nonzero:
            v[v == 0] = 7
    return v


def _cxx_int_div(a, b):
    q = abs(int(a)) // abs(int(b))
    return -q if (a < 0) != (b < 0) else q


_PAIRS = [(np.int8, np.uint8), (np.int32, np.uint32), (np.int16, np.int64), (np.uint64, np.int64),
          (np.uint8, np.float32), (np.int32, np.float64), (np.float32, np.float64), (np.bool_, np.int16),
          (np.uint16, np.int32)]


def _check_binary(op, expect_fn, pairs, rng, result=None, **kw):
    for ta, tb in pairs:
        a = _operands(ta, rng, **kw.get("a", {}))
        b = _operands(tb, rng, **kw.get("b", {}))
        A_, B_ = host(a), host(b)
        o = oracle_for(A_, B_)
        e = core.Func(A.OP[op], A_.dev, B_.dev)
        rt = result(ta, tb) if result else _usual(ta, tb)
        assert xt.dtype_name(o.info(e)[0]) == np.dtype(rt).name, (op, ta, tb)
        got = o.assign(e)
        want = np.array([expect_fn(x, y, rt) for x, y in zip(a.astype(rt) if rt != np.bool_ else a,
                                                              b.astype(rt) if rt != np.bool_ else b)], dtype=rt)
        if result is not None and rt == np.bool_:
            want = np.array([expect_fn(x, y, rt) for x, y in zip(a, b)], dtype=np.bool_)
        assert np.array_equal(got, want), (op, ta, tb)


def _wrap_int(v, rt):
    return np.array(v).astype(np.int64).astype(rt) if v >= 0 or np.issubdtype(rt, np.signedinteger) else \
        np.array(int(v) % (1 << (8 * np.dtype(rt).itemsize)), dtype=np.uint64).astype(rt)


def test_arithmetic_opcodes_closed_form():
    rng = np.random.default_rng(31)
    for op, f in (("PLUS", np.add), ("MINUS", np.subtract), ("MULTIPLIES", np.multiply)):
        _check_binary(op, lambda x, y, rt, f=f: f(np.array(x, rt), np.array(y, rt)), _PAIRS, rng)

    def div(x, y, rt):
        if np.issubdtype(rt, np.floating):
            return np.array(x, rt) / np.array(y, rt)
        return _wrap_int(_cxx_int_div(x, y), rt)

    def mod(x, y, rt):
        return _wrap_int(int(x) - _cxx_int_div(x, y) * int(y), rt)

    _check_binary("DIVIDES", div, _PAIRS, rng, b={"nonzero": True})
    int_pairs = [p for p in _PAIRS if not any(np.issubdtype(t, np.floating) for t in p)]
    _check_binary("MODULUS", mod, [p for p in int_pairs if np.bool_ not in p], rng, b={"nonzero": True})


def test_compare_and_logical_opcodes_closed_form():
    rng = np.random.default_rng(32)
    boolr = lambda ta, tb: np.bool_  # noqa: E731
    cmps = {"LESS": np.less, "LESS_EQUAL": np.less_equal, "GREATER": np.greater,
            "GREATER_EQUAL": np.greater_equal, "EQUAL_TO": np.equal, "NOT_EQUAL_TO": np.not_equal}
    for op, f in cmps.items():
        # compared after the usual arithmetic conversions: a negative int is
        # converted to unsigned against an unsigned operand
        for ta, tb in _PAIRS:
            a, b = _operands(ta, rng), _operands(tb, rng)
            b[:8] = a[:8].astype(tb)  # some equal pairs
            A_, B_ = host(a), host(b)
            o = oracle_for(A_, B_)
            e = core.Func(A.OP[op], A_.dev, B_.dev)
            assert xt.dtype_name(o.info(e)[0]) == "bool"
            ct = _usual(ta, tb)
            assert np.array_equal(o.assign(e), f(a.astype(ct), b.astype(ct))), (op, ta, tb)
    for op, f in (("LOGICAL_AND", np.logical_and), ("LOGICAL_OR", np.logical_or)):
        _check_binary(op, lambda x, y, rt, f=f: f(x != 0, y != 0), _PAIRS, rng, result=boolr)
    # -1 < 1u is false in C++ (int converts to unsigned)
    m1, one = host(np.array([-1], np.int32)), host(np.array([1], np.uint32))
    assert not oracle_for(m1, one).assign(m1.dev < one.dev)[0]


def test_bitwise_and_shift_opcodes_closed_form():
    rng = np.random.default_rng(33)
    int_pairs = [(np.int8, np.uint8), (np.int32, np.uint32), (np.int16, np.int64), (np.uint64, np.int64),
                 (np.bool_, np.int16), (np.uint16, np.int32)]
    for op, f in (("BITWISE_AND", np.bitwise_and), ("BITWISE_OR", np.bitwise_or), ("BITWISE_XOR", np.bitwise_xor)):
        _check_binary(op, lambda x, y, rt, f=f: f(np.array(x, rt), np.array(y, rt)), int_pairs, rng)
    # shifts take the promoted type of the left operand only ([expr.shift])
    lr = lambda ta, tb: _promote(ta)  # noqa: E731
    _check_binary("LEFT_SHIFT", lambda x, y, rt: np.left_shift(np.array(x, rt), np.array(y, rt)),
                  int_pairs, rng, result=lr, a={"shift": True}, b={"shift": True})
    _check_binary("RIGHT_SHIFT", lambda x, y, rt: np.right_shift(np.array(x, rt), np.array(y, rt)),
                  int_pairs, rng, result=lr, b={"shift": True})


def test_unary_and_select_opcodes_closed_form():
    rng = np.random.default_rng(34)
    for t in (np.int8, np.uint8, np.int32, np.uint32, np.int64, np.float32, np.float64, np.bool_):
        a = _operands(t, rng)
        A_ = host(a)
        o = oracle_for(A_)
        pt = _promote(t)
        assert np.array_equal(o.assign(-A_.dev), -(a.astype(pt))) if t != np.bool_ else True
        assert np.array_equal(o.assign(+A_.dev), a.astype(pt))
        assert xt.dtype_name(o.info(+A_.dev)[0]) == np.dtype(pt).name
        assert np.array_equal(o.assign(xt.logical_not(A_.dev)), a == 0)
        if not np.issubdtype(t, np.floating):
            assert np.array_equal(o.assign(~A_.dev), ~(a.astype(pt)))
    # where(c, a, b) -> common type of a and b; maximum / minimum select
    for ta, tb in _PAIRS:
        a, b = _operands(ta, rng), _operands(tb, rng)
        c = rng.integers(0, 2, a.shape).astype(np.bool_)
        A_, B_, C_ = host(a), host(b), host(c)
        o = oracle_for(A_, B_, C_)
        ct = np.result_type(np.array(0, ta), np.array(0, tb)) if False else _common(ta, tb)
        assert np.array_equal(o.assign(xt.where(C_.dev, A_.dev, B_.dev)),
                              np.where(c, a.astype(ct), b.astype(ct))), (ta, tb)
        assert np.array_equal(o.assign(xt.maximum(A_.dev, B_.dev)),
                              np.where(a.astype(ct) > b.astype(ct), a.astype(ct), b.astype(ct))), (ta, tb)
        assert np.array_equal(o.assign(xt.minimum(A_.dev, B_.dev)),
                              np.where(a.astype(ct) < b.astype(ct), a.astype(ct), b.astype(ct))), (ta, tb)
    # static_cast: float -> int truncates toward zero; int64 -> int8 wraps; int -> float rounds
    f = np.array([-3.75, -2.5, -0.5, 0.5, 2.5, 3.75, 100.9], np.float64)
    F = host(f)
    assert np.array_equal(oracle_for(F).assign(xt.cast(F.dev, "int32")), np.trunc(f).astype(np.int32))
    w = np.array([127, 128, 255, 256, -129, 1 << 40], np.int64)
    W = host(w)
    assert np.array_equal(oracle_for(W).assign(xt.cast(W.dev, "int8")), w.astype(np.int8))
    big = np.array([(1 << 24) + 1, (1 << 53) + 1], np.int64)
    Bg = host(big)
    assert np.array_equal(oracle_for(Bg).assign(xt.cast(Bg.dev, "float32")), big.astype(np.float32))


def _common(ta, tb):
    """std::common_type_t of two arithmetic types: the usual arithmetic
    conversions, except that two operands of one type keep it (bool, int8)."""
    ta, tb = np.dtype(ta).type, np.dtype(tb).type
    return ta if ta == tb else _usual(ta, tb)


def test_rounding_and_math_opcodes_closed_form():
    q = np.array([-2.5, -1.5, -0.75, -0.5, -0.25, 0.0, 0.25, 0.5, 0.75, 1.5, 2.5, 3.5], np.float64)
    Q = host(q)
    o = oracle_for(Q)
    assert np.array_equal(o.assign(xt.floor(Q.dev)), np.floor(q))
    assert np.array_equal(o.assign(xt.ceil(Q.dev)), np.ceil(q))
    assert np.array_equal(o.assign(xt.trunc(Q.dev)), np.trunc(q))
    # std::round: halfway cases away from zero (numpy's round is to even)
    assert np.array_equal(o.assign(xt.round(Q.dev)), np.sign(q) * np.floor(np.abs(q) + 0.5))
    assert np.array_equal(o.assign(xt.fabs(Q.dev)), np.abs(q))
    iu = host(np.array([0, 3, 250], np.uint8))
    assert np.array_equal(oracle_for(iu).assign(xt.abs(iu.dev)), np.array([0, 3, 250]))  # identity on unsigned
    ii = host(np.array([-7, 0, 9], np.int32))
    assert np.array_equal(oracle_for(ii).assign(xt.abs(ii.dev)), np.array([7, 0, 9], np.int32))
    rng = np.random.default_rng(35)
    x = rng.uniform(0.05, 0.95, 200)
    X = host(x)
    o = oracle_for(X)
    import math
    # Python's math module calls the same platform libm the oracle's std::
    # functions do, so every one-to-one mapping is pinned bit for bit;
    # exp2 and cbrt (absent from math before Python 3.11) against numpy
    # within 4 ulp (glibc's cbrt is up to 3 ulp off).  A different function,
    # or an argument taken from the wrong operand, is off by far more.
    from helpers import ulp_diff
    libm = {"exp": math.exp, "expm1": math.expm1, "log": math.log, "log2": math.log2, "log10": math.log10,
            "log1p": math.log1p, "sqrt": math.sqrt, "sin": math.sin, "cos": math.cos, "tan": math.tan,
            "tanh": math.tanh, "sinh": math.sinh, "cosh": math.cosh, "asin": math.asin, "acos": math.acos,
            "atan": math.atan, "erf": math.erf}
    for name, f in libm.items():
        got = o.assign(getattr(xt, name)(X.dev))
        assert np.array_equal(got, np.array([f(v) for v in x])), name
    for name, f in {"exp2": np.exp2, "cbrt": np.cbrt}.items():
        assert ulp_diff(o.assign(getattr(xt, name)(X.dev)), f(x)).max() <= 4, name
    y = rng.uniform(0.5, 3.0, 200)
    Y = host(y)
    o = oracle_for(X, Y)
    for name, f in {"pow": math.pow, "atan2": math.atan2, "fmod": math.fmod}.items():
        got = o.assign(getattr(xt, name)(X.dev, Y.dev))
        assert np.array_equal(got, np.array([f(a, b) for a, b in zip(x, y)])), name
    # math.hypot is CPython's own algorithm, not libm's: within 1 ulp
    got = o.assign(xt.hypot(X.dev, Y.dev))
    assert ulp_diff(got, np.array([math.hypot(a, b) for a, b in zip(x, y)])).max() <= 1
    # std::exp(int) computes in double
    n = host(np.array([0, 1, 2], np.int32))
    assert xt.dtype_name(oracle_for(n).info(xt.exp(n.dev))[0]) == "float64"


# ---------------------------------------------------------------- immediate reducers
# test_xreducer.cpp:704-771: sum(a, axes, evaluation_strategy::immediate) equals
# the lazy sum on arange data (exact in double) and on column-major shorts.
IMMEDIATE_AXES_3 = [None, [1], [0, 2], [1, 2]]
IMMEDIATE_AXES_5 = [None, [1], [0, 2], [1, 2], [1, 3, 4], [0, 1, 4], [0, 1, 3], [0, 2, 3], [1, 2, 3]]


@pytest.mark.parametrize("shape,axes_list", [((3, 3, 3), IMMEDIATE_AXES_3), ((4, 3, 6, 2, 7), IMMEDIATE_AXES_5)])
def test_immediate_equals_lazy_kat(shape, axes_list):
    a = host(np.arange(np.prod(shape), dtype=np.float64).reshape(shape))
    o = oracle_for(a)
    for axes in axes_list:
        lazy = o.reduce(xt.sum(a.dev, axes))
        imm = o.reduce(xt.sum(a.dev, axes, evaluation_strategy="immediate"))
        assert np.array_equal(lazy, imm), axes
        assert np.array_equal(imm, np.sum(a.host, axis=None if axes is None else tuple(axes)))


@pytest.mark.parametrize("shape,axes_list", [((1, 5, 3), [[0, 2]]), ((1, 5, 1, 2, 3), [[0, 1, 2], [2, 3], [1, 3]])])
def test_immediate_equals_lazy_column_major_shorts(shape, axes_list):
    rng = np.random.default_rng(41)
    a = host(np.asfortranarray(rng.integers(-1000, 1000, shape).astype(np.int16)))
    o = oracle_for(a)
    for axes in axes_list:
        lazy = o.reduce(xt.sum(a.dev, axes))
        imm = o.reduce(xt.sum(a.dev, axes, evaluation_strategy="immediate"))
        assert np.array_equal(lazy, imm), axes
        assert np.array_equal(imm, np.sum(a.host.astype(np.int32), axis=tuple(axes)))


def test_immediate_fold_order():
    """reduce_immediate's order, checked against explicit loops in float32
    (where order shows in the rounding):
    * every axis: one std::accumulate over the storage (reducers/xreducer.hpp:354-360)
      -- the lazy form folds each row and merges the rows instead;
    * the leading axis strided (sum over axis 0): each output folded down the axis;
    * the leading axis contiguous with an outer reduced axis ({0, 2} of a 3-d
      array): each run accumulated from 0, then merged into its output;
    * xt::initial merged behind the value; with no axes it is not applied."""
    rng = np.random.default_rng(42)
    x = rng.normal(size=(9, 300)).astype(np.float32) * np.float32(1e3)
    X = host(x)
    o = oracle_for(X)
    acc = np.float32(0)
    for v in x.reshape(-1):
        acc = np.float32(acc + v)
    imm = o.reduce(xt.sum(X.dev, None, evaluation_strategy="immediate"))[()]
    assert imm == acc
    lazy = o.reduce(xt.sum(X.dev))[()]
    racc = np.float32(0)
    for r in range(9):
        row = np.float32(0)
        for v in x[r]:
            row = np.float32(row + v)
        racc = np.float32(racc + row)
    assert lazy == racc and lazy != imm  # the two orders differ on this input
    col = o.reduce(xt.sum(X.dev, [0], evaluation_strategy="immediate"))
    want = np.zeros(300, np.float32)
    for r in range(9):
        want = (want + x[r]).astype(np.float32)
    assert np.array_equal(col, want)
    y = rng.normal(size=(4, 5, 60)).astype(np.float32) * np.float32(1e3)
    Y = host(y)
    got = oracle_for(Y).reduce(xt.sum(Y.dev, [0, 2], evaluation_strategy="immediate"))
    for j in range(5):
        out = None
        for i in range(4):
            run = np.float32(0)
            for v in y[i, j]:
                run = np.float32(run + v)
            out = run if out is None else np.float32(out + run)
        assert got[j] == out
    got = oracle_for(Y).reduce(xt.sum(Y.dev, [2], initial=np.float32(7), evaluation_strategy="immediate"))
    lz = oracle_for(Y).reduce(xt.sum(Y.dev, [2], evaluation_strategy="immediate"))
    assert np.array_equal(got, (lz + np.float32(7)).astype(np.float32))
    n = host(np.array([1, 2, 3], np.int32))
    assert np.array_equal(oracle_for(n).reduce(xt.sum(n.dev, [], initial=5, evaluation_strategy="immediate")),
                          np.array([1, 2, 3]))

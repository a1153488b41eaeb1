"""GPU parity of cumsum / cumprod (xg_scan) against the oracle.

Integer scans bit-exact; float sums within a relative error of the
long-double reference (compensated float32 accumulation: rel <= 2e-7 on
each output; float64: 1e-14).
"""

import json
import os

import numpy as np
import pytest

import xtensor_amd as xt

from helpers import Dual, oracle_for

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_arrays.json")))


def kat(test, name, dtype):
    r = KATS[f"test_xaccumulator.cpp:{test}:{name}"]
    return np.array(r["values"], dtype=dtype).reshape(r["shape"])


def sync():
    import torch
    torch.cuda.synchronize()


def test_kat_one_d_short_to_int():
    """test_xaccumulator.cpp:23-34."""
    a = xt.from_numpy(np.array([1, 2, 3, 4], np.int16))
    for axis in (None, 0):
        r = xt.cumsum(a, axis).to_numpy()
        assert r.dtype == np.int32 and np.array_equal(r, [1, 3, 6, 10])


def test_kat_four_d_every_axis():
    """test_xaccumulator.cpp:47-152."""
    T = "xaccumulator.four_d"
    a = xt.from_numpy(kat(T, "arg_0", np.float64))
    assert np.array_equal(xt.cumsum(a).to_numpy(), kat(T, "expected", np.float64))
    for ax in range(4):
        assert np.array_equal(xt.cumsum(a, ax).to_numpy(), kat(T, f"expected_{ax}", np.float64)), ax
    assert np.array_equal(xt.cumsum(a, -1).to_numpy(), kat(T, "expected_3", np.float64))


def test_kat_cumprod_and_misc():
    """test_xaccumulator.cpp:35-45, 154-179, 196-210."""
    T = "xaccumulator.cumprod"
    a = xt.from_numpy(kat(T, "arg_0", np.int64))
    assert np.array_equal(xt.cumprod(a).to_numpy(), kat(T, "expected", np.int64))
    assert np.array_equal(xt.cumprod(a, 0).to_numpy(), kat(T, "expected_0", np.int64))
    assert np.array_equal(xt.cumprod(a, 1).to_numpy(), kat(T, "expected_1", np.int64))
    arr = xt.from_numpy(np.array([[5.0, 6.0, 7.0]]))
    assert np.array_equal(xt.cumsum(arr, 0).to_numpy(), [[5.0, 6.0, 7.0]])
    assert np.array_equal(xt.cumsum(xt.transpose(arr), 1).to_numpy(), [[5.0], [6.0], [7.0]])
    x = xt.from_numpy(np.array([[1.0, 2, 3], [4, 5, 6]]))
    assert np.array_equal(xt.cumsum(x, 0).to_numpy(), [[1, 2, 3], [5, 7, 9]])
    e = xt.from_numpy(np.ones((3, 4, 0)))
    assert xt.cumsum(e).shape == (0,)
    assert xt.cumsum(e, 2).shape == (3, 4, 0)
    assert xt.cumsum(e, 1).shape == (3, 4, 0)


@pytest.mark.parametrize("rows,cols", [(8192, 4096), (64, 100000), (5, 3)])
def test_f32_cumsum_rows(rows, cols):
    rng = np.random.default_rng(2)
    x = rng.normal(1, 1, (rows, cols)).astype(np.float32)
    X = Dual.of(x)
    got = xt.cumsum(X.dev, 1)
    sync()
    g = got.to_numpy()
    ref = oracle_for(X).scan(X.dev, 1, high_precision=True) if rows * cols <= 1 << 22 else None
    if ref is None:  # large case: compare against a float64 numpy scan (exact enough as reference)
        ref = np.cumsum(x.astype(np.float64), axis=1)
    den = np.maximum(np.abs(ref), 1.0)
    err = np.max(np.abs(g - ref) / den)
    assert err <= 2e-7, err


@pytest.mark.parametrize("dtype", ["int32", "int64", "float32", "float64"])
def test_direct_rows_scan_lengths(dtype):
    """Contiguous rows scans around the one-wave / four-wave tile edges
    (1024 elements of 4-byte types, 512 of 8-byte types per wave tile, four
    wave tiles per block tile), sums and products: integers exact, floats
    against the oracle's long-double scan (rel 2e-7 for float32, 1e-14 for
    float64)."""
    rng = np.random.default_rng(11)
    lengths = [1, 3, 511, 512, 513, 1023, 1024, 1025, 2048, 4095, 4096, 4097, 9000]
    for n in lengths:
        rows = 3
        if dtype.startswith("int"):
            x = rng.integers(-50, 50, (rows, n)).astype(dtype)
        else:
            x = rng.normal(0.5, 1, (rows, n)).astype(dtype)
        X = Dual.of(x)
        o = oracle_for(X)
        got = xt.cumsum(X.dev, 1).to_numpy()
        if dtype.startswith("int"):
            assert np.array_equal(got, o.scan(X.dev, 1)), n
            p = rng.choice(np.array([-1, 1, 1, 1], dtype=dtype), (rows, n))
            P = Dual.of(p)
            gp = xt.cumprod(P.dev, 1).to_numpy()
            assert np.array_equal(gp, np.cumprod(p, axis=1)), n
        else:
            ref = o.scan(X.dev, 1, high_precision=True)
            err = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0))
            assert err <= (2e-7 if dtype == "float32" else 1e-14), (n, err)


def test_scan_strided_axis_and_ints():
    rng = np.random.default_rng(3)
    x = rng.integers(-1000, 1000, (300, 7, 33), dtype=np.int32)
    X = Dual.of(x)
    o = oracle_for(X)
    for ax in (0, 1, 2, None):
        got = xt.cumsum(X.dev, ax).to_numpy()
        assert np.array_equal(got, o.scan(X.dev, ax)), ax
    got = xt.cumsum(X.dev, 0, dtype="int64").to_numpy()
    assert got.dtype == np.int64 and np.array_equal(got, np.cumsum(x.astype(np.int64), axis=0))


def test_scan_of_expression_f64():
    rng = np.random.default_rng(4)
    a = rng.uniform(-1, 1, (40, 513))
    A_ = Dual.of(a)
    e = A_.dev * 2 + 1
    got = xt.cumsum(e, 1).to_numpy()
    ref = oracle_for(A_).scan(e, 1, high_precision=True)
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1)) <= 1e-14


@pytest.mark.parametrize("dtype", ["int8", "uint8", "int16", "uint16", "uint32", "uint64", "bfloat16"])
def test_scans_and_reductions_across_dtypes(dtype):
    """cumsum / cumprod / sum / prod / amax / mean over each axis of a small
    array of each less common type, against the oracle: the result type
    follows xtensor's promotion (the oracle decides it), integers are exact
    (wrapping as C++ does), bf16 data is compared in float32 accumulation
    against the oracle's long-double result within 2e-7 relative."""
    rng = np.random.default_rng(21)
    shape = (37, 70)
    if dtype == "bfloat16":
        X = Dual.of(rng.normal(0.5, 1, shape).astype(np.float32), "bfloat16")
    elif dtype.startswith("u"):
        X = Dual.of(rng.integers(0, 9, shape).astype(dtype))
    else:
        X = Dual.of(rng.integers(-9, 9, shape).astype(dtype))
    o = oracle_for(X)
    for ax in (0, 1, None):
        for kind in ("cumsum", "cumprod"):
            got = getattr(xt, kind)(X.dev, ax).to_numpy()
            if dtype == "bfloat16":
                if kind == "cumprod":
                    continue
                ref = o.scan(X.dev, ax, kind, high_precision=True)
                assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)) <= 2e-7, (kind, ax)
            else:
                ref = o.scan(X.dev, ax, kind)
                assert got.dtype == ref.dtype and np.array_equal(got, ref), (kind, ax)
        axes = None if ax is None else [ax]
        for red in (xt.sum, xt.prod, xt.amax, xt.mean):
            r = red(X.dev, axes)
            got = xt.eval(r).to_numpy()
            if dtype == "bfloat16" and red in (xt.sum, xt.mean):
                ref = o.reduce(r, high_precision=True)
                assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)) <= 2e-7, (red.__name__, ax)
            elif dtype == "bfloat16" and red is xt.prod:
                continue  # products of 37-70 bf16 values leave float range: not a parity case
            else:
                ref = o.reduce(r)
                assert got.dtype == ref.dtype, (red.__name__, ax, got.dtype, ref.dtype)
                assert np.array_equal(got, ref) or (got.dtype.kind == "f" and np.allclose(got, ref, rtol=1e-14, atol=0)), \
                    (red.__name__, ax)

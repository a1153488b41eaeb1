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


def tiled_form(name):
    """The tiled scan a kernel name belongs to, whether it ran in three
    launches or in one pass ("_1p")."""
    return name[:-3] if name.endswith("_1p") else name


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


def test_scan_of_expression_along_strided_axis_evaluates_first():
    """cumsum of an expression over 2^20 elements or more evaluates the
    expression row-major into a temporary of the result type first
    (accumulator_impl copies before it scans) and scans that with the one-pass
    tiled kernel (strided axis) or the direct rows kernel (contiguous axis),
    instead of the generic kernels; smaller ones keep the column kernel.  float64 (double literals)
    against the long-double reference, int64 exact; eager and prepared."""
    rng = np.random.default_rng(44)
    a = rng.uniform(-1, 1, (2048, 1024)).astype(np.float32)
    A_ = Dual.of(a)
    got = xt.cumsum(A_.dev * 2.0 + 1.0, 0)
    sync()
    assert xt.last_kernel() == "scan_tiles_1p"
    ref = _scan_ref(a.astype(np.float64) * 2.0 + 1.0, 0, "cumsum")
    _check_scan(got.to_numpy(), ref, "float64", "expr axis 0")
    b = rng.integers(-100, 100, (3, 700, 520)).astype(np.int32)
    B = xt.from_numpy(b)
    got = xt.cumsum(B * 3 - 1, 1)
    sync()
    assert xt.last_kernel() == "scan_tiles_1p"
    ref = np.cumsum(b.astype(np.int64) * 3 - 1, axis=1)
    assert np.array_equal(got.to_numpy().astype(np.int64), ref)
    out = xt.empty((2048, 1024), "float64")
    plan = xt.plan(out, xt.Scan("cumsum", A_.dev * 2.0 + 1.0, 0))
    for _ in range(2):
        plan.run()
        sync()
        assert xt.last_kernel() == "scan_call"  # the plan keeps its descriptors and runs as xg_scan
        _check_scan(out.to_numpy(), _scan_ref(a.astype(np.float64) * 2.0 + 1.0, 0, "cumsum"), "float64", "plan")
    got = xt.cumsum(A_.dev * 2.0 + 1.0, 1)  # the contiguous axis: the direct rows kernel after the copy
    sync()
    assert xt.last_kernel().startswith("scan_rows_direct")
    _check_scan(got.to_numpy(), _scan_ref(a.astype(np.float64) * 2.0 + 1.0, 1, "cumsum"), "float64", "expr axis 1")
    small = rng.uniform(-1, 1, (300, 300)).astype(np.float32)
    got = xt.cumsum(xt.from_numpy(small) * 2.0, 0)
    sync()
    assert xt.last_kernel() == "scan_cols"
    _check_scan(got.to_numpy(), _scan_ref(small.astype(np.float64) * 2.0, 0, "cumsum"), "float64", "small")


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


def _scan_ref(x, axis, kind):
    """numpy's sequential scan, for floats in long double (x86 80-bit: its
    own rounding stays far below the float64 bar), rounded to float64."""
    f = np.cumsum if kind == "cumsum" else np.cumprod
    if x.dtype.kind == "f":
        return f(x.astype(np.longdouble), axis=axis).astype(np.float64)
    return f(x, axis=axis, dtype=x.dtype)  # C++: an int scan stays int (wrapping)


def _check_scan(got, ref, dtype, what, kind="cumsum", length=1):
    """Sums: rel 2e-7 (float32, compensated) / 1e-14 (float64).  Products are
    not compensated: a product of n factors carries up to about n rounding
    errors of the type (n * 2^-24 or n * 2^-53 relative) in any order, the
    sequential one included."""
    if np.dtype(dtype).kind == "f":
        err = np.max(np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1.0)) if ref.size else 0.0
        if kind == "cumprod":
            bar = length * (2.0 ** -24 if dtype == "float32" else 2.0 ** -53)
        else:
            bar = 2e-7 if dtype == "float32" else 1e-14
        assert err <= bar, (what, err, bar)
    else:
        assert got.dtype == ref.dtype and np.array_equal(got, ref), what


def _tiled_input(rng, shape, dtype, kind):
    if kind == "cumprod":  # products that stay in range: +-1 for integers, near 1 for floats
        if np.dtype(dtype).kind == "f":
            return rng.uniform(0.999, 1.001, shape).astype(dtype)
        if np.dtype(dtype).kind == "u":  # products of 3s wrap modulo 2^32, as in C++
            return rng.choice(np.array([1, 1, 1, 3], dtype=dtype), shape)
        return rng.choice(np.array([-1, 1, 1, 1], dtype=dtype), shape)
    if np.dtype(dtype).kind == "f":
        return rng.normal(1.0, 1.0, shape).astype(dtype)
    hi = 1000 if np.dtype(dtype).kind == "i" else 2000
    lo = -1000 if np.dtype(dtype).kind == "i" else 0
    return rng.integers(lo, hi, shape).astype(dtype)


# (shape, tiled): the tiled kernels need rows of whole 16-byte vectors
TILED_SHAPES = [((8192, 4096), True), ((2048, 512), True), ((1000, 300), True), ((77, 8), True), ((3, 1000, 132), True),
                ((130, 2, 260), True), ((77, 5), False), ((130, 2, 257), False)]


@pytest.mark.parametrize("dtype", ["float32", "float64", "int32", "int64", "uint32"])
@pytest.mark.parametrize("kind", ["cumsum", "cumprod"])
def test_tiled_column_scans(dtype, kind):
    """Scans along a strided axis whose other axis is contiguous (cumsum(X, 0)
    of a row-major X, and the middle axis of 3-d arrays) run the tiled
    kernels: chunk totals, carries across chunks, scan from the carry.  Shapes
    cover whole and partial column tiles (256 columns of 4-byte types, 128 of
    8-byte types), partial chunks and rows shorter than one vector.  Integers
    exact; floats against a float64 scan (rel 2e-7 float32, 1e-14 float64)."""
    rng = np.random.default_rng(31)
    for shape, tiled in TILED_SHAPES:
        if kind == "cumprod" and shape[0] * (shape[-1] if len(shape) == 2 else shape[1]) > 4_000_000:
            continue
        axis = 0 if len(shape) == 2 else 1
        x = _tiled_input(rng, shape, dtype, kind)
        X = xt.from_numpy(x)
        got = getattr(xt, kind)(X, axis)
        sync()
        assert (tiled_form(xt.last_kernel()) == "scan_tiles") == tiled, (shape, xt.last_kernel())
        _check_scan(got.to_numpy(), _scan_ref(x, axis, kind), dtype, (kind, shape), kind, shape[axis])


@pytest.mark.parametrize("dtype", ["float32", "float64", "int32", "int64"])
def test_tiled_transposed_scans(dtype):
    """cumsum(transpose(X), 1): the input is contiguous across scan lines and
    the output along them; the tiled kernel goes through an LDS transpose."""
    rng = np.random.default_rng(32)
    for shape, tiled in [((8192, 4096), True), ((1000, 300), True), ((36, 8), True), ((8, 2000), True),
                         ((33, 7), False), ((5, 2000), False)]:
        x = _tiled_input(rng, shape, dtype, "cumsum")
        X = xt.from_numpy(x)
        got = xt.cumsum(xt.transpose(X), 1)
        sync()
        assert (tiled_form(xt.last_kernel()) == "scan_tiles_trans") == tiled, (shape, xt.last_kernel())
        _check_scan(got.to_numpy(), _scan_ref(x.T, 1, "cumsum"), dtype, shape)


def test_tiled_scan_store_conversion_and_plans():
    """A tiled scan into an output of another type of the same size (float32
    sums stored as int32, static_cast), and the same scan as a prepared plan
    replayed on two inputs, exact against the oracle."""
    rng = np.random.default_rng(33)
    x = rng.integers(-50, 50, (600, 300)).astype(np.float32)
    X = Dual.of(x)
    out = xt.empty((600, 300), "int32")
    xt.plan(out, xt.Scan("cumsum", X.dev, 0)).run()
    sync()
    assert tiled_form(xt.last_kernel()) == "scan_tiles"
    ref = oracle_for(X).scan(X.dev, 0)
    assert np.array_equal(out.to_numpy(), ref.astype(np.int32))
    y = rng.integers(-50, 50, (600, 300)).astype(np.float32)
    Y = xt.from_numpy(y)
    o2 = xt.empty((600, 300), "float32")
    p = xt.plan(o2, xt.Scan("cumsum", Y, 0))
    for _ in range(2):
        p.run()
        sync()
        assert np.array_equal(o2.to_numpy(), np.cumsum(y.astype(np.float64), axis=0).astype(np.float32))


# (shape, axis, transposed input) for the one-pass form: 1 to 16 column
# chains (outer x tiles; 1, 2, 4, 8 and 16 ticket counters), partial column
# tiles, partial last chunks, a 3-d outer dim and chains of up to 2048 chunks
# (many look-back windows)
ONE_PASS_CASES = [((8192, 4096), 0, False), ((300, 2100), 0, False), ((4, 200, 520), 1, False),
                  ((20000, 2048), 0, False), ((65, 2304), 0, False), ((130, 2560), 0, False),
                  ((131072, 200), 0, False), ((3, 5000, 256), 1, False), ((1000, 512), 0, False),
                  ((8192, 4096), 1, True), ((600, 2100), 1, True), ((100, 2052), 1, True),
                  ((30000, 100), 1, True)]


def _one_pass_run(x, axis, trans, kind):
    X = xt.from_numpy(x)
    src = xt.transpose(X) if trans else X
    got = getattr(xt, kind)(src, axis)
    sync()
    return got.to_numpy(), xt.last_kernel()


@pytest.mark.parametrize("dtype", ["float32", "int32", "uint32", "float64", "int64"])
@pytest.mark.parametrize("kind", ["cumsum", "cumprod"])
def test_tiled_scans_one_pass(dtype, kind):
    """Strided-axis scans over two or more chunks of the axis run in one
    pass: every chunk of 64 rows is read once, its column totals are
    published and each column looks back across the preceding chunks for its
    carry (32-bit accumulators: value and flag in one word; 64-bit ones:
    value, wait, then one flag per chunk and tile).  Checked against the
    sequential reference (integers exact, float sums rel 2e-7 / 1e-14,
    products n ulps) and against the three-launch form of the same scan (the
    knob that turns the one pass off)."""
    from xtensor_amd import _abi
    lib = _abi.lib()
    rng = np.random.default_rng(35)
    for shape, axis, trans in ONE_PASS_CASES:
        if kind == "cumprod" and np.prod(shape) > 3_000_000:
            continue
        x = _tiled_input(rng, shape, dtype, kind)
        src = x.T if trans else x
        got, name = _one_pass_run(x, axis, trans, kind)
        want = "scan_tiles_trans_1p" if trans else "scan_tiles_1p"
        assert name == want, (shape, name)
        _check_scan(got, _scan_ref(src, axis, kind), dtype, (kind, shape), kind, src.shape[axis])
        try:
            lib.xg_debug_scan_single_pass(0)
            three, name3 = _one_pass_run(x, axis, trans, kind)
        finally:
            lib.xg_debug_scan_single_pass(1)
        assert name3 == want[:-3], (shape, name3)
        if np.dtype(dtype).kind == "f":
            scale = np.maximum(np.abs(three.astype(np.float64)), 1.0)
            ulp = 2.0 ** -24 if dtype == "float32" else 2.0 ** -53
            bar = (2e-7 if dtype == "float32" else 1e-14) if kind == "cumsum" else 2 * src.shape[axis] * ulp
            assert np.max(np.abs(got.astype(np.float64) - three) / scale) <= bar, (kind, shape)
        else:
            assert np.array_equal(got, three), (kind, shape)


def test_tiled_scan_one_pass_repeat_and_plan():
    """The one pass clears its ticket counter and publications before every
    launch: a prepared plan replayed on new inputs, and the same plan after a
    scan of another shape reused the allocator's memory, stay exact (integer
    data)."""
    rng = np.random.default_rng(36)
    out = xt.empty((3000, 2560), "int32")
    xs = [rng.integers(-1000, 1000, (3000, 2560)).astype(np.int32) for _ in range(3)]
    X = xt.empty((3000, 2560), "int32")
    p = xt.plan(out, xt.Scan("cumsum", X, 0))
    for x in xs:
        xt.assign(X, xt.from_numpy(x))
        for _ in range(2):
            p.run()
            sync()
            assert xt.last_kernel() == "scan_tiles_1p"
            assert np.array_equal(out.to_numpy(), np.cumsum(x, axis=0, dtype=np.int32))
        small = xt.cumsum(xt.from_numpy(x[:700, :2048]), 0)
        sync()
        assert np.array_equal(small.to_numpy(), np.cumsum(x[:700, :2048], axis=0, dtype=np.int32))


@pytest.fixture(params=["scan_chunks3", "scan_lookback"])
def long_row_form(request):
    """Few long rows run in chunks of 4 x 16 x 64 x 16 bytes, either as the
    three-launch reduce-then-scan (the default) or as the one-pass look-back
    (the knob that selects it)."""
    from xtensor_amd import _abi
    lib = _abi.lib()
    lib.xg_debug_scan_flat_single_pass(1 if request.param == "scan_lookback" else 0)
    try:
        yield request.param
    finally:
        lib.xg_debug_scan_flat_single_pass(0)


@pytest.mark.parametrize("dtype", ["float32", "float64", "int32", "int64"])
def test_lookback_long_rows(dtype, long_row_form):
    """Few rows much longer than one chunk (16384 elements of 4-byte types,
    8192 of 8-byte types) run the single-pass look-back scan: the flattened
    cumsum of a large array, and cumsum along axis 1 of a few long rows.
    Lengths around the chunk edges and one past 2^25.  Integers exact; floats
    against a float64 scan (rel 2e-7 for float32, 1e-14 for float64); the
    float64 and int64 cases use the flag + value publication, the others the
    packed word."""
    rng = np.random.default_rng(41)
    for n in [16385, 32768, 32769, 300_001, (1 << 25) + 7]:
        x = _tiled_input(rng, (n,), dtype, "cumsum")
        if np.dtype(dtype).kind != "f":
            x = (x % 7 - 3).astype(dtype)
        X = xt.from_numpy(x)
        got = xt.cumsum(X)
        sync()
        assert xt.last_kernel() == long_row_form, (n, xt.last_kernel())
        _check_scan(got.to_numpy(), _scan_ref(x, 0, "cumsum"), dtype, n)
    x = _tiled_input(rng, (3, 1_000_004), dtype, "cumsum")  # rows of whole 16-byte vectors
    X = xt.from_numpy(x)
    got = xt.cumsum(X, 1)
    sync()
    assert xt.last_kernel() == long_row_form
    _check_scan(got.to_numpy(), _scan_ref(x, 1, "cumsum"), dtype, "3 rows")


def test_chunked_scan_past_summed_carries():
    """A row of more than 4096 chunks of 16384 elements: the reduce-then-scan
    form merges the chunk totals into carries in a launch of its own instead
    of in every chunk's workgroup.  int32, exact (values in -3..3)."""
    rng = np.random.default_rng(43)
    n = 4097 * 16384 + 5
    x = rng.integers(-3, 4, n).astype(np.int32)
    got = xt.cumsum(xt.from_numpy(x))
    sync()
    assert xt.last_kernel() == "scan_chunks3"
    assert np.array_equal(got.to_numpy(), np.cumsum(x, dtype=np.int32))


def test_lookback_cumprod_repeat_and_expression(long_row_form):
    """cumprod of +-1 integers and of values near 1 through the look-back
    kernel; the same scan twice in a row (the chunk flags are cleared before
    every launch); the flattened scan of an expression (evaluated into a
    temporary that shares the workspace block with the chunk flags)."""
    rng = np.random.default_rng(42)
    p = rng.choice(np.array([-1, 1, 1, 1], np.int32), 200_000)
    P = xt.from_numpy(p)
    for _ in range(2):
        got = xt.cumprod(P).to_numpy()
        assert xt.last_kernel() == long_row_form
        assert np.array_equal(got, np.cumprod(p))
    f = rng.uniform(0.9999, 1.0001, 500_000).astype(np.float64)
    got = xt.cumprod(xt.from_numpy(f)).to_numpy()
    ref = _scan_ref(f, 0, "cumprod")
    assert np.max(np.abs(got - ref) / np.abs(ref)) <= 1e-13
    a = rng.normal(0, 1, (300, 1000)).astype(np.float32)
    A_ = Dual.of(a)
    e = A_.dev * 2.0 + 1.0  # float64 values (double literals)
    got = xt.cumsum(e).to_numpy()
    assert xt.last_kernel() == long_row_form
    ref = _scan_ref((a.astype(np.float64) * 2.0 + 1.0).ravel(), 0, "cumsum")
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)) <= 1e-14

"""GPU parity of elementwise assignment (xg_assign) against the oracle.

Bar: bit-exact for integer work and for + - * / compare / select / casts
(both sides round every operator once: -ffp-contract=off); <= 2 ulp where
a transcendental (exp, log, ...) is involved, since the device library and
the host libm are different correctly-rounded-to-1-ulp implementations.
"""

import json
import os

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd import _abi as A

from helpers import Dual, bf16_bits_to_f32, oracle_for, same_bits, to_bf16_bits, ulp_diff
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

KATS = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "kat_arrays.json")))


def kat(file, test, name, dtype=np.int64):
    r = KATS[f"{file}:{test}:{name}"]
    return np.array(r["values"], dtype=dtype).reshape(r["shape"])


def sync():
    import torch
    torch.cuda.synchronize()


def check_equal_or_nan(got, ref):
    assert got.shape == ref.shape and got.dtype == ref.dtype
    if got.dtype.kind == "f":
        both_nan = np.isnan(got) & np.isnan(ref)
        g = np.where(both_nan, 0, got)
        r = np.where(both_nan, 0, ref)
        assert same_bits(g, r), f"{np.count_nonzero(g != r)} mismatches"
    else:
        assert np.array_equal(got, ref)


# ---------------------------------------------------------------- broadcast + exp (f64)
def make_ab(rows=2048, cols=1024, seed=0):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (rows, cols))
    b = rng.uniform(-1, 1, (cols,))
    return a, b


@pytest.mark.parametrize("rows,cols", [(2048, 1024), (33, 17), (1, 1)])
def test_mul_add_exp_broadcast_f64(rows, cols):
    """c = a*b + exp(-a*a), b broadcast over rows (f64).

    exp is the only inexact operator: the device's exp(-a*a) is within 2 ulp
    of the host libm's, and c is bit-exact given that exp (products and sum
    are IEEE operations, each rounded once).  Where a*b nearly cancels
    exp(-a*a) an ulp of exp is many ulps of c, so c itself is compared through
    its exp term."""
    a_h, b_h = make_ab(rows, cols)
    A_, B_ = Dual.of(a_h), Dual.of(b_h)
    c = xt.empty((rows, cols), "float64")
    xt.noalias(c).assign(A_.dev * B_.dev + xt.exp(-A_.dev * A_.dev))
    sync()
    used = xt.last_kernel()
    got = c.to_numpy()
    o = oracle_for(A_, B_)
    ex_dev = xt.eval(xt.exp(-A_.dev * A_.dev)).to_numpy()
    ex_ref = o.assign(xt.exp(-A_.dev * A_.dev))
    u = ulp_diff(ex_dev, ex_ref)
    assert u.max() <= 2, f"exp: max ulp {u.max()}"
    assert same_bits(a_h * b_h + ex_dev, got)
    ref = o.assign(A_.dev * B_.dev + xt.exp(-A_.dev * A_.dev))
    # the ULP histogram of exp against the host libm, and how many elements
    # of c differ from the host evaluation because of it
    hist = np.bincount(u.ravel().astype(np.int64), minlength=3)
    assert hist.sum() == u.size and len(hist) == 3, hist
    differ = int(np.count_nonzero(got != ref))
    assert differ <= int(hist[1] + hist[2]), (differ, hist)  # only where exp itself differs
    print(f"kernel={used}: exp ulp histogram vs host libm {hist.tolist()} (0/1/2 ulp), "
          f"c differs from the host in {differ} of {got.size} elements")


def test_spec_and_dynamic_agree_bitwise():
    """The compiled specialisation and the generic interpreter produce the same bits."""
    a_h, b_h = make_ab(257, 1024, seed=3)
    a, b = xt.from_numpy(a_h), xt.from_numpy(b_h)
    c1 = xt.empty(a_h.shape, "float64")
    c2 = xt.empty(a_h.shape, "float64")
    xt.noalias(c1).assign(a * b + xt.exp(-a * a))
    k1 = xt.last_kernel()
    A.lib().xg_debug_force_dynamic(1)
    try:
        xt.noalias(c2).assign(a * b + xt.exp(-a * a))
        k2 = xt.last_kernel()
    finally:
        A.lib().xg_debug_force_dynamic(0)
    sync()
    assert k1 == "Spec_mul_add_exp_f64" and k2 == "dynamic"
    assert same_bits(c1.to_numpy(), c2.to_numpy())


@pytest.mark.parametrize("shape", [(257, 1024), (3, 5), (1000, 1001), (64, 4096)])
def test_units_per_thread_variants_agree(shape):
    """One or four work units per thread (the large-launch form) give the same bits."""
    a_h, b_h = make_ab(*shape, seed=4)
    a, b = xt.from_numpy(a_h), xt.from_numpy(b_h)
    outs = []
    try:
        for upt in (1, 4):
            A.lib().xg_debug_units_per_thread(upt)
            c = xt.empty(a_h.shape, "float64")
            xt.noalias(c).assign(a * b + xt.exp(-a * a))
            if shape[1] % 2 == 0:  # vector-aligned rows run the compiled program
                assert xt.last_kernel() == "Spec_mul_add_exp_f64"
            x = xt.from_numpy(np.random.default_rng(5).normal(0, 1, (3, 256, 56, 56)).astype(np.float32))
            y = xt.zeros(x.shape, "float32")
            params = [xt.from_numpy(np.full((1, 256, 1, 1), v, np.float32)) for v in (0.1, 1.5, 0.9, -0.2)]
            vx = x[:, :, xt.range_(0, 56, 2), :]
            xt.noalias(y[:, :, xt.range_(0, 56, 2), :]).assign(
                xt.clip((vx - params[0]) / xt.sqrt(params[1] + 1e-5) * params[2] + params[3], -3, 3))
            assert xt.last_kernel() == "Spec_affine_clip_f32"
            sync()
            outs.append((c.to_numpy(), y.to_numpy()))
    finally:
        A.lib().xg_debug_units_per_thread(0)
    assert same_bits(outs[0][0], outs[1][0]) and same_bits(outs[0][1], outs[1][1])


def test_full_size_workloads_run_compiled_programs():
    """At full size (operands above the non-temporal threshold, so their loads
    and stores take the streaming form) the strided-view affine workload and
    the int32 bit-mix still match their compiled programs, and the streaming
    and cached forms give the same bits."""
    rng = np.random.default_rng(8)
    x = xt.from_numpy(rng.normal(0, 1, (64, 256, 56, 56)).astype(np.float32))
    params = [xt.from_numpy(np.full((1, 256, 1, 1), v, np.float32)) for v in (0.1, 1.5, 0.9, -0.2)]
    xi = xt.from_numpy(rng.integers(-2**20, 2**20, (64, 256, 56, 56), dtype=np.int32))
    outs = []
    try:
        for nt in (1, 0):
            A.lib().xg_debug_nontemporal(nt, nt)
            y = xt.zeros(x.shape, "float32")
            vx = x[:, :, xt.range_(0, 56, 2), :]
            xt.noalias(y[:, :, xt.range_(0, 56, 2), :]).assign(
                xt.clip((vx - params[0]) / xt.sqrt(params[1] + 1e-5) * params[2] + params[3], -3, 3))
            assert xt.last_kernel() == "Spec_affine_clip_f32"
            yi = xt.empty(xi.shape, "int32")
            xt.noalias(yi).assign((xi * 3 + (xi >> 2)) & 0xFFFF)
            assert xt.last_kernel() == "Spec_bitmix_i32"
            s = xt.sum(x, [3])
            r = xt.eval(s)
            assert xt.last_kernel() == "RSpec_sum_f32"
            sync()
            outs.append((y.to_numpy(), yi.to_numpy(), r.to_numpy()))
    finally:
        A.lib().xg_debug_nontemporal(1, 1)
    for u, v in zip(outs[0], outs[1]):
        assert same_bits(u, v)


# ---------------------------------------------------------------- affine + clip through a strided view
def affine_inputs(n=4, c=256, h=56, w=56, seed=1, dtype=np.float32):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, c, h, w)).astype(dtype)
    mean = rng.normal(0, 0.1, (1, c, 1, 1)).astype(dtype)
    var = rng.uniform(0.5, 2, (1, c, 1, 1)).astype(dtype)
    gamma = rng.uniform(0.5, 1.5, (1, c, 1, 1)).astype(dtype)
    beta = rng.normal(0, 0.1, (1, c, 1, 1)).astype(dtype)
    return x, mean, var, gamma, beta


@pytest.mark.parametrize("n,dtype", [(4, np.float32), (64, np.float32), (2, np.float64)])
def test_affine_clip_strided_view(n, dtype):
    """y[:, :, 0:56:2, :] = clip((x - mean) / sqrt(var + 1e-5) * gamma + beta, -3, 3),
    x read through the same view; y's other rows stay zero."""
    x, mean, var, gamma, beta = affine_inputs(n, dtype=dtype)
    X, M, Vv, G, Bt = (Dual.of(v) for v in (x, mean, var, gamma, beta))
    XV = X.view(xt.all_(), xt.all_(), xt.range_(0, 56, 2), xt.all_())
    y = xt.zeros(x.shape, x.dtype)
    yv = xt.view(y, xt.all_(), xt.all_(), xt.range_(0, 56, 2), xt.all_())

    def expr(xv, m, v, g, b):
        return xt.clip((xv - m) / xt.sqrt(v + 1e-5) * g + b, -3, 3)

    xt.noalias(yv).assign(expr(XV.dev, M.dev, Vv.dev, G.dev, Bt.dev))
    sync()
    used = xt.last_kernel()
    ref = np.zeros_like(x)
    oracle_for(XV, M, Vv, G, Bt).assign(expr(XV.dev, M.dev, Vv.dev, G.dev, Bt.dev), out=ref[:, :, 0:56:2, :])
    got = y.to_numpy()
    assert same_bits(got, ref)
    assert not np.any(got[:, :, 1::2, :])
    print("kernel", used)
    if dtype == np.float32:
        assert used == "Spec_affine_clip_f32"


def test_bitmix_int32():
    """y = (x*3 + (x>>2)) & 0xFFFF on int32 (bit-exact)."""
    rng = np.random.default_rng(1)
    x_h = rng.integers(-2**20, 2**20, (16, 256, 56, 56), dtype=np.int32)
    X = Dual.of(x_h)
    y = xt.empty(x_h.shape, "int32")
    xt.noalias(y).assign((X.dev * 3 + (X.dev >> 2)) & 0xFFFF)
    sync()
    assert xt.last_kernel() == "Spec_bitmix_i32"
    ref = oracle_for(X).assign((X.dev * 3 + (X.dev >> 2)) & 0xFFFF)
    assert np.array_equal(y.to_numpy(), ref)


# ---------------------------------------------------------------- reference KAT replays on the device
BV = "test_extended_broadcast_view.cpp"


def test_kat_broadcasting_on_device():
    """test_extended_broadcast_view.cpp:114-500."""
    T = "xtest_extended.broadcasting"
    a1, a2 = xt.from_numpy(kat(BV, T, "py_a1")), xt.from_numpy(kat(BV, T, "py_a2"))
    b1, b2 = xt.from_numpy(kat(BV, T, "py_b1")), xt.from_numpy(kat(BV, T, "py_b2"))
    b3 = xt.from_numpy(kat(BV, T, "py_b3", np.float64))
    assert np.array_equal(xt.eval(a1 + a2).to_numpy(), kat(BV, T, "py_a1_a2"))
    assert np.array_equal(xt.eval(a1 + b1).to_numpy(), kat(BV, T, "py_b1_a1"))
    assert np.array_equal(xt.eval(a1 + b2).to_numpy(), kat(BV, T, "py_a1_b2"))
    np.testing.assert_allclose(xt.eval(a2 + b3).to_numpy(), kat(BV, T, "py_a2_b3", np.float64), rtol=2e-15)


def test_kat_negative_slices_on_device():
    """test_extended_broadcast_view.cpp:31-112: reversed views, and `view += view`
    with overlap (operator= evaluates into a temporary)."""
    T = "xtest_extended.negative_slices_twod"
    a = xt.from_numpy(kat(BV, T, "py_a", np.float64))
    for keys, name in [((xt.range_(None, -2), xt.range_(None, None, -1)), "py_av0"),
                       ((xt.range_(None, None, -1), xt.range_(None, None, -1)), "py_av1"),
                       ((xt.range_(1, -3), xt.range_(-3, 2, -1)), "py_av2"),
                       ((xt.range_(-1, -4, -1), xt.range_(-3, 1, -2)), "py_av3"),
                       ((xt.range_(-5, -2), xt.range_(-3, 10)), "py_av5")]:
        got = xt.eval(xt.view(a, *keys)).to_numpy()
        assert np.array_equal(got, kat(BV, T, name, np.float64)), name
    T = "xtest_extended.negative_slices_math"
    a0 = xt.from_numpy(kat(BV, T, "py_a0"))
    a0[0:-2] += a0[:3:-1]
    assert np.array_equal(a0.to_numpy(), kat(BV, T, "py_a"))
    a0[::-2] += a0[::2]
    assert np.array_equal(a0.to_numpy(), kat(BV, T, "py_at"))


def test_kat_broadcast_into_view_on_device():
    """test_extended_broadcast_view.cpp:811-953."""
    T = "xtest_extended.broadcast_into_view"
    a = xt.from_numpy(kat(BV, T, "py_a0"))
    a[:, :, 2] = xt.from_numpy(kat(BV, T, "py_b1"))
    assert np.array_equal(a.to_numpy(), kat(BV, T, "py_a2"))
    a[:, 0, None, :] = xt.from_numpy(kat(BV, T, "py_ar2"))
    assert np.array_equal(a.to_numpy(), kat(BV, T, "py_a3"))
    a[:, 2, :] = xt.from_numpy(np.arange(7, dtype=np.int64))
    assert np.array_equal(a.to_numpy(), kat(BV, T, "py_a4"))
    for keys, name in [((1,), "py_ax1"), ((slice(0, 2), slice(0, 2)), "py_ax2"),
                       ((slice(None), slice(0, 2)), "py_ax4"), ((1, slice(0, 2, 2)), "py_ax3")]:
        a = xt.from_numpy(kat(BV, T, "py_a0"))
        a[keys] = -1
        assert np.array_equal(a.to_numpy(), kat(BV, T, name)), name


def test_kat_broadcast_funcs_view_on_device():
    """test_extended_broadcast_view.cpp:971-1031."""
    T = "xtest_extended.broadcast_funcs_view"
    a0 = xt.from_numpy(kat(BV, T, "py_a0"))
    a1 = xt.from_numpy(kat(BV, T, "py_a1"))
    a4 = xt.from_numpy(kat(BV, T, "py_a4").reshape(3, 1))
    a = xt.from_numpy(kat(BV, T, "py_a0"))
    a[0:2] = a0 * 2 + a4
    assert np.array_equal(a.to_numpy(), kat(BV, T, "py_a0_cp1"))
    a[0] *= a1 * 7 + a4
    assert np.array_equal(a.to_numpy(), kat(BV, T, "py_a0_cp2"))
    x0 = xt.from_numpy(kat(BV, T, "py_x0"))
    a[0, :, ::2] = x0[1]
    assert np.array_equal(a.to_numpy(), kat(BV, T, "py_a0_cp3"))
    a[0] = xt.eval(a0 * a0 + 3)[0]
    assert np.array_equal(a.to_numpy()[0], kat(BV, T, "py_x1"))


# ---------------------------------------------------------------- fuzz: random trees, all dtypes
DTYPES = ["bool", "int8", "uint8", "int16", "uint16", "int32", "uint32", "int64", "uint64", "float32", "float64"]
EXACT_BIN = ["PLUS", "MINUS", "MULTIPLIES", "LESS", "LESS_EQUAL", "GREATER", "GREATER_EQUAL", "EQUAL_TO",
             "NOT_EQUAL_TO", "MAXIMUM", "MINIMUM", "LOGICAL_AND", "LOGICAL_OR", "BITWISE_AND", "BITWISE_OR",
             "BITWISE_XOR"]
EXACT_UN = ["NEGATE", "IDENTITY", "LOGICAL_NOT", "ABS", "FLOOR", "CEIL", "TRUNC", "ROUND"]


def random_values(rng, dt, shape):
    if dt == "bool":
        return rng.integers(0, 2, shape).astype(np.bool_)
    if dt.startswith("float"):
        return rng.uniform(-50, 50, shape).astype(dt)
    info = np.iinfo(dt)
    lo, hi = max(info.min, -100), min(info.max, 100)
    return rng.integers(lo, hi, shape, dtype=dt)


def random_tree(rng, depth, leaves):
    if depth == 0 or rng.random() < 0.2:
        if rng.random() < 0.15:
            return xt.scalar(int(rng.integers(1, 5)), ["int32", "float64", "uint8", "int64"][rng.integers(4)])
        dt = DTYPES[rng.integers(len(DTYPES))]
        shape = [(6, 40), (40,), (1, 40), (6, 1), ()][rng.integers(5)]
        d = Dual.of(random_values(rng, dt, shape))
        leaves.append(d)
        return d.dev
    k = rng.random()
    if k < 0.3:
        return xt.Func(A.OP[EXACT_UN[rng.integers(len(EXACT_UN))]], random_tree(rng, depth - 1, leaves))
    if k < 0.4:
        return xt.cast(random_tree(rng, depth - 1, leaves), ["int32", "float32", "float64", "int64", "uint8"][rng.integers(5)])
    if k < 0.9:
        return xt.Func(A.OP[EXACT_BIN[rng.integers(len(EXACT_BIN))]], random_tree(rng, depth - 1, leaves),
                       random_tree(rng, depth - 1, leaves))
    return xt.where(random_tree(rng, depth - 1, leaves) > 0, random_tree(rng, depth - 1, leaves),
                    random_tree(rng, depth - 1, leaves))


@pytest.mark.parametrize("seed", range(12))
def test_random_trees_bit_exact(seed):
    """Random trees of exact operators over every dtype, generic interpreter and
    (where a signature matches) compiled kernels: bit-exact against the oracle."""
    rng = np.random.default_rng(1000 + seed)
    done = 0
    while done < 15:
        leaves = []
        e = random_tree(rng, 3, leaves)
        if not isinstance(e, xt.Func):
            continue
        try:
            dt, shape = xt.expr_info(e)
        except (TypeError, xt.BroadcastError):
            continue
        out = xt.empty(shape, dt)
        try:
            xt.noalias(out).assign(e)
        except NotImplementedError:
            continue
        sync()
        ref = Oracle({d.dev: d.host for d in leaves}).assign(e)
        got = out.to_numpy()
        check_equal_or_nan(got, ref)
        done += 1


@pytest.mark.parametrize("src", DTYPES)
@pytest.mark.parametrize("dst", DTYPES)
def test_store_conversions(src, dst):
    """Assigning into a container of another type is static_cast (utils/xutils.hpp:617-646)."""
    rng = np.random.default_rng(7)
    v = random_values(rng, src, (3, 37))
    if src.startswith("float") and dst in ("uint8", "uint16", "uint32", "uint64", "bool"):
        v = np.abs(v)  # negative float -> unsigned is undefined in C++
    X = Dual.of(v)
    out = xt.empty(v.shape, dst)
    xt.noalias(out).assign(X.dev)
    sync()
    ref = np.zeros(v.shape, np.dtype(dst))
    oracle_for(X).assign(X.dev, out=ref)
    check_equal_or_nan(out.to_numpy(), ref)


@pytest.mark.parametrize("fn", ["exp", "log", "sqrt", "sin", "cos", "tanh", "exp2", "log2", "log10", "expm1",
                                "log1p", "cbrt", "atan", "erf"])
@pytest.mark.parametrize("dt", ["float32", "float64"])
def test_transcendentals_within_2ulp(fn, dt):
    rng = np.random.default_rng(3)
    v = rng.uniform(0.01, 10, 4096).astype(dt)
    X = Dual.of(v)
    e = getattr(xt, fn)(X.dev)
    got = xt.eval(e).to_numpy()
    sync()
    ref = oracle_for(X).assign(e)
    u = ulp_diff(got, ref)
    if fn == "cbrt" and dt == "float64":
        # the host libm's double cbrt is itself up to 3 ulp from the exact cube
        # root on these inputs; the device value is checked against the exact
        # root (60-digit decimal) instead: within 1 ulp
        import decimal
        decimal.getcontext().prec = 60
        third = decimal.Decimal(1) / decimal.Decimal(3)
        worst = 0.0
        for xv, gv in zip(v[:1024], got[:1024]):
            t = decimal.Decimal(float(xv)) ** third
            worst = max(worst, abs(float((decimal.Decimal(float(gv)) - t) / decimal.Decimal(float(np.spacing(gv))))))
        assert worst <= 1.0, f"cbrt/float64: {worst} ulp from the exact root"
        assert u.max() <= 4
        return
    assert u.max() <= 2, f"{fn}/{dt}: max {u.max()} ulp"


def test_bf16_storage():
    """bf16 leaves load as float; a float result stored into bf16 rounds to nearest even."""
    rng = np.random.default_rng(9)
    v = rng.normal(size=(17, 64)).astype(np.float32)
    X = Dual.of(v, "bfloat16")
    out = xt.empty(v.shape, "bfloat16")
    xt.noalias(out).assign(X.dev * 3.0 + 1.0)
    out32 = xt.empty(v.shape, "float32")
    xt.noalias(out32).assign(X.dev * 3.0 + 1.0)
    sync()
    o = oracle_for(X)
    ref32 = o.assign(X.dev * 3.0 + 1.0, dtype="float32")
    assert same_bits(out32.to_numpy(), ref32)
    ref16 = np.zeros(v.shape, np.uint16)
    o.assign(X.dev * 3.0 + 1.0, out=ref16, out_code=A.BFLOAT16)
    assert same_bits(out.to_numpy(), bf16_bits_to_f32(ref16))


def test_transposed_and_strided_operands():
    rng = np.random.default_rng(4)
    a = rng.normal(size=(96, 130))
    b = rng.normal(size=(130, 96))
    A_, B_ = Dual.of(a), Dual.of(b)
    BT = B_.transpose()
    AV = A_.view(xt.range_(None, None, -1), xt.range_(1, None, 1))
    BV_ = BT.view(xt.range_(None, None, -1), xt.range_(1, None, 1))
    out = xt.empty(AV.dev.shape, "float64")
    xt.noalias(out).assign(AV.dev * 2 - BV_.dev)
    sync()
    ref = oracle_for(AV, BV_).assign(AV.dev * 2 - BV_.dev)
    assert same_bits(out.to_numpy(), ref)


def test_broadcast_error_on_assign():
    a = xt.zeros((3, 4))
    with pytest.raises(xt.BroadcastError):
        xt.noalias(a).assign(xt.zeros((5,)))

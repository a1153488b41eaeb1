"""GPU parity of xg_assign_reduce2d: out = e with the reductions of its rows
and / or columns in the same pass.  The reference is the three statements
    out = e; s = reduce(out, {1}); m = reduce(out, {0})
(core/xassign.hpp:208-226, reducers/xreducer.hpp:1778-1828), run through the
library's own single calls and checked against the oracle / numpy float64.

Bars: out bit-identical to the plain assignment (same program); max / min /
integer results bit-exact; float sums within 2e-7 (float32) / 1e-14 (float64)
relative of the float64 / long-double sum of the stored values."""

import numpy as np
import pytest

import xtensor_amd as xt

from helpers import Dual, oracle_for, same_bits

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def compile_small():
    """The fused kernel is compiled at run time; let it be compiled for the
    small shapes here too (by default only launches of 2^20 elements or more
    compile)."""
    from xtensor_amd import _abi as A
    A.lib().xg_debug_jit(2)
    yield
    A.lib().xg_debug_jit(1)


def sync():
    import torch
    torch.cuda.synchronize()


def rel(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.longdouble)
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)))


def test_config1_pipeline_in_one_pass():
    """c = a * b + exp(-a * a) (float64 [2048, 1024], b broadcast over rows,
    seed 0), s = sum(c, {1}), m = amax(c, {0}): one fused launch and the
    column merge."""
    r = np.random.default_rng(0)
    a = r.uniform(-1, 1, (2048, 1024))
    b = r.uniform(-1, 1, 1024)
    A_, B_ = Dual.of(a), Dual.of(b)
    e = A_.dev * B_.dev + xt.exp(-A_.dev * A_.dev)
    c = xt.empty((2048, 1024), "float64")
    s = xt.empty((2048,), "float64")
    m = xt.empty((1024,), "float64")
    xt.assign_reduce(c, e, rows=("sum", s), cols=("amax", m))
    sync()
    assert xt.last_kernel() == "assign_rc"
    c2 = xt.empty((2048, 1024), "float64")
    xt.noalias(c2).assign(e)
    sync()
    got_c = c.to_numpy()
    assert same_bits(got_c, c2.to_numpy())
    assert same_bits(m.to_numpy(), got_c.max(0))
    assert rel(s.to_numpy(), got_c.astype(np.longdouble).sum(1)) <= 1e-14
    # against the oracle's xtensor evaluation of the same three statements:
    # the device exp is within 1 ulp of the host's (values below 1), so c
    # differs by at most 2 ulp of 1 (c itself can be near 0)
    ref_c = oracle_for(A_, B_).assign(e)
    assert np.max(np.abs(got_c - ref_c)) <= 4.5e-16
    assert rel(s.to_numpy(), ref_c.sum(1)) <= 1e-14


@pytest.mark.parametrize("shape", [(300, 77), (512, 256), (64, 2050), (3, 4096), (1000, 1)])
@pytest.mark.parametrize("dtype", ["float32", "float64", "int32"])
def test_rows_and_columns_against_single_calls(shape, dtype):
    """x * w - 1 with w broadcast over the rows: uneven rows (a partial unit
    at every row's end), rows of up to 512 units, and past them or of one
    column (the three single calls run instead); int32 with wrap-around."""
    rng = np.random.default_rng(7)
    R, C = shape
    if dtype == "int32":
        x = rng.integers(-2**30, 2**30, shape, dtype=np.int32)
        w = rng.integers(1, 4, C, dtype=np.int32)
    else:
        x = rng.normal(1, 1, shape).astype(dtype)
        w = rng.uniform(0.5, 1.5, C).astype(dtype)
    X, W = xt.from_numpy(x), xt.from_numpy(w)
    e = X * W - 1
    out = xt.empty(shape, dtype)
    s = xt.empty((R,), dtype)
    mn = xt.empty((C,), dtype)
    xt.assign_reduce(out, e, rows=("sum", s), cols=("amin", mn))
    sync()
    used = xt.last_kernel()
    V = 2 if dtype == "float64" else 4
    assert used == ("assign_rc" if R >= 2 and C >= 2 and (C + V - 1) // V <= 512 else "assign_then_reduce")
    ref_out = xt.empty(shape, dtype)
    xt.noalias(ref_out).assign(e)
    ref_s = xt.eval(xt.sum(ref_out, [1]))
    ref_m = xt.eval(xt.amin(ref_out, [0]))
    sync()
    o = out.to_numpy()
    assert same_bits(o, ref_out.to_numpy())
    assert same_bits(mn.to_numpy(), ref_m.to_numpy())
    if dtype == "int32":
        assert same_bits(s.to_numpy(), ref_s.to_numpy())
    else:
        tol = 2e-7 if dtype == "float32" else 1e-14
        ref = o.astype(np.longdouble).sum(1)
        den = np.maximum(np.abs(ref), np.abs(o).astype(np.longdouble).sum(1) * 1e-3 + 1e-30)
        assert float(np.max(np.abs(s.to_numpy() - ref) / den)) <= tol, used


def test_one_side_only_initial_and_keep_dims():
    rng = np.random.default_rng(9)
    x = rng.normal(0, 1, (257, 132)).astype(np.float32)  # rows of whole 4-element units
    X = xt.from_numpy(x)
    out = xt.empty(x.shape, "float32")
    mx = xt.empty((257, 1), "float32")  # keep_dims-shaped row output
    xt.assign_reduce(out, xt.abs(X), rows=("amax", mx, {"keep_dims": True}))
    sync()
    assert xt.last_kernel() == "assign_rc"
    assert same_bits(mx.to_numpy().ravel(), np.abs(x).max(1))
    cs = xt.empty((132,), "float32")
    xt.assign_reduce(out, xt.abs(X), cols=("sum", cs, {"initial": 5.0}))
    sync()
    ref = np.abs(x).astype(np.float64).sum(0) + 5.0
    assert np.max(np.abs(cs.to_numpy() - ref) / ref) <= 2e-7
    ci = xt.empty((132,), "int64")
    xi = rng.integers(-1000, 1000, (257, 132), dtype=np.int32)
    outi = xt.empty(xi.shape, "int32")
    xt.assign_reduce(outi, xt.from_numpy(xi) * 2, cols=("sum", ci, {"dtype": "int64"}))
    sync()
    assert same_bits(outi.to_numpy(), xi * 2)
    assert ci.to_numpy().tolist() == (xi.astype(np.int64) * 2).sum(0).tolist()


def test_fallbacks_give_the_single_call_results():
    """Reductions the single pass does not take (mean; a bf16 output whose
    values are converted before folding; a strided output) run as the three
    calls, with the same results."""
    rng = np.random.default_rng(11)
    x = rng.normal(0, 1, (128, 96)).astype(np.float32)
    X = xt.from_numpy(x)
    out = xt.empty(x.shape, "float32")
    mean = xt.empty((128,), "float64")
    xt.assign_reduce(out, X + 1, rows=("mean", mean))
    sync()
    assert xt.last_kernel() == "assign_then_reduce"
    assert np.max(np.abs(mean.to_numpy() - (x.astype(np.float64) + 1).mean(1))) <= 1e-6
    ob = xt.empty(x.shape, "bfloat16")
    sb = xt.empty((128,), "float32")
    xt.assign_reduce(ob, X, rows=("sum", sb))
    sync()
    ref = xt.eval(xt.sum(ob, [1]))
    sync()
    assert same_bits(sb.to_numpy(), ref.to_numpy())
    big = xt.zeros((128, 192), "float32")
    view = big[:, xt.range_(0, 192, 2)]
    sv = xt.empty((128,), "float32")
    xt.assign_reduce(view, X, rows=("sum", sv))
    sync()
    assert xt.last_kernel() == "assign_then_reduce"
    assert same_bits(big.to_numpy()[:, ::2], x)


def test_bad_arguments():
    X = xt.from_numpy(np.ones((4, 4), np.float32))
    out = xt.empty((4, 4), "float32")
    with pytest.raises(Exception):
        xt.assign_reduce(out, X, rows=("sum", xt.empty((5,), "float32")))


def test_prepared_single_pass_and_graph_replays():
    """The single pass as a prepared plan (xg_plan_assign_reduce2d): run with
    new input values in place between runs, alone, through the fused entry
    point and replayed as a HIP graph, each time equal to the direct call;
    a plan the single pass does not take is refused."""
    import torch
    r = np.random.default_rng(13)
    a = r.uniform(-1, 1, (640, 384))
    b = r.uniform(-1, 1, 384)
    A_, B_ = xt.from_numpy(a), xt.from_numpy(b)
    e = A_ * B_ + xt.exp(-A_ * A_)
    c, s, m = xt.empty((640, 384), "float64"), xt.empty((640,), "float64"), xt.empty((384,), "float64")
    pl = xt.plan(c, e, rows=("sum", s), cols=("amax", m))
    g = xt.PlanGraph([pl])
    for rep, run in enumerate((pl.run, lambda: xt.execute_many([pl]), g.replay, g.replay)):
        A_.storage.view(torch.float64).mul_(0.9).add_(0.01 * rep)
        run()
        c2, s2, m2 = xt.empty((640, 384), "float64"), xt.empty((640,), "float64"), xt.empty((384,), "float64")
        xt.assign_reduce(c2, e, rows=("sum", s2), cols=("amax", m2))
        sync()
        assert same_bits(c.to_numpy(), c2.to_numpy()), rep
        assert same_bits(s.to_numpy(), s2.to_numpy()), rep
        assert same_bits(m.to_numpy(), m2.to_numpy()), rep
    with pytest.raises(NotImplementedError):
        xt.plan(c, e, rows=("mean", s))

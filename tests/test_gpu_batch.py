"""GPU parity of the batched row templates (xg_batch_*).

Per instance the batch replaces xtensor's expression sequence
    m = amax(x, {1}, keep_dims); e = exp(x - m); out = e / sum(e, {1}, keep_dims)
(test/test_xmath.cpp:985-1027) and the row reducers sum / amax (core/xmath.hpp:1803-1823).

Bars:
  * amax: bit-exact (order independent for non-NaN data); int32 sum: bit-exact
    with two's-complement wrap (sum(int32) is int32 in xtensor).
  * float32 softmax: rel <= 1e-6 per element against the float64 evaluation
    of the same float32 inputs, and rel <= 4e-6 against the oracle's float32
    xtensor evaluation (sequential float32 sum, host expf).
  * float64 softmax: rel <= 1e-13 against the long-double-accumulated oracle.
  * float sums: |err| <= 2e-7 * sum|x| (float32), 1e-15 * sum|x| (float64).
  * a row's result is bit-identical alone and inside any batch, under any
    tuning, and across graph replays.
"""

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd.batch import Batch
from xtensor_amd.workloads import uneven_shapes, uneven_values

from helpers import Dual, oracle_for, same_bits

pytestmark = pytest.mark.gpu


def sync():
    import torch
    torch.cuda.synchronize()


def softmax64(x):
    x = np.asarray(x, np.float64)
    if x.shape[1] == 0:
        return x.copy()
    z = np.exp(x - x.max(axis=1, keepdims=True))
    return z / z.sum(axis=1, keepdims=True)


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    if ref.size == 0:
        return 0.0
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)))


def make(shapes, dtype=np.float32, seed=5, op="softmax"):
    xs = uneven_values(shapes, seed=seed, dtype=np.float64)
    if np.dtype(dtype).kind == "i":
        rng = np.random.default_rng(seed)
        xs = [rng.integers(-2**31, 2**31, size=s, dtype=np.int64).astype(np.int32) for s in shapes]
    else:
        xs = [x.astype(dtype) for x in xs]
    dx = [xt.from_numpy(x) for x in xs]
    if op == "softmax":
        outs = [xt.zeros(x.shape, str(np.dtype(dtype))) for x in xs]
    else:
        outs = [xt.zeros((x.shape[0],), str(np.dtype(dtype))) for x in xs]
    return xs, dx, outs


SHAPES_SMALL = [(1, 1), (3, 2), (5, 4), (7, 5), (2, 8), (9, 9), (4, 16), (6, 17), (3, 32), (8, 33),
                (2, 64), (5, 65), (3, 100), (2, 127), (4, 128), (3, 129), (2, 500), (2, 1000),
                (3, 2048), (2, 2049), (2, 5000), (1, 1), (33, 3), (70, 1), (130, 7), (64, 64)]


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_softmax_uneven_vs_fp64(dtype):
    xs, dx, outs = make(SHAPES_SMALL, dtype)
    b = Batch("softmax", dx, outs)
    b.launch()
    sync()
    tol = 1e-6 if dtype == np.float32 else 1e-13
    worst = 0.0
    for x, o in zip(xs, outs):
        got = o.to_numpy()
        assert got.dtype == dtype
        worst = max(worst, rel_err(got, softmax64(x)))
    assert worst <= tol, worst


def test_softmax_f32_vs_oracle_xtensor_order():
    """The oracle evaluates xtensor's three-step expression in float32."""
    xs, dx, outs = make(SHAPES_SMALL[:18], np.float32)
    b = Batch("softmax", dx, outs)
    b.launch()
    sync()
    for x, o in zip(xs, outs):
        d = Dual.of(x)
        m = xt.amax(d.dev, [1], keep_dims=True)
        e = xt.exp(d.dev - m)
        ref = oracle_for(d).assign(e / xt.sum(e, [1], keep_dims=True))
        assert ref.dtype == np.float32
        assert rel_err(o.to_numpy(), ref) <= 4e-6


@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32])
def test_row_amax_exact(dtype):
    xs, dx, outs = make(SHAPES_SMALL, dtype, op="amax")
    Batch("amax", dx, outs).launch()
    sync()
    for x, o in zip(xs, outs):
        assert same_bits(o.to_numpy(), x.max(axis=1))


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_row_sum_float(dtype):
    xs, dx, outs = make(SHAPES_SMALL, dtype, op="sum")
    Batch("sum", dx, outs).launch()
    sync()
    k = 2e-7 if dtype == np.float32 else 1e-15
    for x, o in zip(xs, outs):
        x64 = x.astype(np.float64)
        ref = x64.sum(axis=1)
        bound = k * np.abs(x64).sum(axis=1) + 1e-300
        assert np.all(np.abs(o.to_numpy().astype(np.float64) - ref) <= bound)


def test_row_sum_int32_wraps():
    xs, dx, outs = make(SHAPES_SMALL, np.int32, op="sum")
    Batch("sum", dx, outs).launch()
    sync()
    for x, o in zip(xs, outs):
        ref = (x.astype(np.int64).sum(axis=1) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
        assert same_bits(o.to_numpy(), ref)


def test_empty_rows_and_columns():
    """sum over an empty axis is 0; amax over it is lowest() (the reducer's init)."""
    shapes = [(0, 5), (3, 0), (2, 3)]
    xs, dx, outs = make(shapes, np.float32, op="sum")
    Batch("sum", dx, outs).launch()
    xs2, dx2, outs2 = make(shapes, np.float32, op="amax")
    Batch("amax", dx2, outs2).launch()
    xs3, dx3, outs3 = make(shapes, np.float32, op="softmax")
    Batch("softmax", dx3, outs3).launch()
    sync()
    assert np.array_equal(outs[1].to_numpy(), np.zeros(3, np.float32))
    assert np.array_equal(outs2[1].to_numpy(), np.full(3, np.finfo(np.float32).min))
    assert rel_err(outs3[2].to_numpy(), softmax64(xs3[2])) <= 1e-6


def test_padded_rows_and_strided_outputs():
    rng = np.random.default_rng(11)
    base = rng.normal(0, 2, (40, 300)).astype(np.float32)
    d = xt.from_numpy(base)
    xv = d[xt.range_(0, 40, 2), xt.range_(0, 257)]         # row stride 600, 257 cols
    ob = xt.zeros((40, 300), "float32")
    ov = ob[xt.range_(0, 40, 2), xt.range_(0, 257)]
    Batch("softmax", [xv], [ov]).launch()
    sums = xt.zeros((40,), "float32")
    sv = sums[xt.range_(0, 40, 2)]                          # output stride 2
    Batch("sum", [xv], [sv]).launch()
    sync()
    ref = softmax64(base[0:40:2, :257])
    assert rel_err(ov.to_numpy(), ref) <= 1e-6
    full = ob.to_numpy()
    assert np.all(full[1::2] == 0) and np.all(full[:, 257:] == 0)   # nothing outside the views written
    s = sums.to_numpy()
    assert np.all(s[1::2] == 0)
    x64 = base[0:40:2, :257].astype(np.float64)
    assert np.all(np.abs(s[0::2] - x64.sum(1)) <= 2e-7 * np.abs(x64).sum(1))


def test_same_alone_and_in_the_mix():
    """Every instance gives bit-identical results alone, in a batch, and under any tuning."""
    shapes = uneven_shapes(300, seed=7, max_rows=64, max_cols=3000)
    xs, dx, outs = make(shapes, np.float32)
    b = Batch("softmax", dx, outs)
    b.launch()
    sync()
    mixed = [o.to_numpy() for o in outs]
    for t in ({"block": 64, "blocks_per_cu": 1, "task_elems": 64, "dyn_percent": 0},
              {"block": 256, "blocks_per_cu": 16, "task_elems": 32768, "dyn_percent": 100},
              {"block": 128, "blocks_per_cu": 3, "task_elems": 1000, "dyn_percent": 37}):
        outs2 = [xt.zeros(x.shape, "float32") for x in xs]
        b2 = Batch("softmax", dx, outs2)
        b2.tuning = t
        b2.launch()
        sync()
        for a, o in zip(mixed, outs2):
            assert same_bits(a, o.to_numpy())
    for k in (0, 17, 123, 299):
        alone = xt.zeros(xs[k].shape, "float32")
        Batch("softmax", [dx[k]], [alone]).launch()
        sync()
        assert same_bits(alone.to_numpy(), mixed[k])


def test_graph_replays_match_eager():
    """Replays of the instantiated graph, with eager work and new input
    values between them, each equal a direct launch on the same inputs."""
    import torch
    shapes = uneven_shapes(64, seed=9, max_rows=100, max_cols=600)
    xs, dx, outs = make(shapes, np.float32)
    ref_outs = [xt.zeros(x.shape, "float32") for x in xs]
    g = Batch("softmax", dx, outs)
    eager = Batch("softmax", dx, ref_outs)
    for rep in range(4):
        for d in dx:   # new values in place (eager torch work between replays)
            t = d.storage.view(torch.float32)
            t.mul_(0.5).add_(float(rep))
        g.execute()
        eager.launch()
        sync()
        for a, b in zip(outs, ref_outs):
            assert same_bits(a.to_numpy(), b.to_numpy())


def test_autotune_keeps_results():
    shapes = uneven_shapes(200, seed=3, max_rows=128, max_cols=2048)
    xs, dx, outs = make(shapes, np.float32)
    b = Batch("softmax", dx, outs)
    b.launch()
    sync()
    before = [o.to_numpy() for o in outs]
    t = b.autotune()
    assert t["block"] in (128, 256) and 0 <= t["dyn_percent"] <= 100
    b.execute()
    sync()
    for a, o in zip(before, outs):
        assert same_bits(a, o.to_numpy())
    assert b.info()["elements"] == sum(r * c for r, c in shapes)


def test_matches_plain_path():
    """Batch result vs the plain per-instance path (amax / fused exp-sum / divide)."""
    shapes = uneven_shapes(24, seed=12, max_rows=40, max_cols=900)
    xs, dx, outs = make(shapes, np.float32)
    Batch("softmax", dx, outs).launch()
    for x, d, o in zip(xs, dx, outs):
        m = xt.eval(xt.amax(d, [1], keep_dims=True))
        s = xt.eval(xt.sum(xt.exp(d - m), [1], keep_dims=True))
        p = xt.zeros(x.shape, "float32")
        xt.noalias(p).assign(xt.exp(d - m) / s)
        sync()
        assert rel_err(p.to_numpy(), o.to_numpy()) <= 2e-6


def test_type_and_shape_errors():
    x = xt.from_numpy(np.zeros((2, 3), np.int32))
    with pytest.raises(Exception, match="dtype"):
        Batch("softmax", [x], [xt.zeros((2, 3), "int32")])
    y = xt.from_numpy(np.zeros((2, 3), np.float32))
    with pytest.raises(ValueError):
        Batch("softmax", [y], [xt.zeros((2, 4), "float32")])
    with pytest.raises(ValueError):
        Batch("sum", [y], [xt.zeros((3,), "float32")])


def test_rows_far_apart_use_one_row_per_wave():
    """Rows 2^25 elements apart do not fit a row group's 32-bit offsets; the
    batch then gives each row a whole wave, with bit-identical results."""
    n = 1 << 25
    big = xt.zeros((n + 8,), "float64")
    rng = np.random.default_rng(21)
    vals = rng.normal(0, 2, (2, 4))
    import torch
    t = big.storage.view(torch.float64)
    t[0:4] = torch.from_numpy(vals[0])
    t[n:n + 4] = torch.from_numpy(vals[1])
    far = xt.DeviceArray(big.storage, "float64", (2, 4), (n, 1), 0)
    near = xt.from_numpy(vals)
    o_far = xt.zeros((2, 4), "float64")
    o_near = xt.zeros((2, 4), "float64")
    Batch("softmax", [far], [o_far]).launch()
    Batch("softmax", [near], [o_near]).launch()
    s_far = xt.zeros((2,), "float64")
    Batch("sum", [far], [s_far]).launch()
    sync()
    assert same_bits(o_far.to_numpy(), o_near.to_numpy())
    assert rel_err(o_far.to_numpy(), softmax64(vals)) <= 1e-13
    assert np.all(np.abs(s_far.to_numpy() - vals.sum(1)) <= 1e-15 * np.abs(vals).sum(1))


def test_full_uneven_batch_4096_vs_fp64():
    """The whole uneven batch the benchmark times: 4096 float32 instances,
    rows ~ log-uniform [1,512], cols ~ log-uniform [1,2048] (seed 3), x ~
    N(0,2), packed at 256-byte boundaries in one buffer per side.  Every
    element within rel 1e-6 of the float64 softmax of the same inputs, under
    the default tuning and the autotuned one (bit-identical to each other)."""
    import torch
    shapes = uneven_shapes(4096, seed=3)
    xs = uneven_values(shapes, seed=3)
    offs, o = [], 0
    for r, c in shapes:
        offs.append(o)
        o += (r * c + 63) // 64 * 64
    host = np.zeros(o, np.float32)
    for off, x in zip(offs, xs):
        host[off:off + x.size] = x.ravel()
    xin = xt.from_numpy(host)
    xout = xt.zeros((o,), "float32")
    ins = [xt.DeviceArray(xin.storage, "float32", s, None, off) for off, s in zip(offs, shapes)]
    outs = [xt.DeviceArray(xout.storage, "float32", s, None, off) for off, s in zip(offs, shapes)]
    b = Batch("softmax", ins, outs)
    b.launch()
    sync()
    first = xout.to_numpy()
    worst = 0.0
    for off, x in zip(offs, xs):
        worst = max(worst, rel_err(first[off:off + x.size].reshape(x.shape), softmax64(x)))
    assert worst <= 1e-6, worst
    xout.storage.view(torch.float32).zero_()
    b.autotune()
    b.execute()
    sync()
    assert same_bits(first, xout.to_numpy())
    assert b.info()["elements"] == sum(r * c for r, c in shapes)

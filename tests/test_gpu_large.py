"""Elementwise assignments past 2^31 elements (offsets that need 64 bits):
the work is cut along the first dimension into pieces with 32-bit offsets,
each run by a compiled kernel.  Checked on the device against the same
arithmetic done by torch (integer addition, exact), on every element."""

import pytest

import xtensor_amd as xt

pytestmark = pytest.mark.gpu


def wrap(t, shape):
    import torch
    return xt.DeviceArray(t.view(torch.uint8), str(t.dtype).replace("torch.", ""), shape)


def test_flat_int8_past_2_31():
    import torch
    n = (1 << 31) + (1 << 21) + 3  # three pieces, the last one still large enough to be compiled
    x = torch.randint(-100, 100, (n,), dtype=torch.int8, device="cuda")
    y = torch.empty_like(x)
    X, Y = wrap(x, (n,)), wrap(y, (n,))
    xt.noalias(Y).assign(X + 3)
    torch.cuda.synchronize()
    assert xt.last_kernel() != "dynamic"
    ref = (x.to(torch.int32) + 3).to(torch.int8)
    assert torch.equal(y, ref)
    # the same as a prepared plan, run alone and through the fused entry point
    pl = xt.plan(Y, X + 3)
    for run in (lambda: pl.run(), lambda: xt.execute_many([pl])):
        y.zero_()
        run()
        torch.cuda.synchronize()
        assert torch.equal(y, ref)
    del x, y, ref, pl
    torch.cuda.empty_cache()


def test_rows_int16_with_broadcast_row_past_2_31():
    import torch
    r, c = 3, (1 << 30) + 5
    x = torch.randint(-1000, 1000, (r, c), dtype=torch.int16, device="cuda")
    b = torch.randint(-1000, 1000, (c,), dtype=torch.int16, device="cuda")
    y = torch.empty_like(x)
    X, B, Y = wrap(x, (r, c)), wrap(b, (c,)), wrap(y, (r, c))
    xt.noalias(Y).assign(X - B)
    torch.cuda.synchronize()
    assert xt.last_kernel() != "dynamic"
    ref = (x.to(torch.int32) - b.to(torch.int32))
    got = y.to(torch.int32)
    assert torch.equal(got, ref) or torch.equal(y, ref.to(torch.int16))
    del x, b, y, ref, got
    torch.cuda.empty_cache()


def test_reductions_over_a_dimension_past_2_31():
    """sum / amax over a reduced dimension of more than 2^31 elements: pieces
    of 2^30 reduce into a temporary, and a second reduction merges them.
    Integer sums into int64 and maxima are exact against torch; a float32
    sum of uniform(0, 1) data is within 2e-7 of torch's float64 sum."""
    import torch
    n = (1 << 31) + (1 << 21) + 3
    x = torch.randint(-100, 100, (n,), dtype=torch.int8, device="cuda")
    X = wrap(x, (n,))
    got = xt.eval(xt.sum(X, None, dtype="int64")).to_numpy()
    assert int(got) == int(x.sum(dtype=torch.int64).item())
    got = xt.eval(xt.amax(X)).to_numpy()
    assert int(got) == int(x.max().item())
    # into an output narrower than the int64 result: the pieces stay int64, the store casts
    o32 = xt.empty((), "int32")
    xt.noalias(o32).assign(xt.sum(X, None, dtype="int64"))
    assert int(o32.to_numpy()) == int(x.sum(dtype=torch.int64).item())
    del x, X
    c = (1 << 31) + 5
    x2 = torch.randint(-100, 100, (2, c), dtype=torch.int8, device="cuda")
    got = xt.eval(xt.sum(wrap(x2, (2, c)), [1], dtype="int64")).to_numpy()
    assert got.tolist() == x2.sum(1, dtype=torch.int64).tolist()
    del x2
    m = (1 << 31) + 7
    f = torch.rand(m, dtype=torch.float32, device="cuda")
    got = float(xt.eval(xt.sum(wrap(f, (m,)), None)).to_numpy())
    ref = float(f.double().sum().item())
    assert abs(got - ref) / ref <= 2e-7, (got, ref)
    del f
    torch.cuda.empty_cache()


def test_reductions_in_pieces_merge_every_kind():
    """prod / amin / mean / norms through the merge of the pieces, xt::initial
    applied once, keep_dims outputs, and reduced dimensions that only pass
    2^31 elements once merged: exact against torch's integer arithmetic."""
    import torch
    n = (1 << 31) + (1 << 21) + 3
    x = torch.randint(-100, 100, (n,), dtype=torch.int8, device="cuda")
    X = wrap(x, (n,))
    total = int(x.sum(dtype=torch.int64).item())
    assert int(xt.eval(xt.amin(X)).to_numpy()) == int(x.min().item())
    assert int(xt.eval(xt.sum(X, None, initial=5, dtype="int64")).to_numpy()) == total + 5
    # mean<void>(int8): the sum in int, divided in double by the whole count
    assert float(xt.eval(xt.mean(X)).to_numpy()) == total / n
    assert int(xt.eval(xt.norm_l1(X)).to_numpy()) == int(x.to(torch.int64).abs().sum().item())
    assert int(xt.eval(xt.norm_l0(X)).to_numpy()) == int((x != 0).sum().item())
    assert int(xt.eval(xt.norm_linf(X)).to_numpy()) == int(x.to(torch.int16).abs().max().item())
    sq = int((x.to(torch.int64) ** 2).sum().item())
    assert int(xt.eval(xt.norm_sq(X)).to_numpy()) == sq
    del x, X
    # prod of +-1 values: the sign of the count of -1s, exact
    s = torch.randint(0, 2, (n,), dtype=torch.int8, device="cuda") * 2 - 1
    S = wrap(s, (n,))
    neg = int((s < 0).sum().item())
    assert int(xt.eval(xt.prod(S, None, dtype="int64")).to_numpy()) == (-1 if neg % 2 else 1)
    assert int(xt.eval(xt.prod(S, None, initial=3, dtype="int64")).to_numpy()) == (-3 if neg % 2 else 3)
    del s, S
    # keep_dims over the long axis of [2, 2^31 + 5]
    c = (1 << 31) + 5
    x2 = torch.randint(-100, 100, (2, c), dtype=torch.int8, device="cuda")
    got = xt.eval(xt.sum(wrap(x2, (2, c)), [1], keep_dims=True, dtype="int64")).to_numpy()
    assert got.shape == (2, 1)
    assert got[:, 0].tolist() == x2.sum(1, dtype=torch.int64).tolist()
    got = xt.eval(xt.amin(wrap(x2, (2, c)), [1], keep_dims=True)).to_numpy()
    assert got[:, 0].tolist() == x2.min(1).values.tolist()
    del x2
    # full reductions whose dimensions merge past 2^31: [2^31 + 5, 2] and [2, 2^30]
    for shape in (((1 << 31) + 5, 2), (2, 1 << 30)):
        x3 = torch.randint(-100, 100, shape, dtype=torch.int8, device="cuda")
        got = int(xt.eval(xt.sum(wrap(x3, shape), None, dtype="int64")).to_numpy())
        assert got == int(x3.sum(dtype=torch.int64).item()), shape
        del x3
    torch.cuda.empty_cache()


def test_float_moments_and_norm_in_pieces():
    """mean, norm_l2 and variance / stddev of 2^31 + 7 float32 values: the
    divisor is the whole count, the square root follows the merge; within
    stated bounds of torch's float64 evaluation."""
    import torch
    m = (1 << 31) + 7
    f = torch.rand(m, dtype=torch.float32, device="cuda")
    F = wrap(f, (m,))
    fd = f.double()
    mean_ref = float(fd.mean().item())
    assert abs(float(xt.eval(xt.mean(F)).to_numpy()) - mean_ref) / mean_ref <= 2e-7
    l2_ref = float(fd.norm().item())
    assert abs(float(xt.eval(xt.norm_l2(F)).to_numpy()) - l2_ref) / l2_ref <= 2e-7
    var_ref = float(fd.var(unbiased=False).item())
    got = float(xt.eval(xt.variance(F)).to_numpy())
    assert abs(got - var_ref) / var_ref <= 1e-6, (got, var_ref)
    got = float(xt.eval(xt.stddev(F)).to_numpy())
    assert abs(got - var_ref ** 0.5) / var_ref ** 0.5 <= 1e-6
    del f, F, fd
    torch.cuda.empty_cache()


def test_scans_past_2_31():
    """cumsum / cumprod along an axis of more than 2^31 elements, and the
    flattened cumsum of such an array: pieces of 2^30 are scanned in turn and
    from the second on the previous piece's last value is folded in.  Inputs
    with closed-form scans (steps of 1 at every third element, signs flipping
    every 2^20 elements), checked on every element with torch arithmetic."""
    import torch

    def step(msg):
        print("  step:", msg, flush=True)

    n = (1 << 31) + (1 << 21) + 3
    i = torch.arange(n, device="cuda", dtype=torch.int64)
    x = (i % 3 == 0).to(torch.int32)
    ref = (i // 3 + 1).to(torch.int32)  # cumsum: the count of multiples of 3 up to i
    X = wrap(x, (n,))
    for axis in (0, None):
        step(f"cumsum int32 over {n} elements, axis {axis}")
        got = xt.cumsum(X, axis)
        assert got.shape == (n,) and got.dtype == "int32"
        assert torch.equal(got.storage.view(torch.int32), ref), axis
        del got
    del ref, x, X
    step("cumprod of signs")
    s = torch.where(i % (1 << 20) == 0, -1, 1).to(torch.int32)
    ref = torch.where((i >> 20) % 2 == 0, -1, 1).to(torch.int32)  # one flip at each multiple of 2^20
    got = xt.cumprod(wrap(s, (n,)), 0)
    assert torch.equal(got.storage.view(torch.int32), ref)
    del got, s, ref, i
    torch.cuda.empty_cache()
    step("cumsum along axis 1 of [2, 2^31 + 8]")
    c = (1 << 31) + 8  # rows of whole 16-byte vectors, 64-bit row offsets
    j = torch.arange(c, device="cuda", dtype=torch.int64)
    x2 = torch.stack([(j % 3 == 0), (j % 5 == 0)]).to(torch.int32)
    got = xt.cumsum(wrap(x2, (2, c)), 1)
    g = got.storage.view(torch.int32).view(2, c)
    assert torch.equal(g[0], (j // 3 + 1).to(torch.int32))
    assert torch.equal(g[1], (j // 5 + 1).to(torch.int32))
    del got, g, x2, j
    torch.cuda.empty_cache()


def test_prepared_plans_past_2_31():
    """Prepared reduction and scan plans over a dimension past 2^31 elements
    keep their descriptors and run as the direct calls (pieces) on each
    execution: the same results as xg_reduce / xg_scan, twice in a row."""
    import torch
    n = (1 << 31) + (1 << 21) + 3
    x = torch.randint(-100, 100, (n,), dtype=torch.int8, device="cuda")
    X = wrap(x, (n,))
    s = xt.empty((), "int64")
    pl = xt.plan(s, xt.sum(X, None, dtype="int64"))
    ref = int(x.sum(dtype=torch.int64).item())
    for _ in range(2):
        s.storage.zero_()
        pl.run()
        torch.cuda.synchronize()
        assert int(s.to_numpy()) == ref
    del pl, s
    i = torch.arange(n, device="cuda", dtype=torch.int64)
    y = (i % 3 == 0).to(torch.int32)
    del x, X
    out = xt.empty((n,), "int32")
    ps = xt.plan(out, xt.Scan("cumsum", wrap(y, (n,)), 0))
    ps.run()
    torch.cuda.synchronize()
    assert torch.equal(out.storage.view(torch.int32), (i // 3 + 1).to(torch.int32))
    del ps, out, y, i
    torch.cuda.empty_cache()

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

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

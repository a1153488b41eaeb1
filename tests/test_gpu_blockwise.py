"""Blockwise reducers (reducers/xblockwise_reducer.hpp) against the
single-pass reducers, replaying test/test_xblockwise_reducer.cpp:103-174:
input int [21, 10, 5] = iota from -5 with +10 on odd flat positions, chunks
[5, 4, 2], six axis sets; results equal (allclose for variance / stddev) and
of the same value type.  A host memmap streams through the device chunk by
chunk."""

import numpy as np
import pytest

import xtensor_amd as xt

pytestmark = pytest.mark.gpu


def sync():
    import torch
    torch.cuda.synchronize()


def kat_input():
    a = np.arange(21 * 10 * 5, dtype=np.int64) - 5
    a[1::2] += 10
    return a.astype(np.int32).reshape(21, 10, 5)


AXES = [[0], [1], [2], [0, 1], [0, 2], [0, 1, 2]]
EXACT = ["sum", "prod", "amin", "amax", "mean", "norm_l0", "norm_l1", "norm_sq", "norm_linf"]


@pytest.mark.parametrize("keep", [False, True])
@pytest.mark.parametrize("kind", EXACT + ["norm_l2", "variance", "stddev"])
def test_matches_single_pass(kind, keep):
    h = kat_input()
    d = xt.from_numpy(h)
    for axes in AXES:
        for src in (h, d):  # host input (streamed) and device input
            got = getattr(xt.blockwise, kind)(src, (5, 4, 2), axes, keep_dims=keep)
            ref = xt.eval(xt.Reducer(kind, d, axes, keep))
            sync()
            g, r = got.to_numpy(), ref.to_numpy()
            assert got.dtype == ref.dtype and g.shape == r.shape, (kind, axes)
            if kind in ("variance", "stddev", "norm_l2"):
                np.testing.assert_allclose(g, r, rtol=1e-12)
            else:
                assert np.array_equal(g, r), (kind, axes)


def test_memmap_float_stream(tmp_path):
    rng = np.random.default_rng(0)
    h = rng.normal(1, 1, (4096, 1000)).astype(np.float32)
    p = tmp_path / "big.f32"
    h.tofile(p)
    mm = np.memmap(p, dtype=np.float32, mode="r", shape=h.shape)
    s0 = xt.blockwise.sum(mm, (1000, 1000), [0])
    m1 = xt.blockwise.mean(mm, (700, 256), [1])
    v = xt.blockwise.variance(mm, (1024, 1000), None, ddof=1)
    mx = xt.blockwise.amax(mm, (333, 333))
    sync()
    h64 = h.astype(np.float64)
    np.testing.assert_allclose(s0.to_numpy(), h64.sum(0), rtol=2e-7)
    # mean<void> of float32 is a float32 sum (rounded once) divided in double
    np.testing.assert_allclose(m1.to_numpy(), h64.mean(1), rtol=2e-7)
    np.testing.assert_allclose(v.to_numpy()[()], h64.var(ddof=1), rtol=1e-10)
    assert mx.to_numpy()[()] == h.max()

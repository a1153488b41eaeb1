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


def test_against_the_oracle():
    """Blockwise reductions against the oracle's restatement of the whole
    reduction (reducers/xreducer.hpp fold order), not against the library's
    own single pass: integers and extrema exactly (int32 KAT input, six axis
    sets), float32 sums / means / norms of a host array streamed in chunks
    within 2e-7 of the oracle's long-double reference."""
    from helpers import oracle_for, Dual
    ints = Dual.of(kat_input())
    oi = oracle_for(ints)
    for kind in ("sum", "prod", "amax", "amin", "norm_l1", "norm_linf", "norm_sq"):
        for axes in AXES:
            got = getattr(xt.blockwise, kind)(ints.host, (5, 4, 2), axes).to_numpy()
            ref = oi.reduce(xt.Reducer(kind, ints.dev, axes))
            assert got.dtype == ref.dtype and np.array_equal(got, ref), (kind, axes)
    rng = np.random.default_rng(7)
    f = Dual.of(rng.normal(1.0, 1.0, (600, 500)).astype(np.float32))
    of = oracle_for(f)
    for kind in ("sum", "mean", "norm_l2", "amax", "amin"):
        for axes in ([0], [1], None):
            got = getattr(xt.blockwise, kind)(f.host, (128, 96), axes).to_numpy()
            r = xt.Reducer(kind, f.dev, axes)
            if kind in ("amax", "amin"):
                assert np.array_equal(got, of.reduce(r)), (kind, axes)
            else:
                ref = of.reduce(r, high_precision=True)
                err = np.max(np.abs(got.astype(np.float64) - ref) / np.abs(ref))
                assert err <= 2e-7, (kind, axes, err)


def test_stream_overlap_many_chunks():
    """A host array in many chunks (the two upload buffers are reused many
    times while reductions are in flight): every chunk's data arrives intact
    (exact integer sums against numpy)."""
    rng = np.random.default_rng(8)
    h = rng.integers(-1000, 1000, (4096, 777)).astype(np.int32)
    got = xt.blockwise.sum(h, (100, 100), [0], dtype="int64").to_numpy()
    assert np.array_equal(got, h.astype(np.int64).sum(0))
    got1 = xt.blockwise.sum(h, (97, 777), [1], dtype="int64").to_numpy()
    assert np.array_equal(got1, h.astype(np.int64).sum(1))

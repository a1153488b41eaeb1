"""A sequence of plans replayed as one HIP graph (xt.PlanGraph) against the
same plans run eagerly: the config-1 expression, its row sum and column
amax, plus a scan.  Replays run several times with new input values written
by eager work between them, and every replay is compared bit for bit."""

import numpy as np
import pytest

import xtensor_amd as xt

pytestmark = pytest.mark.gpu


def same_bits(a, b):
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_plan_graph_replays_match_eager():
    import torch
    rng = np.random.default_rng(0)
    a = xt.from_numpy(rng.uniform(-1, 1, (2048, 1024)))
    b = xt.from_numpy(rng.uniform(-1, 1, (1024,)))

    def outputs():
        return (xt.empty((2048, 1024), "float64"), xt.empty((2048,), "float64"), xt.empty((1024,), "float64"),
                xt.empty((2048, 1024), "float64"))

    def plans(c, s, m, cs):
        return [xt.plan(c, a * b + xt.exp(-a * a)), xt.plan(s, xt.sum(c, [1])), xt.plan(m, xt.amax(c, [0])),
                xt.plan(cs, xt.Scan("cumsum", c, 1))]

    g_out, e_out = outputs(), outputs()
    g = xt.PlanGraph(plans(*g_out))
    eager = plans(*e_out)
    ta = a.storage.view(torch.float64)
    for rep in range(4):
        ta.mul_(0.5).add_(0.125 * rep)   # new values in place, eager torch work between replays
        g.replay()
        for p in eager:
            p.run()
        torch.cuda.synchronize()
        for x, y in zip(g_out, e_out):
            assert same_bits(x.to_numpy(), y.to_numpy()), rep


def test_plan_graph_one_pass_and_chunked_scans():
    """The scans whose launches clear their own counters and flags first: the
    one-pass strided-axis scan (ticket counters, look-back publications;
    int32, exact) and the two-launch scan of a long row (chunk totals, then
    the chunks from their carries; float32, fixed merge order), captured in
    one graph and replayed four times with new inputs written between the
    replays, every replay bit for bit against the same plans run eagerly."""
    import torch
    rng = np.random.default_rng(1)
    xi = xt.from_numpy(rng.integers(-1000, 1000, (3000, 2560)).astype(np.int32))
    xf = xt.from_numpy(rng.normal(1.0, 1.0, (1 << 22,)).astype(np.float32))

    def outputs():
        return xt.empty((3000, 2560), "int32"), xt.empty((1 << 22,), "float32")

    def plans(oi, of):
        return [xt.plan(oi, xt.Scan("cumsum", xi, 0)), xt.plan(of, xt.Scan("cumsum", xf, 0))]

    g_out, e_out = outputs(), outputs()
    g = xt.PlanGraph(plans(*g_out))
    eager = plans(*e_out)
    ti, tf = xi.storage.view(torch.int32), xf.storage.view(torch.float32)
    for rep in range(4):
        ti.add_(rep - 1)
        tf.mul_(0.75).add_(0.25 * rep)
        g.replay()
        for p in eager:
            p.run()
        torch.cuda.synchronize()
        assert xt.last_kernel() == "scan_chunks3"
        for x, y in zip(g_out, e_out):
            assert same_bits(x.to_numpy(), y.to_numpy()), rep
    assert np.array_equal(g_out[0].to_numpy(), np.cumsum(xi.to_numpy(), axis=0, dtype=np.int32))


def test_plan_graph_needs_plans():
    with pytest.raises(ValueError):
        xt.PlanGraph([])

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


def test_plan_graph_needs_plans():
    with pytest.raises(ValueError):
        xt.PlanGraph([])

"""GPU tests of the fused many-plan launch and the request server.

Bars (same as the single-expression paths):
  * affine_clip and bitmix: bit-exact against the oracle / numpy, and
    bit-identical whether a request runs alone, fused with others, or on the
    plain per-request path (they are exact functions of each element).
  * row_sum: |err| <= 2e-7 * sum|x| (float32), 1e-15 * sum|x| (float64), int32
    bit-exact with wrap-around.
  * row_softmax: rel <= 1e-6 (float32) / 1e-13 (float64) against float64.
  * A request's result is bit-identical alone and in any mix (row reductions
    are ordered by row length only).
"""

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd.serve import Server

from helpers import Dual, oracle_for, same_bits

pytestmark = pytest.mark.gpu


def sync():
    import torch
    torch.cuda.synchronize()


def rel_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    if ref.size == 0:
        return 0.0
    return float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)))


def softmax64(x):
    x = np.asarray(x, np.float64)
    z = np.exp(x - x.max(axis=1, keepdims=True))
    return z / z.sum(axis=1, keepdims=True)


def affine_ref(x, mean, var, gamma, beta, dtype):
    """clip((x - mean) / sqrt(var + 1e-5) * gamma + beta, -3, 3) with xtensor's
    types: x - mean in the input type, the rest in double, stored narrowed."""
    C = mean.shape[0]
    sh = (1, C, 1, 1)
    d = (x - mean.reshape(sh)).astype(dtype)
    y = d.astype(np.float64) / np.sqrt(var.reshape(sh).astype(np.float64) + 1e-5)
    y = y * gamma.reshape(sh).astype(np.float64) + beta.reshape(sh).astype(np.float64)
    return np.clip(y, -3.0, 3.0).astype(dtype)


# ---------------------------------------------------------------- fused plans
def test_execute_many_matches_single_plans():
    rng = np.random.default_rng(40)
    C = 16
    params = [xt.from_numpy(rng.uniform(0.5, 1.5, (1, C, 1, 1)).astype(np.float32)) for _ in range(4)]
    singles, fused, outs_a, outs_b = [], [], [], []
    for n in (1, 2, 3, 7, 1, 5):
        x = xt.from_numpy(rng.normal(0, 2, (n, C, 8, 8)).astype(np.float32))
        ya, yb = xt.zeros(x.shape, "float32"), xt.zeros(x.shape, "float32")
        e = lambda xx: xt.clip((xx - params[0]) / xt.sqrt(params[1] + 1e-5) * params[2] + params[3], -3, 3)  # noqa
        singles.append(xt.plan(ya, e(x)))
        fused.append(xt.plan(yb, e(x)))
        outs_a.append(ya)
        outs_b.append(yb)
    for n in (100, 4097, 64):
        xi = xt.from_numpy(rng.integers(-2**20, 2**20, n, dtype=np.int32))
        ya, yb = xt.zeros((n,), "int32"), xt.zeros((n,), "int32")
        singles.append(xt.plan(ya, (xi * 3 + (xi >> 2)) & 0xFFFF))
        fused.append(xt.plan(yb, (xi * 3 + (xi >> 2)) & 0xFFFF))
        outs_a.append(ya)
        outs_b.append(yb)
    # a plan with no compiled program, and a reduction, in the same list
    a = xt.from_numpy(rng.normal(0, 1, (33, 17)))
    ya, yb = xt.zeros((33, 17), "float64"), xt.zeros((33, 17), "float64")
    singles.append(xt.plan(ya, xt.sin(a) * 2 + a))
    fused.append(xt.plan(yb, xt.sin(a) * 2 + a))
    outs_a.append(ya)
    outs_b.append(yb)
    sa, sb = xt.zeros((33,), "float64"), xt.zeros((33,), "float64")
    singles.append(xt.plan(sa, xt.sum(a, [1])))
    fused.append(xt.plan(sb, xt.sum(a, [1])))
    outs_a.append(sa)
    outs_b.append(sb)
    for p in singles:
        p.run()
    xt.execute_many(fused)
    assert xt.last_kernel() == "multi"
    sync()
    for u, v in zip(outs_a, outs_b):
        assert same_bits(u.to_numpy(), v.to_numpy())


# ---------------------------------------------------------------- server
def make_requests(rng, count, max_elems=200_000):
    reqs = []
    for k in range(count):
        tpl = ["affine_clip", "row_sum", "row_softmax", "bitmix"][k % 4]
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(max_elems))))
        if tpl == "affine_clip":
            dt = [np.float32, np.float64][rng.integers(2)]
            N = max(1, n // 1024)
            x = rng.normal(0, 2, (N, 16, 8, 8)).astype(dt)
            ps = [rng.normal(0, 0.1, 16), rng.uniform(0.5, 2, 16), rng.uniform(0.5, 1.5, 16), rng.normal(0, 0.1, 16)]
            ps = [p.astype(dt) for p in ps]
            reqs.append(dict(tpl=tpl, x=x, params=ps, out_shape=x.shape, dt=dt))
        elif tpl == "bitmix":
            x = rng.integers(-2**20, 2**20, n, dtype=np.int32)
            reqs.append(dict(tpl=tpl, x=x, params=[], out_shape=x.shape, dt=np.int32))
        else:
            dts = [np.float32, np.float64, np.int32] if tpl == "row_sum" else [np.float32, np.float64]
            dt = dts[rng.integers(len(dts))]
            cols = int(min(n, [1000, 2048][rng.integers(2)]))
            rows = max(1, n // cols)
            if dt == np.int32:
                x = rng.integers(-2**31, 2**31, (rows, cols), dtype=np.int64).astype(np.int32)
            else:
                x = rng.normal(0, 2, (rows, cols)).astype(dt)
            reqs.append(dict(tpl=tpl, x=x, params=[], out_shape=(rows,) if tpl == "row_sum" else x.shape, dt=dt))
    return reqs


def run_server(reqs, held=False, **cfg):
    """Run the requests through one server.  With `held`, launches are held
    back until every request is queued, so the fused grouping does not
    depend on submission timing."""
    for r in reqs:
        r.setdefault("dx", xt.from_numpy(r["x"]))
        r.setdefault("dps", [xt.from_numpy(p) for p in r["params"]])
    outs = [xt.zeros(r["out_shape"], str(np.dtype(r["dt"]))) for r in reqs]
    sync()
    with Server(**cfg) as srv:
        if held:
            srv.hold(True)
        for k, (r, o) in enumerate(zip(reqs, outs)):
            assert srv.submit(r["tpl"], r["dx"], o, r["dps"], tag=k) == 0
        if held:
            srv.hold(False)
        srv.drain(60_000)
        comps = srv.poll(len(reqs) + 10)
        stats = srv.stats()
    sync()
    return [o.to_numpy() for o in outs], comps, stats


def check_reference(r, got):
    x = r["x"]
    if r["tpl"] == "affine_clip":
        assert same_bits(got, affine_ref(x, *r["params"], r["dt"]))
    elif r["tpl"] == "bitmix":
        ref = ((x.astype(np.int64) * 3 + (x.astype(np.int64) >> 2)) & 0xFFFF).astype(np.int32)
        assert same_bits(got, ref)
    elif r["tpl"] == "row_sum":
        if r["dt"] == np.int32:
            ref = (x.astype(np.int64).sum(axis=1) & 0xFFFFFFFF).astype(np.uint32).view(np.int32)
            assert same_bits(got, ref)
        else:
            x64 = x.astype(np.float64)
            k = 2e-7 if r["dt"] == np.float32 else 1e-15
            assert np.all(np.abs(got.astype(np.float64) - x64.sum(1)) <= k * np.abs(x64).sum(1) + 1e-300)
    else:
        assert rel_err(got, softmax64(x)) <= (1e-6 if r["dt"] == np.float32 else 1e-13)


def test_server_results_alone_in_mix_and_plain():
    rng = np.random.default_rng(41)
    reqs = make_requests(rng, 48)
    mixed, comps, stats = run_server(reqs, held=True, streams=4, max_batch=16)
    assert stats["completed"] == len(reqs) and stats["rejected"] == 0
    assert len(comps) == len(reqs) and all(c["status"] == 0 for c in comps)
    assert max(c["batch_size"] for c in comps) > 1   # something was fused
    assert stats["launches"] < len(reqs)
    for c in comps:
        assert c["arrival_ns"] <= c["admit_ns"] <= c["launch_ns"] <= c["done_ns"]
    for r, got in zip(reqs, mixed):
        check_reference(r, got)
    # a different mix: no hold, other stream and batch limits
    mixed2, _, _ = run_server(reqs, streams=2, max_batch=64)
    for a, b in zip(mixed, mixed2):
        assert same_bits(a, b)
    # each request alone
    for k in range(0, len(reqs), 5):
        alone, _, _ = run_server([reqs[k]], streams=1, max_batch=1)
        assert same_bits(alone[0], mixed[k])
    # the plain per-request path (xtensor's sequence of assign / reduce calls)
    plain, comps_p, stats_p = run_server(reqs, streams=1, max_batch=1, plain=True)
    assert all(c["status"] == 0 and c["batch_size"] == 1 for c in comps_p)
    for r, a, b in zip(reqs, mixed, plain):
        if r["tpl"] in ("affine_clip", "bitmix") or r["dt"] == np.int32:
            assert same_bits(a, b)
        elif r["tpl"] == "row_softmax":
            assert rel_err(a, b) <= (2e-6 if r["dt"] == np.float32 else 1e-13)
        check_reference(r, b)


def test_oracle_parity_small_affine():
    """A small affine_clip request against the oracle's evaluation of the expression."""
    rng = np.random.default_rng(42)
    x = rng.normal(0, 2, (2, 16, 8, 8)).astype(np.float32)
    ps = [rng.normal(0, 0.1, 16), rng.uniform(0.5, 2, 16), rng.uniform(0.5, 1.5, 16), rng.normal(0, 0.1, 16)]
    ps = [p.astype(np.float32) for p in ps]
    r = dict(tpl="affine_clip", x=x, params=ps, out_shape=x.shape, dt=np.float32)
    got, _, _ = run_server([r], streams=2, max_batch=4)
    d = Dual.of(x)
    pd = [Dual.of(p.reshape(1, 16, 1, 1)) for p in ps]
    e = xt.clip((d.dev - pd[0].dev) / xt.sqrt(pd[1].dev + 1e-5) * pd[2].dev + pd[3].dev, -3, 3)
    ref = oracle_for(d, *pd).assign(e, dtype="float32")
    assert same_bits(got[0], ref)


def test_budget_admission_and_queue_limits():
    import torch
    rng = np.random.default_rng(43)
    n = 1 << 24  # 16M float32: a few milliseconds of device work per request
    x = xt.from_numpy(rng.normal(0, 1, (n // 1024, 1024)).astype(np.float32))
    req_bytes = 2 * n * 4
    outs = [xt.zeros(x.shape, "float32") for _ in range(6)]
    with Server(budget_bytes=req_bytes, streams=4, max_batch=8, queue_capacity=2) as srv:
        with pytest.raises(MemoryError):
            big = xt.zeros((n // 512, 1024), "float32")
            srv.submit("row_softmax", xt.DeviceArray(big.storage, "float32", big.shape), big, tag=99)
        codes = [srv.submit("row_softmax", x, o, tag=k) for k, o in enumerate(outs)]
        srv.drain(60_000)
        st = srv.stats()
        comps = srv.poll()
    torch.cuda.synchronize()
    assert codes[0] == 0 and 8 in codes          # later submissions found the queue full
    assert st["peak_inflight_bytes"] <= req_bytes  # never more than the budget admitted
    ok = [k for k, c in enumerate(codes) if c == 0]
    assert sorted(c["tag"] for c in comps) == ok
    ref = softmax64(x.to_numpy()[:64])
    for k in ok:
        assert rel_err(outs[k].to_numpy()[:64], ref) <= 1e-6


def test_bad_requests_rejected():
    with Server(budget_bytes=1 << 30, streams=1, max_batch=1) as srv:
        xi = xt.zeros((4, 4), "int32")
        with pytest.raises(TypeError):
            srv.submit("row_softmax", xi, xt.zeros((4, 4), "int32"))
        xf = xt.zeros((8,), "float32")
        with pytest.raises(TypeError):
            srv.submit("bitmix", xf, xt.zeros((8,), "float32"))
        assert srv.stats()["rejected"] == 2


@pytest.mark.parametrize("plain", [False, True])
def test_expression_and_reduction_requests(plain):
    """Requests of any expression ("expr": noalias(out) = e, strided views and
    broadcasts included) and any reduction ("reduce": any axes, variance
    too), mixed with each other: each result is bit-identical to the same
    expression evaluated alone through the single-call entry points, on the
    fused server and on the plain per-request path."""
    rng = np.random.default_rng(44)
    a = xt.from_numpy(rng.normal(0, 1, (96, 130)).astype(np.float32))
    b = xt.from_numpy(rng.normal(0, 1, (130,)).astype(np.float32))
    big = xt.from_numpy(rng.normal(1, 1, (257, 512)))
    ints = xt.from_numpy(rng.integers(-2**20, 2**20, (3000,), dtype=np.int32))
    exprs = [
        lambda: a * b + 1.5,
        lambda: xt.clip(a.T * 2.0, -1, 1),
        lambda: (ints * 3 + (ints >> 2)) & 0xFFFF,
        lambda: xt.exp(-big * big),
        lambda: big[:, xt.range_(0, 512, 2)] * 2.0,
        lambda: ints * 7 - 1,
        lambda: a * b + 1.5,  # the same program again (fused with the first)
    ]
    reds = [
        lambda: xt.sum(big, [0]),
        lambda: xt.amax(a.T, [1]),
        lambda: xt.variance(big, [1]),
        lambda: xt.sum(ints, None, dtype="int64"),
        lambda: xt.mean(a * b, [0, 1]),
    ]
    reqs = []
    for k in range(3):
        for f in exprs:
            reqs.append(("expr", f))
        for f in reds:
            reqs.append(("reduce", f))
    refs, outs = [], []
    for kind, f in reqs:
        e = f()
        ref = xt.eval(e)
        refs.append(ref)
        outs.append(xt.empty(ref.shape, ref.dtype))
    sync()
    cfg = dict(budget_bytes=1 << 30, streams=4, max_batch=8)
    if plain:
        cfg.update(streams=1, max_batch=1, plain=True)
    with Server(**cfg) as srv:
        srv.hold(True)
        for k, (kind, f) in enumerate(reqs):
            assert srv.submit_expr(outs[k], f(), tag=k) == 0
        srv.hold(False)
        srv.drain(60_000)
        comps = srv.poll(1000)
        st = srv.stats()
    assert len(comps) == len(reqs) and all(c["status"] == 0 for c in comps), comps
    if not plain:
        assert st["launches"] < len(reqs)  # the programs repeated across requests were fused
    sync()
    for k, (o, r) in enumerate(zip(outs, refs)):
        assert same_bits(o.to_numpy(), r.to_numpy()), (k, reqs[k][0])


def test_expression_requests_rejected_or_failed():
    """A malformed request is refused at submission; one whose output does not
    match completes with an error status and leaves the server running."""
    a = xt.from_numpy(np.ones((8, 8), np.float32))
    with Server(budget_bytes=1 << 30, streams=2, max_batch=4) as srv:
        with pytest.raises(Exception):
            srv.submit_expr(xt.empty((8, 8), "float32"), a + xt.from_numpy(np.ones((3,), np.float32)), tag=1)
        assert srv.submit_expr(xt.empty((5,), "float32"), xt.sum(a, [0]), tag=2) == 0
        good = xt.empty((8, 8), "float32")
        assert srv.submit_expr(good, a * 2, tag=3) == 0
        srv.drain(30_000)
        comps = {c["tag"]: c["status"] for c in srv.poll(10)}
    assert comps[2] != 0 and comps[3] == 0
    sync()
    assert np.array_equal(good.to_numpy(), np.full((8, 8), 2, np.float32))

#!/usr/bin/env python3
"""Benchmark of the fused broadcast-expression path on one MI355X.

Workload (one "step" = one launch of the plain one-kernel-per-expression
path over the full batch):
    x [64, 256, 56, 56] float32 ~ N(0, 1) (seed 1), per-channel mean / var /
    gamma / beta [1, 256, 1, 1];
    y[:, :, 0:56:2, :] = clip((x[:, :, 0:56:2, :] - mean) / sqrt(var + 1e-5)
                              * gamma + beta, -3, 3)
Inputs are resident in HBM; three rotating input/output sets (1.2 GB) keep
the working set above the 256 MB last-level (MALL) cache.

Reported: whole-job GB/s of algorithmic bytes (touched x elements read once,
per-channel parameters once, touched y elements written once), the kernel's
own GB/s from HIP events on the launch stream against the 8 TB/s HBM peak,
and the CPU baseline: the same workload restated in typed C++ (the scalar
single-threaded stepper loop xtensor runs for it), timed on this host.
"""

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic.json")


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by the
    factor a plain copy of known size gives), or None if not measured."""
    try:
        with open(TRAFFIC_FILE) as f:
            entry = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    return None if entry is None else int(entry["bytes_per_launch"])
N, C, H, W = 64, 256, 56, 56


def algorithmic_bytes(n=N):
    touched = n * C * (H // 2) * W
    return touched * 4 * 2 + 4 * C * 4  # x read + y written (touched elements), 4 parameter vectors


def make_inputs(np, n=N, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((n, C, H, W), dtype=np.float32)
    mean = rng.normal(0, 0.1, (1, C, 1, 1)).astype(np.float32)
    var = rng.uniform(0.5, 2, (1, C, 1, 1)).astype(np.float32)
    gamma = rng.uniform(0.5, 1.5, (1, C, 1, 1)).astype(np.float32)
    beta = rng.normal(0, 0.1, (1, C, 1, 1)).astype(np.float32)
    return x, mean, var, gamma, beta


def affine_expr(xt, xv, mean, var, gamma, beta):
    return xt.clip((xv - mean) / xt.sqrt(var + 1e-5) * gamma + beta, -3, 3)


def time_plans_region(torch, plans, steps, warmup):
    """Run `steps` plans round-robin back to back; returns (wall seconds, mean
    kernel ms) where the kernel time is one pair of HIP events around the
    whole timed region on the launch stream, divided by `steps`.  Launches
    run back to back (a kernel trace shows ~0.1 us between them), so this is
    the launches' average duration as a kernel trace reports it; an event
    pair around every launch would add the event packets' own time (measured
    4-7 us a launch on the headline)."""
    for i in range(warmup):
        plans[i % len(plans)].run()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for i in range(steps):
        plans[i % len(plans)].run()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_time(e1) / steps


def view_variants(np, torch, xt):
    """The headline's read and write sides apart: copies and the affine clip
    between the strided view x[:, :, 0:56:2, :] (rows of 56 floats every 112)
    and a dense [64, 256, 28, 56] array, each over three operand sets (as the
    timed loop rotates).  Returns [(name, plans, algorithmic bytes)]."""
    x_h, mean_h, var_h, gamma_h, beta_h = make_inputs(np)
    params = [xt.from_numpy(v) for v in (mean_h, var_h, gamma_h, beta_h)]
    view = (xt.all_(), xt.all_(), xt.range_(0, H, 2), xt.all_())
    dshape = (N, C, H // 2, W)
    x0 = xt.from_numpy(x_h)
    d0 = xt.from_numpy(np.ascontiguousarray(x_h[:, :, 0:H:2, :]))
    sets = []
    for _ in range(3):
        x = xt.empty(x_h.shape, "float32")
        x.storage.copy_(x0.storage)
        y = xt.zeros(x_h.shape, "float32")
        d = xt.empty(dshape, "float32")
        d.storage.copy_(d0.storage)
        e = xt.zeros(dshape, "float32")
        sets.append((xt.view(x, *view), xt.view(y, *view), d, e))
    del x0, d0
    nb = N * C * (H // 2) * W * 4 * 2
    out = []
    for name, make in (
            ("copy_dense_to_dense", lambda xv, yv, d, e: xt.plan(e, d)),
            ("copy_view_to_dense", lambda xv, yv, d, e: xt.plan(e, xv)),
            ("copy_dense_to_view", lambda xv, yv, d, e: xt.plan(yv, d)),
            ("copy_view_to_view", lambda xv, yv, d, e: xt.plan(yv, xv)),
            ("affine_clip_dense_to_dense", lambda xv, yv, d, e: xt.plan(e, affine_expr(xt, d, *params))),
            ("affine_clip_view_to_dense", lambda xv, yv, d, e: xt.plan(e, affine_expr(xt, xv, *params))),
            ("affine_clip_dense_to_view", lambda xv, yv, d, e: xt.plan(yv, affine_expr(xt, d, *params))),
            ("affine_clip_view_to_view", lambda xv, yv, d, e: xt.plan(yv, affine_expr(xt, xv, *params)))):
        out.append((name, [make(*st) for st in sets], nb + (4 * C * 4 if name.startswith("affine") else 0)))
    return out


def view_decomposition(np, torch, xt, steps=24):
    """Kernel GB/s (algorithmic bytes over the launches' average duration) of
    every view variant; the view's read side and write side apart."""
    res = {}
    for name, plans, nb in view_variants(np, torch, xt):
        _, kms = time_plans_region(torch, plans, steps, 6)
        res[name] = {"kernel_ms": round(kms, 4), "GBps": round(nb / (kms * 1e-3) / 1e9, 1),
                     "kernel": xt.last_kernel()}
    return res


def _cpu_info():
    """CPU model and logical CPU count of this host (the GPU box's host cores)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return model, n


def _timed(f, args, budget_s):
    """Seconds per repetition of f(*args, reps), reps sized to about budget_s."""
    t1 = f(*args, 1)
    reps = max(1, int(budget_s / max(t1, 1e-4)))
    return f(*args, reps) / reps, reps


def cpu_baseline(np, budget_s=5.0, per_s=1.5):
    """The CPU restatements of xtensor's path for every workload, timed on
    this host (oracle/cpu_baseline.cpp; xtensor itself is not buildable here).
    Headline `value`: the affine + clip workload (xtensor's scalar stepper
    loop, one core) on a sample of the images.  `per_config`: every other
    workload, GB/s of the same algorithmic bytes as the GPU figures, with
    the threads each used (OpenMP over all cores where xtensor's TBB loop
    would run, one core where xtensor's loop is single-threaded)."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_build", "libcpubase.so"))
    for name in ("xb_affine_clip_f32", "xb_bitmix_i32", "xb_mul_add_exp_f64", "xb_row_sum_f32", "xb_col_sum_f32",
                 "xb_col_sum_f32_immediate", "xb_cumsum_rows_f32", "xb_cumsum_cols_f32", "xb_softmax_f32",
                 "xb_row_sum_f64", "xb_col_amax_f64", "xb_amax_all_f32", "xb_norm_l2_all_f32",
                 "xb_row_sum_i32_i64", "xb_col_sum_i32_i64"):
        getattr(lib, name).restype = ctypes.c_double
    threads = int(lib.xb_threads())
    model, ncpu = _cpu_info()
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    sz = ctypes.c_size_t
    n = 8  # sample: 8 of the 64 images per pass, passes repeated for ~budget_s
    x, mean, var, gamma, beta = make_inputs(np, n)
    y = np.zeros_like(x)
    t, reps = _timed(lib.xb_affine_clip_f32, (P(x), P(mean), P(var), P(gamma), P(beta), P(y), n, C, H, W), budget_s)
    gbs = algorithmic_bytes(n) / t / 1e9
    per = {}

    def rec(key, nbytes, sec, thr, what):
        per[key] = {"GBps": round(nbytes / sec / 1e9, 3), "ms": round(sec * 1e3, 3), "threads": thr, "loop": what}

    # config 1: c = a*b + exp(-a*a) (strided loop over rows, threads), s = sum(c, {1}), m = amax(c, {0})
    rng = np.random.default_rng(0)
    a = rng.uniform(-1, 1, (2048, 1024))
    b = rng.uniform(-1, 1, (1024,))
    c = np.empty_like(a)
    sec, _ = _timed(lib.xb_mul_add_exp_f64, (P(a), P(b), P(c), 2048, 1024), per_s)
    rec("mul_add_exp_f64_2048x1024", 2 * a.nbytes + b.nbytes, sec, threads, "strided loop, OpenMP over rows")
    s1 = np.empty(2048)
    sec, _ = _timed(lib.xb_row_sum_f64, (P(c), P(s1), 2048, 1024), per_s)
    rec("sum_axis1_of_c_f64", c.nbytes + s1.nbytes, sec, 1, "lazy fold per row")
    m1 = np.empty(1024)
    sec, _ = _timed(lib.xb_col_amax_f64, (P(c), P(m1), 2048, 1024), per_s)
    rec("amax_axis0_of_c_f64", c.nbytes + m1.nbytes, sec, 1, "immediate: rows folded into the output")
    del a, c
    # config 2 int32 variant over the full tensor
    xi = np.random.default_rng(1).integers(-2**20, 2**20, (N * C * H * W,), dtype=np.int32)
    yi = np.empty_like(xi)
    sec, _ = _timed(lib.xb_bitmix_i32, (P(xi), P(yi), sz(xi.size)), per_s)
    rec("int32_bitmix", 2 * xi.nbytes, sec, threads, "linear loop, OpenMP")
    del xi, yi
    # config 3: [8192, 4096]
    R, K = 8192, 4096
    X = np.random.default_rng(2).normal(1.0, 1.0, (R, K)).astype(np.float32)
    o_r, o_c = np.empty(R, np.float32), np.empty(K, np.float32)
    sec, _ = _timed(lib.xb_row_sum_f32, (P(X), P(o_r), R, K), per_s)
    rec("sum_axis1_f32", X.nbytes + o_r.nbytes, sec, 1, "lazy fold per row (immediate: the same accumulate)")
    sl, _ = _timed(lib.xb_col_sum_f32, (P(X), P(o_c), R, K), per_s)
    si, _ = _timed(lib.xb_col_sum_f32_immediate, (P(X), P(o_c), R, K), per_s)
    rec("sum_axis0_f32", X.nbytes + o_c.nbytes, min(sl, si), 1,
        "immediate: rows added into the output" if si <= sl else "lazy fold per column")
    per["sum_axis0_f32"]["lazy_GBps"] = round((X.nbytes + o_c.nbytes) / sl / 1e9, 3)
    one = np.empty(1, np.float32)
    sec, _ = _timed(lib.xb_amax_all_f32, (P(X), P(one), sz(X.size)), per_s)
    rec("amax_all_f32", X.nbytes + 4, sec, 1, "accumulate over the storage")
    oned = np.empty(1, np.float64)
    sec, _ = _timed(lib.xb_norm_l2_all_f32, (P(X), P(oned), sz(X.size)), per_s)
    rec("norm_l2_all_f32", X.nbytes + 8, sec, 1, "sum of squares in double over the storage")
    Y = np.empty_like(X)
    sec, _ = _timed(lib.xb_cumsum_rows_f32, (P(X), P(Y), R, K), per_s)
    rec("cumsum_axis1_f32", 2 * X.nbytes, sec, 1, "copy, then the recurrence along each row")
    sec, _ = _timed(lib.xb_cumsum_cols_f32, (P(X), P(Y), R, K), per_s)
    rec("cumsum_axis0_f32", 2 * X.nbytes, sec, 1, "copy, then the linear sweep one row apart")
    del X, Y
    Xi = np.random.default_rng(2).integers(-1000, 1001, (R, K)).astype(np.int32)
    o64r, o64c = np.empty(R, np.int64), np.empty(K, np.int64)
    sec, _ = _timed(lib.xb_row_sum_i32_i64, (P(Xi), P(o64r), R, K), per_s)
    rec("sum_int64_axis1_i32", Xi.nbytes + o64r.nbytes, sec, 1, "lazy fold per row")
    sec, _ = _timed(lib.xb_col_sum_i32_i64, (P(Xi), P(o64c), R, K), per_s)
    rec("sum_int64_axis0_i32", Xi.nbytes + o64c.nbytes, sec, 1, "immediate: rows added into the output")
    del Xi
    # config 4: row softmax per instance, every 16th instance of the uneven batch
    from xtensor_amd.workloads import uneven_shapes, uneven_values
    shapes = uneven_shapes(4096, seed=3)
    xs = uneven_values(shapes, seed=3)
    sample = list(range(0, len(shapes), 16))
    t_all, elems = 0.0, 0
    for i in sample:
        r_, c_ = shapes[i]
        xin = np.ascontiguousarray(xs[i])
        xo = np.empty_like(xin)
        tmp = np.empty(r_, np.float32)
        t_all += lib.xb_softmax_f32(P(xin), P(xo), P(tmp), r_, c_, 1)
        elems += r_ * c_
    per["uneven_batch_softmax_f32"] = {"GBps": round(8 * elems / t_all / 1e9, 3),
                                       "inst_per_s": round(len(sample) / t_all, 1), "threads": 1,
                                       "loop": f"amax, exp, sum, divide per row; {len(sample)} of 4096 instances"}
    return {"value": round(gbs, 3), "unit": "GB/s", "cores": 1, "kind": "port",
            "sample": f"{n} of {N} images x {reps} passes ({t * reps:.1f} s), scalar stepper-order loop, "
                      "restatement of xtensor's path (the reference itself is not buildable here)",
            "cpu_model": model, "nproc": ncpu, "omp_threads": threads, "per_config": per}


def make_uneven_batch(np, torch, xt, n_inst=4096):
    """The uneven batch: instances packed at 256-byte boundaries in one HBM
    buffer per side; returns the prepared Batch and its buffers."""
    from xtensor_amd.batch import Batch
    from xtensor_amd.workloads import uneven_shapes, uneven_values

    shapes = uneven_shapes(n_inst, seed=3)
    xs = uneven_values(shapes, seed=3)
    elems = sum(r * c for r, c in shapes)
    offs, o = [], 0
    for r, c in shapes:
        offs.append(o)
        o += (r * c + 63) // 64 * 64
    xin = xt.empty((o,), "float32")
    xout = xt.zeros((o,), "float32")
    host = np.zeros(o, np.float32)
    for off, x in zip(offs, xs):
        host[off:off + x.size] = x.ravel()
    xin.storage.copy_(torch.from_numpy(host.view(np.uint8)))
    ins = [xt.DeviceArray(xin.storage, "float32", s, None, off) for off, s in zip(offs, shapes)]
    outs = [xt.DeviceArray(xout.storage, "float32", s, None, off) for off, s in zip(offs, shapes)]
    return Batch("softmax", ins, outs), ins, outs, offs, xs, xout, elems


def uneven_batch(np, torch, xt, steps=20, n_inst=4096):
    """4096 float32 instances, rows ~ log-uniform [1,512], cols ~ log-uniform
    [1,2048] (seed 3), x ~ N(0,2); row softmax of every instance.  Batched:
    one autotuned persistent launch replayed as a HIP graph.  Plain: three
    prepared plans per instance (amax keep_dims, fused exp-sum, normalise)."""
    b, ins, outs, offs, xs, xout, elems = make_uneven_batch(np, torch, xt, n_inst)
    tuning = b.autotune()
    b.execute()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record(stream)
    for _ in range(steps):
        b.execute()
    e1.record(stream)
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) / steps
    kms = e0.elapsed_time(e1) / steps
    got = np.frombuffer(xout.storage.cpu().numpy().tobytes(), np.float32)
    worst = 0.0
    for off, x in zip(offs, xs):
        x64 = x.astype(np.float64)
        z = np.exp(x64 - x64.max(axis=1, keepdims=True))
        ref = z / z.sum(axis=1, keepdims=True)
        g = got[off:off + x.size].reshape(x.shape)
        worst = max(worst, float(np.max(np.abs(g - ref) / ref)))

    # plain path: 3 prepared plans per instance
    plans = []
    for x, y in zip(ins, outs):
        r = x.shape[0]
        m = xt.empty((r, 1), "float32")
        s = xt.empty((r, 1), "float32")
        plans.append(xt.plan(m, xt.amax(x, [1], keep_dims=True)))
        plans.append(xt.plan(s, xt.sum(xt.exp(x - m), [1], keep_dims=True)))
        plans.append(xt.plan(y, xt.exp(x - m) / s))
    for p in plans[:300]:
        p.run()
    torch.cuda.synchronize()
    reps = 2
    t0 = time.perf_counter()
    for _ in range(reps):
        for p in plans:
            p.run()
    torch.cuda.synchronize()
    plain = (time.perf_counter() - t0) / reps
    # the same plans enqueued from C: one xg_plan_execute per plan in a C loop
    plain_c = None
    loop_lib = os.path.join(os.path.dirname(os.path.abspath(__file__)), "xtensor_amd", "lib", "libxgloop.so")
    if os.path.exists(loop_lib):
        import ctypes

        ll = ctypes.CDLL(loop_lib)
        ll.xgl_time_plans.restype = ctypes.c_int32
        ll.xgl_time_plans.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p,
                                      ctypes.POINTER(ctypes.c_double)]
        handles = (ctypes.c_void_p * len(plans))(*[p.handle.value for p in plans])
        secs = ctypes.c_double()
        st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        rc = ll.xgl_time_plans(handles, len(plans), reps, st, ctypes.byref(secs))
        if rc != 0:
            raise RuntimeError(f"xgl_time_plans failed with status {rc}")
        plain_c = secs.value / reps
    nbytes = 8 * elems
    return {
        "instances": n_inst, "elements": elems, "algorithmic_bytes": nbytes,
        "batched_ms": round(wall * 1e3, 4), "batched_kernel_ms": round(kms, 4),
        "batched_inst_per_s": round(n_inst / wall, 1), "batched_GBps": round(nbytes / (kms * 1e-3) / 1e9, 1),
        "plain_ms": round(plain * 1e3, 3), "plain_inst_per_s": round(n_inst / plain, 1),
        "plain_launches": len(plans), "speedup_vs_plain": round(plain / wall, 1),
        "plain_from_c_ms": None if plain_c is None else round(plain_c * 1e3, 3),
        "plain_from_c_inst_per_s": None if plain_c is None else round(n_inst / plain_c, 1),
        "speedup_vs_plain_from_c": None if plain_c is None else round(plain_c / wall, 1),
        "max_rel_err_vs_fp64": worst, "tuning": tuning, "tasks": b.info()["tasks"],
    }


def reductions(np, torch, xt, steps=12, copies=3):
    """Axis reductions and the scan on X float32 [8192, 4096] ~ N(1, 1) (seed 2),
    its bfloat16 copy, its transpose view and an int32 copy U[-1000, 1000];
    plus the fused broadcast expression of a 2048 x 1024 float64 pair.  Each
    result: kernel time, GB/s of algorithmic bytes (input read once + output
    written) and the relative error against float64 (int: exact).  The large
    inputs rotate over `copies` device copies (with their own outputs), so
    consecutive launches never find their operand in the 256 MB MALL."""
    rng = np.random.default_rng(2)
    R, Cc = 8192, 4096
    X = rng.normal(1.0, 1.0, (R, Cc)).astype(np.float32)
    Xi = rng.integers(-1000, 1001, (R, Cc), dtype=np.int32)
    dX = [xt.from_numpy(X) for _ in range(copies)]
    dXb = [xt.from_numpy(X, "bfloat16") for _ in range(copies + 1)]  # 4 x 67 MB > MALL
    dXi = [xt.from_numpy(Xi) for _ in range(copies)]
    X64 = X.astype(np.float64)
    Xb64 = torch.from_numpy(X).to(torch.bfloat16).to(torch.float64).numpy()
    out = {}

    def run(name, make, srcs, ref, in_bytes, exact=False, dtype=None):
        plans, outs = [], []
        for src in srcs:
            e = make(src)
            if isinstance(e, xt.Reducer):
                dt, shp = xt.reduce_info(e)
                o = xt.empty(shp, dtype or xt.dtype_name(dt))
            else:
                o = xt.empty(xt.expr_info(e)[1], dtype or xt.dtype_name(xt.expr_info(e)[0]))
            plans.append(xt.plan(o, e))
            outs.append(o)
        _, kms = time_plans_region(torch, plans, steps, len(plans))
        ref = np.asarray(ref, np.float64)
        err = 0.0
        for o in outs:
            got = o.to_numpy().astype(np.float64)
            if exact:
                err = max(err, 0.0 if np.array_equal(got, ref) else float("inf"))
            else:
                err = max(err, float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300))))
        o = outs[0]
        nbytes = in_bytes + o.size * o.itemsize
        out[name] = {"kernel_ms": round(kms, 4), "GBps": round(nbytes / (kms * 1e-3) / 1e9, 1),
                     "max_rel_err_vs_fp64": err, "kernel": xt.last_kernel()}
        del outs, plans

    nb = R * Cc * 4
    run("sum_axis0_f32", lambda d: xt.sum(d, [0]), dX, X64.sum(0), nb)
    run("sum_axis1_f32", lambda d: xt.sum(d, [1]), dX, X64.sum(1), nb)
    run("mean_axis0_f32", lambda d: xt.mean(d, [0]), dX, X64.mean(0), nb)
    run("mean_axis1_f32", lambda d: xt.mean(d, [1]), dX, X64.mean(1), nb)
    run("amax_all_f32", lambda d: xt.amax(d), dX, X64.max(), nb, exact=True)
    run("norm_l2_all_f32", lambda d: xt.norm_l2(d), dX, np.sqrt((X64 * X64).sum()), nb)
    run("sum_axis0_transposed_f32", lambda d: xt.sum(d.T, [0]), dX, X64.sum(1), nb)
    run("sum_axis1_transposed_f32", lambda d: xt.sum(d.T, [1]), dX, X64.sum(0), nb)
    run("sum_axis0_bf16", lambda d: xt.sum(d, [0]), dXb, Xb64.sum(0), nb // 2)
    run("sum_axis1_bf16", lambda d: xt.sum(d, [1]), dXb, Xb64.sum(1), nb // 2)
    run("sum_int64_axis0_i32", lambda d: xt.sum(d, [0], dtype="int64"), dXi, Xi.astype(np.int64).sum(0), nb,
        exact=True)
    run("sum_int64_axis1_i32", lambda d: xt.sum(d, [1], dtype="int64"), dXi, Xi.astype(np.int64).sum(1), nb,
        exact=True)
    # cumsum along axis 1 (scan plans, one output per input copy)
    css = [xt.empty((R, Cc), "float32") for _ in dX]
    pcs = [xt.plan(cs, xt.Scan("cumsum", d, 1)) for cs, d in zip(css, dX)]
    _, kms = time_plans_region(torch, pcs, steps, len(pcs))
    ref = np.cumsum(X64, axis=1)
    got = css[0].to_numpy().astype(np.float64)
    out["cumsum_axis1_f32"] = {"kernel_ms": round(kms, 4), "GBps": round(2 * nb / (kms * 1e-3) / 1e9, 1),
                               "max_abs_err_over_rowsum_abs": float(np.max(np.abs(got - ref) /
                                                                    np.abs(X64).sum(1, keepdims=True))),
                               "kernel": xt.last_kernel()}
    del css, pcs

    def scan(name, views, ref, axis):
        """a scan plan per input copy; error against a float64 scan relative
        to max(|ref|, 1)"""
        outs = [xt.empty(v.shape, "float32") for v in views]
        plans = [xt.plan(o, xt.Scan("cumsum", v, axis)) for o, v in zip(outs, views)]
        _, skms = time_plans_region(torch, plans, steps, len(plans))
        kname = xt.last_kernel()
        got = outs[0].to_numpy().astype(np.float64)
        out[name] = {"kernel_ms": round(skms, 4), "GBps": round(2 * nb / (skms * 1e-3) / 1e9, 1),
                     "max_rel_err_vs_fp64": float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0))),
                     "kernel": kname}

    # the scan axis strided: cumsum(X, 0) and cumsum(transpose(X), 1) (tiled
    # kernels), and the flattened cumsum(X) (one row of 2^25 elements: the
    # look-back kernel; planned as the scan of the 1-d view of X's storage)
    scan("cumsum_axis0_f32", dX, np.cumsum(X64, axis=0), 0)
    scan("cumsum_axis1_transposed_f32", [d.T for d in dX], np.cumsum(X64.T, axis=1), 1)
    scan("cumsum_flat_f32", [xt.DeviceArray(d.storage, "float32", (R * Cc,), None, 0) for d in dX],
         np.cumsum(X64.ravel()), 0)
    del dX, dXb, dXi
    # the broadcast expression of 2048 x 1024 float64 and its reductions
    r1 = np.random.default_rng(0)
    a = r1.uniform(-1, 1, (2048, 1024))
    b = r1.uniform(-1, 1, 1024)
    da, db = xt.from_numpy(a), xt.from_numpy(b)
    c = xt.empty((2048, 1024), "float64")
    pc = xt.plan(c, da * db + xt.exp(-da * da))
    _, kms = time_plans_region(torch, [pc], steps, 2)
    kname = xt.last_kernel()
    # the host libm's exp (math.exp), element by element, as xtensor's
    # non-SIMD path calls std::exp; the device exp's ULP histogram against it
    # and the count of c elements that differ from the host evaluation
    ex_host = np.fromiter(map(math.exp, (-a * a).ravel()), np.float64, a.size).reshape(a.shape)
    c_ref = a * b + ex_host
    ex_dev = xt.eval(xt.exp(-da * da)).to_numpy()
    ulp = np.abs(ex_dev.view(np.int64) - ex_host.view(np.int64))  # positive values: plain bit distance
    got_c = c.to_numpy()
    out["mul_add_exp_f64_2048x1024"] = {
        "kernel_ms": round(kms, 4), "GBps": round((2 * 2048 * 1024 + 1024) * 8 / (kms * 1e-3) / 1e9, 1),
        "max_rel_err_vs_host_libm": float(np.max(np.abs(got_c - c_ref) / np.maximum(np.abs(c_ref), 1e-300))),
        "exp_ulp_histogram_vs_host_libm": {str(k): int(v) for k, v in enumerate(np.bincount(ulp.ravel()))},
        "c_elements_differing_from_host_libm": int(np.count_nonzero(got_c != c_ref)),
        "elements": int(a.size),
        "kernel": kname}
    # the 16 MB result c stays cache-resident between these small launches,
    # as it would in the pipeline c -> sum(c) -> amax(c)
    run("sum_axis1_of_c_f64", lambda d: xt.sum(d, [1]), [c], c_ref.sum(1), 2048 * 1024 * 8)
    run("amax_axis0_of_c_f64", lambda d: xt.amax(d, [0]), [c], c.to_numpy().max(0), 2048 * 1024 * 8, exact=True)
    # the whole pipeline c, s = sum(c, {1}), m = amax(c, {0}): three plan
    # launches per step, or one replay of the same plans as a HIP graph
    s1, m1 = xt.empty((2048,), "float64"), xt.empty((1024,), "float64")
    seq = [pc, xt.plan(s1, xt.sum(c, [1])), xt.plan(m1, xt.amax(c, [0]))]

    def per_step(fn, n=50):
        for _ in range(5):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(n):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / n

    eager_ms = per_step(lambda: [p.run() for p in seq])
    graph = xt.PlanGraph(seq)
    graph_ms = per_step(graph.replay)
    # the same three results from one pass over c (xg_plan_assign_reduce2d):
    # the fused kernel writes c and the row sums, a second launch merges the
    # column maxima of its row bands
    pf = xt.plan(c, da * db + xt.exp(-da * da), rows=("sum", s1), cols=("amax", m1))
    pf.run()
    torch.cuda.synchronize()
    fused_ok = (np.array_equal(c.to_numpy(), got_c) and np.array_equal(m1.to_numpy(), got_c.max(0)) and
                float(np.max(np.abs(s1.to_numpy() - c_ref.sum(1)) / np.abs(c_ref.sum(1)))) <= 1e-14)
    fused_ms = per_step(pf.run)
    fgraph = xt.PlanGraph([pf])
    fgraph_ms = per_step(fgraph.replay)
    out["pipeline_c_sum_amax_f64"] = {"launches": len(seq), "eager_ms_per_step": round(eager_ms, 4),
                                      "graph_replay_ms_per_step": round(graph_ms, 4),
                                      "one_pass_ms_per_step": round(fused_ms, 4),
                                      "one_pass_graph_replay_ms_per_step": round(fgraph_ms, 4),
                                      "one_pass_matches": bool(fused_ok)}
    del seq, graph, pc, c, da, db, pf, fgraph
    out.update(fused_expr_reduce_large(np, torch, xt, steps))
    return out


def fused_expr_reduce_large(np, torch, xt, steps, rows=16384, cols=8192):
    """The config-1 expression fused into its axis reductions far past the
    MALL: float64 a [16384, 8192] (1 GiB, uniform(-1, 1), seed 0) and b [8192];
    sum(a*b + exp(-a*a), {1}) and amax(a*b + exp(-a*a), {0}), each one pass
    over a (c is never written), over three copies of a.  Error against the
    float64 evaluation in numpy (exp from numpy, sums pairwise)."""
    r = np.random.default_rng(0)
    a = r.uniform(-1, 1, (rows, cols))
    b = r.uniform(-1, 1, cols)
    da0 = xt.from_numpy(a)
    das = [da0]
    for _ in range(2):
        d = xt.empty(a.shape, "float64")
        d.storage.copy_(da0.storage)
        das.append(d)
    db = xt.from_numpy(b)
    c = a * b + np.exp(-a * a)
    res = {}
    for name, make, oshape, ref in (
            ("fused_sum_axis1_mul_add_exp_f64_16384x8192", lambda d: xt.sum(d * db + xt.exp(-d * d), [1]), (rows,),
             c.sum(1)),
            ("fused_amax_axis0_mul_add_exp_f64_16384x8192", lambda d: xt.amax(d * db + xt.exp(-d * d), [0]), (cols,),
             c.max(0))):
        outs = [xt.empty(oshape, "float64") for _ in das]
        plans = [xt.plan(o, make(d)) for o, d in zip(outs, das)]
        _, kms = time_plans_region(torch, plans, steps, len(plans))
        kname = xt.last_kernel()
        got = outs[0].to_numpy()
        nbytes = a.nbytes + b.nbytes + outs[0].size * 8
        res[name] = {"kernel_ms": round(kms, 4), "GBps": round(nbytes / (kms * 1e-3) / 1e9, 1),
                     "frac": round(nbytes / (kms * 1e-3) / 8e12, 4),
                     "max_rel_err_vs_fp64": float(np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300))),
                     "kernel": kname}
        del outs, plans
    del das, da0
    torch.cuda.empty_cache()
    return res


SERVE_TEMPLATES = {  # template -> dtypes it is defined for
    "affine_clip": ("float32", "float64"),
    "row_sum": ("float32", "float64", "int32"),
    "row_softmax": ("float32", "float64"),
    "bitmix": ("int32",),
}


def serve_requests(np, n_req, seed=4):
    """Request mix: template uniform over the four, total elements log-uniform
    in [1e3, 1e8], dtype uniform over the template's dtypes; shapes
    affine_clip [N, 16, 8, 8], rows x 1024 (sum) / 2048 (softmax), bitmix [n]."""
    rng = np.random.default_rng(seed)
    names = list(SERVE_TEMPLATES)
    out = []
    for _ in range(n_req):
        tpl = names[rng.integers(len(names))]
        dts = SERVE_TEMPLATES[tpl]
        dt = dts[rng.integers(len(dts))]
        n = int(np.exp(rng.uniform(np.log(1e3), np.log(1e8))))
        if tpl == "affine_clip":
            shape, oshape = (max(1, n // 1024), 16, 8, 8), None
        elif tpl == "bitmix":
            shape, oshape = (n,), None
        else:
            cols = min(n, 1024 if tpl == "row_sum" else 2048)
            shape = (max(1, n // cols), cols)
            oshape = (shape[0],) if tpl == "row_sum" else None
        out.append((tpl, dt, shape, oshape or shape))
    return out


def serving(np, torch, xt, rates=(1000, 10000, 50000), n_req=3000, seed=4):
    """Served request stream: Poisson arrivals at each rate (seed 4), operands
    resident in HBM, 64 GB admission budget; continuous batching (4 streams,
    up to 64 requests per launch) vs the plain per-request path (one stream,
    each request alone through the per-expression calls).

    Every request reads its own slice of a resident input pool and writes its
    own slice of an output pool: consecutive requests of one (template, dtype)
    take consecutive regions, the pools are several GB (far past the 256 MB
    MALL), so a request's operands are not cache-resident from an earlier
    one.  Request descriptors are built before the timed stream, so the
    submission loop is the arrival clock and one library call per request."""
    from xtensor_amd.serve import Server, now_ns

    reqs = serve_requests(np, n_req, seed)
    esz = {"float32": 4, "float64": 8, "int32": 4}
    align = 64  # elements: slices start 256-byte aligned
    rnd = lambda n: (n + align - 1) // align * align  # noqa: E731
    # input pools per (template, dtype): requests take consecutive regions,
    # wrapping at 4x the largest request (at least 1 GB)
    maxn, pool_n = {}, {}
    for tpl, dt, shape, _ in reqs:
        maxn[(tpl, dt)] = max(maxn.get((tpl, dt), 0), rnd(int(np.prod(shape))))
    for key, n in maxn.items():
        pool_n[key] = max(4 * n, (1 << 30) // esz[key[1]])
    in_off, cur = [], {}
    for tpl, dt, shape, _ in reqs:
        n, c = rnd(int(np.prod(shape))), cur.get((tpl, dt), 0)
        if c + n > pool_n[(tpl, dt)]:
            c = 0
        in_off.append(c)
        cur[(tpl, dt)] = c + n
    # output pools per dtype, regions taken in request order, at most 16 GiB
    # each: two requests share output memory only with 16 GiB of that
    # dtype's outputs between them (peak_inflight_GB, inputs and outputs
    # together, says how much is in flight at once)
    out_total = {}
    for _, dt, _, oshape in reqs:
        out_total[dt] = out_total.get(dt, 0) + rnd(int(np.prod(oshape)))
    out_pool_n = {dt: min(n, (16 << 30) // esz[dt]) for dt, n in out_total.items()}
    out_off, ocur, wraps = [], {}, 0
    for _, dt, _, oshape in reqs:
        n, c = rnd(int(np.prod(oshape))), ocur.get(dt, 0)
        if c + n > out_pool_n[dt]:
            c, wraps = 0, wraps + 1
        out_off.append(c)
        ocur[dt] = c + n
    gen = torch.Generator(device="cuda").manual_seed(seed)
    pools, params = {}, {}
    for (tpl, dt), n in pool_n.items():
        if dt == "int32":
            t = torch.randint(-2**20, 2**20, (n,), dtype=torch.int32, device="cuda", generator=gen)
        else:
            t = torch.randn(n, dtype=getattr(torch, dt), device="cuda", generator=gen) * 2
        pools[(tpl, dt)] = t
        if tpl == "affine_clip":
            pd = getattr(torch, dt)
            params[dt] = [torch.rand(16, dtype=pd, device="cuda", generator=gen) * s + o
                          for s, o in ((0.2, -0.1), (1.5, 0.5), (1.0, 0.5), (0.2, -0.1))]
    opools = {dt: torch.empty(n, dtype=getattr(torch, dt), device="cuda") for dt, n in sorted(out_pool_n.items())}
    nbytes = [(int(np.prod(shape)) + int(np.prod(oshape))) * esz[dt] for _, dt, shape, oshape in reqs]

    def view(t, shape):
        return xt.DeviceArray(t.view(torch.uint8), str(t.dtype).replace("torch.", ""), shape)

    pviews = {dt: [view(p, (16,)) for p in ps] for dt, ps in params.items()}
    prepared = []
    for k, (tpl, dt, shape, oshape) in enumerate(reqs):
        x = view(pools[(tpl, dt)][in_off[k]:in_off[k] + int(np.prod(shape))], shape)
        o = view(opools[dt][out_off[k]:out_off[k] + int(np.prod(oshape))], oshape)
        prepared.append(Server.prepare(tpl, x, o, pviews[dt] if tpl == "affine_clip" else [], tag=k))
    pool_GB = (sum(t.numel() * t.element_size() for t in pools.values()) +
               sum(t.numel() * t.element_size() for t in opools.values())) / 1e9
    result = {"requests_per_rate": n_req, "budget_GB": 64, "mean_request_MB": round(float(np.mean(nbytes)) / 1e6, 2),
              "operands": "distinct slices per request", "pool_GB": round(pool_GB, 1),
              "output_pool_wraps": wraps}

    def stream(mode, gaps):
        cfg = dict(budget_bytes=64 << 30, queue_capacity=1 << 20)
        cfg.update(dict(streams=4, max_batch=64) if mode == "batched" else dict(streams=1, max_batch=1, plain=True))
        torch.cuda.synchronize()
        with Server(**cfg) as srv:
            t0 = now_ns() + 2_000_000
            arr = t0 + np.cumsum(gaps * 1e9).astype(np.int64)
            arr_l = arr.tolist()
            for k in range(n_req):
                while now_ns() < arr_l[k]:
                    pass
                srv.submit_prepared(prepared[k], arr_l[k])
            srv.drain(600_000)
            comps = srv.poll(n_req + 16)
            st = srv.stats()
        return arr, comps, st

    # one untimed pass of the whole request list per mode first: the first
    # request of each shape class pays one-time set-up (run-time compiled
    # kernels, workspace growth) that a running server has behind it
    for mode in ("batched", "plain"):
        stream(mode, np.zeros(n_req))
    for rate in rates:
        gaps = np.random.default_rng(seed + rate).exponential(1.0 / rate, n_req)
        for mode in ("batched", "plain"):
            arr, comps, st = stream(mode, gaps)
            lat = np.array([c["done_ns"] - c["arrival_ns"] for c in comps]) / 1e6
            span = (max(c["done_ns"] for c in comps) - int(arr[0])) / 1e9
            result[f"{mode}_{rate}"] = {
                "req_per_s": round(n_req / span, 1), "GBps": round(sum(nbytes) / span / 1e9, 1),
                "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
                "launches": st["launches"], "peak_inflight_GB": round(st["peak_inflight_bytes"] / 1e9, 2),
                "arrival_window_ms": round((int(arr[-1]) - int(arr[0])) / 1e6, 2),
                "ok": int(sum(c["status"] == 0 for c in comps)),
            }
            print(f"served stream {mode} at {rate} req/s: {result[f'{mode}_{rate}']}", file=sys.stderr, flush=True)
    del pools, opools
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the served-stream leg")
    ap.add_argument("--profile-batch", action="store_true",
                    help="profiling pass for rocprofv3: the uneven softmax batch autotuned as in the timed run "
                         "(trial launches), then one launch and 3 graph replays")
    ap.add_argument("--profile-reduce", action="store_true",
                    help="profiling pass for rocprofv3: a 256 MiB float32 copy (counter calibration), then "
                         "every reduction and scan of the bench line on [8192, 4096] (float32, its bf16 copy, "
                         "its transpose, int32) and the 2048 x 1024 float64 expression with its row sum and "
                         "column amax, 3 launches each over three input copies")
    ap.add_argument("--profile-view", action="store_true",
                    help="profiling pass for rocprofv3: a 256 MiB float32 copy (counter calibration, 3 "
                         "launches), then each view variant (copies and the affine clip between the strided "
                         "view and a dense array) 1 + 6 launches rotating over three sets, in a fixed order")
    ap.add_argument("--profile", action="store_true",
                    help="profiling pass for rocprofv3: a 256 MiB float32 copy (counter calibration, 3 "
                         "launches) and the workload kernel as in the timed loop (6 + 12 launches rotating "
                         "over the three sets), nothing else")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (single-GPU build)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    import xtensor_amd as xt
    from xtensor_amd import _abi as A

    torch.cuda.set_device(0)
    if args.profile_batch:
        b, *_rest, elems = make_uneven_batch(np, torch, xt)
        b.autotune()   # as the timed run: its trial launches come first, the last three launches are replays
        b.launch()
        torch.cuda.synchronize()
        for _ in range(3):
            b.execute()
        torch.cuda.synchronize()
        print(json.dumps({"profile_pass": "uneven_batch_softmax_f32", "elements": elems,
                          "algorithmic_bytes_each": 8 * elems, "tuning": b.tuning, "tasks": b.info()["tasks"]}))
        return
    if args.profile_view:
        src = xt.empty((64 << 20,), "float32")
        src.storage.zero_()
        dst = xt.empty((64 << 20,), "float32")
        cp = xt.plan(dst, src)
        variants = view_variants(np, torch, xt)
        torch.cuda.synchronize()
        for _ in range(3):
            cp.run()
        order = []
        for name, plans, nb in variants:
            for i in range(7):
                plans[i % 3].run()
            order.append({"variant": name, "launches": 7, "algorithmic_bytes": nb, "kernel": xt.last_kernel()})
            torch.cuda.synchronize()
        print(json.dumps({"profile_pass": "view_decomposition", "copy_bytes_each": 2 * (256 << 20),
                          "calibration_launches": 3, "order": order}))
        return
    if args.profile_reduce:
        src = xt.empty((64 << 20,), "float32")
        src.storage.zero_()
        dst = xt.empty((64 << 20,), "float32")
        cp = xt.plan(dst, src)
        for _ in range(3):
            cp.run()
        X = np.random.default_rng(2).normal(1.0, 1.0, (8192, 4096)).astype(np.float32)
        dX = [xt.from_numpy(X) for _ in range(3)]
        dXb = [xt.from_numpy(X, "bfloat16") for _ in range(3)]
        dXi = [xt.from_numpy(np.random.default_rng(2).integers(-1000, 1001, (8192, 4096), dtype=np.int32))
               for _ in range(3)]
        R_, C_ = X.shape
        flat = lambda d: xt.DeviceArray(d.storage, "float32", (R_ * C_,), None, 0)  # noqa: E731
        names = {}
        for name, srcs, make, oshape, odt in (
                ("sum_axis0_f32", dX, lambda d: xt.sum(d, [0]), (4096,), "float32"),
                ("sum_axis1_f32", dX, lambda d: xt.sum(d, [1]), (8192,), "float32"),
                ("mean_axis0_f32", dX, lambda d: xt.mean(d, [0]), (4096,), "float64"),
                ("mean_axis1_f32", dX, lambda d: xt.mean(d, [1]), (8192,), "float64"),
                ("amax_all_f32", dX, lambda d: xt.amax(d), (), "float32"),
                ("norm_l2_all_f32", dX, lambda d: xt.norm_l2(d), (), "float64"),
                ("sum_axis0_transposed_f32", dX, lambda d: xt.sum(d.T, [0]), (8192,), "float32"),
                ("sum_axis1_transposed_f32", dX, lambda d: xt.sum(d.T, [1]), (4096,), "float32"),
                ("sum_axis0_bf16", dXb, lambda d: xt.sum(d, [0]), (4096,), "float32"),
                ("sum_axis1_bf16", dXb, lambda d: xt.sum(d, [1]), (8192,), "float32"),
                ("sum_int64_axis0_i32", dXi, lambda d: xt.sum(d, [0], dtype="int64"), (4096,), "int64"),
                ("sum_int64_axis1_i32", dXi, lambda d: xt.sum(d, [1], dtype="int64"), (8192,), "int64"),
                ("cumsum_axis1_f32", dX, lambda d: xt.Scan("cumsum", d, 1), (8192, 4096), "float32"),
                ("cumsum_axis0_f32", dX, lambda d: xt.Scan("cumsum", d, 0), (8192, 4096), "float32"),
                ("cumsum_axis1_transposed_f32", dX, lambda d: xt.Scan("cumsum", d.T, 1), (4096, 8192), "float32"),
                ("cumsum_flat_f32", dX, lambda d: xt.Scan("cumsum", flat(d), 0), (R_ * C_,), "float32")):
            outs = [xt.empty(oshape, odt) for _ in srcs]
            plans = [xt.plan(o, make(d)) for o, d in zip(outs, srcs)]
            for p in plans:
                p.run()
            names[name] = xt.last_kernel()
            torch.cuda.synchronize()
            del outs, plans
        r1 = np.random.default_rng(0)
        a = xt.from_numpy(r1.uniform(-1, 1, (2048, 1024)))
        b = xt.from_numpy(r1.uniform(-1, 1, 1024))
        c = xt.empty((2048, 1024), "float64")
        s1, m1 = xt.empty((2048,), "float64"), xt.empty((1024,), "float64")
        for name, p in (("mul_add_exp_f64_2048x1024", xt.plan(c, a * b + xt.exp(-a * a))),
                        ("sum_axis1_of_c_f64", xt.plan(s1, xt.sum(c, [1]))),
                        ("amax_axis0_of_c_f64", xt.plan(m1, xt.amax(c, [0]))),
                        ("one_pass_c_sum_amax_f64", xt.plan(c, a * b + xt.exp(-a * a), rows=("sum", s1),
                                                            cols=("amax", m1)))):
            for _ in range(3):
                p.run()
            names[name] = xt.last_kernel()
            torch.cuda.synchronize()
        print(json.dumps({"profile_pass": "reductions_f32_8192x4096", "copy_bytes_each": 2 * (256 << 20),
                          "input_bytes": X.nbytes, "kernels": names}))
        return
    x_h, mean_h, var_h, gamma_h, beta_h = make_inputs(np)
    params = [xt.from_numpy(v) for v in (mean_h, var_h, gamma_h, beta_h)]
    x0 = xt.from_numpy(x_h)
    sets = []
    for r in range(3):
        x = xt.empty(x_h.shape, "float32")
        x.storage.copy_(x0.storage)
        y = xt.zeros(x_h.shape, "float32")
        sets.append((x, y))
    del x0
    view = (xt.all_(), xt.all_(), xt.range_(0, H, 2), xt.all_())
    plans = [xt.plan(xt.view(y, *view), affine_expr(xt, xt.view(x, *view), *params)) for x, y in sets]
    plans[0].run()
    kernel_name = xt.last_kernel()
    torch.cuda.synchronize()

    if args.profile:
        src = xt.empty((64 << 20,), "float32")
        src.storage.zero_()
        dst = xt.empty((64 << 20,), "float32")
        cp = xt.plan(dst, src)
        for _ in range(3):
            cp.run()
        # the timed loop's steady state: launches rotating over the three sets
        time_plans_region(torch, plans, 12, 6)
        torch.cuda.synchronize()
        print(json.dumps({"profile_pass": True, "copy_bytes_each": 2 * (256 << 20),
                          "workload_algorithmic_bytes": algorithmic_bytes(), "kernel": kernel_name}))
        return

    nbytes = algorithmic_bytes()
    # kernel time: HIP events around the timed region on the launch stream,
    # over the launches; whole-job value: the same region on the host clock
    # (both synchronisations inside it)
    wall, kms = time_plans_region(torch, plans, args.steps, args.warmup)
    value = nbytes * args.steps / wall / 1e9
    achieved = nbytes / (kms * 1e-3) / 1e9

    extra = {"kernel": kernel_name, "kernel_ms": round(kms, 5)}
    # the same kernel with one and with four work units per thread
    per_upt = {}
    for upt in (1, 4):
        A.lib().xg_debug_units_per_thread(upt)
        try:
            _, ukms = time_plans_region(torch, plans, max(10, args.steps // 2), 3)
        finally:
            A.lib().xg_debug_units_per_thread(0)
        per_upt[str(upt)] = round(nbytes / (ukms * 1e-3) / 1e9, 1)
    extra["GBps_by_units_per_thread"] = per_upt
    # the strided view's read side and write side apart
    extra["view_decomposition"] = view_decomposition(np, torch, xt)
    print("view decomposition done", file=sys.stderr, flush=True)
    # the same workload on the generic interpreter (no compiled specialisation)
    A.lib().xg_debug_force_dynamic(1)
    try:
        dplans = [xt.plan(xt.view(y, *view), affine_expr(xt, xt.view(x, *view), *params)) for x, y in sets]
        _, dkms = time_plans_region(torch, dplans, max(5, args.steps // 5), 2)
        extra["generic_interpreter_GBps"] = round(nbytes / (dkms * 1e-3) / 1e9, 1)
    finally:
        A.lib().xg_debug_force_dynamic(0)
    # a program outside the ahead-of-time table (the same workload with a
    # float literal, so every operation is float32): compiled at run time on
    # first use, then launched like a table program
    t0 = time.perf_counter()
    f32lit = xt.scalar(1e-5, "float32")
    jplans = [xt.plan(xt.view(y, *view), xt.clip((xt.view(x, *view) - params[0]) / xt.sqrt(params[1] + f32lit)
                                                  * params[2] + params[3], -3, 3)) for x, y in sets]
    jplans[0].run()
    torch.cuda.synchronize()
    compile_s = time.perf_counter() - t0
    jkernel = xt.last_kernel()
    _, jkms = time_plans_region(torch, jplans, max(10, args.steps // 2), 3)
    extra["runtime_compiled_f32_literal"] = {"kernel": jkernel, "GBps": round(nbytes / (jkms * 1e-3) / 1e9, 1),
                                             "first_call_s": round(compile_s, 2)}
    del jplans
    # ceilings of the access pattern on this GPU: a plain copy through the
    # same views, and a contiguous copy of the same number of bytes (both
    # rotating over the three sets, through the compiled copy program)
    touched = N * C * (H // 2) * W
    cplans = [xt.plan(xt.view(y, *view), xt.view(x, *view)) for x, y in sets]
    _, ckms = time_plans_region(torch, cplans, max(10, args.steps // 2), 3)
    extra["view_copy_GBps"] = round(8 * touched / (ckms * 1e-3) / 1e9, 1)
    extra["view_copy_kernel"] = xt.last_kernel()
    fplans = [xt.plan(xt.DeviceArray(y.storage, "float32", (touched,), None, 0),
                      xt.DeviceArray(x.storage, "float32", (touched,), None, 0)) for x, y in sets]
    _, fkms = time_plans_region(torch, fplans, max(10, args.steps // 2), 3)
    extra["contiguous_copy_same_bytes_GBps"] = round(8 * touched / (fkms * 1e-3) / 1e9, 1)
    del cplans, fplans
    # int32 variant: y = (x*3 + (x>>2)) & 0xFFFF over the full array
    xi = xt.from_numpy(np.random.default_rng(1).integers(-2**20, 2**20, (N, C, H, W), dtype=np.int32))
    yi = xt.empty((N, C, H, W), "int32")
    ip = [xt.plan(yi, (xi * 3 + (xi >> 2)) & 0xFFFF)]
    _, ikms = time_plans_region(torch, ip, max(5, args.steps // 5), 2)
    extra["int32_bitmix_GBps"] = round(N * C * H * W * 8 / (ikms * 1e-3) / 1e9, 1)
    del xi, yi, ip
    extra["reductions_8192x4096"] = reductions(np, torch, xt)
    print("reductions done", file=sys.stderr, flush=True)
    extra["uneven_batch_softmax_f32"] = uneven_batch(np, torch, xt)
    print("uneven batch done", file=sys.stderr, flush=True)
    if not args.no_serve:
        extra["served_stream"] = serving(np, torch, xt)

    line = {
        "metric": "fused expr+axis-reduce GB/s (% of 8 TB/s); batched inst/s + p99 latency",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (f64 intermediates)",
        "data": "synthetic",
        "config": {"workload": "affine_clip_f32_strided_view", "x": [N, C, H, W],
                   "view": "[:, :, 0:56:2, :]", "params": [1, C, 1, 1], "algorithmic_bytes": nbytes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": measured_traffic(kernel_name)},
        "extra": extra,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(np)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

"""Run-time compiled specialisations (hipRTC) against the generic
interpreter and the oracle.

A program outside the ahead-of-time table gets the same kernel templates
instantiated on its own constexpr program.  Both forms run the same
per-operator device code, so elementwise results are bit-identical to the
interpreter's; reductions are checked to the bars of the reduction tests
(their work split differs from the interpreter's).
"""

import ctypes as C
import os
import subprocess
import sys

import numpy as np
import pytest

import xtensor_amd as xt
from xtensor_amd import _abi as A

from helpers import Dual, oracle_for, same_bits, ulp_diff

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def sync():
    import torch
    torch.cuda.synchronize()


def stats():
    out = (C.c_int64 * 4)()
    msg = A.lib().xg_debug_jit_stats(out)
    return list(out), (msg or b"").decode()


class jit_mode:
    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        A.lib().xg_debug_jit(self.mode)

    def __exit__(self, *a):
        A.lib().xg_debug_jit(1)


def interp_and_jit(make_out, expr_fn):
    """Evaluate with the interpreter (forced) and with a compiled kernel."""
    outs = []
    A.lib().xg_debug_force_dynamic(1)
    try:
        o = make_out()
        xt.noalias(o).assign(expr_fn())
        assert xt.last_kernel() == "dynamic"
        sync()
        outs.append(o.to_numpy())
    finally:
        A.lib().xg_debug_force_dynamic(0)
    with jit_mode(2):
        o = make_out()
        xt.noalias(o).assign(expr_fn())
        kern = xt.last_kernel()
        sync()
        outs.append(o.to_numpy())
    return outs, kern


# name: (shape, dtype, expression, ulp bar against the host libm)
CASES = {
    "tanh_mul_add_f32": ((257, 1000), np.float32, lambda x, y: xt.tanh(x) * y + 3, 2),
    "where_compare_f64": ((64, 333), np.float64, lambda x, y: xt.where(x > y, x - y, y * 0.5), 0),
    "int_mix_i32": ((100, 1024), np.int32, lambda x, y: (x * 7 - (y >> 3)) ^ (x & 0x5555), 0),
    # log1p within 2 ulp on each side, then one correctly rounded division
    "log1p_abs_f32_bcast": ((33, 4, 129), np.float32,
                            lambda x, y: xt.log1p(xt.abs(x)) / (y + xt.scalar(2.5, "float32")), 4),
    "cast_u8_f32": ((512, 77), np.float32, lambda x, y: xt.cast(x * 10.0 + 100.0, "uint8"), 0),
}


@pytest.mark.parametrize("name", list(CASES))
def test_jit_elementwise_matches_interpreter_and_oracle(name):
    shape, dt, fn, bar = CASES[name]
    rng = np.random.default_rng(hash(name) % 1000)
    if np.issubdtype(dt, np.integer):
        xh = rng.integers(-2**20, 2**20, shape, dtype=dt)
        yh = rng.integers(-2**20, 2**20, shape[-1:], dtype=dt)
    else:
        xh = rng.normal(0, 2, shape).astype(dt)
        yh = rng.uniform(0.5, 2.0, shape[-1:]).astype(dt)
    x, y = Dual.of(xh), Dual.of(yh)
    res_dt, res_shape = xt.expr_info(fn(x.dev, y.dev))
    (a, b), kern = interp_and_jit(lambda: xt.empty(res_shape, xt.dtype_name(res_dt)), lambda: fn(x.dev, y.dev))
    assert kern.startswith("jit:"), kern
    assert same_bits(a, b)
    ref = oracle_for(x, y).assign(fn(x.dev, y.dev))
    if ref.dtype.kind == "f":
        assert int(np.max(ulp_diff(b, ref))) <= bar
    else:
        assert np.array_equal(b, ref)


def test_jit_strided_view_operands():
    """A compiled kernel over strided views on both sides (rows of a view)."""
    rng = np.random.default_rng(3)
    x = Dual.of(rng.normal(size=(40, 64, 30)).astype(np.float32))
    y = Dual.of(np.zeros((40, 64, 30), np.float32))
    xv = x.view(xt.all_(), xt.range_(0, 63, 3), xt.all_())  # 21 rows, like range(1, 64, 3)
    ref = oracle_for(xv).assign(xt.sqrt(xt.abs(xv.dev)) - 1)
    with jit_mode(2):
        xt.noalias(xt.view(y.dev, xt.all_(), xt.range_(1, 64, 3), xt.all_())).assign(xt.sqrt(xt.abs(xv.dev)) - 1)
        assert xt.last_kernel().startswith("jit:")
    sync()
    got = y.dev.to_numpy()[:, 1:64:3, :]
    assert same_bits(got, ref)


@pytest.mark.parametrize("axes,shape", [([1], (300, 1000)), ([0], (300, 1000)), (None, (123, 4567)),
                                        ([1], (20, 30, 40)), ([0, 2], (20, 30, 40))])
def test_jit_reductions(axes, shape):
    rng = np.random.default_rng(7)
    x = Dual.of(rng.normal(1, 1, shape))
    o = oracle_for(x)
    for make in (lambda e: xt.sum(xt.exp(-e * e) + e, axes), lambda e: xt.amax(xt.abs(e - 1.0), axes)):
        with jit_mode(2):
            got = xt.eval(make(x.dev))
            kern = xt.last_kernel()
        sync()
        assert kern.startswith("jit:"), kern
        ref = o.reduce(make(x.dev), high_precision=True)
        g = got.to_numpy()
        assert float(np.max(np.abs(g - ref) / np.maximum(np.abs(ref), 1e-300))) <= 1e-13


def test_jit_int_reduction_exact():
    h = np.random.default_rng(1).integers(-1000, 1000, (257, 999), dtype=np.int32)
    x = Dual.of(h)
    r = xt.sum(x.dev * 3 - (x.dev >> 1), [1], dtype="int64")
    with jit_mode(2):
        got = xt.eval(r)
        assert xt.last_kernel().startswith("jit:")
    sync()
    assert np.array_equal(got.to_numpy(), oracle_for(x).reduce(r))


def test_jit_cache_and_threshold():
    """A second evaluation of a program reuses the loaded kernel; in the
    default mode small launches stay on the interpreter and large ones
    compile."""
    xs = xt.from_numpy(np.ones((64, 64), np.float32))
    xl = xt.from_numpy(np.ones((1024, 2048), np.float32))
    e = lambda x: xt.cosh(x) * 2 - xt.sinh(x)  # noqa: E731
    with jit_mode(1):
        o = xt.empty((64, 64), "float32")
        xt.noalias(o).assign(e(xs))
        assert xt.last_kernel() == "dynamic"
        before, _ = stats()
        ol = xt.empty((1024, 2048), "float32")
        xt.noalias(ol).assign(e(xl))
        assert xt.last_kernel().startswith("jit:")
        mid, _ = stats()
        xt.noalias(ol).assign(e(xl))
        after, msg = stats()
    sync()
    assert mid[0] + mid[1] == before[0] + before[1] + 1  # one compile (or disk hit)
    assert after[0] + after[1] == mid[0] + mid[1] and after[2] > mid[2], msg
    np.testing.assert_allclose(ol.to_numpy(), np.cosh(1.0) * 2 - np.sinh(1.0), rtol=1e-6)


def test_jit_disk_cache_across_processes(tmp_path):
    """A second process finds the compiled kernel in the disk cache."""
    code = (
        "import sys, ctypes as C, numpy as np; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import xtensor_amd as xt; from xtensor_amd import _abi as A; import torch\n"
        "A.lib().xg_debug_jit(2)\n"
        "x = xt.from_numpy(np.arange(4096, dtype=np.float32).reshape(64, 64))\n"
        "o = xt.empty((64, 64), 'float32'); xt.noalias(o).assign(xt.atan(x) * 0.25 + x)\n"
        "torch.cuda.synchronize(); s = (C.c_int64 * 4)(); A.lib().xg_debug_jit_stats(s)\n"
        "print(xt.last_kernel(), list(s))\n" % (ROOT, os.path.join(ROOT, "tests")))
    env = dict(os.environ, XG_JIT_CACHE=str(tmp_path))
    r1 = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env)
    assert r1.returncode == 0, r1.stderr
    r2 = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120, env=env)
    assert r2.returncode == 0, r2.stderr
    k1, s1 = r1.stdout.strip().split(" ", 1)
    k2, s2 = r2.stdout.strip().split(" ", 1)
    assert k1 == k2 and k1.startswith("jit:")
    assert eval(s1)[0] == 1 and eval(s2)[0] == 0 and eval(s2)[1] == 1

""".npy files to and from device memory against the reference's own fixtures
(tests/golden/xnpy/: the files of test/files/xnpy_files/ that
test/test_xnpy.cpp reads; expected arrays from that test via
tests/golden/kat_arrays.json).

CPU: header parsing (no device access).  GPU: loads equal the reference
arrays exactly; dumps are byte-identical to the reference's files
(test_xnpy.cpp:133-169 compares dump_npy output to them the same way);
strided, Fortran-order and large (chunked) round trips.
"""

import json
import os

import numpy as np
import pytest

import xtensor_amd as xt

HERE = os.path.dirname(os.path.abspath(__file__))
NPY = os.path.join(HERE, "golden", "xnpy")
KATS = json.load(open(os.path.join(HERE, "golden", "kat_arrays.json")))


def kat(test, name, dtype):
    r = KATS[f"test_xnpy.cpp:xnpy.{test}:{name}"]
    return np.array(r["values"], dtype=dtype).reshape(r["shape"])


def f(name):
    return os.path.join(NPY, name + ".npy")


# ---------------------------------------------------------------- headers (CPU)
@pytest.mark.parametrize("name,dtype,shape,fortran", [
    ("double.le", "float64", (3, 3, 3), False), ("double_fortran.le", "float64", (3, 3, 3), True),
    ("bool.le", "bool", (3, 3, 3), False), ("bool_fortran.le", "bool", (3, 3, 3), True),
    ("int.le", "int32", (5,), False), ("unsignedlong.le", "uint64", (5,), False),
    ("unsignedlong_fortran.le", "uint64", (5,), False),
])
def test_header(name, dtype, shape, fortran):
    assert xt.npy_info(f(name)) == (dtype, shape, fortran)


def test_big_endian_and_bad_files_raise(tmp_path):
    """A '>f8' file has no dtype on this host: load_npy<double> throws its
    "Cast error" (io/xnpy.hpp:588-596); a wrong magic string throws too.
    (The reference's double.be.npy itself holds '<f8' data.)"""
    assert xt.npy_info(f("double.be")) == ("float64", (3, 3, 3), False)
    be = tmp_path / "be.npy"
    np.save(be, np.arange(4, dtype=">f8"))
    with pytest.raises(Exception, match="Cast error"):
        xt.npy_info(str(be))
    bad = tmp_path / "bad.npy"
    bad.write_bytes(b"\x93NUMPX\x01\x00" + b" " * 120)
    with pytest.raises(Exception, match="magic"):
        xt.npy_info(str(bad))
    with pytest.raises(Exception):
        xt.npy_info(str(tmp_path / "missing.npy"))


# ---------------------------------------------------------------- device (GPU)
def sync():
    import torch
    torch.cuda.synchronize()


@pytest.mark.gpu
def test_load_reference_files():
    """test_xnpy.cpp:39-84."""
    darr = kat("load", "darr", np.float64)
    barr = kat("load", "barr", np.bool_)
    iarr = kat("load", "iarr1d", np.int32)
    # the files hold more digits than the test's literals: the reference
    # compares with isclose; the exact bits are numpy's reading of the file
    d = xt.load_npy(f("double.le"))
    assert d.dtype == "float64" and np.array_equal(d.to_numpy(), np.load(f("double.le")))
    assert np.allclose(d.to_numpy(), darr, rtol=1e-5, atol=1e-8)
    df = xt.load_npy(f("double_fortran.le"))
    assert df.strides == (1, 3, 9) and np.array_equal(df.to_numpy(), np.load(f("double_fortran.le")))
    assert np.allclose(df.to_numpy(), darr, rtol=1e-5, atol=1e-8)
    b = xt.load_npy(f("bool.le"), "bool")
    assert np.array_equal(b.to_numpy(), barr)
    assert np.array_equal(xt.load_npy(f("bool_fortran.le")).to_numpy(), barr)
    assert np.array_equal(xt.load_npy(f("int.le")).to_numpy(), iarr)
    with pytest.raises(RuntimeError, match="Cast error"):
        xt.load_npy(f("int.le"), "int64")
    # a loaded array is an ordinary operand (test_xnpy.cpp:171-179 casts one)
    c = xt.eval(xt.cast(d * 2.0 + 1.0, "int8"))
    sync()
    assert np.array_equal(c.to_numpy(), (darr * 2.0 + 1.0).astype(np.int8))


@pytest.mark.gpu
def test_dump_is_byte_identical(tmp_path):
    """test_xnpy.cpp:133-169: dump_npy of barr and ularr equals the files."""
    barr = xt.from_numpy(kat("dump", "barr", np.bool_))
    ularr = xt.from_numpy(kat("dump", "ularr", np.uint64))
    p0, p1 = str(tmp_path / "b.npy"), str(tmp_path / "u.npy")
    xt.dump_npy(p0, barr)
    xt.dump_npy(p1, ularr)
    assert open(p0, "rb").read() == open(f("bool.le"), "rb").read()
    assert open(p1, "rb").read() == open(f("unsignedlong.le"), "rb").read()
    back = xt.load_npy(p1)
    assert np.array_equal(back.to_numpy(), kat("dump", "ularr", np.uint64))


@pytest.mark.gpu
def test_dump_views_and_expressions(tmp_path):
    h = np.arange(6 * 7 * 8, dtype=np.float32).reshape(6, 7, 8)
    a = xt.from_numpy(h)
    v = xt.view(a, xt.range_(1, 6, 2), xt.all_(), xt.range_(7, None, -3))  # strided, reversed
    p = str(tmp_path / "v.npy")
    xt.dump_npy(p, v)
    assert np.array_equal(np.load(p), h[1:6:2, :, 7::-3])
    xt.dump_npy(p, xt.sqrt(a) + 1.0)
    r = np.load(p)
    assert r.dtype == np.float64 and np.array_equal(r, np.sqrt(h).astype(np.float64) + 1.0)
    z = xt.from_numpy(np.zeros((0, 3), np.int16))
    xt.dump_npy(p, z)
    assert np.load(p).shape == (0, 3)
    s = xt.from_numpy(np.array(3.5))
    xt.dump_npy(p, s)
    assert np.load(p).shape == () and np.load(p)[()] == 3.5


@pytest.mark.gpu
def test_large_round_trip_and_strided_destination(tmp_path):
    """100 MB through the chunked staging both ways; a load into a strided view."""
    rng = np.random.default_rng(1)
    h = rng.normal(size=(5000, 5000)).astype(np.float32)
    p = str(tmp_path / "big.npy")
    np.save(p, h)
    a = xt.load_npy(p)
    sync()
    assert np.array_equal(a.to_numpy(), h)
    q = str(tmp_path / "big2.npy")
    xt.dump_npy(q, a)
    assert open(q, "rb").read()[128:] == open(p, "rb").read()[128:]
    # into every other column of a wider array
    wide = xt.zeros((5000, 10000), "float32")
    dst = xt.view(wide, xt.all_(), xt.range_(0, 10000, 2))
    from xtensor_amd import _abi as A
    import ctypes as C
    t = dst.to_abi()
    assert A.lib().xg_npy_load(p.encode(), C.byref(t), None) == 0
    sync()
    w = wide.to_numpy()
    assert np.array_equal(w[:, ::2], h) and not w[:, 1::2].any()
    with pytest.raises(Exception):
        xt.load_npy(p, "float64")

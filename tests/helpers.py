"""Test helpers: paired device/host arrays and comparison utilities."""

from __future__ import annotations

import numpy as np

import xtensor_amd as xt
from xtensor_amd import _abi as A
from xtensor_amd.core import DeviceArray, _Placeholder

from oracle.pyoracle import Oracle


class Dual:
    """A device array and the host array holding the same values.

    Views are taken on both sides independently: the device view by the
    front end's own stride arithmetic, the host view by numpy slicing.
    """

    def __init__(self, dev: DeviceArray, host: np.ndarray):
        self.dev = dev
        self.host = host

    @classmethod
    def of(cls, host: np.ndarray, dtype=None, device: bool = True) -> "Dual":
        host = np.asarray(host)
        if dtype == "bfloat16":
            bits = to_bf16_bits(host)
            dev = xt.from_numpy(host, "bfloat16") if device else _Placeholder(A.BFLOAT16, host.shape)
            return cls(dev, bits)
        if device:
            dev = xt.from_numpy(host)
        else:
            dev = _Placeholder(xt.dtype_code(host.dtype), host.shape)
        return cls(dev, host)

    def view(self, *keys) -> "Dual":
        np_keys = tuple(k.as_slice() if isinstance(k, xt.range_) else
                        (slice(None) if isinstance(k, type(xt.all_())) else
                         (None if isinstance(k, type(xt.newaxis())) else k)) for k in keys)
        return Dual(xt.view(self.dev, *keys), self.host[np_keys])

    def transpose(self) -> "Dual":
        return Dual(xt.transpose(self.dev), self.host.T)


def oracle_for(*duals: Dual) -> Oracle:
    return Oracle({d.dev: d.host for d in duals})


def to_bf16_bits(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even) and return the bit patterns."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    nan = (u & 0x7FFFFFFF) > 0x7F800000
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    r[nan] = ((u[nan] >> 16) | 0x40).astype(np.uint16)
    return r.reshape(np.shape(x))


def bf16_bits_to_f32(b: np.ndarray) -> np.ndarray:
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def ulp_diff(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Distance in units in the last place between equal-typed float arrays."""
    a = np.asarray(a)
    b = np.asarray(b)
    if a.dtype == np.float64:
        ia = a.view(np.int64)
        ib = b.view(np.int64)
        ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
        ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
        return np.abs(ia - ib)
    ia = a.astype(np.float32).view(np.int32).astype(np.int64)
    ib = b.astype(np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -0x80000000 - ia, ia)
    ib = np.where(ib < 0, -0x80000000 - ib, ib)
    return np.abs(ia - ib)


def same_bits(a: np.ndarray, b: np.ndarray) -> bool:
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and a.tobytes() == b.tobytes()

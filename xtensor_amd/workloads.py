"""Synthetic workload generators shared by the benchmark and the tests.

Every generator is seeded and uses numpy's PCG64 `default_rng`, so a
workload is fixed by its parameters.
"""

from __future__ import annotations

import numpy as np


def log_uniform_int(rng: np.random.Generator, lo: int, hi: int, n: int) -> np.ndarray:
    """Integers in [lo, hi] whose logarithm is uniform."""
    x = np.exp(rng.uniform(np.log(lo), np.log(hi + 1), n))
    return np.clip(np.floor(x).astype(np.int64), lo, hi)


def uneven_shapes(n: int = 4096, seed: int = 3, max_rows: int = 512, max_cols: int = 2048) -> list:
    """Shapes (rows, cols) of the uneven batch: rows ~ log-uniform int
    [1, max_rows], cols ~ log-uniform int [1, max_cols]."""
    rng = np.random.default_rng(seed)
    rows = log_uniform_int(rng, 1, max_rows, n)
    cols = log_uniform_int(rng, 1, max_cols, n)
    return [(int(r), int(c)) for r, c in zip(rows, cols)]


def uneven_values(shapes, seed: int = 3, sigma: float = 2.0, dtype=np.float32) -> list:
    """Per-instance values ~ N(0, sigma) drawn after the shapes' stream."""
    rng = np.random.default_rng(seed + 1000)
    return [rng.normal(0.0, sigma, size=s).astype(dtype) for s in shapes]

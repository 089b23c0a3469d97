"""Synthetic inputs of the extended workloads (seeded, numpy default_rng)."""
from __future__ import annotations

import numpy as np


def batch_boxes(seed: int = 0, count: int = 64, lo: int = 16, hi: int = 96) -> np.ndarray:
    """(count, 3) box sides in elements, each uniform on [lo, hi]."""
    return np.random.default_rng(seed).integers(lo, hi, size=(count, 3), endpoint=True).astype(np.int32)


def kappa_field(num_elements: int, seed: int, sigma: float = 0.5) -> np.ndarray:
    """Per-element conductivity exp(z), z ~ N(0, sigma^2) i.i.d."""
    return np.exp(np.random.default_rng(seed).normal(0.0, sigma, num_elements))


def multi_rhs_sources(n: int, k: int = 16, seed: int = 1) -> np.ndarray:
    """(n, k) right-hand sides, i.i.d. N(0, 1) per node."""
    return np.random.default_rng(seed).standard_normal((n, k))

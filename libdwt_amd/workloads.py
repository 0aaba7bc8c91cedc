"""Deterministic synthetic inputs for the transform workloads.

All generators use numpy's PCG64 (``np.random.default_rng(seed)``) and are
evaluated in float64 before the final cast, so an input is a pure function of
its parameters and seed on every host.

* :func:`sinusoid_image` -- sum of 8 random 2D sinusoids (amplitude U[0,1],
  spatial frequencies U[0,0.25] cycles/pixel on each axis, phase U[0,2*pi))
  plus N(0, 0.05^2) pixel noise.
* :func:`random_walk` -- cumulative sums of N(0,1) steps, one signal per row.
* :func:`uneven_sizes` -- image sides drawn log-uniform in a range.
"""
from __future__ import annotations

import math

import numpy as np


def sinusoid_image(width: int, height: int, seed: int = 0, n_waves: int = 8,
                   noise: float = 0.05, dtype=np.float32) -> np.ndarray:
    """(height, width) image = sum_k A_k sin(2 pi (fx_k x + fy_k y) + phi_k) + noise."""
    rng = np.random.default_rng(seed)
    amp = rng.uniform(0.0, 1.0, n_waves)
    fx = rng.uniform(0.0, 0.25, n_waves)
    fy = rng.uniform(0.0, 0.25, n_waves)
    ph = rng.uniform(0.0, 2.0 * math.pi, n_waves)
    x = np.arange(width, dtype=np.float64)
    y = np.arange(height, dtype=np.float64)
    img = np.zeros((height, width), dtype=np.float64)
    for k in range(n_waves):
        ax = 2.0 * math.pi * fx[k] * x + ph[k]
        ay = 2.0 * math.pi * fy[k] * y
        # sin(ax + ay) = sin(ax) cos(ay) + cos(ax) sin(ay), as two outer products
        img += amp[k] * (np.cos(ay)[:, None] * np.sin(ax)[None, :]
                         + np.sin(ay)[:, None] * np.cos(ax)[None, :])
    img += rng.normal(0.0, noise, (height, width))
    return img.astype(dtype)


def random_walk(count: int, length: int, seed: int = 0, dtype=np.float32) -> np.ndarray:
    """(count, length) array; row i is a random walk of N(0,1) increments."""
    rng = np.random.default_rng(seed)
    steps = rng.standard_normal((count, length))
    return np.cumsum(steps, axis=1).astype(dtype)


def uneven_sizes(n: int, lo: int, hi: int, seed: int = 0) -> list[tuple[int, int]]:
    """n (width, height) pairs, each side log-uniform in [lo, hi] (integers)."""
    rng = np.random.default_rng(seed)
    ll, lh = math.log(lo), math.log(hi + 1)
    w = np.floor(np.exp(rng.uniform(ll, lh, n))).astype(int)
    h = np.floor(np.exp(rng.uniform(ll, lh, n))).astype(int)
    return [(int(min(max(a, lo), hi)), int(min(max(b, lo), hi))) for a, b in zip(w, h)]


def ceil_log2(x: int) -> int:
    """Smallest r with 2**r >= x (0 for x <= 1)."""
    r = 0
    while (1 << r) < x:
        r += 1
    return r


def level_sizes(width: int, height: int, levels: int) -> list[tuple[int, int]]:
    """Input (W_j, H_j) of each forward level j = 0..levels-1, W_j = ceil(W / 2^j)."""
    return [(-(-width // (1 << j)), -(-height // (1 << j))) for j in range(levels)]


def algorithmic_bytes(width: int, height: int, levels: int, elem: int = 4) -> int:
    """Per-level streaming model for one direction: sum_j (read + write) of the
    level's W_j x H_j input, elem bytes per coefficient."""
    return sum(2 * elem * w * h for w, h in level_sizes(width, height, levels))

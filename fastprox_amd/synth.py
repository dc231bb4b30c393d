"""Synthetic images of the benchmark configurations, with a fixed draw order
(numpy default_rng(seed)) so that every run sees the same arrays.

rectangles: each image is the sum of n axis-aligned rectangles (for each
  rectangle: the two row corners, then the two column corners, integers
  uniform in [0, extent], then the intensity U[0, 1]) plus i.i.d.
  N(0, sigma^2) noise drawn after the rectangles of that image.
disks: each image is the sum of n disks (centre row and column uniform in the
  frame, radius U[rmin, rmax], intensity U[0, 1]) plus N(0, sigma^2) noise.
boxes3d: rectangles in three dimensions.
"""

from __future__ import annotations

import hashlib

import numpy as np


def rectangles(B, shape, n=12, sigma=0.1, seed=0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    out = np.zeros((B,) + tuple(shape), np.float64)
    for b in range(B):
        img = out[b]
        for _ in range(n):
            sl = []
            for L in shape:
                lo, hi = np.sort(rng.integers(0, L + 1, 2))
                sl.append(slice(int(lo), int(hi)))
            img[tuple(sl)] += rng.uniform(0.0, 1.0)
        img += rng.normal(0.0, sigma, size=shape)
    return out.astype(dtype)


def disks(B, shape, n=20, rmin=8, rmax=96, sigma=0.15, seed=0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    H, W = shape
    yy, xx = np.mgrid[0:H, 0:W]
    out = np.zeros((B, H, W), np.float64)
    for b in range(B):
        for _ in range(n):
            cy, cx = rng.uniform(0, H), rng.uniform(0, W)
            r = rng.uniform(rmin, rmax)
            out[b][(yy - cy) ** 2 + (xx - cx) ** 2 <= r * r] += rng.uniform(0.0, 1.0)
        out[b] += rng.normal(0.0, sigma, size=(H, W))
    return out.astype(dtype)


def digest(a):
    """A short hash of an array's bytes, to record which inputs a run saw."""
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()[:16]

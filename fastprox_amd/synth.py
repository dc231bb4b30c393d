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


class Request:
    """One request of a serving trace: arrival time (s), dimension k, side,
    step, gap tolerance, and the seed that picks its source image and crop."""

    __slots__ = ("t", "k", "side", "gamma", "tol", "src")

    def __init__(self, t, k, side, gamma, tol, src):
        self.t, self.k, self.side, self.gamma, self.tol, self.src = t, k, side, gamma, tol, src

    def shape(self):
        return (self.side,) * self.k


def serving_trace(seconds=60.0, rate=200.0, seed=0, p2d=0.8, side2d=(128, 1024), side3d=(48, 128),
                  gammas=(0.01, 0.3), tols=(1e-4, 1e-5, 1e-6)):
    """Poisson arrivals at `rate` per second for `seconds`: each request is,
    with probability p2d, a square 2D image of side U[side2d] (integers),
    else a cube of side U[side3d]; gamma log-uniform in `gammas`; the gap
    tolerance uniform over `tols`. Draw order per request: the gap to the
    next arrival, the class, the side, gamma, the tolerance, the source."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    lg = (np.log(gammas[0]), np.log(gammas[1]))
    while True:
        t += float(rng.exponential(1.0 / rate))
        if t >= seconds:
            return out
        if rng.uniform() < p2d:
            k, side = 2, int(rng.integers(side2d[0], side2d[1] + 1))
        else:
            k, side = 3, int(rng.integers(side3d[0], side3d[1] + 1))
        gamma = float(np.exp(rng.uniform(lg[0], lg[1])))
        tol = float(tols[int(rng.integers(0, len(tols)))])
        src = int(rng.integers(0, 1 << 30))
        out.append(Request(t, k, side, gamma, tol, src))


def serving_sources(n2d=8, n3d=4, side2d=1024, side3d=128, seed=0, dtype=np.float32):
    """The source images a trace's requests crop from: 2D images of 16
    rectangles and 3D volumes of 16 boxes, both with N(0, 0.1^2) noise."""
    a = rectangles(n2d, (side2d, side2d), n=16, sigma=0.1, seed=seed, dtype=dtype)
    b = rectangles(n3d, (side3d,) * 3, n=16, sigma=0.1, seed=seed + 1, dtype=dtype)
    return a, b


def crop_of(req, src2d, src3d):
    """The request's crop: source src % n, offset from the seed's high bits
    along each axis (a view of the source array or tensor)."""
    pool = src2d if req.k == 2 else src3d
    img = pool[req.src % pool.shape[0]]
    sl = []
    h = req.src >> 4
    for a in range(req.k):
        room = img.shape[a] - req.side + 1
        off = h % room
        h //= room
        sl.append(slice(off, off + req.side))
    return img[tuple(sl)]

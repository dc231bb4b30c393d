"""The CPU oracle (oracle/tv_oracle.c) pinned against the reference: its
golden outputs (bitwise for isotropic TV in float64), the reference library
itself where it is built here, the reference's CVXPY fixtures, closed forms
and invariants."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import aniso_certificate, cvxpy_cases, golden, image, iso_certificate

G = golden()
ISO = sorted(k for k, v in G.items() if bool(v["iso"]))
ANISO = sorted(k for k, v in G.items() if not bool(v["iso"]))
ONE_D = {"iso_9x1", "iso_1x9", "iso_4x5_oneaxis"}  # single active axis: tv_iso.cpp:381-398


def _images(c):
    x = c["x"]
    k = len(c["w"])
    return x.reshape((-1,) + x.shape[-k:])


@pytest.mark.parametrize("name", ISO)
def test_iso_matches_reference_golden_bitwise(name):
    """tv_iso.cpp:333-543 restated: the same iterations, z, dual, gap and
    certified error as the reference library, to the bit (float64); for a
    float32 input the reference rounds its double iterate, so do we."""
    c = G[name]
    k = len(c["w"])
    xs = _images(c)
    B = xs.shape[0]
    gam = np.broadcast_to(c["gamma"], (B,))
    zr = c["z"].reshape(xs.shape)
    dr = c["dual"].reshape((k, B) + xs.shape[1:])
    warm = c.get("warm")
    for b in range(B):
        w0 = None if warm is None else warm.reshape((k, B) + xs.shape[1:])[:, b].astype(np.float64)
        r = oracle.tv_iso(xs[b].astype(np.float64), float(gam[b]), c["w"], alpha=c.get("spatial"),
                          tol=float(c["tol"]), max_iter=int(c["max_iter"]), warm=w0)
        assert r.iters == int(c["iters"][b])
        dt = xs.dtype
        if np.isnan(c["gaps"][b]):
            assert np.isnan(r.gap) and np.isnan(r.z).all() and np.isnan(r.dual).all()
            continue
        if name in ONE_D:  # the exact 1D prox (Condat): equal to rounding, not bits
            scale = 1 + np.abs(xs[b]).max()
            np.testing.assert_allclose(r.z.astype(dt), zr[b], rtol=0, atol=4e-16 * scale)
            np.testing.assert_allclose(r.dual.astype(dt), dr[:, b], rtol=0, atol=4e-15 * scale)
        else:
            np.testing.assert_array_equal(r.z.astype(dt), zr[b])
            np.testing.assert_array_equal(r.dual.astype(dt), dr[:, b])
        assert r.gap == c["gaps"][b]
        ze = oracle.zerr(r.gap / 4.0 ** r.scale_exp if r.scale_exp else r.gap, xs[b].astype(np.float64), dt,
                         r.scale_exp, r.gap / 4.0 ** r.scale_exp)
        assert ze == pytest.approx(float(c["zerrs"][b]), rel=1e-14, abs=0)
        assert r.converged == (int(c["status"]) == 0 or r.gap <= float(c["tol"]) * 0.5 * (xs[b] ** 2).sum())


def test_iso_matches_live_reference_on_random_problems():
    """Random shapes, weights, steps and tolerances against the reference
    library built from its sources (skipped where it is not)."""
    refbind = pytest.importorskip("refbind")
    if not refbind.available():
        pytest.skip("reference library not built here")
    rng = np.random.default_rng(7)
    for t in range(24):
        k = 2 if t % 4 else 3
        shape = tuple(int(v) for v in rng.integers(1, 14 if k == 3 else 40, size=k))
        w = rng.uniform(0.2, 2.0, size=k)
        if t % 5 == 0:
            w[rng.integers(k)] = 0.0
        gamma = float(rng.uniform(0.02, 1.5))
        tol = float(10 ** rng.uniform(-12, -4))
        x = image(shape, 100 + t, scale=float(rng.uniform(0.05, 1.0)))
        z, dual, gaps, its, zerrs, status, _ = refbind.Op(True, k, w).solve(x, gamma, tol=tol, max_iter=20000)
        r = oracle.tv_iso(x, gamma, w, tol=tol, max_iter=20000)
        assert r.iters == its[0], (shape, w, gamma, tol)
        if sum(1 for a in range(k) if w[a] > 0 and shape[a] > 1) == 1:  # the exact 1D branch
            np.testing.assert_allclose(r.z, z, rtol=0, atol=4e-16 * (1 + np.abs(x).max()))
            continue
        np.testing.assert_array_equal(r.z, z)
        np.testing.assert_array_equal(r.dual, dual)
        assert r.gap == gaps[0]


@pytest.mark.parametrize("name", ISO)
def test_iso_certificate_recomputed(name):
    """The gap recomputed from (z, dual) by the definitions equals the
    reported one; the dual is feasible and z = x - D^T W u."""
    c = G[name]
    k = len(c["w"])
    xs = _images(c).astype(np.float64)
    B = xs.shape[0]
    gam = np.broadcast_to(c["gamma"], (B,))
    for b in range(B):
        r = oracle.tv_iso(xs[b], float(gam[b]), c["w"], alpha=c.get("spatial"), tol=float(c["tol"]),
                          max_iter=int(c["max_iter"]))
        if np.isnan(r.gap) or r.scale_exp or name == "iso_huge_gamma":  # that gap is the clamped radius's
            continue
        g = iso_certificate(xs[b], r.z, r.dual, float(gam[b]), list(c["w"]), c.get("spatial"))
        assert g == pytest.approx(r.gap, rel=1e-6, abs=1e-12 * (1 + (xs[b] ** 2).sum()))


@pytest.mark.parametrize("name", ANISO)
def test_aniso_device_method_within_the_certificate(name):
    """The device method (_tv_torch.py:229-283) against the reference's ADBPG
    solution: |z - z_ref|^2 <= 2 (gap + gap_ref), both gaps recomputed here
    (the reference's own check, tests/python/test_device_tv.py:96-110)."""
    c = G[name]
    x = c["x"]
    w = list(c["w"])
    beta = c.get("spatial")
    r = oracle.tv_aniso_dev(x, float(c["gamma"]), w, beta=beta, tol=1e-10, max_iter=400000)
    assert r.converged
    g = aniso_certificate(x, r.z, r.dual, float(c["gamma"]), w, beta)
    assert g == pytest.approx(r.gap, rel=1e-6, abs=1e-12)
    # The reference's reported gap, as its own test uses it (its polished z
    # is set in closed form, so its dual is the iteration's, not x - z).
    gref = float(c["gaps"][0])
    assert ((r.z - c["z"]) ** 2).sum() <= 2 * (g + gref) * 1.0001 + 1e-20


CVX = cvxpy_cases()


@pytest.mark.parametrize("case", ["tv_iso_2d", "tv_iso_3d", "tv_iso_2d_pixel", "tv_iso_3d_pixel"])
def test_iso_matches_cvxpy(case):
    """The reference's fixture test (tests/python/test_fixtures.py:64-77):
    worst |z - z_cvx| / (1 + max|x|) < 1e-7 at tol 1e-14."""
    probs = [p for p in CVX if str(p["case"]) == case]
    assert probs
    worst = 0.0
    for p in probs:
        pr = p["param"]
        r = oracle.tv_iso(p["x"], float(p["gamma"]), pr["weights"], alpha=pr.get("pixel_weights"), tol=1e-14,
                          max_iter=1_000_000)
        assert r.converged or case.endswith("_pixel")
        worst = max(worst, np.abs(r.z - p["z"]).max() / (1 + np.abs(p["x"]).max()))
    assert worst < 1e-7


@pytest.mark.parametrize("case", ["tv_aniso_2d", "tv_aniso_3d", "tv_aniso_2d_edge", "tv_aniso_3d_edge"])
def test_aniso_device_method_matches_cvxpy(case):
    """The device method to a tight gap against the CVXPY solutions; the
    distance is certified by the gap (|z - z*| <= sqrt(2 gap))."""
    probs = [p for p in CVX if str(p["case"]) == case]
    assert probs
    for p in probs:
        pr = p["param"]
        beta = pr.get("edge_weights")
        if beta is not None:
            beta = beta.reshape((len(pr["weights"]),) + p["x"].shape)
        r = oracle.tv_aniso_dev(p["x"], float(p["gamma"]), pr["weights"], beta=beta, tol=1e-16,
                                max_iter=300_000)
        bound = np.sqrt(2 * max(r.gap, 0.0)) + 1e-7 * (1 + np.abs(p["x"]).max())
        assert np.sqrt(((r.z - p["z"]) ** 2).sum()) <= bound


def test_tv1d_kkt_and_closed_forms():
    """The exact 1D TV prox: the optimality conditions of prox_{lam TV}
    (u_i = sum_{j<=i} (x_j - z_j) in [-lam, lam], u_{n-1} = 0, u_i = -lam
    on an upward jump and +lam on a downward one) and closed forms."""
    rng = np.random.default_rng(3)
    for t in range(400):
        n = int(rng.integers(1, 50))
        lam = float(rng.choice([0.0, 1e-3, 0.1, 1.0, 10.0]))
        y = rng.normal(size=n) if t % 3 else np.repeat(rng.normal(size=3), 20)[:n]
        z = oracle.tv1d(y, lam)
        u = np.cumsum(y - z)
        tol = 1e-9 * (1 + np.abs(y).max())
        assert abs(u[-1]) <= tol
        assert (np.abs(u) <= lam + tol).all()
        d = np.diff(z)
        assert (np.abs(u[:-1][d > tol] + lam) <= tol).all()
        assert (np.abs(u[:-1][d < -tol] - lam) <= tol).all()
    assert np.array_equal(oracle.tv1d(np.full(7, 2.5), 3.0), np.full(7, 2.5))
    y = np.array([0.0, 0.0, 1.0, 1.0])
    np.testing.assert_allclose(oracle.tv1d(y, 10.0), np.full(4, 0.5), rtol=0, atol=1e-15)
    np.testing.assert_allclose(oracle.tv1d(y, 0.25), [0.125, 0.125, 0.875, 0.875], rtol=0, atol=1e-15)


def test_scaling_is_exact():
    """prox(2^k x, 2^k gamma) = 2^k prox(x, gamma) to the bit when the image
    is rescaled (image_util.hpp:42-65; tests/python/test_tv_scaling.py)."""
    x = image((12, 14), 5)
    base = oracle.tv_iso(x, 0.3, tol=1e-9, max_iter=100000)
    for k in (-100, 90):
        r = oracle.tv_iso(x * 2.0 ** k, 0.3 * 2.0 ** k, tol=1e-9, max_iter=100000)
        assert r.iters == base.iters and r.scale_exp != 0
        np.testing.assert_array_equal(r.z, base.z * 2.0 ** k)


def test_special_values():
    x = image((6, 7), 1)
    r = oracle.tv_iso(x, 0.0)
    assert np.array_equal(r.z, x) and r.iters == 0 and not r.dual.any()
    xn = x.copy()
    xn[2, 2] = np.nan
    r = oracle.tv_iso(xn, 0.2)
    assert np.isnan(r.z).all() and np.isnan(r.gap) and r.iters == 0
    xi = x.copy()
    xi[1, 1] = np.inf
    with pytest.raises(ValueError):
        oracle.tv_iso(xi, 0.2)
    with pytest.raises(ValueError):
        oracle.tv_aniso_dev(xi, 0.2)
    # Huge gamma: the constant mean (test_core.cpp:1453-1459).
    r = oracle.tv_iso(x, 1e8, tol=1e-12, max_iter=100000)
    np.testing.assert_allclose(r.z, np.full_like(x, x.mean()), atol=1e-6)


def test_certificate_catches_a_planted_mistake():
    """The independent certificate fails on an oracle output with a planted
    error (a dual entry with a flipped sign), so it can catch such bugs."""
    x = image((10, 12), 2)
    r = oracle.tv_iso(x, 0.3, tol=1e-10, max_iter=100000)
    bad = r.dual.copy()
    i = np.unravel_index(np.argmax(np.abs(bad[0])), bad[0].shape)
    bad[0][i] = -bad[0][i]
    with pytest.raises(AssertionError):
        iso_certificate(x, r.z, bad, 0.3, [1.0, 1.0])

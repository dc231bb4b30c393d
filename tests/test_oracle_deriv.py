"""The derivative oracle (oracle/tv_deriv.c) pinned against the reference:
its DERIVATIVE_BITS known answers (tests/python/test_tv_iso.py:187-211), the
reference library's JVP/VJP bits on the committed fixtures, adjointness and
central finite differences (test_tv_iso.py:104-130) for isotropic TV; for
the device method's anisotropic derivative (region means), adjointness,
finite differences and the closed form on piecewise-constant images."""

from __future__ import annotations

import hashlib

import numpy as np
import pytest

import oracle
from helpers import image, load_npz

D = load_npz("golden_deriv.npz")


def _bits(*arrs):
    h = hashlib.sha256()
    for a in arrs:
        h.update(np.ascontiguousarray(np.asarray(a, np.float64)).tobytes())
    return h.hexdigest()[:24]


# tests/python/test_tv_iso.py:192-195
DERIVATIVE_BITS = {
    "plain": ("5f753c866c11012cf8709f03", "a6f34516ead98a535fd3df17"),
    "weighted": ("4d081384f400a82aa2a80afd", "5af54739eee4d4b2f74f2471"),
}


@pytest.mark.parametrize("name", sorted(DERIVATIVE_BITS))
def test_derivative_bits(name):
    """The reference's known answer: the audit image D(64, 0), solved to 1e-9,
    its JVP and VJP hashed with the CG residual and iterations."""
    yy, xx = np.mgrid[0:64, 0:64] / 64
    x = ((xx - 0.5) ** 2 + (yy - 0.5) ** 2 < 0.1) + 0.5 * (xx > 0.7)
    x = x + 0.1 * np.random.default_rng(0).standard_normal((64, 64))
    dx = np.random.default_rng(1).standard_normal(x.shape)
    if name == "plain":
        w, al = [1.0, 1.0], None
    else:
        w, al = [1.0, 0.6], np.random.default_rng(2).uniform(0.5, 1.5, x.shape)
    r = oracle.tv_iso(x, 0.2, w, alpha=al, tol=1e-9, max_iter=10 ** 5)
    P = 2 + (x.size if al is not None else 0)
    dp = np.random.default_rng(3).standard_normal(P)
    j = oracle.iso_deriv(x, r.z, r.dual, r.gap, r.scale_exp, 0.2, w, al, dx=dx, dgamma=0.3, dparams=dp)
    v = oracle.iso_deriv(x, r.z, r.dual, r.gap, r.scale_exp, 0.2, w, al, g=dx)
    params = np.concatenate([v.gw, v.gspatial.reshape(-1)]) if al is not None else v.gw
    got = (_bits(j.out, [j.residual, j.iters]), _bits(v.out, v.ggamma, params, [v.residual, v.iters]))
    assert got == DERIVATIVE_BITS[name]


@pytest.mark.parametrize("name", sorted(D))
def test_iso_deriv_matches_reference_fixtures_bitwise(name):
    """At the reference's own kept solution (z, dual, gap), the JVP and VJP
    equal the reference library's to the bit, parameter gradients included."""
    c = D[name]
    x, gamma, w = c["x"], float(c["gamma"]), c["w"]
    al = c.get("alpha")
    r = oracle.tv_iso(x, gamma, w, alpha=al, tol=float(c["tol"]), max_iter=10 ** 6)
    assert r.iters == int(c["iters"])
    v = oracle.iso_deriv(x, c["z"], c["dual"], float(c["gap"]), r.scale_exp, gamma, w, al, g=c["g"])
    np.testing.assert_array_equal(v.out, c["gx"])
    assert v.ggamma == float(c["ggamma"])
    params = np.concatenate([v.gw, v.gspatial.reshape(-1)]) if al is not None else v.gw
    np.testing.assert_array_equal(params, c["gparams"])
    j = oracle.iso_deriv(x, c["z"], c["dual"], float(c["gap"]), r.scale_exp, gamma, w, al, dx=c["dx"],
                         dgamma=float(c["dgamma"]), dparams=c["dparams"])
    np.testing.assert_array_equal(j.out, c["jvp"])


def test_iso_deriv_matches_live_reference():
    """Random shapes, weights and tolerances against the reference library
    built from its sources (skipped where it is not)."""
    refbind = pytest.importorskip("refbind")
    if not refbind.available():
        pytest.skip("reference library not built here")
    rng = np.random.default_rng(17)
    for t in range(10):
        k = 3 if t % 4 == 3 else 2
        shape = tuple(int(v) for v in rng.integers(2, 10 if k == 3 else 60, size=k))
        w = rng.uniform(0.4, 1.6, size=k)
        al = rng.uniform(0.3, 1.7, size=shape) if t % 3 == 1 else None
        gamma = float(rng.uniform(0.05, 0.6))
        tol = float(10 ** rng.uniform(-10, -5))
        x = image(shape, 200 + t)
        op = refbind.Op(True, k, w, al)
        z, dual, gaps, its, _, _, ws = op.solve(x, gamma, tol=tol, max_iter=10 ** 6, keep=True)
        g = rng.standard_normal(shape)
        P = k + (x.size if al is not None else 0)
        dp = rng.standard_normal(P)
        gx, gg, gp = op.vjp(ws, g)
        jv = op.jvp(ws, g, 0.3, dp)
        refbind.Op.free_ws(ws)
        r = oracle.tv_iso(x, gamma, w, alpha=al, tol=tol, max_iter=10 ** 6)
        v = oracle.iso_deriv(x, z, dual, gaps[0], r.scale_exp, gamma, w, al, g=g)
        j = oracle.iso_deriv(x, z, dual, gaps[0], r.scale_exp, gamma, w, al, dx=g, dgamma=0.3, dparams=dp)
        np.testing.assert_array_equal(v.out, gx)
        assert v.ggamma == gg
        np.testing.assert_array_equal(np.concatenate([v.gw, v.gspatial.reshape(-1)]) if al is not None else v.gw, gp)
        np.testing.assert_array_equal(j.out, jv)


@pytest.mark.parametrize("shape", [(8, 9), (4, 4, 5)])
def test_iso_deriv_adjoint_and_finite_differences(shape):
    """tests/python/test_tv_iso.py:104-130 on the oracle: <g, J v> equals
    <J^T g, v> to 1e-8, and central differences agree to 3e-3."""
    rng = np.random.default_rng(5)
    k = len(shape)
    x = image(shape, 5)
    w = rng.uniform(0.6, 1.4, k)
    step = 0.4
    r = oracle.tv_iso(x, step, w, tol=1e-13, max_iter=3_000_000)
    dx, dw, dg = rng.normal(size=shape), rng.normal(size=k), 0.2
    g = rng.normal(size=shape)
    j = oracle.iso_deriv(x, r.z, r.dual, r.gap, r.scale_exp, step, w, dx=dx, dgamma=dg, dparams=dw,
                         deriv_tol=1e-10)
    assert j.residual <= 1e-10 and j.iters > 0
    v = oracle.iso_deriv(x, r.z, r.dual, r.gap, r.scale_exp, step, w, g=g, deriv_tol=1e-10)
    lhs = float((g * j.out).sum())
    rhs = float((v.out * dx).sum() + v.ggamma * dg + v.gw @ dw)
    assert abs(lhs - rhs) <= 1e-8 * (1 + abs(lhs))
    h = 1e-5
    zp = oracle.tv_iso(x + h * dx, step + h * dg, w + h * dw, tol=1e-13, max_iter=3_000_000).z
    zm = oracle.tv_iso(x - h * dx, step - h * dg, w - h * dw, tol=1e-13, max_iter=3_000_000).z
    fd = (zp - zm) / (2 * h)
    assert np.abs(fd - j.out).max() <= 3e-3 * max(1.0, np.abs(fd).max())


def test_iso_deriv_planted_mistake_fails_adjointness():
    """A transposed gradient (the VJP's gx of another cotangent) breaks the
    adjointness check, so that check can catch such mistakes."""
    rng = np.random.default_rng(8)
    x = image((9, 10), 8)
    r = oracle.tv_iso(x, 0.3, tol=1e-12, max_iter=10 ** 6)
    dx, g = rng.normal(size=x.shape), rng.normal(size=x.shape)
    j = oracle.iso_deriv(x, r.z, r.dual, r.gap, 0, 0.3, dx=dx, deriv_tol=1e-12)
    v = oracle.iso_deriv(x, r.z, r.dual, r.gap, 0, 0.3, g=g.T.copy().reshape(x.shape), deriv_tol=1e-12)
    assert abs(float((g * j.out).sum()) - float((v.out * dx).sum())) > 1e-6


def _aniso_u(x, gamma, w, beta=None, tol=1e-14):
    r = oracle.tv_aniso_dev(x, gamma, w, beta=beta, tol=tol, max_iter=2_000_000)
    return r, _q_to_u(r.dual, w)


def _q_to_u(q, w):
    """The per-edge dual from the C form q_a = w_a D_a^T u_a (the reverse
    cumulative sum, _tv_torch.py:594-606)."""
    k = q.shape[0]
    u = np.zeros_like(q)
    for a in range(k):
        if w[a] > 0:
            u[a] = -np.cumsum(q[a], axis=a) / w[a]
            sl = [slice(None)] * k
            sl[a] = slice(-1, None)
            u[a][tuple(sl)] = 0.0
    return u


@pytest.mark.parametrize("shape,beta", [((9, 11), False), ((4, 5, 6), False), ((8, 10), True)])
def test_aniso_deriv_adjoint_and_finite_differences(shape, beta):
    rng = np.random.default_rng(12)
    k = len(shape)
    x = image(shape, 12, scale=0.2)
    w = rng.uniform(0.6, 1.4, k)
    be = rng.uniform(0.5, 1.5, (k,) + shape) if beta else None
    gamma = 0.3
    r, u = _aniso_u(x, gamma, w, be)
    dx, g = rng.normal(size=shape), rng.normal(size=shape)
    dw, dg = rng.normal(size=k), 0.2
    db = rng.normal(size=(k,) + shape) if beta else None
    j = oracle.aniso_deriv(x, r.z, u, gamma, w, be, dx=dx, dgamma=dg, dw=dw, dbeta=db)
    v = oracle.aniso_deriv(x, r.z, u, gamma, w, be, g=g)
    lhs = float((g * j.out).sum())
    rhs = float((v.out * dx).sum() + v.ggamma * dg + v.gw @ dw)
    if beta:
        rhs += float((v.gspatial * db).sum())
    assert abs(lhs - rhs) <= 1e-9 * (1 + abs(lhs))
    h = 1e-6
    bp = None if be is None else be + h * db
    bm = None if be is None else be - h * db
    zp = oracle.tv_aniso_dev(x + h * dx, gamma + h * dg, w + h * dw, beta=bp, tol=1e-15, max_iter=3_000_000).z
    zm = oracle.tv_aniso_dev(x - h * dx, gamma - h * dg, w - h * dw, beta=bm, tol=1e-15, max_iter=3_000_000).z
    fd = (zp - zm) / (2 * h)
    assert np.abs(fd - j.out).max() <= 1e-3 * max(1.0, np.abs(fd).max())


def test_aniso_deriv_region_means_closed_form():
    """A two-level step image with a large gamma: every region is flat inside,
    so J^T g averages g over each region of z, and the region count matches
    the distinct values of z."""
    x = np.zeros((12, 14))
    x[:, 7:] = 1.0
    x += 0.01 * np.random.default_rng(3).standard_normal(x.shape)
    w = [1.0, 1.0]
    r, u = _aniso_u(x, 0.05, w)
    g = np.random.default_rng(4).standard_normal(x.shape)
    v = oracle.aniso_deriv(x, r.z, u, 0.05, w, g=g)
    assert v.regions == 2
    left, right = g[:, :7].mean(), g[:, 7:].mean()
    np.testing.assert_allclose(v.out[:, :7], left, rtol=0, atol=1e-14)
    np.testing.assert_allclose(v.out[:, 7:], right, rtol=0, atol=1e-14)


def test_iso_float32_dual_flat_rule():
    """A float32 dual sits on its sphere only to float32 rounding: the C
    kernels' margin (1e-9) then takes about half the sphere's pixels for
    strictly inside and merges regions across jumps, while the float32 rule
    (64 eps margins, _tv_torch.py:192-195, 404-410) keeps the float64
    solution's structure up to jumps below the float32 floor. Pinned against
    the float64 structure of the same problem: the float32-rounded point under
    the float32 rule gives nearly its regions and gradient, under the C rule
    not."""
    from fastprox_amd.synth import rectangles
    x = rectangles(1, (64, 72), n=10, sigma=0.1, seed=4, dtype=np.float64)[0]
    r = oracle.tv_iso(x, 0.1, tol=1e-10, max_iter=10 ** 6)
    g = np.random.default_rng(1).normal(size=x.shape)
    ref = oracle.iso_deriv(x, r.z, r.dual, r.gap, 0, 0.1, g=g, deriv_tol=1e-12, deriv_max_iter=100000)
    x32, z32, u32 = (v.astype(np.float32).astype(np.float64) for v in (x, r.z, r.dual))
    c_rule = oracle.iso_deriv(x32, z32, u32, r.gap, 0, 0.1, g=g, deriv_tol=1e-12, deriv_max_iter=100000)
    f32_rule = oracle.iso_deriv(x32, z32, u32, r.gap, 0, 0.1, g=g, deriv_tol=1e-12, deriv_max_iter=100000,
                                dual_dtype=np.float32)

    def rel(d):
        return np.abs(d.out - ref.out).max() / np.abs(ref.out).max()
    assert abs(f32_rule.regions - ref.regions) <= 0.05 * ref.regions
    assert rel(f32_rule) <= 0.05
    assert c_rule.regions < 0.5 * ref.regions and rel(c_rule) > 0.5
    # the default is restored: the float64 rule again on the float64 point
    again = oracle.iso_deriv(x, r.z, r.dual, r.gap, 0, 0.1, g=g, deriv_tol=1e-12, deriv_max_iter=100000)
    np.testing.assert_array_equal(again.out, ref.out)

"""Shared test helpers: the golden fixtures, test images, and an independent
numpy restatement of the certificates (the checks of the reference's
tests/python/test_device_tv.py:31-49, written from the definitions)."""

from __future__ import annotations

import pathlib

import numpy as np

GOLDEN = pathlib.Path(__file__).resolve().parent / "golden"


def load_npz(name):
    d = np.load(GOLDEN / name)
    out = {}
    for key in d.files:
        case, field = key.split("/", 1)
        out.setdefault(case, {})[field] = d[key]
    return out


def golden():
    return load_npz("golden_tv.npz")


def cvxpy_cases():
    probs = {}
    d = np.load(GOLDEN / "cvxpy_images.npz")
    for key in d.files:
        k, field = key.split("/", 1)
        probs.setdefault(k, {"param": {}})
        if field.startswith("param/"):
            probs[k]["param"][field[6:]] = d[key]
        else:
            probs[k][field] = d[key]
    return [probs[k] for k in sorted(probs)]


def image(shape, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(scale=scale, size=shape) + (np.indices(shape[-2:]).sum(0) // 4 % 2)


def fdiff(z, ax):
    """Forward difference along ax, zero at the last entry."""
    out = np.zeros_like(z)
    sl = [slice(None)] * z.ndim
    sl[ax] = slice(0, -1)
    out[tuple(sl)] = np.diff(z, axis=ax)
    return out


def fdiff_t(p, ax):
    """The adjoint of fdiff: out[i] = p[i-1] - p[i] (p[-1] = 0, p[n-1] unused)."""
    n = p.shape[ax]
    q = np.take(p, range(n - 1), axis=ax)
    pad_b = [(0, 0)] * p.ndim
    pad_a = [(0, 0)] * p.ndim
    pad_b[ax], pad_a[ax] = (1, 0), (0, 1)
    return np.pad(q, pad_b) - np.pad(q, pad_a)


def iso_certificate(x, z, u, gamma, w, alpha=None, rtol=1e-9):
    """The gap of (z, u) for isotropic TV of one image, from the
    definitions; asserts |u_p| <= gamma alpha_p and z = x - D^T W u."""
    k = len(w)
    r = gamma * (np.ones(x.shape) if alpha is None else alpha)
    un = np.sqrt((u ** 2).sum(0))
    assert (un <= r * (1 + rtol) + 1e-300).all(), float((un - r).max())
    zu = x - sum(w[a] * fdiff_t(u[a], a) for a in range(k))
    np.testing.assert_allclose(z, zu, atol=1e-9 * (1 + np.abs(x).max()))
    g = np.stack([w[a] * fdiff(z, a) for a in range(k)])
    return float((r * np.sqrt((g ** 2).sum(0)) - (g * u).sum(0)).sum())


def aniso_certificate(x, z, q, gamma, w, beta=None, rtol=1e-9):
    """The duality gap P(z) - D(q) of (z, q) for anisotropic TV, q in the C
    form (q_a = D_a^T p_a): P(z) = |z - x|^2 / 2 + gamma sum_a w_a beta |D_a
    z|, D(q) = <x, s> - |s|^2 / 2 with s = sum_a q_a; asserts the dual bounds
    |p_a| = |cumsum_a(q_a)| <= gamma w_a beta. When x - z = s (an iterate of
    the dual methods) this is gamma TV(z) - <z, x - z>; a polished point (the
    closed-form means on the regions read off q, tv_aniso.cpp:886-953) is
    certified by the same dual."""
    k = len(w)
    for a in range(k):
        p = -np.cumsum(q[a], axis=a)
        bound = gamma * w[a] * (np.ones(x.shape) if beta is None else beta[a])
        sl = [slice(None)] * x.ndim
        sl[a] = slice(0, -1)
        assert (np.abs(p[tuple(sl)]) <= bound[tuple(sl)] * (1 + rtol) + 1e-12).all()
    b = np.ones((k,) + x.shape) if beta is None else beta
    tv = sum(w[a] * (b[a] * np.abs(fdiff(z, a))).sum() for a in range(k))
    s = q.sum(0)
    return float(0.5 * ((z - x) ** 2).sum() + gamma * tv - ((x * s).sum() - 0.5 * (s * s).sum()))

"""Anisotropic TV by the reference CPU library's method on the device
(fpr_workspace_set_method, FPR_METHOD_ADBPG: exact 1D steps, ADBPG / FDPG,
gap schedule, restarts, polish; src/ops/tv_aniso.cpp:259-616, 886-953),
against the reference's own outputs: the golden vectors it wrote
(tests/golden/golden_tv.npz) and the reference library built from its sources
(oracle/_ref) on the same inputs.

The bar (float64): the same iteration counts, |z - z_ref|_inf <= 1e-9 and
the dual within 1e-9 (the 1D steps, block updates, dual and region means are
the reference's arithmetic; the gap, restart and region sums are FP64 trees
instead of its sequential sums, so results agree to rounding); float32
inputs are solved in FP64 as the reference does, so the float32 bar is the
same one rounded to float32 (|z - z_ref|_inf <= 1e-4 stated, far less seen)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import refbind
from fastprox_amd import Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.synth import disks, rectangles
from helpers import aniso_certificate, golden, image

pytestmark = pytest.mark.gpu

G = golden()
ZTOL = 1e-9


def _gaps_close(gaps, ref, x, k, tol):
    """Gaps agree within 1e-6 relative or 1/1000 of the image's stopping
    target (tol |x|^2 / 2): a gap near the rounding floor of its sums (the
    TV term less <z, d - z>, each of the size of |x|^2) differs from the
    reference's sequential sums by rounding alone."""
    xb = np.asarray(x, np.float64).reshape((-1,) + x.shape[x.ndim - k:])
    target = tol * 0.5 * (xb.reshape(len(xb), -1) ** 2).sum(axis=1)
    gaps, ref = np.asarray(gaps), np.asarray(ref)
    assert (np.abs(gaps - ref) <= np.maximum(1e-6 * np.abs(ref), 1e-3 * target)).all(), (gaps, ref, target)


def _zerrs_close(zerrs, ref, gaps, gaps_ref):
    """The certified error is sqrt(2 gap) / |x| (plus rounding): the
    reference's, moved by the ratio of the gaps."""
    for z, zr, g, gr in zip(zerrs, ref, gaps, gaps_ref):
        if gr > 0:
            assert z == pytest.approx(zr * np.sqrt(max(g, 0.0) / gr), rel=1e-6, abs=1e-15)


def _ours(x, gamma, w, tol, dtype=torch.float64, method="adbpg", max_iter=100000, warm=None):
    xt = torch.tensor(x, dtype=dtype, device="cuda")
    out = torch.empty_like(xt)
    k = len(w)
    ws = Workspace(Op(False, k, w), xt, method=method)
    assert ws.method() == method
    if warm is not None:
        ws.set_dual(torch.tensor(warm, dtype=dtype, device="cuda"))
        ws.set_warm(True)
    st = ws.prox(xt, gamma, out, tol=tol, max_iter=max_iter)
    B = int(np.prod(x.shape[:x.ndim - k])) if x.ndim > k else 1
    gaps, its, zerrs = ws.image_info(B)
    dual = torch.empty((k,) + tuple(xt.shape), dtype=dtype, device="cuda")
    ws.get_dual(dual)
    return st, out.cpu().numpy(), dual.cpu().numpy(), gaps, its, zerrs, ws


def _ref(x, gamma, w, tol, warm=None, max_iter=100000):
    z, dual, gaps, its, zerrs, st, _ = refbind.Op(False, len(w), w).solve(x, gamma, tol=tol, max_iter=max_iter,
                                                                          threads=16, warm=warm)
    return st, z, dual, gaps, its, zerrs


@pytest.mark.parametrize("name", ["aniso_16x18", "aniso_12x13_w", "aniso_3d"])
def test_golden_vectors(cuda, name):
    """The reference's own outputs for its test images: iterations, z, the
    dual, the gap and the certified error."""
    c = G[name]
    w = [float(v) for v in c["w"]]
    print(f"golden {name}", flush=True)
    st, z, q, gaps, its, zerrs, _ = _ours(c["x"], float(c["gamma"]), w, float(c["tol"]),
                                           max_iter=int(c["max_iter"]))
    assert int(its[0]) == int(c["iters"][0])
    assert np.abs(z - c["z"]).max() <= ZTOL
    assert np.abs(q - c["dual"]).max() <= ZTOL
    _gaps_close(gaps, c["gaps"], c["x"], len(w), float(c["tol"]))
    _zerrs_close(zerrs, c["zerrs"], gaps, c["gaps"])


def _cases():
    return [
        ((16, 18), [1.0, 1.0], 0.3, 1e-8),
        ((12, 13), [0.7, 1.4], 0.5, 1e-8),
        ((2, 57, 90), [1.0, 0.5], 0.05, 1e-8),
        ((130, 3), [1.0, 1.0], 0.2, 1e-8),
        ((3, 40, 33), [1.0, 0.5], 0.5, 1e-6),
        ((6, 7, 8), [1.0, 0.5, 1.2], 0.25, 1e-8),     # three active axes: FDPG
        ((2, 9, 10, 11), [1.0, 0.5, 1.2], 0.1, 1e-7),
        ((5, 9, 11), [1.0, 0.0, 0.7], 0.3, 1e-8),      # 3D, two active axes: ADBPG
        ((1, 20, 30), [1.0, 1.0, 1.0], 0.2, 1e-8),     # 3D with a unit extent
        ((4, 30, 31), [0.0, 1.0], 0.4, 1e-8),          # one active axis: one exact pass
        ((1, 300), [1.0, 1.0], 0.7, 1e-8),             # one active axis (a unit extent)
    ]


@pytest.mark.parametrize("case", range(len(_cases())))
def test_matches_the_reference_library(cuda, case):
    shape, w, gamma, tol = _cases()[case]
    k = len(w)
    x = image(shape, 300 + case)
    print(f"case {case}: {shape} w={w} gamma={gamma}", flush=True)
    st_r, z_r, q_r, g_r, it_r, ze_r = _ref(x, gamma, w, tol)
    st, z, q, gaps, its, zerrs, _ = _ours(x, gamma, w, tol)
    assert st == st_r
    np.testing.assert_array_equal(its, it_r)
    assert np.abs(z - z_r).max() <= ZTOL
    assert np.abs(q - q_r).max() <= ZTOL
    _gaps_close(gaps, g_r, x, k, tol)
    _zerrs_close(zerrs, ze_r, gaps, g_r)


@pytest.fixture(scope="module")
def c1():
    x = disks(4, (512, 512), n=20, rmin=8, rmax=96, sigma=0.15, seed=0, dtype=np.float32)
    return x, _ref(x.astype(np.float64), 0.05, [1.0, 0.5], 1e-6), _ref(x, 0.05, [1.0, 0.5], 1e-6)


def test_c1_float64(cuda, c1):
    """The anisotropic configuration's images (20 disks, noise 0.15, w = (1,
    0.5), gamma 0.05, tol 1e-6) at full size."""
    x, ref64, _ = c1
    print("C1 float64: 4 x 512^2 ADBPG", flush=True)
    st_r, z_r, q_r, g_r, it_r, _ = ref64
    st, z, q, gaps, its, _, ws = _ours(x.astype(np.float64), 0.05, [1.0, 0.5], 1e-6)
    assert st == A.FP_OK == st_r
    np.testing.assert_array_equal(its, it_r)
    assert np.abs(z - z_r).max() <= ZTOL
    assert np.abs(q - q_r).max() <= ZTOL
    _gaps_close(gaps, g_r, x, 2, 1e-6)
    for b in range(4):  # the certificate, restated independently
        g = aniso_certificate(x[b].astype(np.float64), z[b], q[:, b], 0.05, [1.0, 0.5])
        assert g == pytest.approx(gaps[b], rel=1e-6, abs=1e-12)


def test_c1_float32(cuda, c1):
    x, _, ref32 = c1
    print("C1 float32: 4 x 512^2 ADBPG", flush=True)
    st_r, z_r, q_r, g_r, it_r, _ = ref32
    st, z, q, gaps, its, _, _ = _ours(x, 0.05, [1.0, 0.5], 1e-6, dtype=torch.float32)
    assert st == A.FP_OK
    np.testing.assert_array_equal(its, it_r)
    assert np.abs(z.astype(np.float64) - z_r.astype(np.float64)).max() <= 1e-4
    assert np.abs(q.astype(np.float64) - q_r.astype(np.float64)).max() <= 1e-4
    _gaps_close(gaps, g_r, x, 2, 1e-6)


def test_ramps_take_the_taut_string(cuda):
    """Ramps with a large step make Condat rescan past its budget
    (tv1d.cpp:264): those fibres are solved by the taut string, as the
    reference's are."""
    n = 400
    r = np.linspace(0.0, 4.0, n)
    x = np.stack([np.add.outer(r, 0.3 * r), np.add.outer(0.5 * r, r[::-1])])
    x += 0.01 * np.random.default_rng(5).standard_normal(x.shape)
    print("ramps 2 x 400^2", flush=True)
    st_r, z_r, q_r, g_r, it_r, _ = _ref(x, 3.0, [1.0, 1.0], 1e-7)
    st, z, q, gaps, its, _, ws = _ours(x, 3.0, [1.0, 1.0], 1e-7)
    _, fallbacks = ws.polished(2)
    assert fallbacks > 0
    np.testing.assert_array_equal(its, it_r)
    assert np.abs(z - z_r).max() <= ZTOL
    assert np.abs(q - q_r).max() <= ZTOL


def test_warm_start_follows_the_reference(cuda):
    """A warm start from a dual (made feasible, then its point and polish
    tried first, tv_aniso.cpp:441-458, 630-675) on a nearby image."""
    x = image((2, 40, 50), 91)
    w = [1.0, 0.8]
    _, _, q0, _, _, _, _ = _ours(x, 0.3, w, 1e-5)
    x2 = x + 0.01 * np.random.default_rng(2).standard_normal(x.shape)
    st_r, z_r, q_r, g_r, it_r, _ = _ref(x2, 0.3, w, 1e-8, warm=q0)
    st, z, q, gaps, its, _, _ = _ours(x2, 0.3, w, 1e-8, warm=q0)
    np.testing.assert_array_equal(its, it_r)
    assert np.abs(z - z_r).max() <= ZTOL
    # an infeasible seed is projected: q_2 = x - mean(x), q_1 = 0 must not certify z = mean(x)
    bad = np.zeros((2,) + x.shape)
    bad[0] = x - x.mean(axis=(1, 2), keepdims=True)
    st_r, z_r, _, _, it_r, _ = _ref(x, 0.3, w, 1e-8, warm=bad)
    st, z, _, _, its, _, _ = _ours(x, 0.3, w, 1e-8, warm=bad)
    np.testing.assert_array_equal(its, it_r)
    assert np.abs(z - z_r).max() <= ZTOL


def test_scales_nan_inf_and_trivial_images(cuda):
    x = image((4, 20, 24), 17)
    x[1] *= 1e100  # solved at the scale 2^-e x (image_util.hpp:56-65)
    x[2] *= 1e-100
    x[3] = 0.0
    w = [1.0, 0.6]
    st_r, z_r, q_r, g_r, it_r, _ = _ref(x, 0.2, w, 1e-8)
    st, z, q, gaps, its, _, _ = _ours(x, 0.2, w, 1e-8)
    np.testing.assert_array_equal(its, it_r)
    for b in range(4):
        s = max(np.abs(x[b]).max(), 1e-300)
        assert np.abs(z[b] - z_r[b]).max() <= ZTOL * s
    _gaps_close(gaps, g_r, x, 2, 1e-8)
    xn = image((2, 10, 12), 3)
    xn[1, 3, 4] = np.nan
    st, z, q, gaps, its, _, _ = _ours(xn, 0.2, w, 1e-8)
    assert np.isnan(z[1]).all() and np.isnan(q[:, 1]).all() and np.isnan(gaps[1])
    assert not np.isnan(z[0]).any()
    xi = xn.copy()
    xi[1, 3, 4] = np.inf
    xt = torch.tensor(xi, device="cuda")
    ws = Workspace(Op(False, 2, w), xt)
    with pytest.raises(Exception):
        ws.prox(xt, 0.2, torch.empty_like(xt), tol=1e-8, max_iter=1000)
    z0 = _ours(x, 0.0, w, 1e-8)[1]
    np.testing.assert_array_equal(z0, x)


def test_torch_entry_point_methods(cuda):
    """fastprox_amd.torch.tv_aniso: the default is dual FISTA, the
    reference's device path (the same bits as a workspace's FISTA solve);
    method="adbpg" is the reference's CPU method (the same bits as a
    workspace's ADBPG solve, the workspaces' default)."""
    from fastprox_amd.torch import tv_aniso
    x = rectangles(3, (64, 80), n=6, sigma=0.1, seed=4, dtype=np.float64)
    xt = torch.tensor(x, device="cuda")
    z_f, info_f = tv_aniso(xt, 0.1, weights=[1.0, 0.5], tol=1e-7, max_iter=100000, return_info=True)
    z_a, info_a = tv_aniso(xt, 0.1, weights=[1.0, 0.5], tol=1e-7, max_iter=100000, return_info=True,
                           method="adbpg")
    _, zw_a, _, _, its_a, _, _ = _ours(x, 0.1, [1.0, 0.5], 1e-7)
    _, zw_f, _, _, its_f, _, _ = _ours(x, 0.1, [1.0, 0.5], 1e-7, method="fista")
    np.testing.assert_array_equal(z_a.cpu().numpy(), zw_a)
    np.testing.assert_array_equal(z_f.cpu().numpy(), zw_f)
    assert info_a.iters == int(its_a.max()) and info_f.iters == int(its_f.max())
    assert info_a.iters < info_f.iters
    ws = Workspace(Op(False, 2, [1.0, 0.5]), xt)
    assert ws.method() == "adbpg"


def test_derivatives_at_the_polished_solution(cuda):
    """The VJP at an ADBPG solution (regions read off its dual) against the
    reference library's fp_vjp at its own solution of the same image."""
    from fastprox_amd.torch import tv_aniso
    x = rectangles(2, (40, 48), n=5, sigma=0.05, seed=8, dtype=np.float64)
    xt = torch.tensor(x, device="cuda", requires_grad=True)
    z = tv_aniso(xt, 0.08, weights=[1.0, 0.7], tol=1e-10, max_iter=100000, method="adbpg")
    g = torch.tensor(np.random.default_rng(1).standard_normal(x.shape), device="cuda")
    (z * g).sum().backward()
    gx = xt.grad.cpu().numpy()
    ref = refbind.Op(False, 2, [1.0, 0.7])
    *_, ws = ref.solve(x, 0.08, tol=1e-10, max_iter=100000, threads=16, keep=True)
    gx_ref, _, _ = ref.vjp(ws, g.cpu().numpy())
    refbind.Op.free_ws(ws)
    assert np.abs(gx - gx_ref).max() <= 1e-8

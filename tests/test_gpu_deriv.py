"""The device derivatives (deriv.hip) against the derivative oracle
(oracle/tv_deriv.c, itself bit-exact with the reference's fp_jvp/fp_vjp) and
the reference's fixtures, through the C ABI (fp_vjp, fp_jvp, fp_vjp_device on
a keep_structure workspace) and through fpr_deriv (the PyTorch autograd
path). The oracle is given the device solve's own (z, dual, gap), so both
read the same structure; the CG sums differ only in their order, and the
results must agree to 1e-8 relative at a tight CG tolerance (the reference's
own bar for its device path against C: 1e-7, test_device_tv.py:113-141)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from fastprox_amd import Op, Workspace
from fastprox_amd.deriv import Structure
from fastprox_amd.torch import solve as tsolve
from fastprox_amd.torch import tv_aniso, tv_iso
from helpers import image, load_npz

pytestmark = pytest.mark.gpu

D = load_npz("golden_deriv.npz")


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(1.0, np.abs(np.asarray(b)).max()))


def _kept(x, gamma, w, alpha, tol, dtype=torch.float64, steps=None):
    k = len(w)
    xt = torch.tensor(x, dtype=dtype, device="cuda")
    op = Op(True, k, w, spatial=alpha)
    ws = Workspace(op, xt, keep_structure=True)
    out = torch.empty_like(xt)
    if steps is None:
        ws.prox(xt, gamma, out, tol=tol, max_iter=10 ** 6)
    else:
        ws.prox(xt, torch.tensor(steps, dtype=dtype, device="cuda"), out, tol=tol, max_iter=10 ** 6)
    B = int(np.prod(x.shape[:x.ndim - k]))
    dual = torch.empty((k,) + tuple(xt.shape), dtype=dtype, device="cuda")
    ws.get_dual(dual)
    gaps, its, _ = ws.image_info(B)
    return op, ws, xt, out, dual, gaps


@pytest.mark.parametrize("name", sorted(D))
def test_iso_vjp_jvp_float64_match_oracle(cuda, name):
    c = D[name]
    x, gamma, w = c["x"], float(c["gamma"]), c["w"]
    al = c.get("alpha")
    k = x.ndim
    print(f"case {name} shape {x.shape}", flush=True)
    op, ws, xt, z, dual, gaps = _kept(x, gamma, w, al, float(c["tol"]))
    zc, dc = z.cpu().numpy(), dual.cpu().numpy()
    e = oracle.tv_iso(x, gamma, w, alpha=al, tol=float(c["tol"]), max_iter=10 ** 6).scale_exp
    # the structure's inputs are the reference's (z to the bit, the 1D branch to rounding)
    if name == "d_oneaxis":
        np.testing.assert_allclose(zc, c["z"], rtol=0, atol=4e-16 * (1 + np.abs(x).max()))
    else:
        np.testing.assert_array_equal(zc, c["z"])
    P = k + (x.size if al is not None else 0)
    for dtol in (1e-12, 0.0):
        ws.set_deriv_tol(dtol, 100000)
        g = torch.tensor(c["g"], device="cuda")
        gx = torch.empty_like(g)
        gp = np.zeros(P)
        gg = ws.vjp(g, gx, gparams=gp)
        res, its = ws.deriv_info()
        ref = oracle.iso_deriv(x, zc, dc, gaps[0], e, gamma, w, al, g=c["g"], deriv_tol=dtol, deriv_max_iter=100000)
        refp = np.concatenate([ref.gw, ref.gspatial.reshape(-1)]) if al is not None else ref.gw
        if dtol:
            assert _rel(gx.cpu().numpy(), ref.out) <= 1e-8
            assert gg == pytest.approx(ref.ggamma, rel=1e-8, abs=1e-8)
            assert _rel(gp, refp) <= 1e-8
            assert res <= 1e-12
        else:
            # The default tolerance: the same stopping rule on the same
            # structure. CG's iterates drift apart with the rounding of the
            # reductions, so the stop can come a few iterations apart; both
            # results lie within the tolerance's error of the solution.
            r1, r2 = _rel(gx.cpu().numpy(), ref.out), _rel(gx.cpu().numpy(), c["gx"])
            print(f"default tol: iters {its} vs {ref.iters}, residual {res:.3e} vs {ref.residual:.3e}, "
                  f"rel diff {r1:.3e} (oracle) {r2:.3e} (reference)", flush=True)
            assert abs(its - ref.iters) <= max(3, ref.iters // 10)
            bound = 100 * max(ref.residual, 1e-10)
            assert r1 <= bound and r2 <= bound
        dx = torch.tensor(c["dx"], device="cuda")
        jv = torch.empty_like(dx)
        ws.jvp(dx, jv, dgamma=float(c["dgamma"]), dparams=np.asarray(c["dparams"]))
        rj = oracle.iso_deriv(x, zc, dc, gaps[0], e, gamma, w, al, dx=c["dx"], dgamma=float(c["dgamma"]),
                              dparams=c["dparams"], deriv_tol=dtol, deriv_max_iter=100000)
        if dtol:
            assert _rel(jv.cpu().numpy(), rj.out) <= 1e-8
        else:
            assert _rel(jv.cpu().numpy(), c["jvp"]) <= 100 * max(rj.residual, 1e-10)


def test_iso_vjp_device_and_adjointness(cuda):
    """fp_vjp_device writes the gamma gradient on the device; <g, J v> =
    <J^T g, v> (test_tv_iso.py:123-125) on a batch with per-image steps."""
    rng = np.random.default_rng(4)
    x = image((3, 40, 44), 31)
    steps = np.array([0.1, 0.25, 0.4])
    w = [1.0, 0.8]
    op, ws, xt, z, dual, gaps = _kept(x, 0.0, w, None, 1e-10, steps=steps)
    ws.set_deriv_tol(1e-12, 100000)
    g = torch.tensor(rng.normal(size=x.shape), device="cuda")
    dx = torch.tensor(rng.normal(size=x.shape), device="cuda")
    dw = rng.normal(size=2)
    gx = torch.empty_like(g)
    gp = np.zeros(2)
    gg_dev = torch.zeros((), dtype=torch.float64, device="cuda")
    ws.vjp(g, gx, gparams=gp, ggamma_dev=gg_dev)
    gg = ws.vjp(g, torch.empty_like(g))
    assert float(gg_dev) == gg
    jv = torch.empty_like(dx)
    ws.jvp(dx, jv, dgamma=0.3, dparams=dw)
    lhs = float((g * jv).sum())
    rhs = float((gx * dx).sum()) + gg * 0.3 + float(gp @ dw)
    assert abs(lhs - rhs) <= 1e-8 * (1 + abs(lhs))
    # per image against the oracle
    zc, dc = z.cpu().numpy(), dual.cpu().numpy()
    for b in range(3):
        r = oracle.iso_deriv(x[b], zc[b], dc[:, b], gaps[b], 0, steps[b], w, g=g[b].cpu().numpy(),
                             deriv_tol=1e-12, deriv_max_iter=100000)
        assert _rel(gx[b].cpu().numpy(), r.out) <= 1e-8


def test_iso_float32_structure_matches_oracle(cuda):
    """float32 solves: the derivative at the device's own float32 (z, dual,
    gap) equals the oracle's at the same point under the float32 flat rule
    (the same flat set and regions, CG in float64), to float32 rounding."""
    x = image((2, 64, 72), 33).astype(np.float32)
    op, ws, xt, z, dual, gaps = _kept(x, 0.12, [1.0, 1.0], None, 1e-6, dtype=torch.float32)
    ws.set_deriv_tol(1e-11, 100000)
    rng = np.random.default_rng(2)
    g = rng.normal(size=x.shape).astype(np.float32)
    gx = torch.empty_like(xt)
    ws.vjp(torch.tensor(g, device="cuda"), gx)
    zc, dc = z.cpu().numpy().astype(np.float64), dual.cpu().numpy().astype(np.float64)
    for b in range(2):
        r = oracle.iso_deriv(x[b].astype(np.float64), zc[b], dc[:, b], gaps[b], 0, 0.12, g=g[b].astype(np.float64),
                             deriv_tol=1e-11, deriv_max_iter=100000, dual_dtype=np.float32)
        assert _rel(gx[b].cpu().numpy(), r.out) <= 2e-6


def test_iso_deriv_repeats_bitwise(cuda):
    """Fixed-order sums: two VJPs of one structure, and of a rebuilt one, are
    bit-identical."""
    x = image((2, 96, 80), 35)
    op, ws, xt, z, dual, gaps = _kept(x, 0.08, [1.0, 1.0], None, 1e-8)
    g = torch.tensor(np.random.default_rng(5).normal(size=x.shape), device="cuda")
    outs = []
    for _ in range(2):
        gx = torch.empty_like(g)
        ws.vjp(g, gx)
        outs.append(gx.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    S1 = Structure(xt, z, dual, gaps, np.zeros(2, np.int32), torch.tensor([0.08, 0.08], dtype=torch.float64),
                   [1.0, 1.0], True)
    gx2 = S1.vjp(g, 0.0, 10000)[0]
    np.testing.assert_array_equal(gx2.cpu().numpy(), outs[0])


def _q_to_u(q, w):
    k = q.shape[0]
    u = np.zeros_like(q)
    for a in range(k):
        if w[a] > 0:
            u[a] = -np.cumsum(q[a], axis=a) / w[a]
            sl = [slice(None)] * k
            sl[a] = slice(-1, None)
            u[a][tuple(sl)] = 0.0
    return u


@pytest.mark.parametrize("shape,beta", [((40, 36), False), ((6, 12, 10), False), ((30, 28), True)])
def test_aniso_deriv_matches_oracle(cuda, shape, beta):
    rng = np.random.default_rng(21)
    k = len(shape)
    x = image((2,) + shape, 21, scale=0.2)
    w = list(rng.uniform(0.6, 1.4, k))
    gamma = torch.tensor([0.2, 0.35], dtype=torch.float64, device="cuda")
    be = rng.uniform(0.5, 1.5, (k, 2) + shape) if beta else None
    xt = torch.tensor(x, device="cuda")
    bet = None if be is None else torch.tensor(be, device="cuda")
    z, dual, gaps, its, zerrs, conv, exps, ue = tsolve(xt, gamma, w, False, 1e-12, 10 ** 6, sw=bet, keep=True)
    S = Structure(xt, z, ue, gaps, exps, gamma, w, False, bet)
    g = torch.tensor(rng.normal(size=x.shape), device="cuda")
    gx, gg, gw, gsw, _ = S.vjp(g, 0.0, 100)
    dx = torch.tensor(rng.normal(size=x.shape), device="cuda")
    dw = rng.normal(size=k)
    dsw = torch.tensor(rng.normal(size=(k, 2) + shape), device="cuda") if beta else None
    dz, _ = S.jvp(dx, torch.tensor([0.3, -0.1], dtype=torch.float64, device="cuda"), dw, 0.0, 100, dsw)
    zc, uc = z.cpu().numpy(), ue.cpu().numpy()
    tot_gw = np.zeros(k)
    for b in range(2):
        bb = None if be is None else be[:, b]
        v = oracle.aniso_deriv(x[b], zc[b], uc[b], float(gamma[b]), w, bb, g=g[b].cpu().numpy())
        assert _rel(gx[b].cpu().numpy(), v.out) <= 1e-12
        assert float(gg[b]) == pytest.approx(v.ggamma, rel=1e-10, abs=1e-12)
        tot_gw += v.gw
        if beta:
            assert _rel(gsw[:, b].cpu().numpy(), v.gspatial) <= 1e-12
        j = oracle.aniso_deriv(x[b], zc[b], uc[b], float(gamma[b]), w, bb, dx=dx[b].cpu().numpy(),
                               dgamma=[0.3, -0.1][b], dw=dw, dbeta=None if not beta else dsw[:, b].cpu().numpy())
        assert _rel(dz[b].cpu().numpy(), j.out) <= 1e-12
    np.testing.assert_allclose(gw.numpy(), tot_gw, rtol=1e-10, atol=1e-12)
    # the C form's dual, summed back, gives the same per-edge dual to rounding
    assert np.abs(_q_to_u(dual[:, 0].cpu().numpy(), w) - uc[0]).max() <= 1e-12


@pytest.mark.parametrize("kernel", [(0, 2, 32), (0, 6, 32), (0, 8, 16), (2, 2, 16)])
def test_kept_iterate_is_the_output_dual(cuda, kernel):
    """The stop can land inside a launch of several iterations: the kept dual
    iterate (copy_dual_iterate) is then the one the output dual and z come
    from, not the launch's start."""
    from fastprox_amd.torch import Batch
    rng = np.random.default_rng(8)
    shape = (44, 70)
    x = image((6,) + shape, 8, scale=0.2)
    w = [0.9, 1.2]
    op = Op(False, 2, w)
    bt = Batch(op, torch.float64, [shape] * 6, kernel=kernel)
    xs = [torch.tensor(v, device="cuda") for v in x]
    zs = [torch.empty_like(v) for v in xs]
    ds = [torch.empty((2,) + shape, dtype=torch.float64, device="cuda") for _ in xs]
    gam = list(rng.uniform(0.05, 0.3, 6))
    bt.solve(xs, zs, ds, gam, tol=1e-11, max_iter=100000)
    ue = torch.empty((6, 2) + shape, dtype=torch.float64, device="cuda")
    bt.copy_dual_iterate(ue)
    for b in range(6):
        assert np.abs(_q_to_u(ds[b].cpu().numpy(), w) - ue[b].cpu().numpy()).max() <= 1e-12


@pytest.mark.parametrize("method", ["fista", "adbpg"])
def test_aniso_workspace_vjp_matches_structure(cuda, method):
    """fp_vjp on an anisotropic keep_structure workspace reads the same
    per-edge dual as the autograd path, for either method."""
    x = image((2, 30, 34), 41, scale=0.2)
    w = [1.0, 0.6]
    xt = torch.tensor(x, device="cuda")
    op = Op(False, 2, w)
    ws = Workspace(op, xt, keep_structure=True, method=method)
    z = torch.empty_like(xt)
    ws.prox(xt, 0.3, z, tol=1e-10, max_iter=10 ** 6)
    g = torch.tensor(np.random.default_rng(1).normal(size=x.shape), device="cuda")
    gx = torch.empty_like(g)
    gp = np.zeros(2)
    gg = ws.vjp(g, gx, gparams=gp)
    gam = torch.tensor([0.3, 0.3], dtype=torch.float64, device="cuda")
    zz, dual, gaps, its, zerrs, conv, exps, ue = tsolve(xt, gam, w, False, 1e-10, 10 ** 6, keep=True, method=method)
    np.testing.assert_array_equal(zz.cpu().numpy(), z.cpu().numpy())
    S = Structure(xt, zz, ue, gaps, exps, gam, w, False)
    gx2, gg2, gw2, _, _ = S.vjp(g, 0.0, 100)
    np.testing.assert_array_equal(gx.cpu().numpy(), gx2.cpu().numpy())
    assert gg == pytest.approx(float(gg2.sum()), rel=1e-12)
    np.testing.assert_allclose(gp, gw2.numpy(), rtol=1e-12)


def test_torch_backward_iso(cuda):
    """tv_iso's autograd backward against the oracle's VJP at the same
    solution (C3's shape in small), and gamma's gradient summed over the
    batch."""
    rng = np.random.default_rng(7)
    x0 = image((2, 3, 32, 30), 7)
    x = torch.tensor(x0, device="cuda", requires_grad=True)
    gamma = torch.tensor(0.1, dtype=torch.float64, device="cuda", requires_grad=True)
    z, info = tv_iso(x, gamma, tol=1e-10, max_iter=10 ** 6, return_info=True, deriv_tol=1e-12)
    c = torch.tensor(rng.normal(size=x0.shape), device="cuda")
    (z * c).sum().backward()
    xb = x0.reshape(6, 32, 30)
    zb = z.detach().cpu().numpy().reshape(6, 32, 30)
    du = info.dual.detach().cpu().numpy().reshape(2, 6, 32, 30)
    cb = c.cpu().numpy().reshape(6, 32, 30)
    gsum = 0.0
    for b in range(6):
        r = oracle.iso_deriv(xb[b], zb[b], du[:, b], info.gaps[b], 0, 0.1, g=cb[b], deriv_tol=1e-12,
                             deriv_max_iter=100000)
        assert _rel(x.grad.cpu().numpy().reshape(6, 32, 30)[b], r.out) <= 1e-8
        gsum += r.ggamma
    assert float(gamma.grad) == pytest.approx(gsum, rel=1e-8)


def test_torch_jvp_and_finite_differences(cuda):
    """Forward mode (torch.autograd.forward_ad) against central differences of
    the solve (test_tv_iso.py:126-130: 3e-3)."""
    import torch.autograd.forward_ad as fwAD
    rng = np.random.default_rng(9)
    x0 = image((2, 10, 11), 9)
    dx0 = rng.normal(size=x0.shape)
    x = torch.tensor(x0, device="cuda")
    dx = torch.tensor(dx0, device="cuda")
    with fwAD.dual_level():
        xd = fwAD.make_dual(x, dx)
        z = tv_iso(xd, 0.4, tol=1e-14, max_iter=3 * 10 ** 6, deriv_tol=1e-12)
        jv = fwAD.unpack_dual(z).tangent
    h = 1e-5
    zp = tv_iso(x + h * dx, 0.4, tol=1e-14, max_iter=3 * 10 ** 6)
    zm = tv_iso(x - h * dx, 0.4, tol=1e-14, max_iter=3 * 10 ** 6)
    fd = ((zp - zm) / (2 * h)).cpu().numpy()
    assert np.abs(fd - jv.cpu().numpy()).max() <= 3e-3 * max(1.0, np.abs(fd).max())


def test_torch_backward_aniso_weights(cuda):
    """tv_aniso with edge weights and axis weights requiring grad: the
    gradients equal the oracle's region-mean VJP."""
    rng = np.random.default_rng(13)
    x0 = image((2, 16, 18), 13, scale=0.2)
    be0 = rng.uniform(0.5, 1.5, (2, 2, 16, 18))
    x = torch.tensor(x0, device="cuda", requires_grad=True)
    wt = torch.tensor([1.0, 0.7], dtype=torch.float64, device="cuda", requires_grad=True)
    be = torch.tensor(be0, device="cuda", requires_grad=True)
    z = tv_aniso(x, 0.25, weights=wt, edge_weights=be, tol=1e-12, max_iter=10 ** 6)
    c = torch.tensor(rng.normal(size=x0.shape), device="cuda")
    (z * c).sum().backward()
    xt = torch.tensor(x0, device="cuda")
    zz, dual, gaps, its, zerrs, conv, exps, ue = tsolve(xt, torch.tensor([0.25, 0.25], dtype=torch.float64,
                                                                          device="cuda"), [1.0, 0.7],
                                                        False, 1e-12, 10 ** 6, sw=torch.tensor(be0, device="cuda"),
                                                        keep=True)
    gw = np.zeros(2)
    for b in range(2):
        v = oracle.aniso_deriv(x0[b], zz[b].cpu().numpy(), ue[b].cpu().numpy(), 0.25, [1.0, 0.7], be0[:, b],
                               g=c[b].cpu().numpy())
        assert _rel(x.grad[b].cpu().numpy(), v.out) <= 1e-12
        assert _rel(be.grad[:, b].cpu().numpy(), v.gspatial) <= 1e-12
        gw += v.gw
    np.testing.assert_allclose(wt.grad.cpu().numpy(), gw, rtol=1e-10, atol=1e-12)


def test_refine_then_derivatives(cuda):
    """fp_workspace_refine (fastprox.h:759-771): the kept call solved again from
    its input and dual to a tighter tolerance is the warm-started solve of the
    oracle to the bit (float64); the derivatives then read the refined point,
    and the first call's z is untouched."""
    x = image((2, 40, 46), 71)
    op, ws, xt, z, dual, gaps = _kept(x, 0.15, [1.0, 1.0], None, 1e-4)
    z_first = z.clone()
    zr = torch.empty_like(z)
    st, gap, its = ws.refine(1e-11, 10 ** 6, out=zr)
    assert st == 0
    assert torch.equal(z, z_first)
    dual_loose = dual.cpu().numpy()
    gaps2, its2, _ = ws.image_info(2)
    d2 = torch.empty_like(dual)
    ws.get_dual(d2)
    for b in range(2):
        r = oracle.tv_iso(x[b], 0.15, tol=1e-11, max_iter=10 ** 6, warm=dual_loose[:, b])
        assert r.iters == its2[b]
        np.testing.assert_array_equal(zr[b].cpu().numpy(), r.z)
        np.testing.assert_array_equal(d2[:, b].cpu().numpy(), r.dual)
    ws.set_deriv_tol(1e-12, 100000)
    g = torch.tensor(np.random.default_rng(3).normal(size=x.shape), device="cuda")
    gx = torch.empty_like(g)
    ws.vjp(g, gx)
    zc, dc = zr.cpu().numpy(), d2.cpu().numpy()
    for b in range(2):
        v = oracle.iso_deriv(x[b], zc[b], dc[:, b], gaps2[b], 0, 0.15, g=g[b].cpu().numpy(), deriv_tol=1e-12,
                             deriv_max_iter=100000)
        assert _rel(gx[b].cpu().numpy(), v.out) <= 1e-8


def test_torch_backward_float32_near_float64(cuda):
    """The float32 autograd path (float32 solve and dual, the float32 flat
    rule) against the float64 one on the same images: the gradients of x and
    gamma agree to a few per cent (the float32 point merges only jumps below
    its rounding floor). The C kernels' 1e-9 margin on a float32 dual would
    merge regions across jumps and miss by the gradient's size."""
    from fastprox_amd.synth import rectangles
    x = rectangles(2, (64, 72), n=10, sigma=0.1, seed=4, dtype=np.float64)
    g = np.random.default_rng(3).normal(size=x.shape)
    out = {}
    for dt, tol in ((torch.float32, 1e-7), (torch.float64, 1e-10)):
        xt = torch.tensor(x, dtype=dt, device="cuda", requires_grad=True)
        gam = torch.tensor(0.1, dtype=dt, device="cuda", requires_grad=True)
        z = tv_iso(xt, gam, tol=tol, max_iter=10 ** 6, deriv_tol=1e-10, deriv_max_iter=100000)
        (z * torch.tensor(g, dtype=dt, device="cuda")).sum().backward()
        out[dt] = (xt.grad.double().cpu().numpy(), float(gam.grad))
    gx32, gg32 = out[torch.float32]
    gx64, gg64 = out[torch.float64]
    assert np.abs(gx32 - gx64).max() <= 0.05 * np.abs(gx64).max()
    assert abs(gg32 - gg64) <= 0.05 * abs(gg64)


def test_c3_vjp_against_central_differences_and_the_reference(cuda):
    """The differentiable configuration at its size: x (16, 3, 256, 256) from
    the rectangle generator with noise 0.1, gamma 0.1, loss mean((z -
    clean)^2). The VJP in x and gamma (float64, tight tolerances) against
    float64 central differences (h = 1e-6, 32 random directions) and against
    the reference library's fp_vjp at the same solution."""
    from fastprox_amd.synth import rectangles
    rng = np.random.default_rng(31)
    shape = (16, 3, 256, 256)
    x0 = rectangles(48, shape[2:], n=12, sigma=0.1, seed=0, dtype=np.float64).reshape(shape)
    clean = torch.tensor(rectangles(48, shape[2:], n=12, sigma=0.0, seed=0, dtype=np.float64).reshape(shape),
                         device="cuda")
    tol = 1e-12

    def loss(x, g, warm=None):
        z, info = tv_iso(x, g, tol=tol, max_iter=10 ** 6, deriv_tol=1e-12, deriv_max_iter=100000, warm_start=warm,
                         return_info=True)
        return ((z - clean) ** 2).mean(), z, info

    x = torch.tensor(x0, device="cuda", requires_grad=True)
    gam = torch.tensor(0.1, dtype=torch.float64, device="cuda", requires_grad=True)
    L0, z, info = loss(x, gam)
    L0.backward()
    gx, gg = x.grad.detach(), float(gam.grad)
    dual = info.dual.detach()
    # The perturbed solves start from the solution's dual (a warm start), so
    # both sides of a difference carry nearly the same solve error.
    h = 1e-6
    errs, ans = [], []
    with torch.no_grad():
        for i in range(32):
            v = torch.tensor(rng.normal(size=shape), device="cuda")
            fd = (float(loss(x + h * v, gam, dual)[0]) - float(loss(x - h * v, gam, dual)[0])) / (2 * h)
            an = float((gx * v).sum())
            errs.append(abs(fd - an))
            ans.append(abs(an))
        fdg = (float(loss(x, gam + h, dual)[0]) - float(loss(x, gam - h, dual)[0])) / (2 * h)
    # Errors on the scale of the directional derivatives: random directions
    # give |<gx, v>| over a range of 40x here, and the prox is only piecewise
    # smooth (a perturbation of 3 M pixels can cross a kink somewhere, which
    # central differences see and the solution's Jacobian does not); the
    # reference's own bar on small images is 3e-3 (test_tv_iso.py:126-130).
    worst = max(errs) / max(ans)
    print(f"worst FD gap over 32 directions {worst:.2e} of max |<gx, v>| (|<gx, v>| {min(ans):.3e}.."
          f"{max(ans):.3e}); gamma {gg:.6e} vs FD {fdg:.6e}", flush=True)
    assert worst <= 3e-3
    assert abs(gg - fdg) <= 1e-3 * abs(fdg)
    # The reference's VJP at the same solution (the same bits of z), on four
    # of the images (each image's derivative is its own; the reference's CPU
    # solve of all 48 to this tolerance takes minutes): the cotangent of the
    # loss, x's gradient and gamma's for those images.
    refbind = pytest.importorskip("refbind")
    if not refbind.available():
        pytest.skip("the reference library is not built")
    xs = x0.reshape((48,) + shape[2:])[:4].copy()
    gzs = ((2.0 / z.numel()) * (z.detach() - clean)).reshape((48,) + shape[2:])[:4].cpu().numpy()
    print("reference solve of 4 images", flush=True)
    rop = refbind.Op(True, 2)
    zr, dual, gaps, its, zerrs, st, rws = rop.solve(xs, 0.1, tol=1e-10, max_iter=10 ** 6, threads=4, keep=True)
    try:
        op, ws, xt, zd, dd, gd = _kept(xs, 0.1, [1.0, 1.0], None, 1e-10)
        np.testing.assert_array_equal(zd.cpu().numpy(), zr)
        print("reference VJP", flush=True)
        rgx, rgg, _ = rop.vjp(rws, gzs, deriv_tol=1e-12, deriv_max_iter=100000)
        ws.set_deriv_tol(1e-12, 100000)
        gxd = torch.empty_like(xt)
        ggd = ws.vjp(torch.tensor(gzs, device="cuda"), gxd)
        assert _rel(gxd.cpu().numpy(), rgx) <= 1e-7
        assert ggd == pytest.approx(rgg, rel=1e-7)
    finally:
        refbind.Op.free_ws(rws)


@pytest.mark.parametrize("shape", [(1, 70), (70, 1), (2, 65), (63, 3), (5, 125), (130, 9)])
def test_iso_vjp_thin_and_strip_edge_shapes(cuda, shape):
    """The fused 2D apply walks 62-column strips down 8-row segments: images a
    pixel wide or tall (one active axis), two or three pixels across, and
    widths just past a strip (63, 65, 125, 130) give the oracle's VJP, with
    per-image steps and a tight CG tolerance."""
    rng = np.random.default_rng(sum(shape))
    x = image((2,) + shape, 40 + shape[0], scale=0.3)
    steps = np.array([0.12, 0.3])
    w = [1.0, 0.7]
    op, ws, xt, z, dual, gaps = _kept(x, 0.0, w, None, 1e-10, steps=steps)
    ws.set_deriv_tol(1e-12, 100000)
    g = torch.tensor(rng.normal(size=x.shape), device="cuda")
    gx = torch.empty_like(g)
    ws.vjp(g, gx)
    zc, dc = z.cpu().numpy(), dual.cpu().numpy()
    for b in range(2):  # values of order 1: no power-of-two rescaling (scale exponent 0)
        r = oracle.iso_deriv(x[b], zc[b], dc[:, b], gaps[b], 0, steps[b], w, g=g[b].cpu().numpy(),
                             deriv_tol=1e-12, deriv_max_iter=100000)
        assert _rel(gx[b].cpu().numpy(), r.out) <= 1e-8, (shape, b)

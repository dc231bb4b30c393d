"""The HIP path against the oracle and the reference's golden outputs, through
the C ABI (fp_prox, fp_prox_steps, the workspace dual and results, and the
uneven batch API). float64 results must have the reference's bits (same
iterations, z and dual); float32 ones must lie within 1e-4 of the float64
reference and meet the gap tolerance."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import oracle
from fastprox_amd import FastproxError, Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.synth import disks, rectangles
from fastprox_amd.torch import Batch
from helpers import aniso_certificate, fdiff, fdiff_t, golden, image, iso_certificate

pytestmark = pytest.mark.gpu

G = golden()
ISO = sorted(k for k, v in G.items() if bool(v["iso"]))
ONE_D = {"iso_9x1", "iso_1x9", "iso_4x5_oneaxis"}
# (kind, tt, rows): 0 streaming, 1 generic, 2 block
KERNELS_2D = [None, (0, 2, 64), (0, 4, 13), (0, 6, 64), (0, 8, 40), (1, 1, 4)]


def _run(c, dev, kernel=None, dtype=torch.float64):
    k = len(c["w"])
    x = torch.tensor(c["x"], dtype=dtype, device=dev)
    out = torch.empty_like(x)
    op = Op(True, k, c["w"], spatial=c.get("spatial"))
    ws = Workspace(op, x)
    if kernel is not None:
        ws.set_kernel(*kernel)
    if "warm" in c:
        ws.set_dual(torch.tensor(c["warm"], dtype=dtype, device=dev))
    g = c["gamma"]
    if np.ndim(g):
        st = ws.prox(x, torch.tensor(g, dtype=dtype, device=dev), out, tol=float(c["tol"]),
                     max_iter=int(c["max_iter"]))
    else:
        st = ws.prox(x, float(g), out, tol=float(c["tol"]), max_iter=int(c["max_iter"]))
    B = int(np.prod(c["x"].shape[:c["x"].ndim - k]))
    dual = torch.empty((k,) + tuple(x.shape), dtype=dtype, device=dev)
    ws.get_dual(dual)
    gaps, its, zerrs = ws.image_info(B)
    return st, out.cpu().numpy(), dual.cpu().numpy(), gaps, its, zerrs, ws


# every float64 fixture on the default kernel; the 2D ones also on each tiled
# kernel (the tiled kernels are 2D; float32 fixtures are checked below)
BITWISE = [(n, kern) for n in ISO if G[n]["x"].dtype == np.float64
           for kern in (KERNELS_2D if len(G[n]["w"]) == 2 else [None])]


@pytest.mark.parametrize("name,kernel", BITWISE,
                         ids=[f"{n}-{'default' if kern is None else kern}" for n, kern in BITWISE])
def test_iso_golden_float64_bitwise(cuda, name, kernel):
    c = G[name]
    st, z, dual, gaps, its, zerrs, _ = _run(c, cuda, kernel)
    assert st == int(c["status"])
    np.testing.assert_array_equal(its, c["iters"])
    ok = ~np.isnan(c["gaps"])
    if name in ONE_D:
        scale = 1 + np.abs(c["x"]).max()
        np.testing.assert_allclose(z, c["z"], rtol=0, atol=4e-16 * scale)
        np.testing.assert_allclose(dual, c["dual"], rtol=0, atol=4e-15 * scale)
    else:
        np.testing.assert_array_equal(z, c["z"])
        np.testing.assert_array_equal(dual, c["dual"])
    # the gaps are the same sums in another order
    np.testing.assert_allclose(gaps[ok], c["gaps"][ok], rtol=1e-9, atol=0)
    np.testing.assert_allclose(zerrs[ok], c["zerrs"][ok], rtol=1e-9, atol=0)
    assert np.isnan(gaps[~ok]).all() and np.isnan(zerrs[~ok]).all()


@pytest.mark.parametrize("kernel", [None, (0, 4, 64), (0, 8, 64), (2, 4, 32), (1, 1, 4)],
                         ids=lambda k: "default" if k is None else f"{k}")
def test_iso_float32_within_1e4_of_the_reference(cuda, kernel):
    """float32 iterates against the reference's float64 solve at the same
    tolerance: |z - z_ref|_inf <= 1e-4 and the gap at most the target."""
    x = rectangles(3, (96, 120), seed=3, dtype=np.float32)
    xt = torch.tensor(x, device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(True, 2), xt)
    if kernel is not None:
        ws.set_kernel(*kernel)
    assert ws.prox(xt, 0.1, out, tol=1e-6, max_iter=100000) == A.FP_OK
    gaps, its, _ = ws.image_info(3)
    z = out.cpu().numpy()
    for b in range(3):
        r = oracle.tv_iso(x[b].astype(np.float64), 0.1, tol=1e-6, max_iter=100000)
        assert np.abs(z[b] - r.z).max() <= 1e-4
        assert abs(int(its[b]) - r.iters) <= 8
        assert gaps[b] <= 1e-6 * 0.5 * (x[b].astype(np.float64) ** 2).sum()


F32 = [n for n in ISO if G[n]["x"].dtype == np.float32]


@pytest.mark.parametrize("kernel", KERNELS_2D, ids=lambda k: "default" if k is None else f"{k}")
@pytest.mark.parametrize("name", F32)
def test_iso_golden_float32_within_1e4(cuda, name, kernel):
    """The reference's float32 fixtures (its solve promotes to float64 and
    rounds z to float32), solved in float32 on the device: the same status,
    |z - z_ref|_inf <= 1e-4, a feasible dual whose recomputed gap meets the
    target, and an iteration count near the reference's."""
    c = G[name]
    k = len(c["w"])
    assert k == 2
    st, z, dual, gaps, its, _, _ = _run(c, cuda, kernel, dtype=torch.float32)
    assert st == int(c["status"])
    np.testing.assert_allclose(z, c["z"], rtol=0, atol=1e-4)
    xs = c["x"].reshape((-1,) + c["x"].shape[-k:]).astype(np.float64)
    zs = z.reshape(xs.shape).astype(np.float64)
    us = np.moveaxis(dual.reshape((k,) + xs.shape), 0, 1).astype(np.float64)
    gamma, w = float(c["gamma"]), c["w"]
    for b in range(xs.shape[0]):
        target = float(c["tol"]) * 0.5 * (xs[b] ** 2).sum()
        assert gaps[b] <= target
        # the certificate of the rounded dual, in float64: z(u) = x - D^T W u
        # and the gap of (z(u), u)
        u = us[b]
        assert (np.sqrt((u ** 2).sum(0)) <= gamma * (1 + 1e-6)).all()
        zu = xs[b] - sum(w[a] * fdiff_t(u[a], a) for a in range(k))
        assert np.abs(zs[b] - zu).max() <= 1e-5
        g = np.stack([w[a] * fdiff(zu, a) for a in range(k)])
        cert = float((gamma * np.sqrt((g ** 2).sum(0)) - (g * u).sum(0)).sum())
        assert cert <= 2 * target
        assert abs(int(its[b]) - int(c["iters"][b])) <= 12


def _aniso_cases():
    rng = np.random.default_rng(11)
    out = []
    for shape, w, gamma in [((16, 18), [1.0, 1.0], 0.3), ((12, 13), [0.7, 1.4], 0.5), ((2, 57, 90), [1.0, 0.5], 0.05),
                            ((130, 3), [1.0, 1.0], 0.2), ((6, 7, 8), [1.0, 0.5, 1.2], 0.25)]:
        out.append((shape, w, gamma, None))
    be = rng.uniform(0.2, 2.0, size=(2, 11, 13))
    be[0, 4, 5] = 0.0
    out.append(((11, 13), [1.0, 0.7], 0.4, be))
    be3 = rng.uniform(0.2, 2.0, size=(3, 5, 6, 7))
    out.append(((5, 6, 7), [1.0, 1.0, 0.5], 0.3, be3))
    return out


@pytest.mark.parametrize("case", range(len(_aniso_cases())))
def test_aniso_float64_bitwise_against_the_oracle(cuda, case):
    shape, w, gamma, beta = _aniso_cases()[case]
    k = len(w)
    x = image(shape, 40 + case)
    xt = torch.tensor(x, device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(False, k, w, spatial=beta), xt, method="fista")
    ws.prox(xt, gamma, out, tol=1e-8, max_iter=100000)
    B = int(np.prod(shape[:len(shape) - k]))
    gaps, its, _ = ws.image_info(B)
    dual = torch.empty((k,) + shape, dtype=torch.float64, device=cuda)
    ws.get_dual(dual)
    img = shape[len(shape) - k:]
    z, q = out.cpu().numpy().reshape((B,) + img), dual.cpu().numpy().reshape((k, B) + img)
    xb = x.reshape((B,) + img)
    for b in range(B):
        r = oracle.tv_aniso_dev(xb[b], gamma, w, beta=beta, tol=1e-8, max_iter=100000)
        assert its[b] == r.iters
        np.testing.assert_array_equal(z[b], r.z)
        np.testing.assert_array_equal(q[:, b], r.dual)
        assert gaps[b] == pytest.approx(r.gap, rel=1e-9, abs=1e-15)


@pytest.mark.parametrize("name", sorted(k for k, v in G.items() if not bool(v["iso"])))
def test_aniso_within_the_certificate_of_the_reference(cuda, name):
    """The device method against the reference's ADBPG solution:
    |z - z_ref|^2 <= 2 (gap + gap_ref) (test_device_tv.py:96-110)."""
    c = G[name]
    w = list(c["w"])
    xt = torch.tensor(c["x"], device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(False, len(w), w, spatial=c.get("spatial")), xt)
    assert ws.prox(xt, float(c["gamma"]), out, tol=1e-10, max_iter=400000) == A.FP_OK
    dual = torch.empty((len(w),) + tuple(xt.shape), dtype=torch.float64, device=cuda)
    ws.get_dual(dual)
    z, q = out.cpu().numpy(), dual.cpu().numpy()
    g = aniso_certificate(c["x"], z, q, float(c["gamma"]), w, c.get("spatial"))
    gaps, _, _ = ws.image_info(1)
    assert g == pytest.approx(gaps[0], rel=1e-6, abs=1e-12)
    assert ((z - c["z"]) ** 2).sum() <= 2 * (g + float(c["gaps"][0])) * 1.0001 + 1e-20


def test_uneven_batch_each_image_as_alone(cuda):
    """One persistent solve over images of different shapes, steps and pixel
    weights: every image has the oracle's bits, and the same bits as when it
    is solved alone."""
    rng = np.random.default_rng(5)
    shapes = [(37, 64), (64, 37), (5, 130), (100, 101), (1, 40), (48, 48)]
    xs, als, gams = [], [], []
    for i, s in enumerate(shapes):
        xs.append(image(s, 60 + i))
        al = rng.lognormal(0.0, 0.3, size=s)
        al[rng.random(s) < 0.05] = 0.0
        als.append(al)
        gams.append(float(np.exp(rng.uniform(np.log(0.02), np.log(0.3)))))
    op = Op(True, 2)
    bt = Batch(op, torch.float64, shapes, weighted=True)
    xt = [torch.tensor(x, device=cuda) for x in xs]
    zt = [torch.empty_like(x) for x in xt]
    dt = [torch.empty((2,) + x.shape, dtype=torch.float64, device=cuda) for x in xt]
    wt = [torch.tensor(a, device=cuda) for a in als]
    assert bt.solve(xt, zt, dt, gams, weights=wt, tol=1e-8, max_iter=100000) == A.FP_OK
    gaps, its, zerrs, conv = bt.image_info()
    for i, s in enumerate(shapes):
        r = oracle.tv_iso(xs[i], gams[i], alpha=als[i], tol=1e-8, max_iter=100000)
        assert its[i] == r.iters, (i, s)
        np.testing.assert_array_equal(zt[i].cpu().numpy(), r.z)
        np.testing.assert_array_equal(dt[i].cpu().numpy(), r.dual)
        one = Batch(op, torch.float64, [s], weighted=True)
        z1, d1 = torch.empty_like(xt[i]), torch.empty_like(dt[i])
        one.solve([xt[i]], [z1], [d1], [gams[i]], weights=[wt[i]], tol=1e-8, max_iter=100000)
        assert torch.equal(z1, zt[i]) and torch.equal(d1, dt[i])
        assert one.image_info()[0][0] == gaps[i]


def test_strided_aliased_and_host_arrays(cuda):
    x = image((3, 40, 50), 7)
    ref = oracle.tv_iso(x[1], 0.2, tol=1e-8, max_iter=100000)
    op = Op(True, 2)
    # a transposed view: the images of x.T's last two axes
    xt = torch.tensor(np.ascontiguousarray(np.swapaxes(x, 1, 2)), device=cuda).transpose(1, 2)
    out = torch.empty((3, 50, 40), dtype=torch.float64, device=cuda).transpose(1, 2)
    ws = Workspace(op, xt)
    ws.prox(xt, 0.2, out, tol=1e-8, max_iter=100000)
    np.testing.assert_array_equal(out[1].cpu().numpy(), ref.z)
    # out aliasing x
    xa = torch.tensor(x, device=cuda)
    ws2 = Workspace(op, xa)
    ws2.prox(xa, 0.2, xa, tol=1e-8, max_iter=100000)
    np.testing.assert_array_equal(xa[1].cpu().numpy(), ref.z)
    # host arrays through the same C ABI
    xh = x.copy()
    zh = np.empty_like(xh)
    ws3 = Workspace(op, xh)
    ws3.prox(xh, 0.2, zh, tol=1e-8, max_iter=100000)
    np.testing.assert_array_equal(zh[1], ref.z)


def test_warm_start_follows_the_reference(cuda):
    c = G["iso_warm"]
    x = torch.tensor(c["x"], device=cuda)
    out = torch.empty_like(x)
    ws = Workspace(Op(True, 2), x)
    ws.prox(x, float(c["gamma"]), out, tol=1e-4, max_iter=100000)
    ws.set_warm(1)
    ws.prox(x, float(c["gamma"]), out, tol=float(c["tol"]), max_iter=100000)
    _, its, _ = ws.image_info(1)
    assert its[0] == int(c["iters"][0])
    np.testing.assert_array_equal(out.cpu().numpy(), c["z"])


def test_special_values_and_errors(cuda):
    op = Op(True, 2)
    x = torch.tensor(image((2, 9, 10), 1), device=cuda)
    x[1, 2, 3] = float("inf")
    out = torch.empty_like(x)
    ws = Workspace(op, x)
    with pytest.raises(FastproxError) as e:
        ws.prox(x, 0.2, out)
    assert e.value.status == A.FP_EINVAL
    with pytest.raises(FastproxError) as e:
        ws.prox(x, -1.0, out)
    assert e.value.status == A.FP_EPARAM
    with pytest.raises(FastproxError) as e:
        ws.prox(x, 0.2, out, tol=-1.0)
    assert e.value.status == A.FP_EPARAM
    empty = torch.empty((0, 5, 6), dtype=torch.float64, device=cuda)
    wse = Workspace(op, empty)
    assert wse.prox(empty, 0.2, torch.empty_like(empty)) == A.FP_OK
    y = torch.tensor(image((12, 12), 2), device=cuda)
    wsy = Workspace(op, y)
    assert wsy.prox(y, 0.5, torch.empty_like(y), tol=1e-14, max_iter=17) == A.FP_ENOTCONV
    _, iters, ok = wsy.info()
    assert iters == 17 and not ok


def _full_size_check(x, z, dual, gamma, w, iso, tol, gaps):
    xd = x.astype(np.float64)
    for b in range(x.shape[0]):
        if iso:
            g = iso_certificate(xd[b], z[b].astype(np.float64), dual[:, b].astype(np.float64), gamma, w,
                                rtol=1e-6 if z.dtype == np.float32 else 1e-9)
        else:
            q = dual[:, b].astype(np.float64)
            g = aniso_certificate(xd[b], xd[b] - q.sum(0), q, gamma, w, rtol=1e-5 if z.dtype == np.float32 else 1e-9)
        target = tol * 0.5 * (xd[b] ** 2).sum()
        assert g <= target * 1.001 + 1e-6 * target if z.dtype == np.float64 else g <= 1.5 * target + 1e-3 * target
        assert gaps[b] <= target


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_full_size_iso_certificate(cuda, dtype):
    """Four 512x512 images (the benchmark's shape): the gap recomputed in
    float64 from (z, dual) meets the tolerance, the dual is feasible and
    z = x - D^T W u."""
    x = rectangles(4, (512, 512), seed=0, dtype=dtype)
    xt = torch.tensor(x, device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(True, 2), xt)
    assert ws.prox(xt, 0.1, out, tol=1e-6, max_iter=100000) == A.FP_OK
    dual = torch.empty((2,) + x.shape, dtype=xt.dtype, device=cuda)
    ws.get_dual(dual)
    gaps, _, _ = ws.image_info(4)
    _full_size_check(x, out.cpu().numpy(), dual.cpu().numpy(), 0.1, [1.0, 1.0], True, 1e-6, gaps)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_full_size_aniso_certificate(cuda, dtype):
    """Anisotropic 512x512 disk images with axis weights (1, 0.5), gamma
    0.05 (the anisotropic configuration's images), certified by the gap."""
    x = disks(2, (512, 512), seed=1, dtype=dtype)
    xt = torch.tensor(x, device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(False, 2, [1.0, 0.5]), xt)
    assert ws.prox(xt, 0.05, out, tol=1e-6, max_iter=100000) == A.FP_OK
    dual = torch.empty((2,) + x.shape, dtype=xt.dtype, device=cuda)
    ws.get_dual(dual)
    gaps, _, _ = ws.image_info(2)
    _full_size_check(x, out.cpu().numpy(), dual.cpu().numpy(), 0.05, [1.0, 0.5], False, 1e-6, gaps)


def test_runs_are_reproducible(cuda):
    x = torch.tensor(rectangles(3, (200, 170), seed=9, dtype=np.float32), device=cuda)
    ws = Workspace(Op(True, 2), x)
    a, b = torch.empty_like(x), torch.empty_like(x)
    ws.prox(x, 0.1, a, tol=1e-6, max_iter=100000)
    ws.prox(x, 0.1, b, tol=1e-6, max_iter=100000)
    assert torch.equal(a, b)


@pytest.mark.parametrize("iso", [True, False])
def test_3d_float64_bitwise(cuda, iso):
    x = image((2, 9, 11, 13), 77)
    w = [1.0, 0.6, 1.3]
    xt = torch.tensor(x, device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(iso, 3, w), xt, method=None if iso else "fista")
    ws.prox(xt, 0.2, out, tol=1e-8, max_iter=100000)
    _, its, _ = ws.image_info(2)
    for b in range(2):
        r = (oracle.tv_iso if iso else oracle.tv_aniso_dev)(x[b], 0.2, w, tol=1e-8, max_iter=100000)
        assert its[b] == r.iters
        np.testing.assert_array_equal(out[b].cpu().numpy(), r.z)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_tail_kernel_switch_keeps_results(cuda, dtype):
    """The switch to the tail kernel's tiling (fpr_workspace_set_tail) in the
    middle of a solve changes no bit: images with different steps finish at
    different times, so the switch happens while others still iterate."""
    x = torch.tensor(image((6, 150, 170), 50), dtype=dtype, device=cuda)
    steps = torch.tensor([0.02, 0.05, 0.1, 0.2, 0.4, 0.8], dtype=dtype, device=cuda)
    outs = []
    for tail in [(4, 32, 0.0), (4, 16, 0.9), (8, 8, 0.5), (6, 40, 0.99)]:
        op = Op(True, 2)
        ws = Workspace(op, x)
        ws.set_kernel(0, 4, 32)
        ws.set_tail(*tail)
        z = torch.empty_like(x)
        ws.prox(x, steps, z, tol=1e-7, max_iter=100000)
        dual = torch.empty((2,) + tuple(x.shape), dtype=dtype, device=cuda)
        ws.get_dual(dual)
        gaps, its, _ = ws.image_info(6)
        outs.append((z.cpu(), dual.cpu(), its))
    for z, dual, its in outs[1:]:
        assert torch.equal(z, outs[0][0]) and torch.equal(dual, outs[0][1])
        np.testing.assert_array_equal(its, outs[0][2])
    with pytest.raises(FastproxError):  # fewer iterations per launch than the main kernel
        Workspace(Op(True, 2), x).set_tail(2, 16, 0.5)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_graph_replays_match_eager_launches(cuda, dtype):
    """Launch chunks replayed from captured hipGraphs (fpr_batch_set_graph)
    give the eager launches' bits on every replay, with eager work of the
    program and synchronisations between the replays, through a tail switch
    and across solves with other inputs."""
    shapes = [(150, 170), (64, 300), (200, 90), (33, 40)]
    rng = np.random.default_rng(71)
    xs = [torch.tensor(image((1,) + s, 70 + i)[0], dtype=dtype, device=cuda) for i, s in enumerate(shapes)]
    gam = [0.05, 0.2, 0.1, 0.4]

    def solve(bt, xl):
        zs = [torch.empty_like(v) for v in xl]
        ds = [torch.empty((2,) + tuple(v.shape), dtype=dtype, device=cuda) for v in xl]
        bt.solve(xl, zs, ds, gam, tol=1e-7, max_iter=100000)
        return [z.cpu() for z in zs], [d.cpu() for d in ds], bt.image_info()[1]

    eager = Batch(Op(True, 2), dtype, shapes, kernel=(0, 4, 32))
    eager.set_tail(8, 8, 0.6)
    eager.set_graph(False)
    graph = Batch(Op(True, 2), dtype, shapes, kernel=(0, 4, 32))
    graph.set_tail(8, 8, 0.6)
    for rep in range(4):
        xl = [v * (1.0 + 0.25 * rep) for v in xs]
        want = solve(eager, xl)
        junk = torch.tensor(rng.normal(size=(256, 256)), dtype=dtype, device=cuda)
        (junk @ junk).sum().item()  # the program's own eager work between replays
        got = solve(graph, xl)
        torch.cuda.synchronize()
        for a, b in zip(want[0] + want[1], got[0] + got[1]):
            assert torch.equal(a, b), f"replay {rep}"
        np.testing.assert_array_equal(want[2], got[2])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("iso", [True, False])
def test_3d_streaming_kernel_matches_generic(cuda, dtype, iso):
    """The 3D streaming kernel (slabs of slices, LDS rows, DPP columns) runs
    the generic kernel's per-voxel arithmetic: the same iterations, z and
    dual, on shapes that cut its tiles every way (thin axes, an inactive
    axis, pixel weights)."""
    rng = np.random.default_rng(60)
    cases = [((3, 37, 29, 70), [1.0, 0.7, 1.3], None), ((2, 1, 9, 130), [1.0, 1.0, 1.0], None),
             ((2, 20, 8, 63), [0.0, 1.0, 0.5], None), ((1, 33, 15, 62), [1.0, 1.0, 1.0], "sw")]
    for shape, w, sw in cases:
        x = torch.tensor(image(shape, 61), dtype=dtype, device=cuda)
        spatial = None
        if sw:
            spatial = rng.uniform(0.5, 1.5, shape[1:] if iso else (3,) + shape[1:])
        outs = []
        for kernel in [(1, 1, 4), (3, 1, 32), (3, 1, 5)]:
            op = Op(iso, 3, w, spatial=spatial)
            ws = Workspace(op, x)
            ws.set_kernel(*kernel)
            z = torch.empty_like(x)
            ws.prox(x, 0.15, z, tol=1e-8, max_iter=100000)
            dual = torch.empty((3,) + tuple(x.shape), dtype=dtype, device=cuda)
            ws.get_dual(dual)
            gaps, its, _ = ws.image_info(shape[0])
            outs.append((z.cpu(), dual.cpu(), its))
        for z, dual, its in outs[1:]:
            np.testing.assert_array_equal(its, outs[0][2])
            assert torch.equal(z, outs[0][0]) and torch.equal(dual, outs[0][1])


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_split_batch_matches_per_image_solves(cuda, dtype):
    """A batch of more than 2048 tiles runs as two halves on two streams
    (Engine::kSplitTiles): every image has the bits of its own plain solve on
    the same kernel, through a tail switch, and the kept duals and results
    are in image order."""
    n, shape = 120, (96, 128)
    rng = np.random.default_rng(77)
    x = torch.tensor(rectangles(n, shape, n=6, sigma=0.05, seed=77), dtype=dtype, device=cuda)
    gam = list(np.exp(rng.uniform(np.log(0.02), np.log(0.3), n)))
    bt = Batch(Op(True, 2), dtype, [shape] * n, kernel=(0, 2, 4))
    bt.set_tail(4, 4, 0.5)
    zs = [torch.empty_like(v) for v in x]
    ds = [torch.empty((2,) + shape, dtype=dtype, device=cuda) for _ in range(n)]
    bt.solve(list(x), zs, ds, gam, tol=1e-7, max_iter=100000)
    gaps, its, zerrs, conv = bt.image_info()
    for i in range(0, n, 7):
        ws = Workspace(Op(True, 2), x[i:i + 1].contiguous())
        ws.set_kernel(0, 2, 4)
        ws.set_tail(4, 4, 0.5)
        z = torch.empty_like(x[i:i + 1])
        ws.prox(x[i:i + 1].contiguous(), float(gam[i]), z, tol=1e-7, max_iter=100000)
        g1, i1, _ = ws.image_info(1)
        d1 = torch.empty((2, 1) + shape, dtype=dtype, device=cuda)
        ws.get_dual(d1)
        assert int(its[i]) == int(i1[0]), f"image {i}"
        assert torch.equal(zs[i], z[0]) and torch.equal(ds[i], d1[:, 0])

"""float32 parity of the kernels the benchmark configurations run, against the
reference library (oracle/_ref, built from the reference's sources; it
promotes float32 input to float64, tv_iso.cpp:342-347) on the same inputs:

- the uneven configuration's images (pixel weights, per-image gamma, sides up
  to 1024) as one float32 weighted batch: the pixel-weighted float32
  streaming kernel;
- float32 volumes of the serving trace's kind (64^3 and 128^3) on the 3D
  streaming kernel;
- the anisotropic configuration's disk images in float32, against the
  reference's ADBPG solution through the certificate;
- a serving run at the serving trace's real sizes (2D sides 128..1024, 3D
  sides 48..128, tol in {1e-4, 1e-5, 1e-6}): each result has the bits of its
  plain solve, and three per class are within 1e-4 of the reference.

The float32 bar (the tolerance stated for float32): |z - z_ref|_inf <= 1e-4,
every gap at its target, iterations within 8 of the reference's (the gap and
restart tests fall every 4 iterations, so float32 rounding may move a
decision by one or two test points)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fastprox_amd import Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.synth import crop_of, disks, rectangles, serving_sources, serving_trace
from helpers import fdiff

pytestmark = pytest.mark.gpu
refbind = pytest.importorskip("refbind")
if not refbind.available():
    pytest.skip("reference library not built (oracle/_ref)", allow_module_level=True)

ZTOL_F32 = 1e-4   # |z - z_ref|_inf for float32 against the float64 reference
ITER_SLACK = 8    # |iters - iters_ref|


def _target(x, tol):
    x = np.asarray(x, np.float64)
    return tol * 0.5 * (x.reshape(x.shape[0], -1) ** 2).sum(axis=1)


def _ref(x64, gamma, tol, iso=True, k=2, w=None, spatial=None):
    z, dual, gaps, its, zerrs, st, _ = refbind.Op(iso, k, w, spatial=spatial).solve(
        x64, gamma, tol=tol, max_iter=100000, threads=16)
    assert st == A.FP_OK
    return z, dual, gaps, its


def test_uneven_batch_float32_weighted_against_the_reference(cuda):
    """Eight images of the uneven configuration (H, W ~ U[64, 1024], 16
    rectangles + N(0, 0.1^2), pixel weights lognormal(0, 0.3), gamma
    log-uniform [0.02, 0.3], tol 1e-6; the bench's generator and draw order)
    as one float32 batch on the table's pixel-weighted streaming kernel."""
    from fastprox_amd.torch import Batch
    n = 8
    rng = np.random.default_rng(0)
    shapes = [(int(rng.integers(64, 1025)), int(rng.integers(64, 1025))) for _ in range(256)]
    gam = [float(np.exp(rng.uniform(np.log(0.02), np.log(0.3)))) for _ in range(256)]
    xs = [rectangles(1, sh, n=16, sigma=0.1, seed=1000 + i)[0] for i, sh in enumerate(shapes[:n])]
    al = [rng.lognormal(0.0, 0.3, size=sh).astype(np.float32) for sh in shapes[:n]]
    print("C2 float32: shapes", shapes[:n], flush=True)
    bt = Batch(Op(True, 2), torch.float32, shapes[:n], weighted=True)
    kind, tt, _ = _batch_kernel(bt)
    assert kind == 0 and tt >= 2, (kind, tt)  # the streaming kernel, several iterations per pass
    xt = [torch.tensor(x, device=cuda) for x in xs]
    zt = [torch.empty_like(x) for x in xt]
    dt = [torch.empty((2,) + x.shape, dtype=torch.float32, device=cuda) for x in xt]
    wt = [torch.tensor(a, device=cuda) for a in al]
    assert bt.solve(xt, zt, dt, gam[:n], weights=wt, tol=1e-6, max_iter=100000) == A.FP_OK
    gaps, its, zerrs, conv = bt.image_info()
    assert conv.all()
    for i in range(n):
        z_ref, _, gaps_ref, its_ref = _ref(xs[i][None].astype(np.float64), gam[i], 1e-6,
                                           spatial=al[i].astype(np.float64))
        err = float(np.abs(zt[i].cpu().numpy().astype(np.float64) - z_ref[0]).max())
        print(f"  image {i} {shapes[i]}: iters {its[i]} vs {its_ref[0]}, |dz| {err:.2e}", flush=True)
        assert err <= ZTOL_F32, (i, err)
        assert gaps[i] <= _target(xs[i][None], 1e-6)[0]
        assert abs(int(its[i]) - int(its_ref[0])) <= ITER_SLACK, (i, its[i], its_ref[0])


@pytest.mark.parametrize("count,side,gamma,tol", [(2, 64, 0.05, 1e-6), (2, 64, 0.2, 1e-5), (1, 128, 0.05, 1e-5)])
def test_volumes_float32_against_the_reference(cuda, count, side, gamma, tol):
    """float32 volumes of 16 boxes + N(0, 0.1^2) (the serving trace's 3D
    kind) on the 3D streaming kernel the table picks."""
    x = rectangles(count, (side,) * 3, n=16, sigma=0.1, seed=11 + side, dtype=np.float32)
    print(f"3D float32: {count} x {side}^3, gamma {gamma}, tol {tol}", flush=True)
    xt = torch.tensor(x, device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(True, 3), xt)
    assert ws.kernel()[0] == 3  # the 3D streaming kernel
    assert ws.prox(xt, gamma, out, tol=tol, max_iter=100000) == A.FP_OK
    gaps, its, _ = ws.image_info(count)
    z_ref, _, gaps_ref, its_ref = _ref(x.astype(np.float64), gamma, tol, k=3)
    err = float(np.abs(out.cpu().numpy().astype(np.float64) - z_ref).max())
    print(f"  iters {list(its)} vs {list(its_ref)}, |dz| {err:.2e}", flush=True)
    assert err <= ZTOL_F32
    assert (gaps <= _target(x, tol)).all()
    assert np.abs(its.astype(int) - its_ref.astype(int)).max() <= ITER_SLACK


def test_aniso_disks_float32_within_the_reference_certificate(cuda):
    """Four of the anisotropic configuration's 512^2 disk images (20 disks,
    radius U[8, 96], + N(0, 0.15^2)) in float32, w (1, 0.5), gamma 0.05, tol
    1e-6: the device method's z is within the certificate of the reference's
    ADBPG solution, |z - z_ref|^2 <= 2 (gap + gap_ref), with our gap
    recomputed in float64 from the float32 outputs (z and the dual in the C
    form); both gaps at their target."""
    x = disks(4, (512, 512), n=20, rmin=8, rmax=96, sigma=0.15, seed=0, dtype=np.float32)
    w = [1.0, 0.5]
    print("C1 float32: 4 x 512^2 aniso", flush=True)
    z_ref, _, gaps_ref, _ = _ref(x.astype(np.float64), 0.05, 1e-6, iso=False, w=w)
    xt = torch.tensor(x, device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(False, 2, w), xt)
    assert ws.prox(xt, 0.05, out, tol=1e-6, max_iter=100000) == A.FP_OK
    gaps, its, _ = ws.image_info(4)
    q = torch.empty((2,) + tuple(xt.shape), dtype=torch.float32, device=cuda)
    ws.get_dual(q)
    z, q = out.cpu().numpy().astype(np.float64), q.cpu().numpy().astype(np.float64)
    x64 = x.astype(np.float64)
    target = _target(x, 1e-6)
    for b in range(4):
        # the certificate of (z, q) for the weighted aniso objective
        # 1/2 |z - x|^2 + gamma sum_a w_a |D_a z|_1, from the definitions
        tv = sum(w[a] * np.abs(fdiff(z[b], a)).sum() for a in range(2))
        g = float(0.05 * tv - (z[b] * (x64[b] - z[b])).sum())
        np.testing.assert_allclose(x64[b] - z[b], q[:, b].sum(0), atol=1e-5)
        d2 = float(((z[b] - z_ref[b]) ** 2).sum())
        print(f"  image {b}: gap {gaps[b]:.3e} (recomputed {g:.3e}), ref {gaps_ref[b]:.3e}, |dz|^2 {d2:.3e}",
              flush=True)
        assert gaps[b] <= target[b] and gaps_ref[b] <= target[b]
        assert d2 <= 2 * (max(g, gaps[b]) + gaps_ref[b]) * 1.0001 + 1e-12, b


def _batch_kernel(bt):
    import ctypes as C

    from fastprox_amd.core import lib
    g, t, r = C.c_int(), C.c_int(), C.c_int()
    assert lib().fpr_batch_get_kernel(bt.h, C.byref(g), C.byref(t), C.byref(r)) == A.FP_OK
    return g.value, t.value, r.value


def test_serve_at_the_trace_sizes(cuda):
    """About 40 requests of the serving trace at its real sizes (2D squares of
    side U[128, 1024], 3D cubes of side U[48, 128], gamma log-uniform [0.01,
    0.3], tol in {1e-4, 1e-5, 1e-6}), submitted at their arrival times
    through the continuous-batching loop: each result has the bits of the
    plain solve of its image alone on the same kernel, and the first three
    requests of each class are within 1e-4 of the reference library's
    solve (iterations within 8)."""
    from fastprox_amd.serve import Server, replay
    a, b = serving_sources()
    src2d, src3d = torch.tensor(a, device=cuda), torch.tensor(b, device=cuda)
    trace = serving_trace(seconds=0.2, rate=200, seed=3)
    sides = {k: [r.side for r in trace if r.k == k] for k in (2, 3)}
    print(f"serving {len(trace)} requests; 2D sides {sides[2]}; 3D sides {sides[3]}", flush=True)
    assert len(trace) >= 30 and len(sides[3]) >= 3 and max(sides[2]) >= 700
    srv = Server(Op(True, 2), Op(True, 3), pool_bytes=4 << 30, slots=64)
    res = replay(srv, trace, src2d, src3d, keep_outputs=True)
    torch.cuda.synchronize()
    assert len(res) == len(trace)
    # replay() submits request i as the i-th submission: ids are in trace order
    res = sorted(res, key=lambda d: d["id"])
    ids0 = res[0]["id"]
    kern = {2: srv.kernel(2), 3: srv.kernel(3)}
    checked = {2: 0, 3: 0}
    for d in res:
        r = trace[d["id"] - ids0]
        x = crop_of(r, src2d, src3d).contiguous()
        ws = Workspace(Op(True, r.k), x[None])
        ws.set_kernel(*kern[r.k])
        z = torch.empty_like(x[None])
        st = ws.prox(x[None], r.gamma, z, tol=r.tol, max_iter=100000)
        gaps, its, _ = ws.image_info(1)
        assert st == A.FP_OK and d["status"] == 0 and d["converged"]
        assert d["iters"] == int(its[0]) and d["gap"] == float(gaps[0]), (d["id"], r.k, r.side)
        assert torch.equal(d["z"], z[0]), (d["id"], r.k, r.side)
        if checked[r.k] < 3:
            x64 = x.cpu().numpy().astype(np.float64)[None]
            z_ref, _, _, its_ref = _ref(x64, r.gamma, r.tol, k=r.k)
            err = float(np.abs(z.cpu().numpy().astype(np.float64) - z_ref).max())
            print(f"  request {d['id']} k={r.k} side={r.side} tol={r.tol:g}: iters {d['iters']} vs "
                  f"{its_ref[0]}, |dz| {err:.2e}", flush=True)
            assert err <= ZTOL_F32
            assert abs(d["iters"] - int(its_ref[0])) <= ITER_SLACK
            checked[r.k] += 1
    assert checked == {2: 3, 3: 3}

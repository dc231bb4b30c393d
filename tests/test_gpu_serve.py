"""The serving loop (serve.cpp) on the GPU: every request's result is the
plain solve's (fp_prox on that image alone, same kernel) to the bit, whatever
else runs beside it; admission waits for memory and slots; kept duals warm
later requests; edge cases (gamma 0, NaN, Inf) come back as the plain solve
reports them."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fastprox_amd import Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.serve import Server, latency_stats, replay
from fastprox_amd.synth import crop_of, serving_sources, serving_trace

pytestmark = pytest.mark.gpu

def _alone(x, gamma, tol, warm=None, kernel=None, max_iter=100000):
    """The plain solve of x alone, on the step kernel the server uses for
    x's class (the iterates do not depend on the kernel; the gaps' partial
    sums follow its tiling)."""
    k = x.dim()
    xc = x.contiguous()
    op = Op(True, k)
    ws = Workspace(op, xc)
    ws.set_kernel(*kernel[k])
    if warm is not None:
        ws.set_dual(warm)
        ws.set_warm(True)
    out = torch.empty_like(xc)
    st = ws.prox(xc, gamma, out, tol=tol, max_iter=max_iter)
    gaps, its, _ = ws.image_info(1)
    dual = torch.empty((k,) + tuple(xc.shape), dtype=xc.dtype, device=xc.device)
    ws.get_dual(dual)
    return out, int(its[0]), float(gaps[0]), st, dual


def _sources():
    a, b = serving_sources(n2d=3, n3d=2, side2d=320, side3d=40)
    return torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda")


def test_serve_same_alone_and_in_the_mix(cuda):
    src2d, src3d = _sources()
    trace = serving_trace(seconds=1.0, rate=60, seed=3, side2d=(24, 320), side3d=(8, 40))
    assert len(trace) >= 30
    srv = Server(Op(True, 2), Op(True, 3), pool_bytes=64 << 20, slots=6)
    outs, reqs = {}, {}
    half = len(trace) // 2
    for j, r in enumerate(trace):
        x = crop_of(r, src2d, src3d)
        out = torch.empty(x.shape, dtype=x.dtype, device="cuda")
        rid = srv.submit(x, r.gamma, tol=r.tol, out=out)
        outs[rid], reqs[rid] = out, (x, r)
        if j == half:  # the second half arrives while the first is on the GPU
            for _ in range(3):
                srv.step()
    res = srv.drain()
    torch.cuda.synchronize()
    kern = {2: srv.kernel(2), 3: srv.kernel(3)}
    assert len(res) == len(trace)
    admitted_late = 0
    first_done = min(d["t_done"] for d in res)
    for d in res:
        x, r = reqs[d["id"]]
        z, its, gap, st, _ = _alone(x, r.gamma, r.tol, kernel=kern)
        print(f"request {d['id']} k={r.k} side={r.side} iters {d['iters']} vs {its}", flush=True)
        assert d["status"] == 0
        assert d["iters"] == its
        assert d["gap"] == gap
        assert torch.equal(outs[d["id"]], z)
        admitted_late += d["t_admit"] >= first_done
    assert admitted_late > 0  # continuous batching: admissions after completions


@pytest.mark.parametrize("mixed", [True, False], ids=["2d+3d", "2d-two-lanes"])
def test_serve_grouped_admission(cuda, mixed):
    """150 requests queued at once are set up in groups of up to 64 (one
    gather, one prologue, one commit per group): each request keeps its own
    tolerance, iteration cap and strides, and its result has the bits of its
    plain solve. Exact 1D images and warm starts, which take the single path,
    are interleaved and keep their order. A server of 2D requests alone with
    a pool of 1 GiB runs two 2D lanes: a warm start goes to the lane that
    keeps its dual, and the results are the same."""
    src2d, src3d = _sources()
    rng = np.random.default_rng(21)
    if mixed:
        srv = Server(Op(True, 2), Op(True, 3), pool_bytes=256 << 20, slots=200)
    else:
        srv = Server(Op(True, 2), None, pool_bytes=1 << 30, slots=200)
    kept = torch.empty((60, 70), device="cuda")
    a = srv.submit(src2d[2][:60, :70], 0.1, tol=1e-5, out=kept, keep_dual=True)
    srv.drain()
    cases = {}
    for j in range(150):
        if j % 37 == 5:
            x = src2d[j % 3][7, 10:10 + 30 + j]  # a row: exact 1D TV
            x = x.reshape(1, -1)
        elif j % 41 == 7:
            x = src2d[2][:60, :70]
        elif j % 5 == 0 and mixed:
            x = src3d[j % 2][:8 + j % 7, 3:14, ::2][:, :, :9]  # strided volume
        else:
            h, w = (int(v) for v in rng.integers(16, 97, size=2))
            c0 = (2 * j) % 120
            x = src2d[j % 3][j:j + h, c0:c0 + 2 * w:2]  # every other column
        gamma = float(np.exp(rng.uniform(np.log(0.02), np.log(0.3))))
        tol = float(rng.choice([1e-4, 1e-5, 1e-6]))
        mi = 7 if j % 11 == 3 else 100000
        out = torch.empty(tuple(x.shape), device="cuda")
        warm = a if j % 41 == 7 else -1
        rid = srv.submit(x, gamma, tol=tol, max_iter=mi, out=out, warm_from=warm)
        cases[rid] = (x, gamma, tol, mi, out, warm >= 0)
    res = srv.drain()
    torch.cuda.synchronize()
    assert len(res) == 150
    kern = {2: srv.kernel(2), 3: srv.kernel(3) if mixed else None}
    _, _, _, _, dual_a = _alone(src2d[2][:60, :70], 0.1, 1e-5, kernel=kern)
    for d in res:
        x, gamma, tol, mi, out, warm = cases[d["id"]]
        z, its, gap, st, _ = _alone(x, gamma, tol, kernel=kern, max_iter=mi,
                                    warm=dual_a if warm else None)
        assert d["iters"] == its, (d["id"], tuple(x.shape), mi)
        assert d["converged"] == (st == A.FP_OK), (d["id"], st)
        assert d["gap"] == gap and torch.equal(out, z), (d["id"], tuple(x.shape))
    assert sum(1 for c in cases.values() if c[3] == 7) >= 10
    assert sum(1 for d in res if not d["converged"]) >= 5  # stopped by their own cap


def test_serve_warm_start_that_cannot_fit_beside_its_dual(cuda):
    """A pool that holds one 200^2 request but not that request and a kept
    dual of its shape: a warm start from the kept dual finds nothing else to
    evict and no request that would free room, so its dual is evicted and it
    runs cold (the plain cold solve's bits) instead of waiting for good."""
    src2d, _ = _sources()
    x = src2d[0][:200, :200]
    srv = Server(Op(True, 2), None, pool_bytes=48 * 200 * 200 + 4096, slots=4)
    kern = {2: srv.kernel(2)}
    a = srv.submit(x, 0.1, tol=1e-5, out=torch.empty((200, 200), device="cuda"), keep_dual=True)
    assert len(srv.drain()) == 1
    x2 = src2d[1][:200, :200]
    out = torch.empty((200, 200), device="cuda")
    srv.submit(x2, 0.1, tol=1e-5, out=out, warm_from=a)
    res = srv.drain()
    assert len(res) == 1 and res[0]["converged"]
    z, its, _, _, _ = _alone(x2, 0.1, 1e-5, kernel=kern)
    assert res[0]["iters"] == its and torch.equal(out, z)


def test_serve_admission_by_memory(cuda):
    """A pool that holds two 200^2 requests at a time: the rest wait, are
    admitted as memory frees, and all finish with the plain results."""
    src2d, src3d = _sources()
    srv = Server(Op(True, 2), None, pool_bytes=2 * 48 * 200 * 200 + 4096, slots=16)
    xs, outs = [], []
    for j in range(7):
        x = src2d[j % 3][j * 10:j * 10 + 200, 5:205]
        out = torch.empty((200, 200), device="cuda")
        srv.submit(x, 0.05 + 0.02 * j, tol=1e-5, out=out)
        xs.append(x)
        outs.append(out)
    res = sorted(srv.drain(), key=lambda d: d["id"])
    kern = {2: srv.kernel(2)}
    assert len(res) == 7
    # never more than two live at once (a request is live from its admission to its completion)
    for d in res:
        live = sum(1 for e in res if e["t_admit"] <= d["t_admit"] < e["t_done"])
        assert live <= 2
    assert any(d["t_admit"] > min(e["t_done"] for e in res) for d in res)
    for j, d in enumerate(res):
        z, its, gap, st, _ = _alone(xs[j], 0.05 + 0.02 * j, 1e-5, kernel=kern)
        assert torch.equal(outs[j], z) and d["iters"] == its


def test_serve_warm_start_from_a_kept_dual(cuda):
    src2d, _ = _sources()
    x = src2d[1][:128, :150]
    srv = Server(Op(True, 2), None, pool_bytes=32 << 20, slots=4)
    za = torch.empty((128, 150), device="cuda")
    a = srv.submit(x, 0.1, tol=1e-5, out=za, keep_dual=True)
    srv.drain()
    zb = torch.empty_like(za)
    srv.submit(x, 0.12, tol=1e-6, out=zb, warm_from=a)
    (rb,) = srv.drain()
    kern = {2: srv.kernel(2)}
    _, _, _, _, dual_a = _alone(x, 0.1, 1e-5, kernel=kern)
    z, its, gap, st, _ = _alone(x, 0.12, 1e-6, warm=dual_a, kernel=kern)
    assert torch.equal(zb, z) and rb["iters"] == its
    srv.release(a)
    with pytest.raises(Exception):
        srv.release(a)


def test_serve_edge_cases(cuda):
    src2d, src3d = _sources()
    srv = Server(Op(True, 2), Op(True, 3), pool_bytes=32 << 20, slots=8)
    x0 = src2d[0][:40, :50].contiguous()
    xn = x0.clone()
    xn[3, 4] = float("nan")
    xi = x0.clone()
    xi[1, 1] = float("inf")
    outs = [torch.empty_like(x0) for _ in range(3)]
    ids = [srv.submit(x0, 0.0, out=outs[0]), srv.submit(xn, 0.1, out=outs[1]), srv.submit(xi, 0.1, out=outs[2])]
    res = {d["id"]: d for d in srv.drain()}
    assert torch.equal(outs[0], x0) and res[ids[0]]["iters"] == 0
    assert torch.isnan(outs[1]).all() and np.isnan(res[ids[1]]["gap"])
    assert res[ids[2]]["status"] != 0


def test_replay_reports_latency(cuda):
    src2d, src3d = _sources()
    trace = serving_trace(seconds=0.5, rate=80, seed=5, side2d=(24, 320), side3d=(8, 40))
    srv = Server(Op(True, 2), Op(True, 3), pool_bytes=128 << 20, slots=32)
    res = replay(srv, trace, src2d, src3d)
    assert len(res) == len(trace)
    st = latency_stats(res)
    assert st["p99_ms"] >= st["p50_ms"] > 0 and st["converged"] == len(trace)


def test_serve_full_slots_keep_the_kept_dual(cuda):
    """Slots, not memory, are what is short: with both slots busy and a
    request queued, the kept dual stays resident, and a later warm start
    from it has the bits of the warm plain solve."""
    src2d, _ = _sources()
    srv = Server(Op(True, 2), None, pool_bytes=64 << 20, slots=2)
    kern = {2: srv.kernel(2)}
    xa = src2d[0][:120, :130]
    a = srv.submit(xa, 0.1, tol=1e-5, out=torch.empty((120, 130), device="cuda"), keep_dual=True)
    assert len(srv.drain()) == 1
    busy = [src2d[1][:300, :300], src2d[2][:300, :300]]
    for x in busy:  # long solves that hold both slots
        srv.submit(x, 0.3, tol=1e-7, out=torch.empty((300, 300), device="cuda"))
    srv.submit(src2d[1][:100, :100], 0.1, tol=1e-5, out=torch.empty((100, 100), device="cuda"))
    xe = src2d[0][1:121, 2:132]
    ze = torch.empty((120, 130), device="cuda")
    e = srv.submit(xe, 0.12, tol=1e-6, out=ze, warm_from=a)
    for _ in range(4):  # rounds in which the queued requests find no slot
        srv.step()
    res = {d["id"]: d for d in srv.drain()}
    torch.cuda.synchronize()
    _, _, _, _, dual_a = _alone(xa, 0.1, 1e-5, kernel=kern)
    z, its, gap, st, _ = _alone(xe, 0.12, 1e-6, warm=dual_a, kernel=kern)
    assert res[e]["iters"] == its and torch.equal(ze, z)
    srv.release(a)  # still kept


def test_serve_broadcast_inputs(cuda):
    """Inputs with a zero stride (expanded views) are gathered from their
    strides: the results equal the plain solves of their contiguous copies."""
    srv = Server(Op(True, 2), None, pool_bytes=32 << 20, slots=4)
    kern = {2: srv.kernel(2)}
    g = torch.Generator(device="cpu").manual_seed(3)
    row = torch.rand((1, 64), generator=g).cuda()
    col = torch.rand((48, 1), generator=g).cuda()
    xs = [row.expand(48, 64), col.expand(48, 64), torch.full((1, 1), 0.7, device="cuda").expand(40, 30)]
    outs = [torch.empty(tuple(x.shape), device="cuda") for x in xs]
    ids = [srv.submit(x, 0.05, tol=1e-6, out=o) for x, o in zip(xs, outs)]
    res = {d["id"]: d for d in srv.drain()}
    torch.cuda.synchronize()
    for x, o, i in zip(xs, outs, ids):
        z, its, gap, st, _ = _alone(x.contiguous(), 0.05, 1e-6, kernel=kern)
        assert res[i]["iters"] == its and torch.equal(o, z), tuple(x.stride())


def test_serve_orders_on_the_callers_stream(cuda):
    """A request submitted on a side stream right after the work that
    produces its input, and read on that stream after the poll that returns
    it, without a host synchronisation in between: the server reads x after
    it is produced and writes z before the stream's later work reads it."""
    src2d, _ = _sources()
    srv = Server(Op(True, 2), None, pool_bytes=64 << 20, slots=4)
    kern = {2: srv.kernel(2)}
    side = torch.cuda.Stream()
    a = torch.randn((2048, 2048), device="cuda")
    torch.cuda.synchronize()
    with torch.cuda.stream(side):
        x = torch.zeros((200, 210), device="cuda")
        for _ in range(20):  # keeps the side stream busy before x is written
            a = (a @ a) * 1e-3
        x.copy_(src2d[0][:200, :210] + 0.0 * a[:200, :210])
        out = torch.empty_like(x)
        rid = srv.submit(x, 0.1, tol=1e-6, out=out)
        res = srv.drain()
        got = out.clone()  # on the side stream, after the poll
    torch.cuda.synchronize()
    z, its, gap, st, _ = _alone(src2d[0][:200, :210], 0.1, 1e-6, kernel=kern)
    assert res[0]["id"] == rid and res[0]["iters"] == its
    assert torch.equal(got, z)

"""The device-memory rule of the boundary and concurrent callers.

- After fp_workspace_create (keep_structure set), no call allocates device or
  pinned memory: fp_prox, fp_prox_steps, fp_vjp, fp_vjp_device, fp_jvp,
  fp_workspace_refine, the dual calls, the structure queries and
  fp_workspace_set_branch leave the library's allocation count
  (fpr_alloc_count, every hipMalloc and hipHostMalloc it makes) unchanged
  (the reference's device rule, docs/spec.md:887; its counter,
  tests/c/test_core.cpp:40-67).
- Several host threads, each with its own workspace and stream, share one
  operator: every result has the bits of the same calls made one after the
  other (the reference's tests/c/test_core.cpp:2690-2760,
  tests/python/test_threads.py:37-52)."""

from __future__ import annotations

import threading

import numpy as np
import pytest
import torch

from fastprox_amd import Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.core import lib
from helpers import image

pytestmark = pytest.mark.gpu


def _count():
    return int(lib().fpr_alloc_count())


@pytest.mark.parametrize("case", ["iso2d_f32", "iso2d_f64", "aniso2d_f64", "iso3d_f32", "iso2d_weighted_f64"])
def test_no_allocation_after_workspace_create(cuda, case):
    iso = not case.startswith("aniso")
    k = 3 if "3d" in case else 2
    dtype = torch.float32 if case.endswith("f32") else torch.float64
    shape = (2, 12, 14, 16) if k == 3 else (3, 40, 50)
    x = torch.tensor(image(shape, 5), dtype=dtype, device=cuda)
    spatial = None
    if "weighted" in case:
        spatial = np.random.default_rng(2).uniform(0.5, 1.5, shape[1:])
    op = Op(iso, k, [1.0, 0.7] if k == 2 else None, spatial=spatial)
    g = torch.tensor(image(shape, 6), dtype=dtype, device=cuda)
    out = torch.empty_like(x)
    gx = torch.empty_like(x)
    dz = torch.empty_like(x)
    dual = torch.empty((k,) + tuple(x.shape), dtype=dtype, device=cuda)
    other = torch.float64 if dtype == torch.float32 else torch.float32
    dual_other = torch.empty((k,) + tuple(x.shape), dtype=other, device=cuda).transpose(-1, -2).contiguous() \
        .transpose(-1, -2)  # strided, of the other dtype
    steps = torch.tensor([0.1 + 0.05 * b for b in range(shape[0])], dtype=dtype, device=cuda)
    gdev = torch.zeros((), dtype=torch.float64, device=cuda)
    labels = torch.empty(tuple(x.shape), dtype=torch.int64, device=cuda)
    sign = torch.empty(tuple(x.shape), dtype=torch.int8, device=cuda)
    states = torch.empty(tuple(x.shape), dtype=torch.int8, device=cuda)
    mask = torch.empty(tuple(x.shape), dtype=torch.int8, device=cuda).transpose(-1, -2).contiguous() \
        .transpose(-1, -2)  # strided
    coef = torch.empty_like(x)
    P = op.nparams
    gp = torch.zeros(P, dtype=torch.float64, device=cuda)
    torch.cuda.synchronize()
    before = _count()
    ws = Workspace(op, x, keep_structure=True)
    n0 = _count()
    assert n0 > before  # the counter sees the workspace's buffers
    print(f"{case}: {n0} allocations after create", flush=True)
    for rep in range(2):
        ws.prox(x, 0.15, out, tol=1e-8, max_iter=100000)
        ws.vjp(g, gx, gparams=gp)
        ws.jvp(g, dz, dgamma=0.5)
        ws.vjp(g, gx, ggamma_dev=gdev)
        ws.get_dual(dual)
        ws.get_dual(dual_other)
        ws.regions(labels, sign)
        ws.states(states)
        ws.constraint(coef)
        ws.kinks(1e-6, mask)
        if not iso:
            ws.set_branch(torch.ones(tuple(x.shape), dtype=torch.int8, device=cuda), 1e-6)
            ws.vjp(g, gx)
            ws.set_branch(None)
        ws.refine(1e-10)
        ws.vjp(g, gx)
        ws.prox(x, steps, out, tol=1e-8, max_iter=100000)  # fp_prox_steps
        ws.jvp(None, dz, dgamma=1.0)
        ws.set_dual(dual)
        ws.set_warm(True)
        ws.prox(x, 0.15, out, tol=1e-8, max_iter=100000)
        ws.set_warm(False)
        torch.cuda.synchronize()
        assert _count() == n0, f"{case}: {_count() - n0} allocations after create (pass {rep})"


def test_threads_share_one_op(cuda):
    """Four threads, each with its own workspace on its own stream, call
    fp_prox, fp_vjp and fp_jvp through one shared operator on their own
    images, three rounds each; every result equals the serial run's bits.
    Thread 0's batch (16 x 512^2 float32, 2,080 tiles) runs as two halves on
    two streams, the second driven by a worker thread of the library."""
    from fastprox_amd.synth import rectangles
    op = Op(True, 2)
    T, rounds = 4, 3
    xs = [torch.tensor(rectangles(16, (512, 512), seed=3), device=cuda)]
    xs += [torch.tensor(image((3, 96, 80), 40 + t), device=cuda) for t in range(1, T)]
    gs = [torch.randn(tuple(x.shape), dtype=x.dtype, device=cuda,
                      generator=torch.Generator(device=cuda).manual_seed(50 + t)) for t, x in enumerate(xs)]
    tols = [1e-6] + [1e-8] * (T - 1)

    def work(t, ws, stream, res):
        with torch.cuda.stream(stream):
            for r in range(rounds):
                out, gx, dz = (torch.empty_like(xs[t]) for _ in range(3))
                gam = 0.05 + 0.03 * r + 0.01 * t
                st = ws.prox(xs[t], gam, out, tol=tols[t], max_iter=100000)
                gg = ws.vjp(gs[t], gx)
                ws.jvp(gs[t], dz)
                stream.synchronize()
                res.append((st, out.cpu(), gx.cpu(), gg, dz.cpu(), ws.image_info(xs[t].shape[0])[1].copy()))

    def make():
        streams = [torch.cuda.Stream() for _ in range(T)]
        wss = [Workspace(op, xs[t], keep_structure=True, stream=streams[t].cuda_stream) for t in range(T)]
        return streams, wss

    streams, wss = make()
    assert wss[0].stats() is not None
    serial = [[] for _ in range(T)]
    for t in range(T):
        work(t, wss[t], streams[t], serial[t])
    streams, wss = make()
    par = [[] for _ in range(T)]
    th = [threading.Thread(target=work, args=(t, wss[t], streams[t], par[t])) for t in range(T)]
    for h in th:
        h.start()
    for h in th:
        h.join()
    torch.cuda.synchronize()
    for t in range(T):
        assert len(par[t]) == rounds
        for a, b in zip(serial[t], par[t]):
            assert a[0] == b[0] == A.FP_OK
            assert torch.equal(a[1], b[1]) and torch.equal(a[2], b[2]) and torch.equal(a[4], b[4])
            assert a[3] == b[3]
            np.testing.assert_array_equal(a[5], b[5])

"""Solves captured into a hipGraph (the reference's device rule for a
capturing stream, docs/spec.md:890): on a stream being captured, fp_prox and
fpr_batch_solve enqueue the whole static schedule that max_iter allows, each
step and epilogue returning at once for the images already finished, with no
host wait or read; the results are read at the first query after the graph
has run. Each graph is replayed several times, with eager work and
synchronisations between the replays, and every replay must give the eager
solve's bits, iterations and gaps."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from fastprox_amd import FastproxError, Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.synth import rectangles
from fastprox_amd.torch import Batch

pytestmark = pytest.mark.gpu

_hip = None
RELAXED = 2  # hipStreamCaptureModeRelaxed


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
        P = C.c_void_p
        for name, args in (("hipStreamBeginCapture", [P, C.c_int]), ("hipStreamEndCapture", [P, C.POINTER(P)]),
                           ("hipGraphInstantiate", [C.POINTER(P), P, P, P, C.c_size_t]),
                           ("hipGraphLaunch", [P, P]), ("hipStreamSynchronize", [P]),
                           ("hipGraphExecDestroy", [P]), ("hipGraphDestroy", [P])):
            f = getattr(_hip, name)
            f.restype = C.c_int
            f.argtypes = args
    return _hip


class Captured:
    """The work fn enqueues on stream (a torch.cuda.Stream), captured once
    and replayed by launch()."""

    def __init__(self, stream, fn):
        H = hip()
        self.s = C.c_void_p(stream.cuda_stream)
        self.g, self.ex = C.c_void_p(), C.c_void_p()
        assert H.hipStreamBeginCapture(self.s, RELAXED) == 0
        try:
            fn()
        finally:
            assert H.hipStreamEndCapture(self.s, C.byref(self.g)) == 0
        assert H.hipGraphInstantiate(C.byref(self.ex), self.g, None, None, 0) == 0

    def launch(self):
        assert hip().hipGraphLaunch(self.ex, self.s) == 0
        assert hip().hipStreamSynchronize(self.s) == 0

    def __del__(self):
        if getattr(self, "ex", None) and self.ex.value:
            hip().hipGraphExecDestroy(self.ex)
        if getattr(self, "g", None) and self.g.value:
            hip().hipGraphDestroy(self.g)


@pytest.mark.parametrize("iso,dtype", [(True, torch.float32), (True, torch.float64), (False, torch.float64)])
def test_captured_workspace_solve_replays_the_eager_bits(cuda, iso, dtype):
    x = torch.tensor(rectangles(4, (96, 112), n=8, sigma=0.1, seed=3), dtype=dtype, device=cuda)
    op = Op(iso, 2, [1.0, 0.7])
    st = torch.cuda.Stream()
    ws = Workspace(op, x, stream=st.cuda_stream, method=None if iso else "fista")
    ref = torch.empty_like(x)
    assert ws.prox(x, 0.1, ref, tol=1e-6, max_iter=400) == A.FP_OK
    torch.cuda.synchronize()
    gaps_e, its_e, zerrs_e = ws.image_info(4)
    out = torch.empty_like(x)
    g = Captured(st, lambda: ws.prox(x, 0.1, out, tol=1e-6, max_iter=400, with_info=False))
    for rep in range(3):
        out.fill_(float("nan"))
        torch.cuda.synchronize()
        _ = (x * 2.0).sum().item()  # eager work between the replays
        g.launch()
        assert torch.equal(out, ref), rep
        gaps, its, zerrs = ws.image_info(4)
        np.testing.assert_array_equal(its, its_e)
        np.testing.assert_array_equal(gaps, gaps_e)
        np.testing.assert_array_equal(zerrs, zerrs_e)
        gap, iters, conv = ws.info()
        assert conv and iters == int(its_e.max())


def test_captured_uneven_batch_replays_the_eager_bits(cuda):
    """The persistent batched solve of images of different shapes, steps and
    pixel weights, captured whole (one graph for the batch)."""
    rng = np.random.default_rng(7)
    shapes = [(64, 80), (120, 96), (37, 200), (150, 150), (1, 90), (200, 64)]
    xs = [torch.tensor(rectangles(1, s, n=6, sigma=0.1, seed=20 + i)[0], device=cuda) for i, s in enumerate(shapes)]
    ws_ = [torch.tensor(rng.lognormal(0.0, 0.3, size=s), dtype=torch.float32, device=cuda) for s in shapes]
    gam = list(np.exp(rng.uniform(np.log(0.02), np.log(0.3), len(shapes))))
    bt = Batch(Op(True, 2), torch.float32, shapes, weighted=True)
    st = torch.cuda.Stream()
    zs_e = [torch.empty_like(v) for v in xs]
    ds_e = [torch.empty((2,) + tuple(v.shape), dtype=torch.float32, device=cuda) for v in xs]
    assert bt.solve(xs, zs_e, ds_e, gam, weights=ws_, tol=1e-6, max_iter=600, stream=st.cuda_stream) == A.FP_OK
    torch.cuda.synchronize()
    gaps_e, its_e, zerrs_e, conv_e = bt.image_info()
    zs = [torch.empty_like(v) for v in xs]
    ds = [torch.empty_like(v) for v in ds_e]
    g = Captured(st, lambda: bt.solve(xs, zs, ds, gam, weights=ws_, tol=1e-6, max_iter=600,
                                      stream=st.cuda_stream))
    for rep in range(3):
        for z in zs:
            z.fill_(float("nan"))
        torch.cuda.synchronize()
        g.launch()
        for i in range(len(shapes)):
            assert torch.equal(zs[i], zs_e[i]), (rep, i)
            assert torch.equal(ds[i], ds_e[i]), (rep, i)
        gaps, its, zerrs, conv = bt.image_info()
        np.testing.assert_array_equal(its, its_e)
        np.testing.assert_array_equal(gaps, gaps_e)
        assert conv.all()


def test_capture_rules(cuda):
    """Under capture: host results, kept points and host arrays are refused
    before any launch; so is the ADBPG method, whose schedule reads the host
    between steps. The stream's capture is still ended cleanly."""
    x = torch.tensor(rectangles(2, (40, 48), n=4, sigma=0.1, seed=5), dtype=torch.float64, device=cuda)
    st = torch.cuda.Stream()
    out = torch.empty_like(x)
    ws = Workspace(Op(True, 2), x, stream=st.cuda_stream)
    ws.prox(x, 0.1, out)
    ad = Workspace(Op(False, 2), x, stream=st.cuda_stream)
    H = hip()
    s = C.c_void_p(st.cuda_stream)
    for bad in (lambda: ws.prox(x, 0.1, out),  # an fp_info to fill
                lambda: ws.prox(x.cpu().numpy(), 0.1, out, with_info=False),
                lambda: ad.prox(x, 0.1, out, with_info=False)):
        assert H.hipStreamBeginCapture(s, RELAXED) == 0
        with pytest.raises(FastproxError):
            bad()
        gr = C.c_void_p()
        H.hipStreamEndCapture(s, C.byref(gr))
        if gr.value:
            H.hipGraphDestroy(gr)
    assert ws.prox(x, 0.1, out) == A.FP_OK  # the workspace works eagerly afterwards

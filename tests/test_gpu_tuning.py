"""The kernel the library picks follows the committed tuning table's
first-match rule over image size and batch size (a lone image, up to 8,
more), and the timing of a batch run as two halves on two streams counts the
step kernels' overlap once (fpr_batch_busy)."""

from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np
import pytest
import torch

from fastprox_amd import Op, Workspace, autotune
from fastprox_amd.core import lib
from fastprox_amd.synth import rectangles
from fastprox_amd.torch import Batch

pytestmark = pytest.mark.gpu

TABLE = pathlib.Path(autotune.__file__).with_name("tuning.txt")


def _expected(n, B, k=2, tsize=4, iso=1, sw=0):
    last = None
    for ln in TABLE.read_text().splitlines():
        v = ln.split("#")[0].split()
        if not v:
            continue
        v = [int(float(x)) for x in v[:8]] + [int(v[11]) if len(v) == 12 else 0]
        if tuple(v[:4]) != (k, tsize, iso, sw):
            continue
        if v[8] and B > v[8]:
            continue
        if not v[8]:
            last = v
        if n <= v[4]:
            return v[5], v[6], v[7]
    return last[5], last[6], last[7]


def _batch_kernel(bt):
    g, t, r = C.c_int(), C.c_int(), C.c_int()
    lib().fpr_batch_get_kernel(bt.h, C.byref(g), C.byref(t), C.byref(r))
    return g.value, t.value, r.value


@pytest.mark.parametrize("B", [1, 2, 8, 9, 64])
@pytest.mark.parametrize("side", [100, 256, 512])
def test_default_kernel_follows_the_table(cuda, side, B):
    x = torch.zeros((B, side, side), device=cuda)
    ws = Workspace(Op(True, 2), x)
    assert ws.kernel() == _expected(side * side, B), (side, B)
    bt = Batch(Op(True, 2), torch.float32, [(side, side)] * B)
    assert _batch_kernel(bt) == _expected(side * side, B), (side, B)


def test_busy_time_counts_the_two_streams_once(cuda):
    def run(n):
        x = torch.tensor(rectangles(n, (96, 128), n=6, sigma=0.05, seed=9, dtype=np.float32), device=cuda)
        bt = Batch(Op(True, 2), torch.float32, [(96, 128)] * n, kernel=(0, 2, 4))
        bt.set_timing(True)
        zs = [torch.empty_like(v) for v in x]
        ds = [torch.empty((2, 96, 128), device=cuda) for _ in range(n)]
        bt.solve(list(x), zs, ds, [0.1] * n, tol=1e-6, max_iter=100000)
        step_ms, _, launches = bt.timing()
        return step_ms, bt.busy_ms(), launches
    # 120 images of 72 tiles: two halves on two streams, whose launches may
    # overlap; the union is at most their sum and at least the larger half's
    step_ms, busy, launches = run(120)
    assert launches > 0 and 0.5 * step_ms * (1 - 1e-3) <= busy <= step_ms * (1 + 1e-6)
    # 20 images: one stream, the union is the sum
    step_ms, busy, launches = run(20)
    assert busy == pytest.approx(step_ms, rel=1e-3)


def test_a_tail_shorter_than_the_main_kernel_is_rejected(cuda):
    """A tail kernel runs at least the main kernel's iterations per launch
    (the launch before the switch was planned for them): fewer is FP_EINVAL,
    and the batch keeps solving with its previous tail."""
    from fastprox_amd import FastproxError
    from fastprox_amd import _capi as A
    bt = Batch(Op(True, 2), torch.float32, [(64, 80)] * 3, kernel=(0, 6, 32))
    with pytest.raises(FastproxError) as err:
        bt.set_tail(4, 16, 0.5)
    assert err.value.status == A.FP_EINVAL
    bt.set_tail(8, 16, 0.5)
    x = torch.tensor(rectangles(3, (64, 80), n=4, sigma=0.05, seed=2, dtype=np.float32), device=cuda)
    zs = [torch.empty_like(v) for v in x]
    ds = [torch.empty((2, 64, 80), device=cuda) for _ in range(3)]
    assert bt.solve(list(x), zs, ds, [0.1] * 3, tol=1e-6, max_iter=100000) == A.FP_OK

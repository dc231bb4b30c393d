"""The 3D streaming kernel with several iterations per launch (kernel kind 3,
TT >= 2: levels pipelined two slices apart, rows exchanged through LDS; float32
without pixel weights) against the one-iteration 3D streaming kernel (TT 1) on
the same inputs: a fixed number of iterations (no gap stop) gives the same
iterate to float32 rounding on shapes that cut its tiles every way (thin axes,
an inactive axis, slabs that do not divide the volume, volumes smaller than a
tile); solves to a gap tolerance take the same number of iterations within the
float32 slack and agree to 1e-4; warm starts and the dual output agree too."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fastprox_amd import Op, Workspace
from fastprox_amd import _capi as A
from helpers import image

pytestmark = pytest.mark.gpu

CASES = [((2, 37, 29, 70), [1.0, 0.7, 1.3]), ((2, 1, 9, 130), [1.0, 1.0, 1.0]),
         ((2, 20, 8, 63), [0.0, 1.0, 0.5]), ((1, 33, 15, 62), [1.0, 1.0, 1.0]),
         ((1, 17, 40, 5), [1.0, 0.0, 1.0]), ((2, 48, 48, 48), [1.0, 1.0, 1.0])]
STEP_TOL = 2e-5    # |dz|, |d dual| after a fixed number of float32 iterations (values of order 1)
SOLVE_TOL = 1e-4   # |dz| of two float32 solves that stop at the same iteration
ITER_SLACK = 8


def _agree(got, want, label):
    """Two float32 solves of the same images to the same tolerance: the same
    stop (within the float32 slack: gap and restart tests fall every four
    iterations, and float32 rounding can move a decision by one test) and
    z within 1e-4 where they stop at the same iteration; where they stop at
    different ones, within what the two certificates allow, |z1 - z2|^2 <=
    (|z1 - z*| + |z2 - z*|)^2 <= 4 (gap1 + gap2)."""
    its1, its2 = got[2].astype(int), want[2].astype(int)
    print(f"{label}: iters {list(its1)} vs {list(its2)}", flush=True)
    assert np.abs(its1 - its2).max() <= ITER_SLACK
    for b in range(len(its1)):
        d = (got[0][b] - want[0][b]).double()
        dz, d2 = float(d.abs().max()), float((d * d).sum())
        print(f"  image {b}: |dz| {dz:.2e}, |dz|^2 {d2:.2e}, gaps {got[3][b]:.2e} {want[3][b]:.2e}", flush=True)
        if its1[b] == its2[b]:
            assert dz <= SOLVE_TOL, (b, dz)
        else:
            assert d2 <= 4 * (got[3][b] + want[3][b]), (b, d2)


def _solve(x, w, iso, kernel, gamma, tol, max_iter, warm_from=None):
    op = Op(iso, 3, w)
    ws = Workspace(op, x)
    ws.set_kernel(*kernel)
    z = torch.empty_like(x)
    if warm_from is not None:
        ws.set_dual(warm_from)
        ws.set_warm(True)
    st = ws.prox(x, gamma, z, tol=tol, max_iter=max_iter)
    assert st in (A.FP_OK, A.FP_ENOTCONV)  # a fixed iteration count ends unconverged
    dual = torch.empty((3,) + tuple(x.shape), dtype=x.dtype, device=x.device)
    ws.get_dual(dual)
    gaps, its, _ = ws.image_info(x.shape[0])
    return z, dual, its, gaps


@pytest.mark.parametrize("iso", [True, False])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_fixed_iterations_match_the_one_iteration_kernel(cuda, iso, case):
    shape, w = CASES[case]
    print(f"s3t fixed iterations: shape {shape} w {w} iso {iso}", flush=True)
    x = torch.tensor(image(shape, 70 + case), dtype=torch.float32, device=cuda)
    for n in (13, 40):
        want = _solve(x, w, iso, (3, 1, 32), 0.15, 1e-30, n)
        assert (want[2] == n).all()
        for kernel in [(3, 2, 32), (3, 2, 5), (3, 2, 7)]:
            got = _solve(x, w, iso, kernel, 0.15, 1e-30, n)
            np.testing.assert_array_equal(got[2], want[2])
            dz = float((got[0] - want[0]).abs().max())
            dq = float((got[1] - want[1]).abs().max())
            print(f"  {kernel} after {n}: |dz| {dz:.2e} |dq| {dq:.2e}", flush=True)
            assert dz <= STEP_TOL and dq <= STEP_TOL, (kernel, n, dz, dq)


@pytest.mark.parametrize("iso", [True, False])
def test_solves_and_warm_starts_match(cuda, iso):
    for case in (0, 2, 5):
        shape, w = CASES[case]
        x = torch.tensor(image(shape, 80 + case), dtype=torch.float32, device=cuda)
        want = _solve(x, w, iso, (3, 1, 32), 0.1, 1e-6, 100000)
        got = _solve(x, w, iso, (3, 2, 16), 0.1, 1e-6, 100000)
        _agree(got, want, f"s3t solve {shape} iso {iso}")
        # a warm start from the solution's dual, on a perturbed input
        x2 = x * 1.05
        want2 = _solve(x2, w, iso, (3, 1, 32), 0.1, 1e-6, 100000, warm_from=want[1])
        got2 = _solve(x2, w, iso, (3, 2, 16), 0.1, 1e-6, 100000, warm_from=want[1])
        _agree(got2, want2, "  warm")


def test_kernel_choice_is_checked(cuda):
    """Several iterations per launch only for float32 volumes without pixel
    weights: float64 or weighted workspaces refuse the choice."""
    x64 = torch.tensor(image((1, 8, 8, 8), 1), dtype=torch.float64, device=cuda)
    ws = Workspace(Op(True, 3), x64)
    with pytest.raises(Exception):
        ws.set_kernel(3, 2, 16)
    x32 = x64.float()
    ws = Workspace(Op(True, 3, spatial=np.ones((8, 8, 8))), x32)
    with pytest.raises(Exception):
        ws.set_kernel(3, 2, 16)
    ws = Workspace(Op(True, 3), x32)
    ws.set_kernel(3, 2, 16)
    assert ws.kernel()[:3] == (3, 2, 16)

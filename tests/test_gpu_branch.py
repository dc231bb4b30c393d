"""fp_workspace_set_branch on the device (anisotropic TV): at an exact kink
(two pixels [0, 1] under gamma w = 1/2 merge with no jump and the edge's dual
on its bound) side -1 joins the edge's sides and +1 keeps them apart, so
fp_jvp and fp_vjp are the region means of the joined or the split regions
(include/fastprox.h:485-512). Errors leave an earlier override in place, and
the next prox call clears it."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fastprox_amd import FastproxError, Op, Workspace
from fastprox_amd import _capi as A

pytestmark = pytest.mark.gpu

KTOL = 1e-4  # the kink test's tolerance (the solve's z and dual are within 1e-7)


def _solved(x, gamma, cuda):
    xt = torch.tensor(x, dtype=torch.float64, device=cuda)
    ws = Workspace(Op(False, 2), xt, keep_structure=True)
    out = torch.empty_like(xt)
    assert ws.prox(xt, gamma, out, tol=1e-14, max_iter=200000, keep_structure=True) == A.FP_OK
    return ws, xt, out


def _jvp(ws, xt, i):
    dx = torch.zeros_like(xt)
    dx.view(-1)[i] = 1.0
    dz = torch.empty_like(xt)
    ws.jvp(dx, dz)
    return dz.cpu().numpy()


def _vjp(ws, xt, i):
    g = torch.zeros_like(xt)
    g.view(-1)[i] = 1.0
    gx = torch.empty_like(xt)
    ws.vjp(g, gx)
    return gx.cpu().numpy()


def test_one_edge_joined_and_apart(cuda):
    x = np.array([[0.0, 1.0]])
    ws, xt, z = _solved(x, 0.5, cuda)
    np.testing.assert_allclose(z.cpu().numpy(), [[0.5, 0.5]], atol=1e-6)
    mask = torch.zeros((1, 2), dtype=torch.int8, device=cuda)
    assert ws.kinks(KTOL, mask) == 1 and int(mask[0, 0]) == 2  # the edge along the last axis
    side = torch.tensor([[-1, 0]], dtype=torch.int8)
    ws.set_branch(side, KTOL)
    np.testing.assert_allclose(_jvp(ws, xt, 0), [[0.5, 0.5]], atol=1e-12)
    np.testing.assert_allclose(_vjp(ws, xt, 1), [[0.5, 0.5]], atol=1e-12)
    ws.set_branch(-side, KTOL)  # +1: apart
    np.testing.assert_allclose(_jvp(ws, xt, 0), [[1.0, 0.0]], atol=1e-12)
    np.testing.assert_allclose(_vjp(ws, xt, 1), [[0.0, 1.0]], atol=1e-12)
    # side 0 at the kink's tail, or no override: the default rule, one of the two
    ws.set_branch(torch.zeros((1, 2), dtype=torch.int8), KTOL)
    d0 = _jvp(ws, xt, 0)
    ws.set_branch(None)
    np.testing.assert_array_equal(_jvp(ws, xt, 0), d0)
    assert np.allclose(d0, [[0.5, 0.5]]) or np.allclose(d0, [[1.0, 0.0]])


def test_rows_of_a_2x2_image(cuda):
    """Two rows [0, 1]: the vertical edges are flat (no jump, dual inside),
    the horizontal ones at the kink; joined, every pixel is one region; apart,
    the two columns."""
    x = np.array([[0.0, 1.0], [0.0, 1.0]])
    ws, xt, z = _solved(x, 0.5, cuda)
    np.testing.assert_allclose(z.cpu().numpy(), 0.5, atol=1e-6)
    mask = torch.zeros((2, 2), dtype=torch.int8, device=cuda)
    assert ws.kinks(KTOL, mask) == 2
    side = torch.zeros((2, 2), dtype=torch.int8, device=cuda)
    side[:, 0] = -1
    ws.set_branch(side, KTOL)  # a device array
    np.testing.assert_allclose(_jvp(ws, xt, 0), 0.25, atol=1e-12)
    side[:, 0] = 1
    ws.set_branch(side.t().contiguous().t(), KTOL)  # strided
    np.testing.assert_allclose(_jvp(ws, xt, 0), [[0.5, 0.0], [0.5, 0.0]], atol=1e-12)
    np.testing.assert_allclose(_vjp(ws, xt, 3), [[0.0, 0.5], [0.0, 0.5]], atol=1e-12)


def test_errors_keep_the_override_and_prox_clears_it(cuda):
    x = np.array([[0.0, 1.0]])
    ws, xt, z = _solved(x, 0.5, cuda)
    ws.set_branch(torch.tensor([[-1, 0]], dtype=torch.int8), KTOL)
    bad = [(torch.tensor([[2, 0]], dtype=torch.int8), KTOL, A.FP_EPARAM),
           (torch.tensor([[-1, 0]], dtype=torch.int8), -1.0, A.FP_EPARAM),
           (torch.tensor([[-1, 0]], dtype=torch.int32), KTOL, A.FP_EDTYPE),
           (torch.tensor([[-1, 0, 0]], dtype=torch.int8), KTOL, A.FP_ESHAPE)]
    for side, tol, status in bad:
        with pytest.raises(FastproxError) as err:
            ws.set_branch(side, tol)
        assert err.value.status == status
    np.testing.assert_allclose(_jvp(ws, xt, 0), [[0.5, 0.5]], atol=1e-12)  # still joined
    ws.set_branch(torch.tensor([[1, 0]], dtype=torch.int8), KTOL)
    apart = _jvp(ws, xt, 0)
    out = torch.empty_like(xt)
    ws.prox(xt, 0.5, out, tol=1e-14, max_iter=200000, keep_structure=True)
    ws.set_branch(None)
    default = _jvp(ws, xt, 0)
    ws.set_branch(torch.tensor([[1, 0]], dtype=torch.int8), KTOL)
    np.testing.assert_array_equal(_jvp(ws, xt, 0), apart)
    ws.prox(xt, 0.5, out, tol=1e-14, max_iter=200000, keep_structure=True)
    np.testing.assert_array_equal(_jvp(ws, xt, 0), default)  # cleared by the prox call


def test_isotropic_tv_has_no_branches(cuda):
    xt = torch.tensor([[0.0, 1.0]], dtype=torch.float64, device=cuda)
    ws = Workspace(Op(True, 2), xt, keep_structure=True)
    out = torch.empty_like(xt)
    ws.prox(xt, 0.5, out, tol=1e-12, keep_structure=True)
    with pytest.raises(FastproxError) as err:
        ws.set_branch(torch.zeros((1, 2), dtype=torch.int8), KTOL)
    assert err.value.status == A.FP_EINVAL

"""The reference library's fp_workspace_set_branch (include/fastprox.h:485-512)
on the exact kinks the device tests use: two pixels [0, 1] under gamma w =
1/2, and two such rows. Side -1 joins the edge's sides (the region mean), +1
keeps them apart; these are the values tests/test_gpu_branch.py asks of the
device. Skipped without the reference build (oracle/_ref)."""

from __future__ import annotations

import numpy as np
import pytest

import refbind

pytestmark = pytest.mark.skipif(not refbind.available(), reason="reference library not built (oracle/_ref)")

KTOL = 1e-4


def _jvp_at(op, ws, shape, i):
    dx = np.zeros(shape)
    dx.flat[i] = 1.0
    return op.jvp(ws, dx)


def test_one_edge():
    op = refbind.Op(False, 2)
    x = np.array([[0.0, 1.0]])
    z, dual, gaps, its, zerrs, st, ws = op.solve(x, 0.5, tol=1e-14, max_iter=100000, keep=True)
    try:
        np.testing.assert_allclose(z, [[0.5, 0.5]], atol=1e-6)
        assert op.set_branch(ws, [[-1, 0]], KTOL) == 0
        np.testing.assert_allclose(_jvp_at(op, ws, x.shape, 0), [[0.5, 0.5]], atol=1e-12)
        assert op.set_branch(ws, [[1, 0]], KTOL) == 0
        np.testing.assert_allclose(_jvp_at(op, ws, x.shape, 0), [[1.0, 0.0]], atol=1e-12)
        gx, _, _ = op.vjp(ws, np.array([[0.0, 1.0]]))
        np.testing.assert_allclose(gx, [[0.0, 1.0]], atol=1e-12)
        assert op.set_branch(ws, [[2, 0]], KTOL) != 0  # out of range
        np.testing.assert_allclose(_jvp_at(op, ws, x.shape, 0), [[1.0, 0.0]], atol=1e-12)  # kept
    finally:
        refbind.Op.free_ws(ws)


def test_two_rows():
    op = refbind.Op(False, 2)
    x = np.array([[0.0, 1.0], [0.0, 1.0]])
    z, dual, gaps, its, zerrs, st, ws = op.solve(x, 0.5, tol=1e-14, max_iter=100000, keep=True)
    try:
        assert op.set_branch(ws, [[-1, 0], [-1, 0]], KTOL) == 0
        np.testing.assert_allclose(_jvp_at(op, ws, x.shape, 0), 0.25, atol=1e-12)
        assert op.set_branch(ws, [[1, 0], [1, 0]], KTOL) == 0
        np.testing.assert_allclose(_jvp_at(op, ws, x.shape, 0), [[0.5, 0.0], [0.5, 0.0]], atol=1e-12)
    finally:
        refbind.Op.free_ws(ws)


def test_isotropic_tv_has_no_branches():
    op = refbind.Op(True, 2)
    x = np.array([[0.0, 1.0]])
    *_, ws = op.solve(x, 0.5, tol=1e-12, keep=True)
    try:
        assert op.set_branch(ws, [[0, 0]], KTOL) == 1  # FP_EINVAL
    finally:
        refbind.Op.free_ws(ws)

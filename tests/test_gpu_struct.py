"""The structure queries and kinks on the device (fp_op_struct_kind,
fp_workspace_units/states/regions/constraint/kinks) against the reference
library built from its sources (oracle/_ref) on the same problems: float64
isotropic solves have the reference's bits, and anisotropic solves by the
reference's method (ADBPG) its z to rounding, so the regions, labels and
kinks must be the reference's exactly. Anisotropic structures of dual FISTA
solves are checked against the oracle's region count and the regions'
defining property, and their kinks against the reference on one-axis
images, where both solve the exact 1D problem."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from fastprox_amd import FastproxError, Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.synth import rectangles

pytestmark = pytest.mark.gpu

refbind = pytest.importorskip("refbind")


def _ref_struct(ws, shape, k, tol):
    """The reference library's labels, signs, region counts and kinks at the
    point kept in its workspace ws."""
    L = refbind.lib()
    B = int(np.prod(shape[:len(shape) - k]))
    lab = np.empty(shape, np.int64)
    sg = np.empty(shape, np.int8)
    nr = np.empty(B, np.int64)
    assert L.fp_workspace_regions(ws, A.numpy_array(lab).ref, A.numpy_array(sg).ref,
                                  nr.ctypes.data_as(C.POINTER(C.c_int64)), B) == A.FP_OK
    m = np.empty(shape, np.int8)
    cnt = C.c_int64()
    assert L.fp_workspace_kinks(ws, float(tol), A.numpy_array(m).ref, C.byref(cnt)) == A.FP_OK
    return lab, sg, nr, m, cnt.value


def _ours(iso, k, w, x, gamma, tol, spatial=None, dtype=torch.float64, method=None):
    xt = torch.tensor(x, dtype=dtype, device="cuda")
    ws = Workspace(Op(iso, k, w, spatial=spatial), xt, keep_structure=True, method=method)
    out = torch.empty_like(xt)
    ws.prox(xt, gamma, out, tol=tol, max_iter=10 ** 6)
    return ws, xt, out


@pytest.mark.parametrize("case", ["2d", "2d_weighted", "3d"])
def test_iso_regions_and_kinks_match_the_reference(cuda, case):
    if not refbind.available():
        pytest.skip("the reference library is not built")
    rng = np.random.default_rng(3)
    if case == "3d":
        x = rectangles(2, (10, 12, 14), n=6, sigma=0.05, seed=9, dtype=np.float64)
        k, w, sp = 3, [1.0, 0.8, 1.2], None
    else:
        x = rectangles(3, (40, 46), n=8, sigma=0.05, seed=8, dtype=np.float64)
        k, w = 2, [1.0, 0.7]
        sp = rng.uniform(0.5, 1.5, size=(40, 46)) if case == "2d_weighted" else None
        if sp is not None:
            sp[5:9, 10:12] = 0.0  # pixels of zero radius: never kinks
    gamma, tol = 0.08, 1e-9
    ws, xt, z = _ours(True, k, w, x, gamma, tol, spatial=sp)
    rop = refbind.Op(True, k, w, spatial=sp)
    zr, dual, gaps, its, zerrs, st, rws = rop.solve(x, gamma, tol=tol, max_iter=10 ** 6, keep=True)
    try:
        np.testing.assert_array_equal(z.cpu().numpy(), zr)  # the same point
        B = x.shape[0]
        assert ws.units() == (B, int(np.prod(x.shape[1:])))
        assert Op(True, k, w).struct_kind() == A.FP_STRUCT_CURVED
        st8 = torch.full(x.shape, 7, dtype=torch.int8, device="cuda")
        assert ws.states(st8) == 0 and int(st8.abs().sum()) == 0
        for ktol in (1e-7, 1e-3):
            lab_r, sg_r, nr_r, m_r, c_r = _ref_struct(rws, x.shape, k, ktol)
            lab = torch.empty(x.shape, dtype=torch.int64, device="cuda")
            sg = torch.empty(x.shape, dtype=torch.int8, device="cuda")
            nr = ws.regions(lab, sg)
            np.testing.assert_array_equal(nr, nr_r)
            np.testing.assert_array_equal(lab.cpu().numpy(), lab_r)
            np.testing.assert_array_equal(sg.cpu().numpy(), sg_r)
            m = torch.empty(x.shape, dtype=torch.int8, device="cuda")
            c = ws.kinks(ktol, m)
            np.testing.assert_array_equal(m.cpu().numpy(), m_r)
            assert c == c_r
        # int32 labels into a strided host array, and a strided device mask
        lab32 = np.full(x.shape[:-1] + (2 * x.shape[-1],), -5, np.int32)[..., ::2]
        ws.regions(lab32)
        np.testing.assert_array_equal(lab32, lab_r)
        mt = torch.zeros(x.shape[:-1] + (3 * x.shape[-1],), dtype=torch.int8, device="cuda")[..., 1::3]
        assert ws.kinks(1e-3, mt) == c_r
        np.testing.assert_array_equal(mt.cpu().numpy(), m_r)
        coef = torch.full(x.shape, 3.0, dtype=torch.float64, device="cuda")
        act = ws.constraint(coef)
        assert not act.any() and not coef.any()
    finally:
        refbind.Op.free_ws(rws)


@pytest.mark.parametrize("method", ["fista", "adbpg"])
def test_aniso_regions_are_the_oracle_structure(cuda, method):
    """The anisotropic regions at a solution of either method: as many as
    the oracle's structure at the same point, numbered in the order of their
    first entries, z constant on each (to the solve's accuracy)."""
    x = rectangles(2, (36, 40), n=8, sigma=0.05, seed=12, dtype=np.float64)
    w = [1.0, 0.6]
    ws, xt, z = _ours(False, 2, w, x, 0.1, 1e-12, method=method)
    assert Op(False, 2, w).struct_kind() == A.FP_STRUCT_PROJECTOR
    lab = torch.empty(x.shape, dtype=torch.int32, device="cuda")
    sg = torch.empty(x.shape, dtype=torch.int8, device="cuda")
    nr = ws.regions(lab, sg)
    from fastprox_amd.torch import solve as tsolve
    gam = torch.tensor([0.1, 0.1], dtype=torch.float64, device="cuda")
    zz, dual, gaps, its, zerrs, conv, exps, ue = tsolve(xt, gam, w, False, 1e-12, 10 ** 6, keep=True, method=method)
    np.testing.assert_array_equal(zz.cpu().numpy(), z.cpu().numpy())
    zc, uc, lb = z.cpu().numpy(), ue.cpu().numpy(), lab.cpu().numpy()
    assert (sg.cpu().numpy() == 1).all()
    for b in range(2):
        d = oracle.aniso_deriv(x[b], zc[b], uc[b], 0.1, w, g=np.ones(x.shape[1:]))
        assert nr[b] == d.regions
        l = lb[b].reshape(-1)
        first = np.unique(l, return_index=True)[1]
        np.testing.assert_array_equal(l[np.sort(first)], np.arange(nr[b]))
        zb = zc[b].reshape(-1)
        spread = np.zeros(nr[b])
        np.maximum.at(spread, l, zb - zb[first[l]])
        assert np.abs(spread).max() <= 1e-8 * (1 + np.abs(zb).max())


@pytest.mark.parametrize("case", ["2d", "2d_b", "3d"])
def test_aniso_adbpg_structure_matches_the_reference(cuda, case):
    """At a solution of the reference's method (FPR_METHOD_ADBPG, the
    workspaces' default: z within rounding of the reference's), the
    anisotropic regions, labels, signs and kinks are the reference library's
    exactly, and the VJP in x, gamma and the axis weights agrees to
    rounding."""
    if not refbind.available():
        pytest.skip("the reference library is not built")
    shape, w, gamma, seed = {"2d": ((3, 40, 46), [1.0, 0.7], 0.08, 8), "2d_b": ((2, 36, 40), [1.0, 0.6], 0.1, 9),
                             "3d": ((2, 10, 12, 14), [1.0, 0.8, 1.2], 0.08, 10)}[case]
    k = len(w)
    x = rectangles(shape[0], shape[1:], n=8, sigma=0.05, seed=seed, dtype=np.float64)
    ws, xt, z = _ours(False, k, w, x, gamma, 1e-9, method="adbpg")
    rop = refbind.Op(False, k, w)
    zr, dual, gaps, its, zerrs, st, rws = rop.solve(x, gamma, tol=1e-9, max_iter=10 ** 6, keep=True)
    try:
        assert np.abs(z.cpu().numpy() - zr).max() <= 1e-12
        for ktol in (1e-7, 1e-3):
            lab_r, sg_r, nr_r, m_r, c_r = _ref_struct(rws, shape, k, ktol)
            lab = torch.empty(shape, dtype=torch.int64, device="cuda")
            sg = torch.empty(shape, dtype=torch.int8, device="cuda")
            nr = ws.regions(lab, sg)
            np.testing.assert_array_equal(nr, nr_r)
            np.testing.assert_array_equal(lab.cpu().numpy(), lab_r)
            np.testing.assert_array_equal(sg.cpu().numpy(), sg_r)
            m = torch.empty(shape, dtype=torch.int8, device="cuda")
            assert ws.kinks(ktol, m) == c_r
            np.testing.assert_array_equal(m.cpu().numpy(), m_r)
        g = np.random.default_rng(1).standard_normal(shape)
        gx_r, gg_r, gp_r = rop.vjp(rws, g)
        gx = torch.empty_like(xt)
        gp = np.zeros(k)
        gg = ws.vjp(torch.tensor(g, device="cuda"), gx, gparams=gp)
        assert np.abs(gx.cpu().numpy() - gx_r).max() <= 1e-12
        assert gg == pytest.approx(gg_r, rel=1e-12, abs=1e-12)
        np.testing.assert_allclose(gp, gp_r[:k], rtol=1e-12, atol=1e-12)
    finally:
        refbind.Op.free_ws(rws)


def test_aniso_kinks_match_the_reference_on_one_axis(cuda):
    """With one active axis the reference solves the exact 1D problem on
    every row and the device method converges to it (z within 1e-7 at a gap
    tolerance of 1e-12): the kinks (edges without a jump whose 1D dual is at
    its bound) are the reference's at tolerances well above that
    difference."""
    if not refbind.available():
        pytest.skip("the reference library is not built")
    x = rectangles(2, (30, 50), n=10, sigma=0.0, seed=14, dtype=np.float64)
    w = [0.0, 1.0]
    gamma = 0.05
    ws, xt, z = _ours(False, 2, w, x, gamma, 1e-12)
    rop = refbind.Op(False, 2, w)
    zr, dual, gaps, its, zerrs, st, rws = rop.solve(x, gamma, tol=1e-12, max_iter=10 ** 6, keep=True)
    try:
        np.testing.assert_allclose(z.cpu().numpy(), zr, rtol=0, atol=1e-7)
        for ktol in (1e-4, 1e-3):
            _, _, _, m_r, c_r = _ref_struct(rws, x.shape, 2, ktol)
            m = torch.empty(x.shape, dtype=torch.int8, device="cuda")
            assert ws.kinks(ktol, m) == c_r
            np.testing.assert_array_equal(m.cpu().numpy(), m_r)
        assert c_r > 0  # the case has kinks to find
    finally:
        refbind.Op.free_ws(rws)


def test_struct_queries_errors_and_nan(cuda):
    x = rectangles(2, (20, 24), n=4, sigma=0.05, seed=3, dtype=np.float64)
    xt = torch.tensor(x, device="cuda")
    plain = Workspace(Op(True, 2), xt)
    plain.prox(xt, 0.1, torch.empty_like(xt), tol=1e-8, max_iter=10 ** 5)
    with pytest.raises(FastproxError):  # no kept point
        plain.regions()
    with pytest.raises(FastproxError):
        plain.kinks(1e-6)
    x[1, 3, 4] = np.nan
    ws, xt, z = _ours(True, 2, [1.0, 1.0], x, 0.1, 1e-8)
    lab = torch.empty(x.shape, dtype=torch.int64, device="cuda")
    sg = torch.empty(x.shape, dtype=torch.int8, device="cuda")
    nr = ws.regions(lab, sg)
    assert nr[1] == -1 and nr[0] > 0
    assert (lab[1] == -1).all() and (sg[1] == 0).all() and (sg[0] == 1).all()
    assert ws.kinks(1e-3) >= 0
    for bad in (torch.empty(x.shape, dtype=torch.float64, device="cuda"),  # dtype
                torch.empty((2, 20, 23), dtype=torch.int64, device="cuda")):  # shape
        with pytest.raises(FastproxError):
            ws.regions(bad)
    with pytest.raises(FastproxError):
        ws.kinks(-1.0)
    with pytest.raises(FastproxError):
        ws.kinks(1e-3, torch.empty(x.shape, dtype=torch.int32, device="cuda"))
    assert A.FP_EINVAL == _raw_units(None)


def _raw_units(ws):
    from fastprox_amd.core import lib
    c = C.c_int64()
    return lib().fp_workspace_units(ws, C.byref(c), None)

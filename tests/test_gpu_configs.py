"""The reference's own CPU configuration at its size, against the reference
library (oracle/_ref) on the same inputs: 8 x 256^2, each image 12
rectangles + N(0, 0.1^2) noise from seed 0, isotropic TV, gamma 0.1, gap tol
1e-6, cold start. float64: the reference's bits (z, dual, iterations);
float32 (solved in float32 here, promoted to float64 by the reference):
|z - z_ref|_inf <= 1e-4 and every gap at its target. And the anisotropic
configuration's images (20 disks, w (1, 0.5), gamma 0.05) on four of its
64 images: within the certificate of the reference's solution."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from fastprox_amd import Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.synth import disks, rectangles
from helpers import aniso_certificate

pytestmark = pytest.mark.gpu
refbind = pytest.importorskip("refbind")
if not refbind.available():
    pytest.skip("reference library not built (oracle/_ref)", allow_module_level=True)


def _ours(x, dtype, gamma, tol, iso=True, w=None):
    xt = torch.tensor(x, dtype=dtype, device="cuda")
    out = torch.empty_like(xt)
    ws = Workspace(Op(iso, 2, w), xt)
    st = ws.prox(xt, gamma, out, tol=tol, max_iter=100000)
    B = x.shape[0]
    gaps, its, _ = ws.image_info(B)
    dual = torch.empty((2,) + tuple(xt.shape), dtype=dtype, device="cuda")
    ws.get_dual(dual)
    return st, out.cpu().numpy(), dual.cpu().numpy(), gaps, its


@pytest.fixture(scope="module")
def c0():
    x = rectangles(8, (256, 256), n=12, sigma=0.1, seed=0, dtype=np.float32)
    z, dual, gaps, its, zerrs, st, _ = refbind.Op(True, 2).solve(x.astype(np.float64), 0.1, tol=1e-6,
                                                                   max_iter=100000, threads=16)
    return x, z, dual, gaps, its


def test_c0_float64_bits(cuda, c0):
    x, z_ref, dual_ref, gaps_ref, its_ref = c0
    print("C0 float64: 8 x 256^2", flush=True)
    st, z, dual, gaps, its = _ours(x.astype(np.float64), torch.float64, 0.1, 1e-6)
    assert st == A.FP_OK
    np.testing.assert_array_equal(its, its_ref)
    np.testing.assert_array_equal(z, z_ref)
    np.testing.assert_array_equal(dual, dual_ref)
    np.testing.assert_allclose(gaps, gaps_ref, rtol=1e-9)


def test_c0_float32_within_1e4(cuda, c0):
    x, z_ref, dual_ref, gaps_ref, its_ref = c0
    print("C0 float32: 8 x 256^2", flush=True)
    st, z, dual, gaps, its = _ours(x, torch.float32, 0.1, 1e-6)
    assert st == A.FP_OK
    assert np.abs(z.astype(np.float64) - z_ref).max() <= 1e-4
    target = 1e-6 * 0.5 * (x.astype(np.float64) ** 2).sum(axis=(1, 2))
    assert (gaps <= target).all()
    assert np.abs(its.astype(int) - its_ref.astype(int)).max() <= 8


def test_c1_images_within_the_certificate(cuda):
    x = disks(4, (512, 512), n=20, rmin=8, rmax=96, sigma=0.15, seed=0, dtype=np.float32).astype(np.float64)
    w = [1.0, 0.5]
    print("C1 float64: 4 x 512^2 aniso", flush=True)
    z_ref, dual_ref, gaps_ref, _, _, st_ref, _ = refbind.Op(False, 2, w).solve(x, 0.05, tol=1e-6, max_iter=100000,
                                                                                threads=16)
    st, z, q, gaps, its = _ours(x, torch.float64, 0.05, 1e-6, iso=False, w=w)
    assert st == A.FP_OK
    for b in range(4):
        g = aniso_certificate(x[b], z[b], q[:, b], 0.05, w)
        assert g == pytest.approx(gaps[b], rel=1e-6, abs=1e-12)
        assert ((z[b] - z_ref[b]) ** 2).sum() <= 2 * (g + gaps_ref[b]) * 1.0001 + 1e-20


def test_c2_images_bitwise_against_the_reference(cuda):
    """The uneven configuration's first four images (H, W ~ U[64, 1024], 16
    rectangles + N(0, 0.1^2), pixel weights lognormal(0, 0.3), gamma
    log-uniform [0.02, 0.3], tol 1e-6; the generator of the bench's uneven
    leg), solved in float64 as one uneven batch: each image has the
    reference library's bits, solved alone with its weighted operator."""
    from fastprox_amd.torch import Batch
    n = 4
    rng = np.random.default_rng(0)
    shapes = [(int(rng.integers(64, 1025)), int(rng.integers(64, 1025))) for _ in range(256)][:n]
    gam = [float(np.exp(rng.uniform(np.log(0.02), np.log(0.3)))) for _ in range(256)][:n]
    xs = [rectangles(1, sh, n=16, sigma=0.1, seed=1000 + i)[0].astype(np.float64) for i, sh in enumerate(shapes)]
    al = [rng.lognormal(0.0, 0.3, size=sh).astype(np.float32).astype(np.float64) for sh in shapes]
    print("C2 float64: shapes", shapes, flush=True)
    bt = Batch(Op(True, 2), torch.float64, shapes, weighted=True)
    xt = [torch.tensor(x, device=cuda) for x in xs]
    zt = [torch.empty_like(x) for x in xt]
    dt = [torch.empty((2,) + x.shape, dtype=torch.float64, device=cuda) for x in xt]
    wt = [torch.tensor(a, device=cuda) for a in al]
    assert bt.solve(xt, zt, dt, gam, weights=wt, tol=1e-6, max_iter=100000) == A.FP_OK
    gaps, its, zerrs, conv = bt.image_info()
    for i in range(n):
        z_ref, dual_ref, _, its_ref, _, _, _ = refbind.Op(True, 2, spatial=al[i]).solve(
            xs[i], gam[i], tol=1e-6, max_iter=100000, threads=16)
        assert int(its[i]) == int(its_ref[0]), i
        np.testing.assert_array_equal(zt[i].cpu().numpy(), z_ref)
        np.testing.assert_array_equal(dt[i].cpu().numpy(), dual_ref)


@pytest.mark.parametrize("gamma,tol", [(0.05, 1e-6), (0.2, 1e-5)])
def test_volumes_bitwise_against_the_reference(cuda, gamma, tol):
    """Two 64^3 volumes of 16 boxes + N(0, 0.1^2) (the serving trace's 3D
    kind), float64, on the 3D streaming kernel: the reference library's bits
    (z, dual, iterations)."""
    x = rectangles(2, (64, 64, 64), n=16, sigma=0.1, seed=7, dtype=np.float64)
    print(f"3D float64: 2 x 64^3, gamma {gamma}, tol {tol}", flush=True)
    xt = torch.tensor(x, device=cuda)
    out = torch.empty_like(xt)
    ws = Workspace(Op(True, 3), xt)
    assert ws.prox(xt, gamma, out, tol=tol, max_iter=100000) == A.FP_OK
    gaps, its, _ = ws.image_info(2)
    dual = torch.empty((3,) + tuple(xt.shape), dtype=torch.float64, device=cuda)
    ws.get_dual(dual)
    z_ref, dual_ref, gaps_ref, its_ref, _, _, _ = refbind.Op(True, 3).solve(x, gamma, tol=tol, max_iter=100000,
                                                                            threads=16)
    np.testing.assert_array_equal(its, its_ref)
    np.testing.assert_array_equal(out.cpu().numpy(), z_ref)
    np.testing.assert_array_equal(dual.cpu().numpy(), dual_ref)
    np.testing.assert_allclose(gaps, gaps_ref, rtol=1e-9)

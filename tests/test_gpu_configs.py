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

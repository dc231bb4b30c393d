"""fastprox_amd.torch.tv_iso / tv_aniso: the entry points of the reference's
fastprox.torch (bindings/python/fastprox/torch.py:305-351) on the GPU."""

from __future__ import annotations

import warnings

import numpy as np
import pytest
import torch

import oracle
from fastprox_amd.torch import ConvergenceWarning, inner_accuracy, tv_aniso, tv_iso
from helpers import aniso_certificate, image, iso_certificate

pytestmark = pytest.mark.gpu


def test_tv_iso_matches_the_oracle(cuda):
    x = image((2, 3, 24, 30), 4)
    z, info = tv_iso(torch.tensor(x, device=cuda), 0.2, tol=1e-9, max_iter=100000, return_info=True)
    assert z.shape == x.shape and info.converged
    xb = x.reshape(6, 24, 30)
    zb = z.cpu().numpy().reshape(6, 24, 30)
    for b in range(6):
        r = oracle.tv_iso(xb[b], 0.2, tol=1e-9, max_iter=100000)
        np.testing.assert_array_equal(zb[b], r.z)
    assert info.dual.shape == (2,) + x.shape
    assert info.iters == max(int(v) for v in info.image_iters)


def test_per_image_gamma_and_pixel_weights(cuda):
    x = image((3, 20, 22), 5)
    gam = torch.tensor([0.05, 0.2, 0.0], dtype=torch.float64, device=cuda)
    al = np.random.default_rng(1).lognormal(0.0, 0.3, size=(3, 20, 22))
    z = tv_iso(torch.tensor(x, device=cuda), gam, pixel_weights=torch.tensor(al, device=cuda), tol=1e-9,
               max_iter=100000)
    for b in range(3):
        r = oracle.tv_iso(x[b], float(gam[b]), alpha=al[b], tol=1e-9, max_iter=100000)
        np.testing.assert_array_equal(z[b].cpu().numpy(), r.z)


def test_tv_aniso_dual_form_and_certificate(cuda):
    x = image((2, 30, 26), 6)
    z, info = tv_aniso(torch.tensor(x, device=cuda), 0.3, weights=[1.0, 0.5], tol=1e-8, max_iter=100000,
                       return_info=True)
    q = info.dual.cpu().numpy()
    for b in range(2):
        g = aniso_certificate(x[b], z[b].cpu().numpy(), q[:, b], 0.3, [1.0, 0.5])
        assert g <= 1e-8 * 0.5 * (x[b] ** 2).sum() * 1.001


def test_warm_start_round_trip(cuda):
    x = torch.tensor(image((30, 40), 8), device=cuda)
    z1, i1 = tv_iso(x, 0.2, tol=1e-5, max_iter=100000, return_info=True)
    z2, i2 = tv_iso(x, 0.2, tol=1e-5, max_iter=100000, warm_start=i1.dual, return_info=True)
    # The seed is projected again (tv_iso.cpp:456-477), as the reference does.
    r = oracle.tv_iso(x.cpu().numpy(), 0.2, tol=1e-5, max_iter=100000, warm=i1.dual.cpu().numpy())
    assert i2.iters == r.iters == 0
    np.testing.assert_array_equal(z2.cpu().numpy(), r.z)
    assert (z1 - z2).abs().max() < 1e-12
    z3, i3 = tv_aniso(x, 0.2, tol=1e-5, max_iter=100000, return_info=True)
    z4, i4 = tv_aniso(x, 0.2, tol=1e-5, max_iter=100000, warm_start=i3.dual, return_info=True)
    assert i4.iters <= 4
    assert (z4 - z3).abs().max() < 1e-2


def test_tolerance_plumbing_and_warnings(cuda):
    x = torch.tensor(image((16, 16), 9), device=cuda)
    with pytest.warns(ConvergenceWarning):
        tv_iso(x, 0.5, tol=1e-14, max_iter=8)
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        tv_iso(x, 0.5, tol=1e-14, max_iter=8, warn=False)
    a, ia = tv_iso(x, 0.3, ztol=1e-4, return_info=True, max_iter=100000)
    b, ib = tv_iso(x, 0.3, tol=1e-8, return_info=True, max_iter=100000)
    assert torch.equal(a, b)
    with inner_accuracy(1e-3):
        c, ic = tv_iso(x, 0.3, return_info=True, max_iter=100000)
    assert torch.equal(a, c)
    with pytest.raises(ValueError):
        tv_iso(x, 0.3, tol=1e-6, ztol=1e-3)


def test_errors_and_empty(cuda):
    x = torch.tensor(image((8, 9), 1), device=cuda)
    with pytest.raises(ValueError):
        tv_iso(x, -1.0)
    with pytest.raises(ValueError):
        tv_iso(x, 0.1, ndim=4)
    with pytest.raises(TypeError):
        tv_iso(x.to(torch.int32), 0.1)
    xi = x.clone()
    xi[0, 0] = float("inf")
    with pytest.raises(ValueError):
        tv_iso(xi, 0.1)
    e = torch.empty((0, 4, 5), dtype=torch.float32, device=cuda)
    assert tv_iso(e, 0.1).shape == (0, 4, 5)
    # an empty batch stays in the graph of every input: zero gradients of their shapes
    eg = e.clone().requires_grad_(True)
    g = torch.tensor(0.1, device=cuda, requires_grad=True)
    w = torch.tensor([1.0, 0.5], device=cuda, requires_grad=True)
    al = torch.ones((4, 5), device=cuda, requires_grad=True)
    z, info = tv_iso(eg, g, weights=w, pixel_weights=al, return_info=True)
    assert info.converged and info.iters == 0 and info.dual.shape == (2, 0, 4, 5)
    z.sum().backward()
    assert eg.grad.shape == e.shape and float(g.grad) == 0.0
    assert torch.equal(w.grad, torch.zeros(2, device=cuda)) and torch.equal(al.grad, torch.zeros((4, 5), device=cuda))
    xn = x.clone()
    xn[2, 2] = float("nan")
    z, info = tv_iso(xn, 0.1, return_info=True)
    assert torch.isnan(z).all() and info.converged

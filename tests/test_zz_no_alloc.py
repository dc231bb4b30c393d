"""fpr_prox_static, the allocation-free fp_prox (its scratch in a fixed
device array of the library), against fp_prox on a workspace with the same
kernel: the same z, iterations and gap, on two calls in a row (the second
on a smaller input), with the library's allocation count unchanged; sizes
beyond its scratch are refused with FP_ENOMEM."""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest
import torch

from fastprox_amd import Op, Workspace
from fastprox_amd import _capi as A
from fastprox_amd.core import as_array, lib
from helpers import image

pytestmark = pytest.mark.gpu


def _static(op, x, gamma, out, tol=1e-8, max_iter=100000):
    o = A.fp_options()
    lib().fp_options_init(C.byref(o))
    o.tol, o.max_iter = tol, max_iter
    info = A.fp_info()
    s = lib().fpr_prox_static(op.h, float(gamma), as_array(x).ref, as_array(out).ref, C.byref(o),
                              C.c_void_p(torch.cuda.current_stream().cuda_stream or None), C.byref(info))
    return s, info.gap, info.iters


def _workspace(op, x, gamma, kernel, tol=1e-8, max_iter=100000):
    ws = Workspace(op, x, kernel=kernel)
    out = torch.empty_like(x)
    st = ws.prox(x, gamma, out, tol=tol, max_iter=max_iter)
    gap, iters, _ = ws.info()
    return st, out, (gap, iters)


CASES = [
    ("iso2d_f64", True, 2, torch.float64, (3, 64, 70), (2, 40, 33), None),
    ("iso2d_f32", True, 2, torch.float32, (4, 96, 80), (1, 50, 60), None),
    ("aniso2d_f64", False, 2, torch.float64, (2, 48, 52), (1, 30, 20), None),
    ("iso3d_f32", True, 3, torch.float32, (2, 20, 24, 30), (1, 9, 17, 12), None),
    ("iso2d_weighted_f64", True, 2, torch.float64, (2, 40, 44), (1, 40, 44), "sw"),
]


@pytest.mark.parametrize("name,iso,k,dtype,shape1,shape2,sw", CASES, ids=[c[0] for c in CASES])
def test_static_solve_matches_the_workspace_solve(cuda, name, iso, k, dtype, shape1, shape2, sw):
    kernel = (3, 1, 16) if k == 3 else (0, 6, 16)
    spatial = np.random.default_rng(4).uniform(0.5, 1.5, shape1[1:]) if sw else None
    op = Op(iso, k, [1.0, 0.7] if k == 2 else None, spatial=spatial)
    before = int(lib().fpr_alloc_count())
    for j, shape in enumerate((shape1, shape2)):
        x = torch.tensor(image(shape, 20 + j), dtype=dtype, device=cuda)
        gamma = 0.12 + 0.05 * j
        out = torch.full_like(x, float("nan"))
        n0 = int(lib().fpr_alloc_count())
        st, gap, its = _static(op, x, gamma, out)
        torch.cuda.synchronize()
        assert int(lib().fpr_alloc_count()) == n0, "fpr_prox_static allocated"
        st_ws, z_ws, inf = _workspace(op, x, gamma, kernel)
        print(f"{name} call {j}: {tuple(shape)} status {st} iters {its} vs {inf[1]}", flush=True)
        assert st == st_ws == A.FP_OK
        assert torch.equal(out, z_ws)
        assert its == inf[1] and gap == inf[0]
    assert int(lib().fpr_alloc_count()) >= before


def test_static_solve_out_aliasing_x(cuda):
    op = Op(True, 2)
    x = torch.tensor(image((2, 50, 60), 30), device=cuda)
    st_ws, z_ws, inf = _workspace(op, x, 0.2, (0, 6, 16))
    y = x.clone()
    st, gap, its = _static(op, y, 0.2, y)
    assert st == A.FP_OK and torch.equal(y, z_ws) and its == inf[1]


def test_static_solve_limits(cuda):
    op = Op(True, 2)
    big = torch.zeros((1, 4096, 4096), dtype=torch.float32, device=cuda)
    out = torch.empty_like(big)
    st, _, _ = _static(op, big, 0.1, out)
    assert st == A.FP_ENOMEM and b"scratch" in lib().fpr_last_error()
    many = torch.zeros((300, 8, 8), dtype=torch.float64, device=cuda)
    st, _, _ = _static(op, many, 0.1, torch.empty_like(many))
    assert st == A.FP_ENOMEM and b"256" in lib().fpr_last_error()
    host = np.zeros((2, 8, 8))
    o = A.fp_options()
    lib().fp_options_init(C.byref(o))
    info = A.fp_info()
    h_out = np.empty_like(host)
    st = lib().fpr_prox_static(op.h, 0.1, A.numpy_array(host).ref, A.numpy_array(h_out).ref, C.byref(o), None,
                               C.byref(info))
    assert st == A.FP_EINVAL
    strided = torch.zeros((2, 16, 16), dtype=torch.float64, device=cuda).transpose(1, 2)
    st, _, _ = _static(op, strided, 0.1, torch.empty((2, 16, 16), dtype=torch.float64, device=cuda))
    assert st == A.FP_EINVAL
    st, _, _ = _static(op, many[:2], -1.0, torch.empty_like(many[:2]))
    assert st == A.FP_EPARAM

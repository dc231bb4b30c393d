"""The derivatives of the PyTorch entry points (_tv_torch.py:286-585
Structure): the structure of a solved batch on the device and its VJP and
JVP, through the C ABI's fpr_deriv handle (libfastprox_rocm: deriv.hip).

Isotropic TV: the derivative systems of the C kernel (src/ops/tv_iso.cpp:
551-631, 672-1067): flat pixels read off the dual, their connected regions,
conjugate gradients on M^T (I + G^T H G) M with the Jacobi preconditioner, by
default to the relative error of z that each image's gap certifies.
Anisotropic TV: the device method's exact region means (_tv_torch.py:
346-362, 528-585).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as A
from .core import Op, _check, lib


def _declare(L):
    if getattr(L, "_deriv_declared", False):
        return
    P, D = C.c_void_p, C.c_double
    L.fpr_deriv_create.restype = C.c_int
    L.fpr_deriv_create.argtypes = [P, C.c_int, C.c_int64, C.POINTER(C.c_int64), P, P, P, P, C.c_int,
                                   C.POINTER(D), C.POINTER(D), C.POINTER(C.c_int), P, C.POINTER(P)]
    L.fpr_deriv_free.restype = None
    L.fpr_deriv_free.argtypes = [P]
    L.fpr_deriv_vjp.restype = C.c_int
    L.fpr_deriv_vjp.argtypes = [P, P, P, P, C.POINTER(D), P, D, C.c_int, P]
    L.fpr_deriv_jvp.restype = C.c_int
    L.fpr_deriv_jvp.argtypes = [P, P, P, C.POINTER(D), P, P, D, C.c_int, P]
    L.fpr_deriv_info.restype = C.c_int
    L.fpr_deriv_info.argtypes = [P, C.POINTER(D), C.POINTER(C.c_int), C.POINTER(C.c_int64)]
    L.fpr_deriv_image_info.restype = C.c_int
    L.fpr_deriv_image_info.argtypes = [P, C.POINTER(D), C.POINTER(C.c_int), C.c_int64]
    L._deriv_declared = True


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else C.c_void_p()


class Structure:
    """What the derivatives need at a solution of a batch x (B, *img):
    z (B, *img); dual: the isotropic dual (k, B, *img) in the C form, or the
    anisotropic per-edge dual (B, k, *img); the per-image gaps and scale
    exponents of the solve (numpy); gamma (B,); w the axis weights; sw the
    spatial weights (pixel (B, *img) or edge (k, B, *img)) or None."""

    def __init__(self, x, z, dual, gaps, exps, gamma, w, iso, sw=None):
        L = lib()
        _declare(L)
        self.k, self.iso = len(w), bool(iso)
        self.B = x.shape[0]
        self.img = tuple(x.shape[1:])
        self.dtype, self.device = x.dtype, x.device
        self.op = Op(iso, self.k, w)
        self.x = x.detach().contiguous()
        self.z = z.detach().contiguous()
        u = dual.detach()
        self.u = (u.transpose(0, 1) if iso else u).contiguous()  # (B, k, *img)
        self.sw = None
        if sw is not None:
            s = sw.detach().to(dtype=x.dtype)
            self.sw = (s if iso else s.transpose(0, 1)).contiguous()  # (B, *img) / (B, k, *img)
        self.gam = np.ascontiguousarray(gamma.detach().double().cpu().numpy().reshape(-1))
        self.gaps = np.ascontiguousarray(np.asarray(gaps, np.float64).reshape(-1))
        self.exps = np.ascontiguousarray(np.asarray(exps, np.int32).reshape(-1))
        shp = np.ascontiguousarray(np.asarray(self.img, np.int64))
        dt = A.FP_F32 if x.dtype == torch.float32 else A.FP_F64
        self.h = C.c_void_p()
        with torch.cuda.device(x.device):
            stream = torch.cuda.current_stream().cuda_stream
            _check(L.fpr_deriv_create(self.op.h, dt, self.B, shp.ctypes.data_as(C.POINTER(C.c_int64)),
                                      _ptr(self.x), _ptr(self.z), _ptr(self.u), _ptr(self.sw), 1,
                                      self.gam.ctypes.data_as(C.POINTER(C.c_double)),
                                      self.gaps.ctypes.data_as(C.POINTER(C.c_double)),
                                      self.exps.ctypes.data_as(C.POINTER(C.c_int)), C.c_void_p(stream or None),
                                      C.byref(self.h)), "fpr_deriv_create")

    def _info(self):
        r, i, n = C.c_double(), C.c_int(), C.c_int64()
        _check(lib().fpr_deriv_info(self.h, C.byref(r), C.byref(i), C.byref(n)), "fpr_deriv_info")
        return r.value, i.value

    def regions(self):
        r, i, n = C.c_double(), C.c_int(), C.c_int64()
        _check(lib().fpr_deriv_info(self.h, C.byref(r), C.byref(i), C.byref(n)), "fpr_deriv_info")
        return n.value

    def image_info(self):
        res, its = np.zeros(self.B), np.zeros(self.B, np.int32)
        _check(lib().fpr_deriv_image_info(self.h, res.ctypes.data_as(C.POINTER(C.c_double)),
                                          its.ctypes.data_as(C.POINTER(C.c_int)), self.B), "fpr_deriv_image_info")
        return res, its

    def vjp(self, gz, tol, max_iter):
        """(gx, d/dgamma per image (B,), d/dw (k,) summed over the batch,
        d/dsw shaped like sw or None, (residual, iterations))."""
        g = gz.detach().to(dtype=self.dtype).contiguous()
        gx = torch.empty_like(g)
        gg = torch.empty(self.B, dtype=torch.float64, device=self.device)
        gw = np.zeros(self.k)
        gsw = None
        if self.sw is not None:
            gsw = torch.empty(self.sw.shape, dtype=torch.float64, device=self.device)
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream().cuda_stream
            _check(lib().fpr_deriv_vjp(self.h, _ptr(g), _ptr(gx), _ptr(gg), gw.ctypes.data_as(C.POINTER(C.c_double)),
                                       _ptr(gsw), float(tol or 0.0), int(max_iter), C.c_void_p(stream or None)),
                   "fpr_deriv_vjp")
        out_sw = None
        if gsw is not None:
            out_sw = (gsw if self.iso else gsw.transpose(0, 1)).to(self.dtype)
        return (gx, gg.to(self.dtype), torch.tensor(gw, dtype=torch.float64), out_sw, self._info())

    def jvp(self, dx, dgamma, dw, tol, max_iter, dsw=None):
        """(dz (B, *img), (residual, iterations)) along dx (or None), dgamma
        (a number, (B,) or None), dw (k numbers or None), dsw (like sw or
        None)."""
        dxc = None if dx is None else dx.detach().to(dtype=self.dtype).contiguous()
        dg = None
        if dgamma is not None:
            dg = torch.as_tensor(dgamma, dtype=torch.float64, device=self.device).detach()
            dg = dg.expand(self.B).contiguous() if dg.numel() == 1 else dg.reshape(self.B).contiguous()
        dwa = None if dw is None else np.ascontiguousarray(np.asarray(dw, np.float64).reshape(-1))
        if dwa is not None and dwa.size == 1:
            dwa = np.repeat(dwa, self.k)
        ds = None
        if dsw is not None:
            if self.sw is None:
                raise ValueError("no spatial weights to differentiate")
            t = dsw.detach().double()
            ds = (t if self.iso else t.transpose(0, 1)).contiguous()
        dz = torch.empty_like(self.z)
        with torch.cuda.device(self.device):
            stream = torch.cuda.current_stream().cuda_stream
            _check(lib().fpr_deriv_jvp(self.h, _ptr(dxc), _ptr(dg),
                                       dwa.ctypes.data_as(C.POINTER(C.c_double)) if dwa is not None else None,
                                       _ptr(ds), _ptr(dz), float(tol or 0.0), int(max_iter),
                                       C.c_void_p(stream or None)), "fpr_deriv_jvp")
        return dz, self._info()

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            try:
                lib().fpr_deriv_free(self.h)
            except Exception:
                pass
            self.h = C.c_void_p()

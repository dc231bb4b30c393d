"""The reference fastprox library built from its sources (oracle/_ref, by
oracle/Makefile.ref) behind a small numpy API, for validating the oracle in
the build container. Absent on machines without /root/reference: callers
skip."""

from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

from fastprox_amd import _capi as A

REF = pathlib.Path(__file__).resolve().parent.parent / "oracle" / "_ref" / "libfastprox_ref.so"
_lib = None


def available():
    return REF.exists()


def lib():
    global _lib
    if _lib is None:
        _lib = A.declare(C.CDLL(str(REF)))
    return _lib


def _check(s, what):
    if s not in (A.FP_OK, A.FP_ENOTCONV):
        raise RuntimeError(f"{what}: {A.STATUS.get(s, s)}")
    return s


class Op:
    def __init__(self, iso, k, w=None, spatial=None):
        L = lib()
        self.iso, self.k = iso, k
        self.h = C.c_void_p()
        keep = []
        wa = None
        if w is not None:
            wv = np.ascontiguousarray(np.atleast_1d(np.asarray(w, np.float64)))
            keep.append(wv)
            wa = A.numpy_array(wv)
        if spatial is None:
            f = L.fp_op_tv_iso if iso else L.fp_op_tv_aniso
            _check(f(k, wa.ref if wa else None, C.byref(self.h)), "op")
        else:
            sv = np.ascontiguousarray(spatial, np.float64)
            keep.append(sv)
            sa = A.numpy_array(sv)
            f = L.fp_op_tv_iso_weighted if iso else L.fp_op_tv_aniso_weighted
            _check(f(k, wa.ref if wa else None, sa.ref, C.byref(self.h)), "op")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.fp_op_free(self.h)

    def solve(self, x, gamma, tol=1e-6, max_iter=1000, threads=1, keep=False, warm=None, steps=None):
        """(z, dual, gaps, iters, zerrs, status, ws) for x (batch..., *img)."""
        L = lib()
        x = np.ascontiguousarray(x)
        out = np.empty_like(x)
        opt = A.fp_options()
        L.fp_options_init(C.byref(opt))
        opt.tol, opt.max_iter, opt.threads, opt.keep_structure = tol, max_iter, threads, int(keep)
        xa, oa = A.numpy_array(x), A.numpy_array(out)
        ws = C.c_void_p()
        _check(L.fp_workspace_create(self.h, xa.ref, C.byref(opt), C.byref(ws)), "workspace")
        if warm is not None:
            wv = np.ascontiguousarray(warm, x.dtype)
            _check(L.fp_workspace_set_dual(ws, A.numpy_array(wv).ref), "set_dual")
        info = A.fp_info()
        if steps is None:
            s = L.fp_prox(self.h, float(gamma), xa.ref, oa.ref, C.byref(opt), ws, C.byref(info))
        else:
            sv = np.ascontiguousarray(steps, x.dtype)
            s = L.fp_prox_steps(self.h, A.numpy_array(sv).ref, xa.ref, oa.ref, C.byref(opt), ws)
        _check(s, "prox")
        B = int(np.prod(x.shape[:x.ndim - self.k]))
        dual = np.empty((self.k,) + x.shape, x.dtype)
        _check(L.fp_workspace_get_dual(ws, A.numpy_array(dual).ref), "dual")
        gaps, its, zerrs = np.empty(B), np.empty(B, np.int32), np.empty(B)
        _check(L.fp_workspace_image_info(ws, gaps.ctypes.data_as(C.POINTER(C.c_double)),
                                         its.ctypes.data_as(C.POINTER(C.c_int)),
                                         zerrs.ctypes.data_as(C.POINTER(C.c_double)), B), "image_info")
        res = (out, dual, gaps, its, zerrs, s, ws if keep else None)
        if not keep:
            L.fp_workspace_free(ws)
        return res

    def vjp(self, ws, g, deriv_tol=0.0, deriv_max_iter=10000):
        """(gx, ggamma, gparams) at the point kept in ws."""
        L = lib()
        g = np.ascontiguousarray(g)
        gx = np.empty_like(g)
        _check(L.fp_workspace_set_deriv_tol(ws, deriv_tol, deriv_max_iter), "deriv_tol")
        P = L.fp_op_nparams(self.h)
        gp = np.zeros(P)
        gg = C.c_double()
        _check(L.fp_vjp(self.h, ws, A.numpy_array(g).ref, A.numpy_array(gx).ref, C.byref(gg),
                        A.numpy_array(gp).ref), "vjp")
        return gx, gg.value, gp

    def jvp(self, ws, dx, dgamma=0.0, dparams=None):
        L = lib()
        dx = np.ascontiguousarray(dx)
        out = np.empty_like(dx)
        dp = None if dparams is None else A.numpy_array(np.ascontiguousarray(dparams, np.float64))
        _check(L.fp_jvp(self.h, ws, A.numpy_array(dx).ref, float(dgamma), dp.ref if dp else None,
                        A.numpy_array(out).ref), "jvp")
        return out

    def set_branch(self, ws, side, tol):
        """fp_workspace_set_branch (None: the default rule); the status."""
        L = lib()
        if side is None:
            return L.fp_workspace_set_branch(ws, None, float(tol))
        sd = np.ascontiguousarray(side, np.int8)
        return L.fp_workspace_set_branch(ws, A.numpy_array(sd).ref, float(tol))

    @staticmethod
    def free_ws(ws):
        lib().fp_workspace_free(ws)

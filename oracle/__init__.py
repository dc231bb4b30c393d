"""Python view of the CPU oracle (oracle/tv_oracle.c). Test infrastructure:
imported only by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker. Every function takes numpy float64 arrays of one image
in C order."""

from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
LIB = HERE / "libtv_oracle.so"
_lib = None


def build():
    subprocess.run(["make", "-s", "-f", str(HERE / "Makefile")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        dp, ip, i64p = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int64)
        L.or_tv_iso.restype = C.c_int
        L.or_tv_iso.argtypes = [C.c_int, i64p, dp, dp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp, dp,
                                dp, ip, ip, ip]
        L.or_tv_aniso_dev.restype = C.c_int
        L.or_tv_aniso_dev.argtypes = [C.c_int, i64p, dp, dp, C.c_double, C.c_double, C.c_int, C.c_int, dp, dp,
                                      dp, dp, ip, ip]
        L.or_tv1d.restype = None
        L.or_tv1d.argtypes = [dp, dp, C.c_int64, C.c_double]
        L.or_iso_gap.restype = C.c_double
        L.or_iso_gap.argtypes = [C.c_int, i64p, dp, dp, C.c_double, dp, dp]
        L.or_aniso_gap.restype = C.c_double
        L.or_aniso_gap.argtypes = [C.c_int, i64p, dp, dp, C.c_double, dp, dp]
        L.or_iso_deriv.restype = C.c_int
        L.or_iso_deriv.argtypes = [C.c_int, i64p, dp, dp, C.c_double, dp, dp, dp, C.c_double, C.c_int, C.c_int,
                                   dp, C.c_double, dp, C.c_double, C.c_int, dp, dp, dp, ip, i64p]
        L.or_deriv_set_eps.restype = None
        L.or_deriv_set_eps.argtypes = [C.c_double, C.c_double]
        L.or_aniso_deriv.restype = C.c_int
        L.or_aniso_deriv.argtypes = [C.c_int, i64p, dp, dp, C.c_double, dp, dp, dp, C.c_double, C.c_double,
                                     C.c_int, dp, C.c_double, dp, dp, dp, dp, i64p]
        L.or_zerr.restype = C.c_double
        L.or_zerr.argtypes = [C.c_double, dp, C.c_int64, C.c_double, C.c_double, C.c_int, C.c_double]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Result:
    def __init__(self, z, dual, gap, iters, converged, scale_exp=0):
        self.z, self.dual, self.gap, self.iters, self.converged = z, dual, gap, iters, converged
        self.scale_exp = scale_exp


def tv_iso(x, gamma, w=None, alpha=None, tol=1e-6, max_iter=1000, warm=None):
    """One image x (k axes) -> Result, as the reference's TVIso on that image."""
    x = _c(x)
    k = x.ndim
    w = np.ones(k) if w is None else _c(np.broadcast_to(np.asarray(w, np.float64), (k,)))
    alpha = _c(alpha)
    shape = (C.c_int64 * k)(*x.shape)
    z = np.empty_like(x)
    dual = np.zeros((k,) + x.shape) if warm is None else _c(warm).copy()
    gap, it, conv, se = C.c_double(), C.c_int(), C.c_int(), C.c_int()
    s = lib().or_tv_iso(k, shape, _p(w), _p(alpha), float(gamma), float(tol), int(max_iter),
                        int(warm is not None), _p(x), _p(z), _p(dual), C.byref(gap), C.byref(it), C.byref(conv),
                        C.byref(se))
    if s == 1:
        raise ValueError("x contains an infinite entry")
    if s:
        raise MemoryError
    return Result(z, dual, gap.value, it.value, bool(conv.value), se.value)


def tv_aniso_dev(x, gamma, w=None, beta=None, tol=1e-6, max_iter=1000, warm=None):
    """One image x -> Result by the device method for anisotropic TV; beta:
    edge weights (k, *x.shape) or None; dual in the C form q (k, *x.shape)."""
    x = _c(x)
    k = x.ndim
    w = np.ones(k) if w is None else _c(np.broadcast_to(np.asarray(w, np.float64), (k,)))
    beta = _c(beta)
    shape = (C.c_int64 * k)(*x.shape)
    z = np.empty_like(x)
    dual = np.zeros((k,) + x.shape) if warm is None else _c(warm).copy()
    gap, it, conv = C.c_double(), C.c_int(), C.c_int()
    s = lib().or_tv_aniso_dev(k, shape, _p(w), _p(beta), float(gamma), float(tol), int(max_iter),
                              int(warm is not None), _p(x), _p(z), _p(dual), C.byref(gap), C.byref(it),
                              C.byref(conv))
    if s == 1:
        raise ValueError("x contains an infinite entry")
    if s:
        raise MemoryError
    return Result(z, dual, gap.value, it.value, bool(conv.value))


def tv1d(y, lam):
    y = _c(y)
    out = np.empty_like(y)
    lib().or_tv1d(_p(y), _p(out), y.size, float(lam))
    return out


def iso_gap(z, u, gamma, w=None, alpha=None):
    z = _c(z)
    k = z.ndim
    w = np.ones(k) if w is None else _c(np.broadcast_to(np.asarray(w, np.float64), (k,)))
    shape = (C.c_int64 * k)(*z.shape)
    return lib().or_iso_gap(k, shape, _p(w), _p(_c(alpha)), float(gamma), _p(z), _p(_c(u)))


def aniso_gap(x, z, gamma, w=None, beta=None):
    x, z = _c(x), _c(z)
    k = z.ndim
    w = np.ones(k) if w is None else _c(np.broadcast_to(np.asarray(w, np.float64), (k,)))
    shape = (C.c_int64 * k)(*z.shape)
    return lib().or_aniso_gap(k, shape, _p(w), _p(_c(beta)), float(gamma), _p(x), _p(z))


def zerr(gap, x, dtype, scale_exp=0, scaled_gap=0.0):
    x = _c(x).reshape(-1)
    f = np.finfo(dtype)
    u = 0.5 * float(f.eps)
    eta = 0.5 * float(f.smallest_subnormal)
    return lib().or_zerr(float(gap), _p(x), x.size, u, eta, int(scale_exp), float(scaled_gap))


class Deriv:
    """A derivative's result: out (gx or dz), the parameter gradients (vjp),
    the CG residual and iterations, and the number of regions."""

    def __init__(self, out, gw=None, gspatial=None, ggamma=None, residual=0.0, iters=0, regions=0):
        self.out, self.gw, self.gspatial, self.ggamma = out, gw, gspatial, ggamma
        self.residual, self.iters, self.regions = residual, iters, regions


def iso_deriv(x, z, u, gap, scale_exp, gamma, w=None, alpha=None, *, g=None, dx=None, dgamma=0.0, dparams=None,
              deriv_tol=0.0, deriv_max_iter=10000, dual_dtype=np.float64):
    """The isotropic derivative at a solve's output (z, dual u, gap, scale
    exponent) of one image x: with g, the VJP (out = J^T g; gw, gspatial,
    ggamma the parameter gradients); else the JVP along (dx, dgamma,
    dparams). dual_dtype float32: the flat test of a float32 dual (the torch
    device path's 64 eps rules) instead of the C kernels'."""
    if np.dtype(dual_dtype) == np.float32:
        e32 = float(np.finfo(np.float32).eps)
        lib().or_deriv_set_eps(e32, max(1e-9, 64 * e32))
        try:
            return iso_deriv(x, z, u, gap, scale_exp, gamma, w, alpha, g=g, dx=dx, dgamma=dgamma, dparams=dparams,
                             deriv_tol=deriv_tol, deriv_max_iter=deriv_max_iter)
        finally:
            lib().or_deriv_set_eps(2.2e-16, 1e-9)
    x, z, u = _c(x), _c(z), _c(u)
    k = x.ndim
    w = np.ones(k) if w is None else _c(np.broadcast_to(np.asarray(w, np.float64), (k,)))
    alpha = _c(alpha)
    shape = (C.c_int64 * k)(*x.shape)
    out = np.empty_like(x)
    P = k + (x.size if alpha is not None else 0)
    acc = np.zeros(P + 1)
    res, its, nr = C.c_double(), C.c_int(), C.c_int64()
    mode = 0 if g is not None else 1
    inp = _c(g) if g is not None else _c(dx)
    dp = None if dparams is None else _c(np.asarray(dparams, np.float64).reshape(-1))
    s = lib().or_iso_deriv(k, shape, _p(w), _p(alpha), float(gamma), _p(x), _p(z), _p(u), float(gap),
                           int(scale_exp), mode, _p(inp), float(dgamma), _p(dp), float(deriv_tol),
                           int(deriv_max_iter), _p(out), _p(acc), C.byref(res), C.byref(its), C.byref(nr))
    if s:
        raise MemoryError
    if mode == 0:
        return Deriv(out, acc[:k].copy(), acc[k:P].reshape(x.shape).copy() if alpha is not None else None,
                     float(acc[P]), res.value, its.value, nr.value)
    return Deriv(out, residual=res.value, iters=its.value, regions=nr.value)


def aniso_deriv(x, z, u, gamma, w=None, beta=None, dtype=np.float64, *, g=None, dx=None, dgamma=0.0, dw=None,
                dbeta=None):
    """The device method's anisotropic derivative at (z, per-edge dual u (k,
    *img)) of one image x: with g, the VJP; else the JVP."""
    x, z, u = _c(x), _c(z), _c(u)
    k = x.ndim
    w = np.ones(k) if w is None else _c(np.broadcast_to(np.asarray(w, np.float64), (k,)))
    beta = _c(beta)
    eps = float(np.finfo(dtype).eps)
    eta = max(1e-9, 64 * eps)
    shape = (C.c_int64 * k)(*x.shape)
    out = np.empty_like(x)
    acc = np.zeros(k + 1 + (k * x.size if beta is not None else 0))
    nr = C.c_int64()
    mode = 0 if g is not None else 1
    inp = _c(g) if g is not None else _c(dx)
    s = lib().or_aniso_deriv(k, shape, _p(w), _p(beta), float(gamma), _p(x), _p(z), _p(u), eta, eps, mode,
                             _p(inp), float(dgamma), _p(_c(dw)), _p(_c(dbeta)), _p(out), _p(acc), C.byref(nr))
    if s:
        raise MemoryError
    if mode == 0:
        return Deriv(out, acc[:k].copy(), acc[k + 1:].reshape((k,) + x.shape).copy() if beta is not None else None,
                     float(acc[k]), regions=nr.value)
    return Deriv(out, regions=nr.value)

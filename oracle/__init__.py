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

"""Operators and workspaces of libfastprox_rocm behind a small Python API.

The calls go through the C ABI (include/fastprox_rocm.h) with ctypes:
torch tensors on the GPU are passed as device arrays (FP_DEVICE_ROCM), numpy
arrays and CPU tensors as host arrays. The library is required: without the
built extension every call raises (there is no CPU fallback)."""

from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _capi as A

_lib = None
_lock = threading.Lock()


def lib():
    """The loaded library; raises OSError when it is not built."""
    global _lib
    with _lock:
        if _lib is None:
            _lib = A.load()
    return _lib


class FastproxError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        L = lib()
        msg = L.fp_strerror(status).decode()
        f = getattr(L, "fpr_last_error", None)
        if f is not None:
            f.restype = C.c_char_p
            detail = f().decode()
            if detail:
                msg += f": {detail}"
        super().__init__(f"{what}: {A.STATUS.get(status, status)} ({msg})" if what else msg)


def _check(status, what, ok=(A.FP_OK,)):
    if status not in ok:
        raise FastproxError(status, what)
    return status


def _is_torch(x):
    t = type(x)
    return t.__module__.startswith("torch") and t.__name__ in ("Tensor", "Parameter")


def as_array(x):
    """An A.Array over a torch tensor (device or host) or a numpy array."""
    if _is_torch(x):
        import torch
        dt = {torch.float32: A.FP_F32, torch.float64: A.FP_F64, torch.int8: A.FP_I8,
              torch.int32: A.FP_I32, torch.int64: A.FP_I64}.get(x.dtype)
        if dt is None:
            raise TypeError(f"unsupported dtype {x.dtype}")
        dev = A.FP_DEVICE_ROCM if x.is_cuda else A.FP_DEVICE_CPU
        return A.Array(x.data_ptr(), dt, tuple(x.shape), tuple(x.stride()), dev)
    x = np.asarray(x)
    return A.numpy_array(x)


class Op:
    """fp_op_tv_iso / fp_op_tv_aniso (include/fastprox.h:322-352): isotropic
    or anisotropic TV of the last k axes, with axis weights and optional
    spatial weights (pixel weights for iso, edge weights (k, *img) for
    aniso)."""

    def __init__(self, iso=True, k=2, weights=None, spatial=None):
        self.iso, self.k = bool(iso), int(k)
        self.h = C.c_void_p()
        L = lib()
        wa = None
        if weights is not None:
            self._w = np.ascontiguousarray(np.atleast_1d(np.asarray(weights, np.float64)))
            wa = A.numpy_array(self._w)
        if spatial is None:
            f = L.fp_op_tv_iso if iso else L.fp_op_tv_aniso
            _check(f(self.k, wa.ref if wa else None, C.byref(self.h)), "fp_op")
        else:
            self._s = np.ascontiguousarray(np.asarray(spatial, np.float64))
            sa = A.numpy_array(self._s)
            f = L.fp_op_tv_iso_weighted if iso else L.fp_op_tv_aniso_weighted
            _check(f(self.k, wa.ref if wa else None, sa.ref, C.byref(self.h)), "fp_op")

    def struct_kind(self):
        """fp_op_struct_kind: A.FP_STRUCT_CURVED (iso) or FP_STRUCT_PROJECTOR (aniso)."""
        k = C.c_int()
        _check(lib().fp_op_struct_kind(self.h, C.byref(k)), "fp_op_struct_kind")
        return k.value

    @property
    def nparams(self):
        return int(lib().fp_op_nparams(self.h))

    @property
    def name(self):
        return lib().fp_op_name(self.h).decode()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None and self.h.value:
            _lib.fp_op_free(self.h)
            self.h = C.c_void_p()


class Workspace:
    """fp_workspace_create (include/fastprox.h:411): the device buffers of
    every solve for arrays shaped like `like`; calls on one workspace must
    not overlap. stream: a HIP stream handle (int) or None."""

    def __init__(self, op, like, keep_structure=False, stream=None, kernel=None, method=None):
        self.op = op
        self.h = C.c_void_p()
        L = lib()
        opt = A.fp_options()
        L.fp_options_init(C.byref(opt))
        opt.keep_structure = int(keep_structure)
        arr = as_array(like)
        _check(L.fp_workspace_create(op.h, arr.ref, C.byref(opt), C.byref(self.h)), "fp_workspace_create")
        self.shape = tuple(like.shape)
        self.keep = bool(keep_structure)
        if stream is not None:
            self.set_stream(stream)
        if kernel is not None:
            self.set_kernel(*kernel)
        if method is not None:
            self.set_method(method)

    def set_method(self, method):
        """fpr_workspace_set_method: "adbpg" (the reference CPU library's
        method, the default for anisotropic TV without edge weights) or
        "fista" (dual FISTA on the edge duals)."""
        _check(lib().fpr_workspace_set_method(self.h, A.METHODS[method]), "set_method")

    def method(self):
        m = C.c_int()
        _check(lib().fpr_workspace_get_method(self.h, C.byref(m)), "get_method")
        return {v: k for k, v in A.METHODS.items()}[m.value]

    def polished(self, n):
        """Per image of the last ADBPG solve, whether its polish certified
        it; and the fibres that took the taut string."""
        p = (C.c_int * max(n, 1))()
        f = C.c_int()
        _check(lib().fpr_workspace_polished(self.h, p, n, C.byref(f)), "polished")
        return [bool(p[i]) for i in range(n)], f.value

    def set_stream(self, stream):
        _check(lib().fp_workspace_set_stream(self.h, C.c_void_p(int(stream) if stream else None)), "set_stream")

    def set_kernel(self, generic, tt=1, ry=4):
        _check(lib().fpr_workspace_set_kernel(self.h, int(generic), int(tt), int(ry)), "set_kernel")

    def set_tail(self, tt, ry, frac):
        """fpr_workspace_set_tail: the streaming kernel (tt, ry) once the live
        tiles fall below frac (0: off)."""
        f = lib().fpr_workspace_set_tail
        f.restype = C.c_int
        f.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double]
        _check(f(self.h, int(tt), int(ry), float(frac)), "set_tail")

    def kernel(self):
        g, t, r = C.c_int(), C.c_int(), C.c_int()
        _check(lib().fpr_workspace_get_kernel(self.h, C.byref(g), C.byref(t), C.byref(r)), "get_kernel")
        return g.value, t.value, r.value  # kind: 0 stream 2D, 1 generic, 2 block 2D, 3 stream 3D

    def stats(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().fpr_workspace_stats(self.h, C.byref(a), C.byref(b)), "stats")
        return a.value, b.value

    def prox(self, x, gamma, out, tol=1e-6, max_iter=1000, keep_structure=None, with_info=True):
        """fp_prox (one gamma) or fp_prox_steps (an array of steps, the batch
        shape). Returns the status (FP_OK or FP_ENOTCONV); raises on errors.
        with_info=False passes no fp_info (the form a stream under hipGraph
        capture takes: the results are read later by image_info/info)."""
        L = lib()
        opt = A.fp_options()
        L.fp_options_init(C.byref(opt))
        opt.tol, opt.max_iter = float(tol), int(max_iter)
        opt.keep_structure = int(self.keep if keep_structure is None else keep_structure)
        xa, oa = as_array(x), as_array(out)
        if isinstance(gamma, (int, float, np.floating, np.integer)):
            info = A.fp_info()
            s = L.fp_prox(self.op.h, float(gamma), xa.ref, oa.ref, C.byref(opt), self.h,
                          C.byref(info) if with_info else None)
        else:
            ga = as_array(gamma)
            s = L.fp_prox_steps(self.op.h, ga.ref, xa.ref, oa.ref, C.byref(opt), self.h)
        return _check(s, "fp_prox", ok=(A.FP_OK, A.FP_ENOTCONV))

    def info(self):
        inf = A.fp_info()
        s = lib().fp_workspace_info(self.h, C.byref(inf))
        _check(s, "fp_workspace_info", ok=(A.FP_OK, A.FP_ENOTCONV))
        return inf.gap, inf.iters, s == A.FP_OK

    def image_info(self, count):
        gaps, its, zerrs = np.empty(count), np.empty(count, np.int32), np.empty(count)
        _check(lib().fp_workspace_image_info(self.h, gaps.ctypes.data_as(C.POINTER(C.c_double)),
                                             its.ctypes.data_as(C.POINTER(C.c_int)),
                                             zerrs.ctypes.data_as(C.POINTER(C.c_double)), int(count)),
               "fp_workspace_image_info")
        return gaps, its, zerrs

    def refine(self, tol, max_iter=100000, out=None):
        """fp_workspace_refine: the kept call solved again from its input and
        dual to tol; returns (status, gap, iters)."""
        inf = A.fp_info()
        s = lib().fp_workspace_refine(self.op.h, self.h, float(tol), int(max_iter),
                                      as_array(out).ref if out is not None else None, C.byref(inf))
        _check(s, "fp_workspace_refine", ok=(A.FP_OK, A.FP_ENOTCONV))
        return s, inf.gap, inf.iters

    # ---- structure of the derivative and kinks (fp_workspace_units ...) ----

    def units(self):
        """(count, size): the images of the kept point and their entries."""
        c, n = C.c_int64(), C.c_int64()
        _check(lib().fp_workspace_units(self.h, C.byref(c), C.byref(n)), "fp_workspace_units")
        return c.value, n.value

    def states(self, out=None):
        """fp_workspace_states into out (int8, x's shape) if given; the number
        of entries not FREE."""
        nf = C.c_int64()
        _check(lib().fp_workspace_states(self.h, as_array(out).ref if out is not None else None, C.byref(nf)),
               "fp_workspace_states")
        return nf.value

    def regions(self, labels=None, sign=None):
        """fp_workspace_regions into labels (int32 or int64) and sign (int8),
        each x's shape, if given; the regions per image (numpy int64)."""
        count = self.units()[0]
        nr = np.zeros(max(count, 1), np.int64)
        _check(lib().fp_workspace_regions(self.h, as_array(labels).ref if labels is not None else None,
                                          as_array(sign).ref if sign is not None else None,
                                          nr.ctypes.data_as(C.POINTER(C.c_int64)), count), "fp_workspace_regions")
        return nr[:count]

    def constraint(self, coef=None):
        """fp_workspace_constraint: coef (x's dtype and shape) if given; the
        per-image active flags (numpy int8)."""
        count = self.units()[0]
        act = np.ones(max(count, 1), np.int8)
        _check(lib().fp_workspace_constraint(self.h, as_array(coef).ref if coef is not None else None,
                                             act.ctypes.data_as(C.POINTER(C.c_int8)), count),
               "fp_workspace_constraint")
        return act[:count]

    def kinks(self, tol, mask=None):
        """fp_workspace_kinks at tol into mask (int8, x's shape) if given; the
        entries at a kink."""
        c = C.c_int64()
        _check(lib().fp_workspace_kinks(self.h, float(tol), as_array(mask).ref if mask is not None else None,
                                        C.byref(c)), "fp_workspace_kinks")
        return c.value

    def set_branch(self, side, tol=0.0):
        """fp_workspace_set_branch (anisotropic TV): side (int8, x's shape,
        values -1, 0, +1) picks, at every kink edge within tol, whether fp_jvp
        and fp_vjp join its sides (-1) or keep them apart (+1); None restores
        the default rule."""
        _check(lib().fp_workspace_set_branch(self.h, as_array(side).ref if side is not None else None, float(tol)),
               "fp_workspace_set_branch")

    def get_dual(self, out):
        _check(lib().fp_workspace_get_dual(self.h, as_array(out).ref), "fp_workspace_get_dual")
        return out

    def set_dual(self, dual):
        _check(lib().fp_workspace_set_dual(self.h, as_array(dual).ref), "fp_workspace_set_dual")

    def set_warm(self, on):
        _check(lib().fp_workspace_set_warm(self.h, int(on)), "fp_workspace_set_warm")

    def set_deriv_tol(self, tol, max_iter=10000):
        _check(lib().fp_workspace_set_deriv_tol(self.h, float(tol), int(max_iter)), "set_deriv_tol")

    def deriv_info(self):
        r, i = C.c_double(), C.c_int()
        _check(lib().fp_workspace_deriv_info(self.h, C.byref(r), C.byref(i)), "deriv_info")
        return r.value, i.value

    def vjp(self, g, gx, gparams=None, ggamma_dev=None):
        """fp_vjp (ggamma returned as a float) or, with ggamma_dev (a 0-d
        device tensor), fp_vjp_device."""
        L = lib()
        gp = as_array(gparams) if gparams is not None else None
        if ggamma_dev is not None:
            _check(L.fp_vjp_device(self.op.h, self.h, as_array(g).ref, as_array(gx).ref,
                                   as_array(ggamma_dev).ref, gp.ref if gp else None), "fp_vjp_device")
            return None
        gg = C.c_double()
        _check(L.fp_vjp(self.op.h, self.h, as_array(g).ref, as_array(gx).ref, C.byref(gg),
                        gp.ref if gp else None), "fp_vjp")
        return gg.value

    def jvp(self, dx, dout, dgamma=0.0, dparams=None):
        dp = as_array(dparams) if dparams is not None else None
        _check(lib().fp_jvp(self.op.h, self.h, as_array(dx).ref if dx is not None else None, float(dgamma),
                            dp.ref if dp else None, as_array(dout).ref), "fp_jvp")
        return dout

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None and self.h.value:
            _lib.fp_workspace_free(self.h)
            self.h = C.c_void_p()

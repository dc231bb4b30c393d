"""PyTorch entry points for image TV on the GPU, with the signatures and
semantics of fastprox.torch.tv_iso / tv_aniso
(bindings/python/fastprox/torch.py:305-351 in the reference).

    z = tv_iso(x, 0.1)                       # x (..., H, W) on cuda, f32/f64
    z, info = tv_aniso(x, 0.05, weights=[1.0, 0.5], tol=1e-6, return_info=True)

Every solve runs in libfastprox_rocm (HIP kernels) on the tensor's device and
torch's current stream; the derivatives (reverse and forward mode) are the
implicit-differentiation systems of the reference, also on the device. There
is no CPU fallback: a CPU tensor or a missing library raises."""

from __future__ import annotations

import contextlib
import contextvars
import ctypes as C
import math
import threading
import warnings

import numpy as np
import torch

from . import IterInfo
from . import _capi as A
from .core import FastproxError, Op, _check, lib


class AccuracyWarning(UserWarning):
    """An iterative prox certified less accuracy than asked for
    (bindings/python/fastprox/__init__.py:602-605)."""


class ConvergenceWarning(AccuracyWarning):
    """A solve stopped at max_iter before its tolerance
    (bindings/python/fastprox/__init__.py:608-612)."""


_ACCURACY = contextvars.ContextVar("fastprox_amd_inner_accuracy", default=None)


@contextlib.contextmanager
def inner_accuracy(target, factor=0.1):
    """Calls in the block that set neither tol nor ztol use ztol = factor *
    target (bindings/python/fastprox/__init__.py:618-645)."""
    target, factor = float(target), float(factor)
    if not (0.0 < target < float("inf") and 0.0 < factor < float("inf")):
        raise ValueError(f"target and factor must be finite and > 0, got {target} and {factor}")
    ztol = factor * target
    _sq_tol(ztol)
    token = _ACCURACY.set(ztol)
    try:
        yield ztol
    finally:
        _ACCURACY.reset(token)


def _sq_tol(ztol):
    """The gap tolerance that certifies a ztol: the gap bound |z - z*|^2 <=
    2 gap with gap <= tol |x|^2 / 2 gives |z - z*| <= sqrt(tol) |x|, so ztol
    asks for tol = ztol^2, which must not underflow."""
    v = float(ztol)
    if not (v >= 0.0 and math.isfinite(v)):
        raise ValueError(f"ztol must be finite and >= 0, got {v}")
    sq = v * v
    if v > 0.0 and sq < np.finfo(np.float64).tiny:
        raise ValueError(f"ztol={v} is below 1e-154: its square, the gap tolerance, underflows")
    return sq


def _call_tol(tol, ztol, default, warn=True):
    """The gap tolerance of one call (the reference's precedence,
    bindings/python/fastprox/__init__.py:663-679): its own tol, or its ztol
    squared; without either, the enclosing inner_accuracy block's ztol
    squared; outside a block, the default. An own tolerance looser than the
    block's draws an AccuracyWarning."""
    if tol is not None and ztol is not None:
        raise ValueError("pass tol or ztol, not both")
    block = _ACCURACY.get()
    if tol is not None:
        own, label = float(tol), f"tol={float(tol):g}"
    elif ztol is not None:
        own, label = _sq_tol(ztol), f"ztol={float(ztol):g}"
    else:
        return default if block is None else _sq_tol(block)
    if warn and block is not None and own > _sq_tol(block):
        warnings.warn(f"{label} certifies less than inner_accuracy's ztol={block:g}", AccuracyWarning,
                      stacklevel=4)
    return own


def _check_floor(x, tol, ztol, warn):
    """__init__.py:82-94: a ztol below the dtype's unit roundoff cannot be certified."""
    if not warn:
        return
    need = ztol if ztol is not None else (_ACCURACY.get() if tol is None else None)
    floor = 0.5 * float(torch.finfo(x.dtype).eps)
    if need is not None and 0.0 < float(need) < floor:
        warnings.warn(f"ztol={float(need):g} is below the unit roundoff {floor:.3g} of {x.dtype}",
                      AccuracyWarning, stacklevel=4)


def tv_iso(x, gamma=1.0, *, ndim=2, weights=None, pixel_weights=None, tol=None, ztol=None, max_iter=1000,
           warm_start=None, return_info=False, deriv_tol=None, deriv_max_iter=10000, warn=True):
    r"""Isotropic TV of the last ``ndim`` axes of ``x`` (torch.py:305-332 in the
    reference): prox of gamma sum_p alpha_p |(w_a D_a z)_p|_2, each image
    stopping when its duality gap is at most tol |x|^2 / 2. ``gamma``: a
    number or one step per image (broadcasting to the batch shape);
    ``pixel_weights`` broadcast to x.shape. Differentiable in x, gamma,
    weights and pixel_weights."""
    t = _call_tol(tol, ztol, 1e-6, warn)
    _check_floor(x, tol, ztol, warn)
    return _device_tv(True, x, gamma, ndim, weights, t, max_iter, warm_start, return_info, deriv_tol,
                      deriv_max_iter, pixel_weights, warn and "fastprox_amd.torch.tv_iso")


def tv_aniso(x, gamma=1.0, *, ndim=2, weights=None, edge_weights=None, tol=None, ztol=None, max_iter=1000,
             warm_start=None, return_info=False, deriv_tol=None, deriv_max_iter=10000, warn=True, method=None):
    r"""Anisotropic TV of the last ``ndim`` axes (torch.py:335-351 in the
    reference); the dual is in the C form q_a with x - z = sum_a q_a.
    ``edge_weights``: (ndim, ...) broadcasting to (ndim, *x.shape).
    ``method``: "fista" (the default), dual FISTA on the edge duals, the
    reference's device path (_tv_torch.py:209-282); or "adbpg", the
    reference CPU library's method (exact 1D steps on every fibre, block-dual
    FISTA, polish: its iteration counts and results; no edge weights)."""
    if method not in (None, "adbpg", "fista"):
        raise ValueError('method must be "adbpg", "fista" or None')
    if method == "adbpg" and edge_weights is not None:
        raise ValueError('method "adbpg" takes no edge_weights')
    t = _call_tol(tol, ztol, 1e-6, warn)
    _check_floor(x, tol, ztol, warn)
    return _device_tv(False, x, gamma, ndim, weights, t, max_iter, warm_start, return_info, deriv_tol,
                      deriv_max_iter, edge_weights, warn and "fastprox_amd.torch.tv_aniso", method)


class _Cfg:
    """The fixed settings of one call, shared by its forward and its
    derivatives (which record their CG residual and iterations in deriv)."""

    def __init__(self, iso, k, w, tol, max_iter, deriv_tol, deriv_max_iter, method=None):
        self.iso, self.k, self.w = iso, k, w
        self.method = method
        self.tol, self.max_iter = float(tol), int(max_iter)
        self.deriv_tol, self.deriv_max_iter = deriv_tol, int(deriv_max_iter)
        self.info = None
        self.deriv = None


def _broadcasts(src, dst):
    """Whether shape src broadcasts to shape dst: src is no longer than dst,
    and each of its trailing extents equals dst's or is 1."""
    if len(src) > len(dst):
        return False
    return all(a == b or a == 1 for a, b in zip(src[::-1], dst[::-1]))


def _as_float(v, x):
    """v (a number, sequence, array or tensor) as a tensor of x's dtype on
    x's device; a tensor stays in the autograd graph."""
    t = v if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v, np.float64))
    return t.to(dtype=x.dtype, device=x.device)


def _require_finite_nonneg(t, name):
    d = t.detach()
    if not bool((torch.isfinite(d) & (d >= 0)).all()):
        raise ValueError(f"{name} must be finite and >= 0")


def _check_input(x, k, tol, max_iter):
    """The image tensor and the solve's limits (checked in the reference's
    order, with its messages)."""
    if not isinstance(x, torch.Tensor):
        raise TypeError("x must be a torch.Tensor")
    if x.dtype not in (torch.float32, torch.float64):
        raise TypeError(f"fastprox operates on float32 and float64 tensors, got {x.dtype}")
    if k not in (2, 3):
        raise ValueError("ndim must be 2 or 3")
    if x.dim() < k:
        raise ValueError(f"needs a tensor with at least {k} dimensions")
    if not (tol >= 0 and max_iter >= 0):
        raise ValueError("tol must be >= 0 and max_iter >= 0")
    if not x.is_cuda:
        raise ValueError("fastprox_amd runs on the GPU: x must be a cuda (ROCm) tensor")


def _axis_weights(weights, k):
    """(the k axis weights as floats, the weights tensor when it is a float
    tensor the gradient flows to, else None). One value is every axis's."""
    if weights is None:
        return [1.0] * k, None
    wt = weights if isinstance(weights, torch.Tensor) and weights.is_floating_point() else None
    raw = weights.detach().cpu().numpy() if isinstance(weights, torch.Tensor) else weights
    vals = [float(v) for v in np.ravel(np.asarray(raw, np.float64))]
    if len(vals) == 1:
        vals *= k
    if len(vals) != k or not all(math.isfinite(v) and v >= 0 for v in vals):
        raise ValueError("weights must be one or ndim finite values >= 0")
    return vals, wt


def _per_image(gamma, x, k):
    """One step per image, shape (B,): gamma is a number or broadcasts to
    the batch shape x.shape[:-k]; a tensor stays in the graph, so its
    gradient sums back to its own shape."""
    batch = tuple(x.shape[:-k])
    g = _as_float(gamma, x)
    if g.numel() == 1:
        g = g.reshape(())
    if not _broadcasts(tuple(g.shape), batch):
        raise ValueError(f"gamma must be a number or broadcast to the batch shape {batch}, "
                         f"got shape {tuple(g.shape)}")
    _require_finite_nonneg(g, "gamma")
    return g.expand(batch).reshape(-1)


def _spatial(sw, x, k, iso):
    """Spatial weights in the layout of the batch solve: pixel weights that
    broadcast to x.shape as (B, *img); edge weights of shape (k, *s), s
    broadcasting to x.shape, as (k, B, *img). A tensor stays in the graph."""
    name = "pixel_weights" if iso else "edge_weights"
    full = tuple(x.shape) if iso else (k,) + tuple(x.shape)
    t = _as_float(sw, x)
    shape = tuple(t.shape)
    fits = True
    if not iso:
        # axis 0 holds the k edge directions; unit axes after it bring the
        # rest to x's rank, aligned on the right
        fits = len(shape) >= 1 and shape[0] == k and len(shape) - 1 <= x.dim()
        if fits:
            shape = (k,) + (1,) * (x.dim() + 1 - len(shape)) + shape[1:]
            t = t.reshape(shape)
    if not (fits and _broadcasts(shape, full)):
        what = "broadcast to" if iso else f"have shape ({k}, ...) broadcasting to"
        raise ValueError(f"{name} must {what} {full}, got shape {tuple(getattr(sw, 'shape', ()))}")
    _require_finite_nonneg(t, name)
    lead = (-1,) if iso else (k, -1)
    return t.expand(full).reshape(lead + tuple(x.shape[-k:]))


def _device_tv(iso, x, gamma, ndim, weights, tol, max_iter, warm_start, return_info, deriv_tol, deriv_max_iter,
               spatial, warn, method=None):
    """tv_iso and tv_aniso: the arguments checked (the reference's rules and
    messages, bindings/python/fastprox/torch.py:363-413), then one solve of
    every image of the batch on the device through _DeviceTV, which also
    takes the empty batch."""
    _check_input(x, ndim, tol, max_iter)
    w, wt = _axis_weights(weights, ndim)
    gb = _per_image(gamma, x, ndim)
    warm = None
    if warm_start is not None:
        warm = torch.as_tensor(warm_start).to(dtype=x.dtype, device=x.device)
        if tuple(warm.shape) != (ndim, *x.shape):
            raise ValueError(f"warm_start must have shape {(ndim, *x.shape)}")
    if deriv_tol is not None:
        if not deriv_tol > 0:
            raise ValueError("deriv_tol must be > 0 or None")
        deriv_tol = float(deriv_tol)
    cfg = _Cfg(iso, ndim, w, tol, max_iter, deriv_tol, deriv_max_iter, method)
    sw = None if spatial is None else _spatial(spatial, x, ndim, iso)
    if x.numel() and bool(torch.isinf(x.detach()).any()):
        raise ValueError("x contains ±Inf; image operators reject infinite values")
    z = _DeviceTV.apply(cfg, x, gb, wt, sw, warm)
    if warn and not cfg.info.converged:
        warnings.warn(f"{warn} stopped at max_iter={max_iter} before its tolerance; the certified relative "
                      f"error of z is {cfg.info.zerr:.3g}", ConvergenceWarning, stacklevel=3)
    return (z, cfg.info) if return_info else z


# ---------------------------------------------------------------------------
# Batches through the C ABI (fpr_batch_*), cached per configuration
# ---------------------------------------------------------------------------

class _FprImage(C.Structure):
    _fields_ = [("x", C.c_void_p), ("z", C.c_void_p), ("dual", C.c_void_p), ("weights", C.c_void_p),
                ("gamma", C.c_double)]


def _declare_batch(L):
    if getattr(L, "_batch_declared", False):
        return
    P = C.c_void_p
    L.fpr_batch_create.restype = C.c_int
    L.fpr_batch_create.argtypes = [P, C.c_int, C.c_int64, C.POINTER(C.c_int64), C.c_int, C.POINTER(P)]
    L.fpr_batch_free.restype = None
    L.fpr_batch_free.argtypes = [P]
    L.fpr_batch_solve.restype = C.c_int
    L.fpr_batch_solve.argtypes = [P, C.POINTER(_FprImage), C.POINTER(A.fp_options), C.c_int, P]
    L.fpr_batch_image_info.restype = C.c_int
    L.fpr_batch_image_info.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_double),
                                       C.POINTER(C.c_int), C.c_int64]
    L.fpr_batch_set_kernel.restype = C.c_int
    L.fpr_batch_set_kernel.argtypes = [P, C.c_int, C.c_int, C.c_int]
    L.fpr_batch_stats.restype = C.c_int
    L.fpr_batch_stats.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fpr_batch_pixel_passes.restype = C.c_int
    L.fpr_batch_pixel_passes.argtypes = [P, C.POINTER(C.c_double)]
    L.fpr_batch_set_timing.restype = C.c_int
    L.fpr_batch_set_timing.argtypes = [P, C.c_int]
    L.fpr_batch_set_graph.restype = C.c_int
    L.fpr_batch_set_graph.argtypes = [P, C.c_int]
    L.fpr_batch_timing.restype = C.c_int
    L.fpr_batch_timing.argtypes = [P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_int)]
    L.fpr_batch_set_tail.restype = C.c_int
    L.fpr_batch_set_tail.argtypes = [P, C.c_int, C.c_int, C.c_double]
    L.fpr_batch_scale_exps.restype = C.c_int
    L.fpr_batch_scale_exps.argtypes = [P, C.POINTER(C.c_int), C.c_int64]
    L.fpr_batch_copy_dual_iterate.restype = C.c_int
    L.fpr_batch_copy_dual_iterate.argtypes = [P, P, P]
    L.fpr_batch_set_method.restype = C.c_int
    L.fpr_batch_set_method.argtypes = [P, C.c_int]
    L.fpr_batch_get_method.restype = C.c_int
    L.fpr_batch_get_method.argtypes = [P, C.POINTER(C.c_int)]
    L._batch_declared = True


class Batch:
    """fpr_batch_create: one persistent solve over images of (possibly)
    different shapes, each with its own step and spatial weights."""

    def __init__(self, op, dtype, shapes, weighted=False, kernel=None, method=None):
        L = lib()
        _declare_batch(L)
        self.op = op
        self.k = op.k
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        flat = np.ascontiguousarray(np.asarray(self.shapes, np.int64).reshape(-1))
        self.h = C.c_void_p()
        dt = A.FP_F32 if dtype == torch.float32 else A.FP_F64
        _check(L.fpr_batch_create(op.h, dt, len(self.shapes), flat.ctypes.data_as(C.POINTER(C.c_int64)),
                                  int(weighted), C.byref(self.h)), "fpr_batch_create")
        self.weighted = bool(weighted)
        if kernel is not None:
            self.set_kernel(*kernel)
        if method is not None:
            self.set_method(method)

    def set_kernel(self, generic, tt=1, ry=4):
        _check(lib().fpr_batch_set_kernel(self.h, int(generic), int(tt), int(ry)), "fpr_batch_set_kernel")

    def set_method(self, method):
        """fpr_batch_set_method: "adbpg" or "fista" (anisotropic batches)."""
        _check(lib().fpr_batch_set_method(self.h, A.METHODS[method]), "fpr_batch_set_method")

    def method(self):
        m = C.c_int()
        _check(lib().fpr_batch_get_method(self.h, C.byref(m)), "fpr_batch_get_method")
        return {v: k for k, v in A.METHODS.items()}[m.value]

    def solve(self, xs, zs, duals, gammas, weights=None, tol=1e-6, max_iter=1000, warm=False, stream=None):
        """Lists of device tensors (contiguous); returns the status."""
        n = len(self.shapes)
        imgs = (_FprImage * max(n, 1))()
        for i in range(n):
            imgs[i].x = xs[i].data_ptr()
            imgs[i].z = zs[i].data_ptr()
            imgs[i].dual = duals[i].data_ptr()
            imgs[i].weights = weights[i].data_ptr() if weights is not None else None
            imgs[i].gamma = float(gammas[i])
        opt = A.fp_options()
        lib().fp_options_init(C.byref(opt))
        opt.tol, opt.max_iter = float(tol), int(max_iter)
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        s = lib().fpr_batch_solve(self.h, imgs, C.byref(opt), int(warm), C.c_void_p(stream or None))
        return _check(s, "fpr_batch_solve", ok=(A.FP_OK, A.FP_ENOTCONV))

    def image_info(self):
        n = len(self.shapes)
        gaps, its, zerrs, conv = np.empty(n), np.empty(n, np.int32), np.empty(n), np.empty(n, np.int32)
        _check(lib().fpr_batch_image_info(self.h, gaps.ctypes.data_as(C.POINTER(C.c_double)),
                                          its.ctypes.data_as(C.POINTER(C.c_int)),
                                          zerrs.ctypes.data_as(C.POINTER(C.c_double)),
                                          conv.ctypes.data_as(C.POINTER(C.c_int)), n), "fpr_batch_image_info")
        return gaps, its, zerrs, conv.astype(bool)

    def stats(self):
        a, b = C.c_int(), C.c_int()
        _check(lib().fpr_batch_stats(self.h, C.byref(a), C.byref(b)), "fpr_batch_stats")
        return a.value, b.value

    def pixel_passes(self):
        """The last solve's pixels times the step launches that iterated them
        (fpr_batch_pixel_passes)."""
        v = C.c_double()
        _check(lib().fpr_batch_pixel_passes(self.h, C.byref(v)), "fpr_batch_pixel_passes")
        return v.value

    def set_tail(self, tt, ry, frac):
        """Once the live tiles drop below frac, continue on the streaming
        kernel (tt, ry) (fpr_batch_set_tail); frac 0 turns it off."""
        _check(lib().fpr_batch_set_tail(self.h, int(tt), int(ry), float(frac)), "fpr_batch_set_tail")

    def scale_exps(self):
        e = np.zeros(len(self.shapes), np.int32)
        _check(lib().fpr_batch_scale_exps(self.h, e.ctypes.data_as(C.POINTER(C.c_int)), len(self.shapes)),
               "fpr_batch_scale_exps")
        return e

    def copy_dual_iterate(self, dst, stream=None):
        """The per-edge (or per-pixel) dual iterate of every image into dst,
        a contiguous (B, k, *img) device tensor of the batch's dtype."""
        if stream is None:
            stream = torch.cuda.current_stream().cuda_stream
        _check(lib().fpr_batch_copy_dual_iterate(self.h, C.c_void_p(dst.data_ptr()), C.c_void_p(stream or None)),
               "fpr_batch_copy_dual_iterate")

    def set_timing(self, on):
        _check(lib().fpr_batch_set_timing(self.h, int(on)), "fpr_batch_set_timing")

    def set_graph(self, on):
        """Replay captured hipGraphs of the launch chunks (fpr_batch_set_graph)."""
        _check(lib().fpr_batch_set_graph(self.h, int(on)), "fpr_batch_set_graph")

    def timing(self):
        """(step-kernel ms, epilogue ms, step launches) of the last solve."""
        a, b, c = C.c_double(), C.c_double(), C.c_int()
        _check(lib().fpr_batch_timing(self.h, C.byref(a), C.byref(b), C.byref(c)), "fpr_batch_timing")
        return a.value, b.value, c.value

    def busy_ms(self):
        """Time at least one step kernel of the last timed solve ran
        (fpr_batch_busy): below the summed step time when the batch's two
        halves ran on two streams at once."""
        a = C.c_double()
        _check(lib().fpr_batch_busy(self.h, C.byref(a)), "fpr_batch_busy")
        return a.value

    def __del__(self):
        try:
            if getattr(self, "h", None) and self.h.value:
                lib().fpr_batch_free(self.h)
                self.h = None
        except Exception:  # interpreter shutdown
            pass


_CACHE = {}
_CACHE_LOCK = threading.Lock()
_CACHE_MAX = 16


def _batch_for(iso, k, w, dtype, device, B, img, weighted, method=None):
    key = (iso, k, tuple(w), dtype, device, B, img, weighted, method)
    with _CACHE_LOCK:
        hit = _CACHE.pop(key, None)
        if hit is None:
            with torch.cuda.device(device):
                op = Op(iso, k, w)
                hit = Batch(op, dtype, [img] * B, weighted=weighted, method=method)
        _CACHE[key] = hit
        while len(_CACHE) > _CACHE_MAX:
            _CACHE.pop(next(iter(_CACHE)))
        return hit


def solve(x, gamma, w, iso, tol, max_iter, warm=None, sw=None, keep=False, method=None):
    """(z, dual (k, B, *img) in the C form, gaps, iters, zerrs, converged)
    for x (B, *img) on the GPU; gamma (B,); warm: a dual (k, B, *img) or None;
    sw: pixel weights (B, *img) or edge weights (k, B, *img). keep: also the
    scale exponents and, anisotropic, the per-edge dual (B, k, *img) that the
    derivatives read."""
    k = len(w)
    B = x.shape[0]
    img = tuple(x.shape[1:])
    xb = x.detach().contiguous()
    z = torch.empty_like(xb)
    dual = torch.empty((B, k) + img, dtype=x.dtype, device=x.device)
    if warm is not None:
        dual.copy_(warm.transpose(0, 1))
    wlist = None
    if sw is not None:
        swd = sw.detach().to(dtype=x.dtype)
        wlist = [swd[b].contiguous() for b in range(B)] if iso else \
            [swd[:, b].contiguous() for b in range(B)]
    bt = _batch_for(iso, k, w, x.dtype, x.device.index or 0, B, img, sw is not None, method)
    g = gamma.detach().double().cpu().numpy().reshape(-1)
    with torch.cuda.device(x.device):
        bt.solve(list(xb), list(z), list(dual), g, weights=wlist, tol=tol, max_iter=max_iter,
                 warm=warm is not None)
    gaps, its, zerrs, conv = bt.image_info()
    if not keep:
        return z, dual.transpose(0, 1), gaps, its, zerrs, conv
    exps = bt.scale_exps()
    uedge = None
    if not iso:
        uedge = torch.empty((B, k) + img, dtype=x.dtype, device=x.device)
        with torch.cuda.device(x.device):
            bt.copy_dual_iterate(uedge)
    return z, dual.transpose(0, 1), gaps, its, zerrs, conv, exps, uedge


# The last backward's (CG residual, CG iterations, regions): the largest over
# the images (0, 0 for anisotropic TV, whose region means are exact).
LAST_DERIV = None


class _DeviceTV(torch.autograd.Function):
    """The solve as an autograd function (the reference's _DeviceTV,
    bindings/python/fastprox/torch.py:466-524): the forward solves every
    image on the device and keeps the solution; reverse and forward mode
    are the implicit-differentiation systems at it (deriv.py), built once
    and shared by both. Inputs: (cfg, x, per-image steps, axis-weight
    tensor or None, spatial weights or None, warm start or None)."""

    @staticmethod
    def forward(ctx, cfg, x, gamma, wt, sw, warm):
        k = cfg.k
        shape = tuple(x.shape)
        img = shape[-k:]
        ctx.cfg, ctx.shape, ctx.wt = cfg, shape, wt
        ctx.sw = None if sw is None else sw.detach()
        ctx.S = None
        ctx.empty = x.numel() == 0
        if ctx.empty:
            cfg.info = IterInfo(0.0, 0, 0, True, torch.zeros((k,) + shape, dtype=x.dtype, device=x.device), 0.0)
            return x.detach().clone()
        xb = x.detach().reshape((-1,) + img)
        wb = None if warm is None else warm.reshape((k, -1) + img)
        z, dual, gaps, its, zerrs, conv, exps, uedge = solve(xb, gamma.detach(), cfg.w, cfg.iso, cfg.tol,
                                                             cfg.max_iter, wb, ctx.sw, keep=True,
                                                             method=cfg.method)
        n = its.size
        cfg.info = IterInfo(float(gaps.sum()), int(its.max()) if n else 0, 0, bool(conv.all()),
                            dual.reshape((k,) + shape), float(zerrs.max()) if n else 0.0, gaps, its)
        ctx.xb, ctx.z, ctx.dual, ctx.gam = xb, z, dual, gamma.detach()
        ctx.gaps, ctx.exps, ctx.uedge = gaps, exps, uedge
        return z.reshape(shape)

    @staticmethod
    def _structure(ctx):
        if ctx.S is None:
            from .deriv import Structure
            cfg = ctx.cfg
            u = ctx.dual if cfg.iso else ctx.uedge
            ctx.S = Structure(ctx.xb, ctx.z, u, ctx.gaps, ctx.exps, ctx.gam, cfg.w, cfg.iso, ctx.sw)
        return ctx.S

    @staticmethod
    def backward(ctx, c):
        if torch.is_grad_enabled():
            raise RuntimeError("fastprox_amd: second-order differentiation through the prox is not supported; "
                               "use create_graph=False")
        need = ctx.needs_input_grad  # (cfg, x, gamma, wt, sw, warm)
        grads = [None] * 6
        if ctx.empty:  # nothing to differentiate: zero gradients of the inputs' shapes
            grads[1] = torch.zeros(ctx.shape, dtype=c.dtype, device=c.device) if need[1] else None
            grads[2] = c.new_zeros((0,)) if need[2] else None
            if ctx.wt is not None and need[3]:
                grads[3] = torch.zeros_like(ctx.wt)
            if ctx.sw is not None and need[4]:
                grads[4] = torch.zeros_like(ctx.sw)
            return tuple(grads)
        cfg = ctx.cfg
        S = _DeviceTV._structure(ctx)
        with torch.no_grad():
            gx, gg, gw, gsw, cfg.deriv = S.vjp(c.reshape(ctx.z.shape).contiguous(), cfg.deriv_tol,
                                               cfg.deriv_max_iter)
        global LAST_DERIV
        LAST_DERIV = (cfg.deriv[0], cfg.deriv[1], S.regions())
        if need[1]:
            grads[1] = gx.reshape(ctx.shape)
        if need[2]:
            grads[2] = gg
        if ctx.wt is not None and need[3]:
            # one weight tensor for every axis takes the sum of the axes' gradients
            gw = gw.to(dtype=ctx.wt.dtype, device=ctx.wt.device)
            grads[3] = (gw.sum() if ctx.wt.numel() == 1 else gw).reshape(ctx.wt.shape)
        if ctx.sw is not None and need[4]:
            grads[4] = gsw
        return tuple(grads)

    @staticmethod
    def jvp(ctx, _cfg, dx, dgamma, dwt, dsw, _dwarm):
        if ctx.empty:
            return torch.zeros(ctx.shape, dtype=ctx.cfg.info.dual.dtype, device=ctx.cfg.info.dual.device)
        cfg = ctx.cfg
        S = _DeviceTV._structure(ctx)
        dw = None
        if dwt is not None:
            tangent = [float(v) for v in dwt.detach().cpu().reshape(-1)]
            dw = tangent if len(tangent) == cfg.k else tangent[:1] * cfg.k
        with torch.no_grad():
            dz, cfg.deriv = S.jvp(None if dx is None else dx.reshape(ctx.z.shape), dgamma, dw, cfg.deriv_tol,
                                  cfg.deriv_max_iter, dsw)
        return dz.reshape(ctx.shape)


__all__ = ["tv_iso", "tv_aniso", "inner_accuracy", "AccuracyWarning", "ConvergenceWarning", "Batch", "solve",
           "FastproxError"]

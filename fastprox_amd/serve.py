"""The serving loop (libfastprox_rocm: serve.cpp) behind a small Python API:
requests of different sizes, steps and tolerances solved by continuous
batching on one GPU, with admission by HBM footprint and per-request
latencies.

    srv = Server(Op(True, 2), Op(True, 3), pool_bytes=8 << 30)
    i = srv.submit(x, gamma=0.1, tol=1e-6)      # x: a cuda tensor (2D or 3D)
    while srv.busy():
        srv.step()
    results = srv.poll()
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _capi as A
from .core import FastproxError, _check, lib


class _Request(C.Structure):
    _fields_ = [("ndim", C.c_int), ("shape", C.c_int64 * 3), ("x", C.c_void_p), ("strides", C.POINTER(C.c_int64)),
                ("z", C.c_void_p), ("gamma", C.c_double), ("tol", C.c_double), ("max_iter", C.c_int),
                ("keep_dual", C.c_int), ("warm_from", C.c_int64), ("stream", C.c_void_p)]


class _Result(C.Structure):
    _fields_ = [("id", C.c_int64), ("status", C.c_int), ("converged", C.c_int), ("iters", C.c_int),
                ("gap", C.c_double), ("zerr", C.c_double), ("t_submit", C.c_double), ("t_admit", C.c_double),
                ("t_done", C.c_double)]


RESULT_FIELDS = ("id", "status", "converged", "iters", "gap", "zerr", "t_submit", "t_admit", "t_done")


def _declare(L):
    if getattr(L, "_serve_declared", False):
        return
    P = C.c_void_p
    L.fpr_server_create.restype = C.c_int
    L.fpr_server_create.argtypes = [P, P, C.c_int, C.c_size_t, C.c_int, C.POINTER(P)]
    L.fpr_server_free.restype = None
    L.fpr_server_free.argtypes = [P]
    L.fpr_server_submit.restype = C.c_int
    L.fpr_server_submit.argtypes = [P, C.POINTER(_Request), C.c_double, C.POINTER(C.c_int64)]
    L.fpr_server_step.restype = C.c_int
    L.fpr_server_step.argtypes = [P, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.fpr_server_poll.restype = C.c_int64
    L.fpr_server_poll.argtypes = [P, C.POINTER(_Result), C.c_int64]
    L.fpr_server_clock.restype = C.c_double
    L.fpr_server_clock.argtypes = [P]
    L.fpr_server_release.restype = C.c_int
    L.fpr_server_release.argtypes = [P, C.c_int64]
    L.fpr_server_stats.restype = C.c_int
    L.fpr_server_stats.argtypes = [P, C.POINTER(C.c_int64)]
    L.fpr_server_get_kernel.restype = C.c_int
    L.fpr_server_get_kernel.argtypes = [P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L._serve_declared = True


class Server:
    """fpr_server_create: op2d / op3d (fastprox_amd.Op, or None) serve 2D
    and 3D requests; each class gets `slots` slots and a device pool of
    pool_bytes."""

    def __init__(self, op2d=None, op3d=None, dtype=torch.float32, pool_bytes=8 << 30, slots=256):
        L = lib()
        _declare(L)
        self.ops = (op2d, op3d)
        self.dtype = dtype
        self.h = C.c_void_p()
        dt = A.FP_F32 if dtype == torch.float32 else A.FP_F64
        _check(L.fpr_server_create(op2d.h if op2d else None, op3d.h if op3d else None, dt, int(pool_bytes),
                                   int(slots), C.byref(self.h)), "fpr_server_create")
        self._held = {}  # inputs and outputs of queued or live requests
        self._busy, self._queued = 0, 0

    def clock(self):
        return float(lib().fpr_server_clock(self.h))

    def submit(self, x, gamma, tol=1e-6, max_iter=100000, out=None, keep_dual=False, warm_from=-1, t_submit=-1.0,
               stream=None):
        """Queues the image x (a 2D or 3D cuda tensor of the server's dtype,
        any strides); out: a contiguous cuda tensor for z, or None. stream
        (default: torch's current stream): x is read after the work queued
        on it so far, and out is written (in place, at any time until then)
        before the work queued on it after the poll() that returns the
        request. Returns the request id."""
        if x.dtype != self.dtype or not x.is_cuda or x.dim() not in (2, 3):
            raise ValueError("x must be a 2D or 3D cuda tensor of the server's dtype")
        r = _Request()
        r.ndim = x.dim()
        for a in range(x.dim()):
            r.shape[a] = x.shape[a]
        st = (C.c_int64 * 3)(*x.stride(), *([0] * (3 - x.dim())))
        r.x = x.data_ptr()
        r.strides = st
        if out is not None:
            if not out.is_contiguous() or tuple(out.shape) != tuple(x.shape) or out.dtype != x.dtype:
                raise ValueError("out must be contiguous and shaped like x")
            r.z = out.data_ptr()
        r.gamma, r.tol, r.max_iter = float(gamma), float(tol), int(max_iter)
        r.keep_dual = int(bool(keep_dual))
        r.warm_from = int(warm_from)
        if stream is None:
            stream = torch.cuda.current_stream(x.device)
        r.stream = C.c_void_p(stream.cuda_stream or None)
        rid = C.c_int64()
        _check(lib().fpr_server_submit(self.h, C.byref(r), float(t_submit), C.byref(rid)), "fpr_server_submit")
        self._held[rid.value] = (x, out, st)
        self._queued += 1
        return rid.value

    def step(self):
        """One round of the loop; (busy, queued)."""
        b, q = C.c_int(), C.c_int()
        _check(lib().fpr_server_step(self.h, C.byref(b), C.byref(q)), "fpr_server_step")
        self._busy, self._queued = b.value, q.value
        return b.value, q.value

    def busy(self):
        return self._busy > 0 or self._queued > 0

    def poll(self, max_results=4096):
        """Finished requests: a list of dicts (RESULT_FIELDS)."""
        buf = (_Result * max_results)()
        n = int(lib().fpr_server_poll(self.h, buf, max_results))
        out = []
        for i in range(n):
            r = buf[i]
            out.append({f: getattr(r, f) for f in RESULT_FIELDS})
            self._held.pop(r.id, None)
        return out

    def release(self, rid):
        _check(lib().fpr_server_release(self.h, int(rid)), "fpr_server_release")

    def launches(self):
        n = C.c_int64()
        _check(lib().fpr_server_stats(self.h, C.byref(n)), "fpr_server_stats")
        return n.value

    def kernel(self, k):
        """(kind, TT, RY) of the step kernel serving k-dimensional requests
        (fpr_server_get_kernel)."""
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        _check(lib().fpr_server_get_kernel(self.h, int(k), C.byref(a), C.byref(b), C.byref(c)),
               "fpr_server_get_kernel")
        return a.value, b.value, c.value

    def drain(self):
        """Steps until every submitted request is done; their results. Raises
        FastproxError if queued requests remain with nothing live or in flight
        after a step (no later step could admit them)."""
        out = []
        while True:
            b, q = self.step()
            out += self.poll()
            if b == 0 and q == 0:
                return out + self.poll()
            if b == 0:
                raise FastproxError(A.FP_ENOMEM, f"fpr_server_step: {q} queued request(s) can never be admitted")

    def __del__(self):
        if getattr(self, "h", None) and self.h.value:
            try:
                lib().fpr_server_free(self.h)
            except Exception:
                pass
            self.h = C.c_void_p()


def replay(srv, trace, src2d, src3d, warmup=0.0, progress=None, keep_outputs=False):
    """Replays a serving trace (synth.serving_trace) against srv in real time:
    each request is submitted at its arrival time (recorded as its submit
    time, so queueing behind a late host loop counts in its latency).
    src2d / src3d: cuda tensors of source images the requests crop from.
    Returns the results (dicts, with the request's class, side and pixels;
    with keep_outputs, also "z": the request's output tensor)."""
    from .synth import crop_of
    import time
    t0 = srv.clock() + warmup
    meta = {}
    res = []
    i, n = 0, len(trace)
    last_report = 0.0
    while True:
        now = srv.clock() - t0
        while i < n and trace[i].t <= now:
            r = trace[i]
            x = crop_of(r, src2d, src3d)
            out = torch.empty(tuple(x.shape), dtype=x.dtype, device=x.device) if keep_outputs else None
            rid = srv.submit(x, r.gamma, tol=r.tol, max_iter=100000, out=out, t_submit=t0 + r.t)
            meta[rid] = (r.k, r.side, r.side ** r.k, out)
            i += 1
        b, q = srv.step()
        for d in srv.poll():
            d["k"], d["side"], d["pixels"], out = meta.pop(d["id"])
            if keep_outputs:
                d["z"] = out
            res.append(d)
        if progress and now - last_report >= 5.0:
            progress(f"served {len(res)} of {n} requests, {i} submitted, t = {now:.1f} s")
            last_report = now
        if b == 0 and q == 0:
            if i >= n:
                return res
            wait = trace[i].t - (srv.clock() - t0)
            if wait > 0:
                time.sleep(min(wait, 0.01))
        elif b == 0:
            raise FastproxError(A.FP_ENOMEM, f"fpr_server_step: {q} queued request(s) can never be admitted")


def latency_stats(results):
    """p50 / p99 / mean / max latency (ms) and throughput of replay results."""
    if not results:
        return {}
    lat = np.array([r["t_done"] - r["t_submit"] for r in results]) * 1e3
    span = max(r["t_done"] for r in results) - min(r["t_submit"] for r in results)
    px = sum(r["pixels"] for r in results)
    out = {"requests": len(results), "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "mean_ms": float(lat.mean()), "max_ms": float(lat.max()), "span_s": float(span),
           "throughput_req_s": len(results) / span if span > 0 else 0.0,
           "throughput_mpx_s": px / span / 1e6 if span > 0 else 0.0,
           "converged": int(sum(1 for r in results if r["converged"])),
           "iters_mean": float(np.mean([r["iters"] for r in results]))}
    for k in (2, 3):
        sub = [r for r in results if r["k"] == k]
        if sub:
            lk = np.array([r["t_done"] - r["t_submit"] for r in sub]) * 1e3
            out[f"{k}d"] = {"requests": len(sub), "p50_ms": float(np.percentile(lk, 50)),
                            "p99_ms": float(np.percentile(lk, 99))}
    return out

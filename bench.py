"""Benchmark: isotropic TV of 512x512 images to duality gap <= 1e-6 on one
MI355X, batched (one persistent solve over the batch) against the plain
per-call solve (one image per fp_prox call), with the dominant kernel's
roofline and the CPU path timed on the same host.

A step is one solve of the whole batch (every image to its gap tolerance).
Prints one JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch B] [--size S]
"""

from __future__ import annotations

import argparse
import json
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
BYTES_PER_PIXEL_ITER = {(2, 4): 36, (2, 8): 72, (3, 4): 52, (3, 8): 104}  # 1 + 4k words


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(x, gamma, tol, max_iter, budget_s=20.0):
    """The reference's CPU solver on a bounded sample of the same images:
    the reference library built from its sources (oracle/_ref) with its own
    thread pool when present, else the oracle's single-threaded restatement."""
    ref = ROOT / "oracle" / "_ref" / "libfastprox_ref.so"
    n_img = x.shape[0]
    if ref.exists():
        sys.path.insert(0, str(ROOT / "tests"))
        import refbind
        threads = min(16, os.cpu_count() or 1)
        op = refbind.Op(True, 2)
        done, t0 = 0, time.perf_counter()
        while done < n_img and time.perf_counter() - t0 < budget_s:
            op.solve(x[done:done + 1].astype(np.float32), gamma, tol=tol, max_iter=max_iter, threads=threads)
            done += 1
        dt = time.perf_counter() - t0
        return {"value": done / dt, "unit": "images/s", "cores": threads, "kind": "reference",
                "sample": f"{done} of the {n_img} images, one fp_prox call each, fp_options.threads={threads}, "
                          f"{dt:.1f} s"}
    import oracle
    done, t0 = 0, time.perf_counter()
    while done < n_img and time.perf_counter() - t0 < budget_s:
        oracle.tv_iso(x[done].astype(np.float64), gamma, tol=tol, max_iter=max_iter)
        done += 1
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "images/s", "cores": 1, "kind": "port",
            "sample": f"{done} of the {n_img} images, the oracle's C restatement, 1 thread, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--gamma", type=float, default=0.1)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iter", type=int, default=100000)
    ap.add_argument("--kernel", default=None, help="'generic', 'stream:TT,SEG' or 'block:TT,RY'")
    ap.add_argument("--plain", type=int, default=8, help="images for the plain per-call rate (0: skip)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="copy 1 GiB five times (the HBM counters' scale for the traffic figure)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied 5 x 1 GiB (read 1 GiB, write 1 GiB per copy)")
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU: --gpus {args.gpus} is not supported (images are independent; "
              f"the multi-GPU form is data-parallel over images, see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

    import torch

    from fastprox_amd import Op
    from fastprox_amd.synth import digest, rectangles
    from fastprox_amd.torch import Batch

    dev = torch.device("cuda:0")
    B, S = args.batch, args.size
    x = rectangles(B, (S, S), n=12, sigma=0.1, seed=0, dtype=np.float32)
    log(f"inputs: {B} x {S}x{S} f32, digest {digest(x)}")
    xs = torch.tensor(x, device=dev)
    zs = torch.empty_like(xs)
    duals = torch.empty((B, 2, S, S), dtype=torch.float32, device=dev)
    op = Op(True, 2)
    kernel = None
    if args.kernel:
        if args.kernel == "generic":
            kernel = (1, 1, 4)
        else:
            kind, vals = args.kernel.split(":")
            kernel = (0 if kind == "stream" else 2,) + tuple(int(v) for v in vals.split(","))
    bt = Batch(op, torch.float32, [(S, S)] * B, kernel=kernel)
    bt.set_timing(True)
    gam = [args.gamma] * B
    xl, zl, dl = list(xs), list(zs), list(duals)

    def step():
        return bt.solve(xl, zl, dl, gam, tol=args.tol, max_iter=args.max_iter)

    for i in range(args.warmup):
        step()
        log(f"warmup {i + 1}/{args.warmup} done")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    step_ms_sum, epi_ms_sum, launches_sum = 0.0, 0.0, 0
    for i in range(args.steps):
        st = step()
        a, b_, c = bt.timing()
        step_ms_sum += a
        epi_ms_sum += b_
        launches_sum += c
        log(f"step {i + 1}/{args.steps} status {st}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gaps, its, zerrs, conv = bt.image_info()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    # Roofline of the step kernel: algorithmic bytes (36 B per pixel-iteration
    # in fp32 2D: read x, u, v; write u, v) over the measured kernel time.
    alg_bytes = float(np.sum(its.astype(np.float64)) * S * S * BYTES_PER_PIXEL_ITER[(2, 4)]) * args.steps
    step_s = step_ms_sum / 1e3
    achieved = alg_bytes / step_s / 1e9 if step_s > 0 else 0.0
    traffic, traffic_src = None, None
    prof = ROOT / "profiles" / "stream2d_iso_f32_512x512_b64_traffic.json"
    if prof.exists() and kernel_name(bt) == "stream:TT=4,SEG=32" and B == 64 and S == 512:
        rec = json.loads(prof.read_text())
        traffic, traffic_src = rec["hbm_bytes_per_dispatch"], str(prof.relative_to(ROOT))
    log(f"batched: {value:.1f} images/s, {ms_per_step:.2f} ms per batch; iters mean {its.mean():.1f} "
        f"max {its.max()}; step kernels {step_ms_sum / args.steps:.2f} ms, epilogues {epi_ms_sum / args.steps:.2f}"
        f" ms per batch, {launches_sum / args.steps:.0f} launches; converged {bool(conv.all())}")

    plain = None
    if args.plain > 0:
        from fastprox_amd import Workspace
        P = min(args.plain, B)
        one = xs[:1]
        out1 = torch.empty_like(one)
        ws = Workspace(op, one)
        ws.prox(xs[:1], args.gamma, out1, tol=args.tol, max_iter=args.max_iter)  # warm-up call
        torch.cuda.synchronize()
        tp0 = time.perf_counter()
        for i in range(P):
            ws.prox(xs[i:i + 1], args.gamma, out1, tol=args.tol, max_iter=args.max_iter)
        torch.cuda.synchronize()
        plain = P / (time.perf_counter() - tp0)
        log(f"plain per-call: {plain:.1f} images/s over {P} images")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(x, args.gamma, args.tol, args.max_iter)
        log(f"cpu baseline: {cpu}")

    res = {
        "metric": "iso-TV 512^2 images/s to gap<=1e-6 (batched persistent solve)",
        "value": value,
        "unit": "images/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"isotropic TV, batch {B} x {S}x{S} float32, 12 rectangles + N(0,0.1^2) "
                               f"noise (seed 0), gamma {args.gamma}, gap tol {args.tol}, cold start",
                   "batch": B, "size": [S, S], "gamma": args.gamma, "tol": args.tol,
                   "kernel": kernel_name(bt),
                   "iters_mean": float(its.mean()), "iters_max": int(its.max()),
                   "input_digest": digest(x)},
        "plain_images_per_s": plain,
        "speedup_vs_plain": (value / plain) if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": {0: "stream2d_kernel", 1: "step1_kernel", 2: "step2d_kernel"}[bt_kernel(bt)[0]],
                     "step_kernel_ms_per_batch": step_ms_sum / args.steps,
                     "epilogue_ms_per_batch": epi_ms_sum / args.steps,
                     "launches_per_batch": launches_sum / args.steps,
                     "alg_bytes_per_batch": alg_bytes / args.steps},
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)


def bt_kernel(bt):
    import ctypes as C

    from fastprox_amd.core import lib
    # the batch's kernel choice through the workspace-independent query
    g, t, r = C.c_int(), C.c_int(), C.c_int()
    L = lib()
    L.fpr_batch_get_kernel(bt.h, C.byref(g), C.byref(t), C.byref(r))
    return g.value, t.value, r.value


def kernel_name(bt):
    kind, tt, ry = bt_kernel(bt)
    return {0: f"stream:TT={tt},SEG={ry}", 1: "generic", 2: f"block:TT={tt},RY={ry}"}[kind]


if __name__ == "__main__":
    main()

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
    ap.add_argument("--legs", default="b256,uneven,aniso,3d,backward,serve512,serve",
                    help="the other configurations to measure after the batched one: any of b256 (the same "
                         "images, 256 per batch), uneven (256 uneven "
                         "images with pixel weights), aniso (64 x 512^2 f32 and f64), 3d (16 x 128^3), "
                         "backward ((16,3,256,256) "
                         "autograd), serve (the request trace); '' for none")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-pool-gib", type=float, default=24.0)
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
              f"the multi-GPU form is data-parallel over images, one engine or server per GPU)", file=sys.stderr)
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
    step_ms_sum, epi_ms_sum, launches_sum, busy_ms_sum = 0.0, 0.0, 0, 0.0
    for i in range(args.steps):
        st = step()
        a, b_, c = bt.timing()
        step_ms_sum += a
        epi_ms_sum += b_
        launches_sum += c
        busy_ms_sum += bt.busy_ms()
        log(f"step {i + 1}/{args.steps} status {st}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gaps, its, zerrs, conv = bt.image_info()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    # Roofline of the step kernel: algorithmic bytes (36 B per pixel-iteration
    # in fp32 2D: read x, u, v; write u, v) over the time the step kernels
    # held the device. The batch runs as two halves on two streams whose
    # launches overlap, so that time is the union of the event-timed launch
    # intervals of both streams (busy), not their sum.
    alg_bytes = float(np.sum(its.astype(np.float64)) * S * S * BYTES_PER_PIXEL_ITER[(2, 4)]) * args.steps
    step_s = busy_ms_sum / 1e3
    achieved = alg_bytes / step_s / 1e9 if step_s > 0 else 0.0
    # HBM bytes per step dispatch from the committed rocprofv3 counter runs
    # of this same workload and kernel choice (FETCH_SIZE and WRITE_SIZE,
    # copy-calibrated), when they match.
    traffic, traffic_src, hbm_gbs = None, None, None
    prof = ROOT / "profiles" / "stream2d_iso_f32_512x512_b64_traffic.json"
    if prof.exists():
        rec = json.loads(prof.read_text())
        if rec.get("kernel_choice") == kernel_name(bt) and rec.get("batch") == B and rec.get("size") == S:
            traffic, traffic_src = rec["hbm_bytes_per_dispatch"], str(prof.relative_to(ROOT))
            if launches_sum and busy_ms_sum:
                hbm_gbs = traffic * launches_sum / (busy_ms_sum * 1e6)
    log(f"batched: {value:.1f} images/s, {ms_per_step:.2f} ms per batch; iters mean {its.mean():.1f} "
        f"max {its.max()}; step kernels {step_ms_sum / args.steps:.2f} ms ({busy_ms_sum / args.steps:.2f} ms "
        f"busy), epilogues {epi_ms_sum / args.steps:.2f}"
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

    torch_pc = None
    if args.plain > 0:
        torch_pc = torch_per_call(xs[:min(4, B)], args.gamma, args.tol, args.max_iter, zs)
        log(f"plain per-call with torch ops: {torch_pc}")

    legs = [v for v in args.legs.split(",") if v]
    extra = {}
    if "b256" in legs and S == 512:
        extra["batched_b256"] = leg_b256(args, plain)
    if "uneven" in legs:
        extra["uneven_batch"] = leg_uneven(args)
    if "aniso" in legs:
        extra["aniso"] = leg_aniso(args)
    if "3d" in legs:
        extra["volumes_3d"] = leg_3d(args)
    if "backward" in legs:
        extra["backward"] = leg_backward(args)
    if "serve512" in legs and S == 512:
        extra["served_512"] = leg_serve512(args, xs, plain)
    if "serve" in legs:
        extra["served"] = leg_serve(args)

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
        "plain_torch_per_call": torch_pc,
        "speedup_vs_torch_per_call": (value / torch_pc["images_per_s"]) if torch_pc else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "hbm_gbs_measured": hbm_gbs,
                     "note": "achieved counts the algorithmic 36 B of every pixel-iteration over the time the "
                             "step kernels held the device (step_kernel_busy_ms: the union of the event-timed "
                             "launch intervals of the batch's two streams, whose launches overlap); the kernel "
                             "runs TT iterations per HBM pass, so achieved can exceed the one-pass peak, while "
                             "hbm_gbs_measured (traffic per dispatch x dispatches over the same time) is the "
                             "HBM rate",
                     "kernel": {0: "stream2d_kernel", 1: "step1_kernel", 2: "step2d_kernel",
                                3: "stream3d_kernel"}[bt_kernel(bt)[0]],
                     "step_kernel_busy_ms_per_batch": busy_ms_sum / args.steps,
                     "step_kernel_ms_per_batch": step_ms_sum / args.steps,
                     "epilogue_ms_per_batch": epi_ms_sum / args.steps,
                     "launches_per_batch": launches_sum / args.steps,
                     "alg_bytes_per_batch": alg_bytes / args.steps},
        "cpu_baseline": cpu,
    }
    res.update(extra)
    print(json.dumps(res), flush=True)


def _sync_time(fn, steps):
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        r = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, r


def torch_per_call(xs, gamma, tol, max_iter, zref):
    """The reference's per-call GPU solve as it runs there: its torch device
    loop (_tv_torch.py:229-283; operators :93-186) restated with torch ops,
    isotropic 2D, one image per call, the restart test and the gap (with a
    host sync) every 4 iterations. A baseline to compare against, not a
    path of this library; its z is checked against the library's."""
    import torch
    n = xs.shape[0]

    def primal(x, v):  # x - G^T v (w = 1)
        z = x.clone()
        z[:, :-1, :].add_(v[0][:, :-1, :])
        z[:, 1:, :].sub_(v[0][:, :-1, :])
        z[:, :, :-1].add_(v[1][:, :, :-1])
        z[:, :, 1:].sub_(v[1][:, :, :-1])
        return z

    def grad(z):
        g = torch.zeros((2,) + tuple(z.shape), dtype=z.dtype, device=z.device)
        torch.sub(z[:, 1:, :], z[:, :-1, :], out=g[0][:, :-1, :])
        torch.sub(z[:, :, 1:], z[:, :, :-1], out=g[1][:, :, :-1])
        return g

    def solve(x):
        tau = 1.0 / 8.0
        u = torch.zeros((2,) + tuple(x.shape), dtype=x.dtype, device=x.device)
        v = u.clone()
        t = torch.ones(1, dtype=x.dtype, device=x.device)
        target = tol * float((x.double() ** 2).sum()) / 2
        it = 0
        while it < max_iter:
            y = v + tau * grad(primal(x, v))
            up = y.div_(torch.hypot(y[0], y[1]).div_(gamma).clamp_(min=1.0))
            it += 1
            tn = (1 + torch.sqrt(1 + 4 * t * t)) / 2
            coef = (t - 1) / tn
            check = it % 4 == 0 or it == max_iter
            if check:
                restart = ((v - up) * (up - u)).sum() > 0
                coef = torch.where(restart, torch.zeros_like(coef), coef)
                tn = torch.where(restart, torch.ones_like(tn), tn)
            v = torch.addcmul(up, coef, up - u)
            u, t = up, tn
            if check:
                z = primal(x, u)
                g = grad(z)
                gap = gamma * torch.hypot(g[0], g[1]).sum() - (g * u).sum()
                if float(gap) <= target:
                    break
        return primal(x, u), it

    solve(xs[:1])  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    its, dz = [], 0.0
    for i in range(n):
        z, it = solve(xs[i:i + 1])
        its.append(it)
        dz = max(dz, float((z[0] - zref[i]).abs().max()))
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    return {"images_per_s": n / sec, "images": n, "iters": its, "max_abs_z_vs_library": dz,
            "method": "the reference's torch device loop restated with torch ops, one image per call"}


def leg_b256(args, plain, B=256, S=512):
    """The headline workload at 256 images per batch (the same generator)."""
    import torch

    from fastprox_amd import Op
    from fastprox_amd.synth import rectangles
    from fastprox_amd.torch import Batch
    x = rectangles(B, (S, S), n=12, sigma=0.1, seed=0, dtype=np.float32)
    xs = torch.tensor(x, device="cuda")
    zs = torch.empty_like(xs)
    ds = torch.empty((B, 2, S, S), dtype=torch.float32, device="cuda")
    bt = Batch(Op(True, 2), torch.float32, [(S, S)] * B)
    xl, zl, dl = list(xs), list(zs), list(ds)

    def run():
        return bt.solve(xl, zl, dl, [args.gamma] * B, tol=args.tol, max_iter=args.max_iter)
    run()
    sec, st = _sync_time(run, 2)
    gaps, its, zerrs, conv = bt.image_info()
    out = {"workload": f"isotropic TV, batch {B} x {S}x{S} float32 (the headline generator, seed 0), gamma "
                       f"{args.gamma}, tol {args.tol}", "images_per_s": B / sec, "ms_per_batch": sec * 1e3,
           "iters_mean": float(its.mean()), "iters_max": int(its.max()), "converged": bool(conv.all()),
           "speedup_vs_plain": (B / sec / plain) if plain else None}
    log(f"batched b256: {out}")
    del xs, zs, ds, bt
    return out


def leg_uneven(args, n=256, seed=0):
    """256 float32 images, H and W uniform in [64, 1024], 16 rectangles +
    N(0, 0.1^2) noise, pixel weights lognormal(0, 0.3), gamma log-uniform in
    [0.02, 0.3], gap tol 1e-6: one persistent solve over the uneven batch."""
    import torch

    from fastprox_amd import Op
    from fastprox_amd.synth import rectangles
    from fastprox_amd.torch import Batch
    rng = np.random.default_rng(seed)
    shapes = [(int(rng.integers(64, 1025)), int(rng.integers(64, 1025))) for _ in range(n)]
    gam = [float(np.exp(rng.uniform(np.log(0.02), np.log(0.3)))) for _ in range(n)]
    dev = torch.device("cuda:0")
    xs, ws, zs, ds = [], [], [], []
    for i, sh in enumerate(shapes):
        xs.append(torch.tensor(rectangles(1, sh, n=16, sigma=0.1, seed=1000 + i)[0], device=dev))
        ws.append(torch.tensor(rng.lognormal(0.0, 0.3, size=sh).astype(np.float32), device=dev))
        zs.append(torch.empty(sh, dtype=torch.float32, device=dev))
        ds.append(torch.empty((2,) + sh, dtype=torch.float32, device=dev))
    px = sum(a * b for a, b in shapes)
    bt = Batch(Op(True, 2), torch.float32, shapes, weighted=True)
    bt.set_timing(True)

    def run():
        return bt.solve(xs, zs, ds, gam, weights=ws, tol=1e-6, max_iter=args.max_iter)
    run()
    sec, st = _sync_time(run, max(1, min(args.steps, 3)))
    gaps, its, zerrs, conv = bt.image_info()
    a, b_, c = bt.timing()
    out = {"workload": f"{n} float32 images, H,W ~ U[64,1024] ({px / 1e6:.1f} Mpx), pixel weights lognormal(0,0.3), "
                       f"gamma log-uniform [0.02,0.3], tol 1e-6, one persistent solve",
           "images_per_s": n / sec, "mpx_per_s": px / sec / 1e6, "ms_per_batch": sec * 1e3,
           "iters_mean": float(its.mean()), "iters_max": int(its.max()), "converged": bool(conv.all()),
           "kernel": kernel_name(bt), "step_kernel_ms": a, "step_kernel_busy_ms": bt.busy_ms(), "epilogue_ms": b_,
           "launches": c}
    log(f"uneven batch: {out}")
    return out


def leg_aniso(args, B=64, S=512):
    """64 x 512^2 disks (20 disks, radius U[8,96], + N(0, 0.15^2)), weights
    (1, 0.5), gamma 0.05, gap tol 1e-6, in float32 and float64."""
    import torch

    from fastprox_amd import Op
    from fastprox_amd.synth import digest, disks
    from fastprox_amd.torch import Batch
    x = disks(B, (S, S), seed=0, dtype=np.float64)
    out = {"workload": f"anisotropic TV, {B} x {S}x{S} disks, w (1, 0.5), gamma 0.05, tol 1e-6 (device method)",
           "input_digest": digest(x)}
    for dt, name in ((torch.float32, "f32"), (torch.float64, "f64")):
        xs = torch.tensor(x, dtype=dt, device="cuda")
        zs = torch.empty_like(xs)
        ds = torch.empty((B, 2, S, S), dtype=dt, device="cuda")
        bt = Batch(Op(False, 2, [1.0, 0.5]), dt, [(S, S)] * B)
        xl, zl, dl = list(xs), list(zs), list(ds)

        def run():
            return bt.solve(xl, zl, dl, [0.05] * B, tol=1e-6, max_iter=args.max_iter)
        run()
        sec, st = _sync_time(run, 1)
        gaps, its, zerrs, conv = bt.image_info()
        out[name] = {"images_per_s": B / sec, "ms_per_batch": sec * 1e3, "iters_mean": float(its.mean()),
                     "iters_max": int(its.max()), "converged": bool(conv.all())}
        log(f"aniso {name}: {out[name]}")
        del xs, zs, ds, bt
    return out


def leg_3d(args, B=16, S=128):
    """16 x 128^3 float32 volumes (16 boxes + N(0, 0.1^2)), isotropic,
    gamma 0.05, gap tol 1e-5: the 3D streaming kernel's throughput."""
    import torch

    from fastprox_amd import Op
    from fastprox_amd.synth import rectangles
    from fastprox_amd.torch import Batch
    x = rectangles(B, (S, S, S), n=16, sigma=0.1, seed=0)
    xs = torch.tensor(x, device="cuda")
    zs = torch.empty_like(xs)
    ds = torch.empty((B, 3, S, S, S), dtype=torch.float32, device="cuda")
    bt = Batch(Op(True, 3), torch.float32, [(S, S, S)] * B)
    bt.set_timing(True)
    xl, zl, dl = list(xs), list(zs), list(ds)

    def run():
        return bt.solve(xl, zl, dl, [0.05] * B, tol=1e-5, max_iter=args.max_iter)
    run()
    sec, st = _sync_time(run, 1)
    gaps, its, zerrs, conv = bt.image_info()
    a, b_, c = bt.timing()
    alg = float(its.sum()) * S ** 3 * BYTES_PER_PIXEL_ITER[(3, 4)]
    out = {"workload": f"isotropic TV, {B} x {S}^3 float32 boxes, gamma 0.05, tol 1e-5", "volumes_per_s": B / sec,
           "ms_per_batch": sec * 1e3, "iters_mean": float(its.mean()), "iters_max": int(its.max()),
           "kernel": kernel_name(bt), "step_kernel_ms": a, "epilogue_ms": b_, "launches": c,
           "alg_gbs": alg / (a / 1e3) / 1e9 if a else None}
    log(f"3d: {out}")
    return out


def leg_backward(args):
    """x (16, 3, 256, 256) float32 (the 256^2 rectangles generator per
    channel, noise 0.1), gamma 0.1 with requires_grad, loss mean((z -
    clean)^2): forward solve (tol 1e-6) and backward (the VJP) times."""
    import torch

    import fastprox_amd.torch as fpt
    from fastprox_amd.synth import rectangles
    from fastprox_amd.torch import tv_iso
    clean = rectangles(48, (256, 256), n=12, sigma=0.0, seed=0).reshape(16, 3, 256, 256)
    noisy = clean + np.random.default_rng(1).normal(0.0, 0.1, size=clean.shape).astype(np.float32)
    xc = torch.tensor(noisy, device="cuda")
    ct = torch.tensor(clean, device="cuda")
    res = {}
    for it in range(3):
        x = xc.clone().requires_grad_(True)
        g = torch.tensor(0.1, device="cuda", requires_grad=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        z, info = tv_iso(x, g, tol=1e-6, max_iter=args.max_iter, return_info=True)
        loss = ((z - ct) ** 2).mean()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        loss.backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        res = {"workload": "x (16,3,256,256) float32, gamma 0.1 requires_grad, loss mean((z - clean)^2), "
                           "solve tol 1e-6, VJP at the default tolerance",
               "forward_ms": (t1 - t0) * 1e3, "backward_ms": (t2 - t1) * 1e3,
               "solve_iters_max": int(info.iters), "cg_residual_max": float(fpt.LAST_DERIV[0]),
               "cg_iters_max": int(fpt.LAST_DERIV[1]), "regions": int(fpt.LAST_DERIV[2]),
               "grad_gamma": float(g.grad), "grad_x_norm": float(x.grad.norm())}
    log(f"backward: {res}")
    return res


def leg_serve(args):
    """The request trace: Poisson arrivals at --serve-rate per second for
    --serve-seconds; 80% 2D float32 squares of side U[128, 1024], 20% 3D
    cubes of side U[48, 128]; gamma log-uniform [0.01, 0.3]; tol from
    {1e-4, 1e-5, 1e-6}; each request a crop of a pre-generated source image."""
    import torch

    from fastprox_amd import Op
    from fastprox_amd.serve import Server, latency_stats, replay
    from fastprox_amd.synth import crop_of, serving_sources, serving_trace
    trace = serving_trace(seconds=args.serve_seconds, rate=args.serve_rate, seed=0)
    a, b = serving_sources()
    src2d, src3d = torch.tensor(a, device="cuda"), torch.tensor(b, device="cuda")
    srv = Server(Op(True, 2), Op(True, 3), pool_bytes=int(args.serve_pool_gib * (1 << 30)), slots=512)
    # warm-up: a few requests, not recorded
    replay(srv, serving_trace(seconds=0.2, rate=100, seed=99), src2d, src3d)
    log(f"serving {len(trace)} requests over {args.serve_seconds:.0f} s ...")
    res = replay(srv, trace, src2d, src3d, progress=log)
    st = latency_stats(res)
    st["workload"] = (f"Poisson {args.serve_rate:.0f} req/s for {args.serve_seconds:.0f} s: 80% 2D f32 side "
                      f"U[128,1024], 20% 3D f32 side U[48,128], gamma log-uniform [0.01,0.3], tol in "
                      f"{{1e-4,1e-5,1e-6}}; continuous batching, 512 slots per class, "
                      f"{args.serve_pool_gib:.0f} GiB pool per class")
    st["launches"] = srv.launches()
    log(f"served: {st}")
    if not args.no_cpu:
        st["cpu"] = serve_cpu_baseline(trace, a, b)
        log(f"served cpu: {st['cpu']}")
    return st


def leg_serve512(args, xs, plain, n=2048):
    """Served capacity at 512^2: n requests of the headline images (each of
    the batch's images n / B times), gamma 0.1, tol 1e-6, all submitted at
    once, so the offered load exceeds what the server completes; images/s
    from the first submission to the last completion, through the same
    continuous-batching loop (slots, pool admission, chunks in flight) as
    the request trace."""
    import torch

    from fastprox_amd import Op
    from fastprox_amd.serve import Server
    B = xs.shape[0]
    srv = Server(Op(True, 2), None, pool_bytes=int(args.serve_pool_gib * (1 << 30)), slots=512)
    outs = [torch.empty_like(xs[0]) for _ in range(n)]

    def run(count):
        t0 = srv.clock()
        for i in range(count):
            srv.submit(xs[i % B], args.gamma, tol=args.tol, max_iter=args.max_iter, out=outs[i], t_submit=t0)
        res = srv.drain()
        return res, srv.clock() - t0
    run(64)  # warm-up, not recorded
    res, sec = run(n)
    lat = np.array([r["t_done"] - r["t_submit"] for r in res]) * 1e3
    ok = all(r["converged"] for r in res) and len(res) == n
    # the steady state: completions between the first and the last quarter
    td = np.sort([r["t_done"] for r in res])
    steady = (n // 2) / (td[3 * n // 4] - td[n // 4])
    out = {"workload": f"{n} requests of the headline 512^2 float32 images, gamma {args.gamma}, tol {args.tol}, "
                       f"all submitted at t = 0 (offered load above capacity), 512 slots, "
                       f"{args.serve_pool_gib:.0f} GiB pool",
           "images_per_s": n / sec, "seconds": sec, "converged": bool(ok),
           "steady_images_per_s": float(steady),
           "steady_speedup_vs_plain": (float(steady) / plain) if plain else None,
           "kernel": "stream:TT={1},SEG={2}".format(*srv.kernel(2)),
           "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
           "speedup_vs_plain": (n / sec / plain) if plain else None}
    log(f"served 512: {out}")
    del srv
    return out


def serve_cpu_baseline(trace, src2d, src3d, budget_s=25.0):
    """The reference CPU solver's latency per request (no queueing) on the
    first requests of the same trace, at 16 threads, while the budget lasts."""
    ref = ROOT / "oracle" / "_ref" / "libfastprox_ref.so"
    if not ref.exists():
        return None
    from fastprox_amd.synth import crop_of
    sys.path.insert(0, str(ROOT / "tests"))
    import refbind
    threads = min(16, os.cpu_count() or 1)
    ops = {2: refbind.Op(True, 2), 3: refbind.Op(True, 3)}
    lat, t_all = [], time.perf_counter()
    for r in trace:
        if time.perf_counter() - t_all > budget_s:
            break
        x = np.ascontiguousarray(crop_of(r, src2d, src3d))
        t0 = time.perf_counter()
        ops[r.k].solve(x, r.gamma, tol=r.tol, max_iter=100000, threads=threads)
        lat.append(time.perf_counter() - t0)
    lat = np.array(lat) * 1e3
    return {"requests": int(lat.size), "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "mean_ms": float(lat.mean()), "cores": threads,
            "kind": "reference", "sample": f"the first {lat.size} requests of the trace, one fp_prox call each"}


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
    return {0: f"stream:TT={tt},SEG={ry}", 1: "generic", 2: f"block:TT={tt},RY={ry}", 3: f"s3:SLICES={ry}"}[kind]


if __name__ == "__main__":
    main()

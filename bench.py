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
# Bytes of one pixel's state pass (read d, u, v; write u, v: 1 + 4k words),
# per (k, dtype bytes). A step launch makes one pass over every image it
# iterates, whatever its iterations; a one-iteration-per-pass kernel makes
# one pass per pixel-iteration.
BYTES_PER_PIXEL_PASS = {(2, 4): 36, (2, 8): 72, (3, 4): 52, (3, 8): 104}
BYTES_PER_PIXEL_ITER = BYTES_PER_PIXEL_PASS


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(x, gamma, tol, max_iter, reps=5, batch=16):
    """The reference's CPU solver on a bounded sample of the same images:
    the reference library built from its sources (oracle/_ref), one fp_prox
    call on the first `batch` images with fp_options.threads = 16 (its pool
    runs the images in parallel), the workspace created outside the timed
    call; the median of `reps` timed calls after one untimed call. Without
    the library, the oracle's single-threaded restatement on one image."""
    ref = ROOT / "oracle" / "_ref" / "libfastprox_ref.so"
    host = host_cpu()
    if ref.exists():
        import ctypes as C
        sys.path.insert(0, str(ROOT / "tests"))
        import refbind
        from fastprox_amd import _capi as A
        L = refbind.lib()
        threads = min(16, host["nproc"] or 1)
        n = min(batch, x.shape[0])
        xs = np.ascontiguousarray(x[:n])
        z = np.empty_like(xs)
        opt = A.fp_options()
        L.fp_options_init(C.byref(opt))
        opt.tol, opt.max_iter, opt.threads = tol, max_iter, threads
        op = refbind.Op(True, 2)
        xa, za = A.numpy_array(xs), A.numpy_array(z)
        ws = C.c_void_p()
        assert L.fp_workspace_create(op.h, xa.ref, C.byref(opt), C.byref(ws)) == A.FP_OK
        info = A.fp_info()
        med, ts, st = _median_time(lambda: L.fp_prox(op.h, float(gamma), xa.ref, za.ref, C.byref(opt), ws,
                                                     C.byref(info)), reps)
        L.fp_workspace_free(ws)
        return {"value": n / med, "unit": "images/s", "cores": threads, "kind": "reference",
                "host": host, "reps": reps, "median_s": med, "times_s": ts,
                "sample": f"the first {n} of the {x.shape[0]} images ({x.dtype}, promoted by the library), one "
                          f"fp_prox call per repetition with fp_options.threads={threads}; median of {reps} after "
                          f"one untimed call"}
    import oracle
    med, ts, _ = _median_time(lambda: oracle.tv_iso(x[0].astype(np.float64), gamma, tol=tol, max_iter=max_iter),
                              reps)
    return {"value": 1.0 / med, "unit": "images/s", "cores": 1, "kind": "port", "host": host, "reps": reps,
            "median_s": med, "times_s": ts,
            "sample": f"image 0 of the {x.shape[0]}, the oracle's C restatement, 1 thread, median of {reps}"}


def host_cpu():
    """The host's CPU model and the logical CPUs this process may use."""
    model = "unknown"
    try:
        for ln in pathlib.Path("/proc/cpuinfo").read_text().splitlines():
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"model": model, "nproc": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
            else os.cpu_count()}


def _median_time(fn, reps):
    """fn() once untimed, then the median wall time of reps timed calls."""
    fn()
    ts = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts, r


def cpu_table(threads_list, reps=5):
    """The reference library (oracle/_ref, built from its sources) on the
    reference's own CPU configuration and the anisotropic one, through its C
    ABI with a workspace created outside the timed call: C0 (8 x 256^2, 12
    rectangles + N(0, 0.1^2), seed 0, iso, gamma 0.1, tol 1e-6) and C1 (64 x
    512^2 disks, aniso w (1, 0.5), gamma 0.05, tol 1e-6, float32 and float64
    inputs, promoted by the reference). Median of reps timed calls after one
    untimed call, per thread count."""
    sys.path.insert(0, str(ROOT / "tests"))
    import ctypes as C

    import refbind
    from fastprox_amd import _capi as A
    from fastprox_amd.synth import digest, disks, rectangles
    if not refbind.available():
        raise SystemExit("oracle/_ref is not built")
    L = refbind.lib()
    out = {"host": host_cpu(), "reps": reps, "rows": []}

    def timed_solve(op, x, gamma, tol, threads):
        xs = np.ascontiguousarray(x)
        z = np.empty_like(xs)
        opt = A.fp_options()
        L.fp_options_init(C.byref(opt))
        opt.tol, opt.max_iter, opt.threads = tol, 100000, threads
        xa, za = A.numpy_array(xs), A.numpy_array(z)
        ws = C.c_void_p()
        assert L.fp_workspace_create(op.h, xa.ref, C.byref(opt), C.byref(ws)) == A.FP_OK
        info = A.fp_info()

        def call():
            return L.fp_prox(op.h, float(gamma), xa.ref, za.ref, C.byref(opt), ws, C.byref(info))
        med, ts, st = _median_time(call, reps)
        B = xs.shape[0]
        gaps, its, zerrs = np.empty(B), np.empty(B, np.int32), np.empty(B)
        L.fp_workspace_image_info(ws, gaps.ctypes.data_as(C.POINTER(C.c_double)),
                                  its.ctypes.data_as(C.POINTER(C.c_int)), zerrs.ctypes.data_as(C.POINTER(C.c_double)),
                                  B)
        L.fp_workspace_free(ws)
        return med, ts, st, gaps, its

    cases = [("C0 8x256^2 f32 iso gamma 0.1 tol 1e-6", rectangles(8, (256, 256), n=12, sigma=0.1, seed=0),
              refbind.Op(True, 2), 0.1)]
    d = disks(64, (512, 512), seed=0, dtype=np.float64)
    cases.append(("C1 64x512^2 f32 aniso w (1,0.5) gamma 0.05 tol 1e-6", d.astype(np.float32),
                  refbind.Op(False, 2, [1.0, 0.5]), 0.05))
    cases.append(("C1 64x512^2 f64 aniso w (1,0.5) gamma 0.05 tol 1e-6", d, refbind.Op(False, 2, [1.0, 0.5]), 0.05))
    for name, x, op, gamma in cases:
        for th in threads_list:
            med, ts, st, gaps, its = timed_solve(op, x, gamma, 1e-6, th)
            row = {"config": name, "threads": th, "median_s": med, "times_s": ts,
                   "images_per_s": x.shape[0] / med, "status": A.STATUS.get(st, st),
                   "iters": [int(v) for v in its], "gaps": [float(v) for v in gaps], "input_digest": digest(x)}
            log(f"{name}, {th} threads: {med:.3f} s median ({x.shape[0] / med:.2f} images/s), iters "
                f"{int(its.min())}..{int(its.max())}")
            out["rows"].append(row)
    return out


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
    ap.add_argument("--legs", default="b256,f64,uneven,aniso,3d,backward,serve512,serve",
                    help="the other configurations to measure after the batched one: any of b256 (the same "
                         "images, 256 per batch), f64 (the headline in float64), uneven (256 uneven "
                         "images with pixel weights), aniso (64 x 512^2 f32 and f64), 3d (16 x 128^3), "
                         "backward ((16,3,256,256) "
                         "autograd), serve (the request trace); '' for none")
    ap.add_argument("--serve-seconds", type=float, default=60.0)
    ap.add_argument("--serve-rate", type=float, default=200.0)
    ap.add_argument("--serve-pool-gib", type=float, default=24.0)
    ap.add_argument("--serve-cpu-requests", type=int, default=512,
                    help="requests in the served CPU baseline's sample (stopped at 120 s of CPU time)")
    ap.add_argument("--cpu-table", default=None,
                    help="only time the reference library on the CPU configurations at these thread counts "
                         "(comma-separated) and print that table as JSON; no GPU is used")
    args = ap.parse_args()
    if args.cpu_table:
        print(json.dumps(cpu_table([int(v) for v in args.cpu_table.split(",")])), flush=True)
        return
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
    step_ms_sum, epi_ms_sum, launches_sum, busy_ms_sum, passes_sum = 0.0, 0.0, 0, 0.0, 0.0
    for i in range(args.steps):
        st = step()
        a, b_, c = bt.timing()
        step_ms_sum += a
        epi_ms_sum += b_
        launches_sum += c
        busy_ms_sum += bt.busy_ms()
        passes_sum += bt.pixel_passes()
        log(f"step {i + 1}/{args.steps} status {st}")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    gaps, its, zerrs, conv = bt.image_info()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = B * args.steps / (t1 - t0)
    # Roofline of the step kernel (HBM-bound: no dense contraction). The
    # algorithmic bytes of one launch are one state pass over every pixel it
    # iterates: 36 B per pixel (read x, u, v; write u, v in float32), whatever
    # its iterations (the fused TT iterations of a launch cost one round trip
    # of the state). achieved: those bytes per launch over the launch's
    # average duration (HIP events around each step launch on the stream it
    # runs on; the launches of the batch's two halves overlap, so each runs
    # longer than alone); achieved_device: the same bytes over the time at
    # least one step kernel ran (the union of the launch intervals of both
    # streams). traffic: the HBM bytes per dispatch the rocprofv3 counters
    # read for this kernel and workload (FETCH_SIZE, WRITE_SIZE,
    # copy-calibrated), committed under profiles/; its rate over the same
    # launch durations is the measured HBM rate.
    bpp = BYTES_PER_PIXEL_PASS[(2, 4)]
    alg_bytes = passes_sum * bpp
    launch_ms = step_ms_sum / launches_sum if launches_sum else 0.0
    alg_per_launch = alg_bytes / launches_sum if launches_sum else 0.0
    achieved = alg_per_launch / (launch_ms * 1e6) if launch_ms > 0 else 0.0
    achieved_device = alg_bytes / (busy_ms_sum * 1e6) if busy_ms_sum > 0 else 0.0
    one_pass_equiv = float(np.sum(its.astype(np.float64)) * S * S * bpp) * args.steps / (busy_ms_sum * 1e6) \
        if busy_ms_sum > 0 else 0.0
    traffic, traffic_src, hbm_gbs, valu_busy = None, None, None, None
    prof = ROOT / "profiles" / "stream2d_iso_f32_512x512_b64_counters.json"
    if prof.exists():
        rec = json.loads(prof.read_text())
        if rec.get("kernel_choice") == kernel_name(bt) and rec.get("batch") == B and rec.get("size") == S:
            traffic, traffic_src = rec["hbm_bytes_per_dispatch"], str(prof.relative_to(ROOT))
            valu_busy = rec.get("valu_busy_frac")
            if launch_ms > 0:
                hbm_gbs = traffic / (launch_ms * 1e6)
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
    if "f64" in legs and S == 512:
        extra["batched_f64"] = leg_f64(args)
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
                     "alg_bytes_per_launch": alg_per_launch, "launch_ms_mean": launch_ms,
                     "achieved_device": achieved_device, "frac_device": achieved_device / HBM_PEAK_GBS,
                     "hbm_gbs_measured": hbm_gbs,
                     "hbm_frac_measured": hbm_gbs / HBM_PEAK_GBS if hbm_gbs else None,
                     "valu_busy_frac": valu_busy,
                     "alg_equiv_one_pass_gbs": one_pass_equiv,
                     "note": "achieved: the algorithmic bytes of a step launch (36 B per pixel it iterates: one "
                             "state pass, read x, u, v and write u, v, whatever its TT iterations) over the "
                             "launch's average duration; achieved_device: the same bytes over the union of the "
                             "launch intervals of the batch's two streams; hbm_gbs_measured: the counters' HBM "
                             "bytes per dispatch (traffic) over the average launch duration; "
                             "alg_equiv_one_pass_gbs: 36 B per pixel-iteration over the union time, the rate a "
                             "one-iteration-per-pass kernel would need (not a fraction of any roof)",
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
    # The same batch as one captured hipGraph: on a capturing stream the solve
    # enqueues the static schedule that max_iter allows (every step/epilogue
    # pair, empty once an image has finished), so max_iter is set to twice
    # the iterations the batch needed; the replay's images equal an eager
    # solve's at that max_iter.
    try:
        out["graph"] = _uneven_graph(bt, xs, zs, ds, gam, ws, 2 * int(its.max()), max(1, min(args.steps, 3)), n)
        log(f"uneven batch, one captured graph: {out['graph']}")
    except Exception as e:  # reported, not fatal
        out["graph"] = {"error": repr(e)}
    return out


def _uneven_graph(bt, xs, zs, ds, gam, ws, max_iter, steps, n):
    import ctypes as C

    import torch
    hip = C.CDLL("libamdhip64.so")
    P = C.c_void_p
    for name, argt in (("hipStreamBeginCapture", [P, C.c_int]), ("hipStreamEndCapture", [P, C.POINTER(P)]),
                       ("hipGraphInstantiate", [C.POINTER(P), P, P, P, C.c_size_t]), ("hipGraphLaunch", [P, P]),
                       ("hipStreamSynchronize", [P]), ("hipGraphExecDestroy", [P]), ("hipGraphDestroy", [P])):
        getattr(hip, name).restype = C.c_int
        getattr(hip, name).argtypes = argt
    bt.set_timing(False)
    st = torch.cuda.Stream()
    s = C.c_void_p(st.cuda_stream)
    ze = [torch.empty_like(z) for z in zs]
    de = [torch.empty_like(d) for d in ds]
    bt.solve(xs, ze, de, gam, weights=ws, tol=1e-6, max_iter=max_iter, stream=st.cuda_stream)
    torch.cuda.synchronize()
    g, ex = C.c_void_p(), C.c_void_p()
    t0 = time.perf_counter()
    assert hip.hipStreamBeginCapture(s, 2) == 0
    try:
        bt.solve(xs, zs, ds, gam, weights=ws, tol=1e-6, max_iter=max_iter, stream=st.cuda_stream)
    finally:
        assert hip.hipStreamEndCapture(s, C.byref(g)) == 0
    assert hip.hipGraphInstantiate(C.byref(ex), g, None, None, 0) == 0
    capture_s = time.perf_counter() - t0
    try:
        assert hip.hipGraphLaunch(ex, s) == 0 and hip.hipStreamSynchronize(s) == 0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            assert hip.hipGraphLaunch(ex, s) == 0
        assert hip.hipStreamSynchronize(s) == 0
        sec = (time.perf_counter() - t0) / steps
        same = all(torch.equal(a, b) for a, b in zip(zs, ze))
        gaps, its, zerrs, conv = bt.image_info()
        launches, _ = bt.stats()
    finally:
        hip.hipGraphExecDestroy(ex)
        hip.hipGraphDestroy(g)
    return {"images_per_s": n / sec, "ms_per_batch": sec * 1e3, "max_iter": max_iter,
            "kernels_in_graph": 2 * launches, "capture_and_instantiate_s": capture_s,
            "equal_to_eager": bool(same), "converged": bool(conv.all())}


def leg_f64(args, B=64, S=512):
    """The headline workload in float64 (the reference computes in FP64,
    tv_iso.cpp:342-347): 64 x 512^2, the headline generator, gamma 0.1, tol
    1e-6; one state pass is 72 B per pixel."""
    import torch

    from fastprox_amd import Op
    from fastprox_amd.synth import rectangles
    from fastprox_amd.torch import Batch
    x = rectangles(B, (S, S), n=12, sigma=0.1, seed=0, dtype=np.float32).astype(np.float64)
    xs = torch.tensor(x, device="cuda")
    zs = torch.empty_like(xs)
    ds = torch.empty((B, 2, S, S), dtype=torch.float64, device="cuda")
    bt = Batch(Op(True, 2), torch.float64, [(S, S)] * B)
    bt.set_timing(True)
    xl, zl, dl = list(xs), list(zs), list(ds)

    def run():
        return bt.solve(xl, zl, dl, [args.gamma] * B, tol=args.tol, max_iter=args.max_iter)
    run()
    sec, st = _sync_time(run, 2)
    gaps, its, zerrs, conv = bt.image_info()
    a, b_, c = bt.timing()
    busy = bt.busy_ms()
    alg = bt.pixel_passes() * BYTES_PER_PIXEL_PASS[(2, 8)]
    out = {"workload": f"isotropic TV, batch {B} x {S}x{S} float64 (the headline images), gamma {args.gamma}, "
                       f"tol {args.tol}", "images_per_s": B / sec, "ms_per_batch": sec * 1e3,
           "iters_mean": float(its.mean()), "iters_max": int(its.max()), "converged": bool(conv.all()),
           "kernel": kernel_name(bt), "step_kernel_ms": a, "step_kernel_busy_ms": busy, "launches": c,
           "alg_gbs_per_launch": alg / (a * 1e6) if a else None,
           "alg_gbs_device": alg / (busy * 1e6) if busy else None,
           "frac_device": alg / (busy * 1e6) / HBM_PEAK_GBS if busy else None}
    log(f"batched f64: {out}")
    del xs, zs, ds, bt
    return out


def leg_aniso(args, B=64, S=512):
    """64 x 512^2 disks (20 disks, radius U[8,96], + N(0, 0.15^2)), weights
    (1, 0.5), gamma 0.05, gap tol 1e-6, in float32 and float64."""
    import torch

    from fastprox_amd import Op
    from fastprox_amd.synth import digest, disks
    from fastprox_amd.torch import Batch
    x = disks(B, (S, S), seed=0, dtype=np.float64)
    out = {"workload": f"anisotropic TV, {B} x {S}x{S} disks, w (1, 0.5), gamma 0.05, tol 1e-6; fista: dual FISTA "
                       f"on the edge duals (the reference's device method), adbpg: the reference CPU library's "
                       f"method (its iteration counts)",
           "input_digest": digest(x)}
    for method in ("fista", "adbpg"):
        for dt, name in ((torch.float32, "f32"), (torch.float64, "f64")):
            xs = torch.tensor(x, dtype=dt, device="cuda")
            zs = torch.empty_like(xs)
            ds = torch.empty((B, 2, S, S), dtype=dt, device="cuda")
            bt = Batch(Op(False, 2, [1.0, 0.5]), dt, [(S, S)] * B, method=method)
            xl, zl, dl = list(xs), list(zs), list(ds)

            def run():
                return bt.solve(xl, zl, dl, [0.05] * B, tol=1e-6, max_iter=args.max_iter)
            run()
            sec, st = _sync_time(run, 2)
            gaps, its, zerrs, conv = bt.image_info()
            key = name if method == "fista" else f"adbpg_{name}"
            out[key] = {"images_per_s": B / sec, "ms_per_batch": sec * 1e3, "iters_mean": float(its.mean()),
                        "iters_max": int(its.max()), "converged": bool(conv.all())}
            log(f"aniso {key}: {out[key]}")
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
    busy = bt.busy_ms()
    # one state pass (52 B per voxel in float32) per launch and voxel iterated
    alg = bt.pixel_passes() * BYTES_PER_PIXEL_PASS[(3, 4)]
    out = {"workload": f"isotropic TV, {B} x {S}^3 float32 boxes, gamma 0.05, tol 1e-5", "volumes_per_s": B / sec,
           "ms_per_batch": sec * 1e3, "iters_mean": float(its.mean()), "iters_max": int(its.max()),
           "kernel": kernel_name(bt), "step_kernel_ms": a, "step_kernel_busy_ms": busy, "epilogue_ms": b_,
           "launches": c, "alg_bytes": alg,
           "alg_gbs_per_launch": alg / (a * 1e6) if a else None,
           "alg_gbs_device": alg / (busy * 1e6) if busy else None,
           "frac_device": alg / (busy * 1e6) / HBM_PEAK_GBS if busy else None}
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
        st["cpu"] = serve_cpu_baseline(trace, a, b, requests=args.serve_cpu_requests)
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


def serve_cpu_baseline(trace, src2d, src3d, budget_s=120.0, requests=None, seed=1):
    """The reference CPU solver's latency per request (no queueing), at 16
    threads, on a random sample of the same trace stratified by class (the
    2D and 3D requests in the trace's proportion, each class in a random
    order), until `requests` are done or the budget ends."""
    ref = ROOT / "oracle" / "_ref" / "libfastprox_ref.so"
    if not ref.exists():
        return None
    from fastprox_amd.synth import crop_of
    sys.path.insert(0, str(ROOT / "tests"))
    import refbind
    threads = min(16, host_cpu()["nproc"] or 1)
    ops = {2: refbind.Op(True, 2), 3: refbind.Op(True, 3)}
    rng = np.random.default_rng(seed)
    by = {k: list(rng.permutation([i for i, r in enumerate(trace) if r.k == k])) for k in (2, 3)}
    n_all = len(trace)
    share3 = len(by[3]) / n_all if n_all else 0.0
    lat = {2: [], 3: []}
    t_all = time.perf_counter()
    done = 0
    while by[2] or by[3]:
        if requests is not None and done >= requests:
            break
        if time.perf_counter() - t_all > budget_s:
            break
        # the class whose share of the sample is furthest below its share of the trace
        k = 3 if by[3] and (not by[2] or len(lat[3]) < share3 * (done + 1)) else 2
        r = trace[by[k].pop()]
        x = np.ascontiguousarray(crop_of(r, src2d, src3d))
        t0 = time.perf_counter()
        ops[r.k].solve(x, r.gamma, tol=r.tol, max_iter=100000, threads=threads)
        lat[k].append(time.perf_counter() - t0)
        done += 1
    both = np.array(lat[2] + lat[3]) * 1e3
    out = {"requests": int(both.size), "p50_ms": float(np.percentile(both, 50)),
           "p99_ms": float(np.percentile(both, 99)), "mean_ms": float(both.mean()), "cores": threads,
           "host": host_cpu(), "kind": "reference",
           "sample": f"{both.size} requests of the trace drawn at random, stratified by class ({len(lat[2])} 2D, "
                     f"{len(lat[3])} 3D), one fp_prox call each, fp_options.threads={threads}"}
    for k in (2, 3):
        if lat[k]:
            a = np.array(lat[k]) * 1e3
            out[f"{k}d"] = {"requests": int(a.size), "p50_ms": float(np.percentile(a, 50)),
                            "p99_ms": float(np.percentile(a, 99))}
    return out


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
    return {0: f"stream:TT={tt},SEG={ry}", 1: "generic", 2: f"block:TT={tt},RY={ry}", 3: f"s3:TT={tt},SLICES={ry}"}[kind]


if __name__ == "__main__":
    main()

"""The step-kernel autotuner: measures the candidate kernels on the device for
each class of problem (dimension, dtype, iso/aniso, spatial weights) and image
size, and writes the tuning table that libfastprox_rocm reads at start-up
(fastprox_amd/tuning.txt; format in csrc/autotune.cpp).

For every class and size a batch of synthetic images (rectangles plus noise,
about 16 Mpx in all: at 512^2 the 64 images of the headline batch, whose
tiles run as two halves on two streams) is solved to a gap tolerance of 1e-6 with each candidate:
the 2D streaming kernel with TT iterations per launch and segments of SEG rows,
with or without a tail kernel; the 3D streaming kernel with slabs of SLICES
slices and one, two or three iterations per launch (several for float32
volumes without pixel weights). The fastest candidate (wall time of the whole solve) is kept. Results
do not depend on the choice (every kernel computes the same iterates).

Small batches are tuned on their own (SMALL_BATCHES: one image, as a plain
per-call solve, and 8 images), over shorter tiles as well: a lone 512^2 image
fills the GPU only with many short tiles, which a full batch does not need.
Their lines carry the batch size as a last column (max_images).

    python -m fastprox_amd.autotune [--out fastprox_amd/tuning.txt] [--quick] [--classes 0,1,...]

--classes measures only those classes (indices into CLASSES) and replaces
their lines in the table at --out, keeping the others.
"""

from __future__ import annotations

import argparse
import pathlib
import sys
import time

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent

SIZES_2D = (128, 256, 512, 1024)
SIZES_3D = (32, 64, 128)
MAIN_2D = [(tt, seg) for tt in (2, 4, 6, 8) for seg in (16, 32, 40, 64)]
TAILS = [(0, 0, 0.0), (4, 16, 0.5), (4, 8, 0.3), (6, 16, 0.5), (6, 16, 0.7), (8, 16, 0.5), (8, 8, 0.3)]
SLICES_3D = (8, 16, 32, 64)
SMALL_BATCHES = (1, 8)
SMALL_2D = [(tt, seg) for tt in (2, 4, 6, 8) for seg in (8, 16, 32)]
SMALL_3D = (4, 8, 16, 32)


def _batch(op, dtype, shapes, weighted):
    from .torch import Batch
    return Batch(op, dtype, shapes, weighted=weighted)


def _time_solve(bt, xs, zs, ds, gam, ws, reps):
    import torch
    bt.solve(xs, zs, ds, gam, weights=ws, tol=1e-6, max_iter=100000)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        bt.solve(xs, zs, ds, gam, weights=ws, tol=1e-6, max_iter=100000)
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def tune_class(k, dtype, iso, sw, sizes, log, quick=False):
    rows = []
    for side in sizes:
        full = max(1, min(256, (16 << 20) // side ** k))
        small = [b for b in SMALL_BATCHES if b < full and not quick]
        for B in small + [full]:
            rows.append(_tune_one(k, dtype, iso, sw, side, B, B < full, log, quick))
    return rows


def _tune_one(k, dtype, iso, sw, side, B, small, log, quick):
    """The best kernel for a batch of B images of side^k; a row of the table
    (max_images = B for a small batch, 0 for the full one)."""
    import torch

    from . import Op
    from .synth import rectangles
    tsize = 4 if dtype == torch.float32 else 8
    n = side ** k
    shape = (side,) * k
    x = rectangles(B, shape, n=12, sigma=0.1, seed=side, dtype=np.float32 if tsize == 4 else np.float64)
    xs = [torch.tensor(v, device="cuda") for v in x]
    zs = [torch.empty_like(v) for v in xs]
    ds = [torch.empty((k,) + shape, dtype=dtype, device="cuda") for _ in xs]
    rng = np.random.default_rng(side)
    ws = None
    if sw:
        wshape = shape if iso else (k,) + shape
        ws = [torch.tensor(rng.lognormal(0.0, 0.3, size=wshape), dtype=dtype, device="cuda") for _ in xs]
    op = Op(iso, k, [1.0] * k if iso else [1.0, 0.5] + ([1.0] if k == 3 else []))
    bt = _batch(op, dtype, [shape] * B, sw)
    gam = [0.1] * B
    best = None
    if k == 2:
        mains = [(2, 32), (4, 32), (8, 32)] if quick else (SMALL_2D if small else MAIN_2D)
        for tt, seg in mains:
            bt.set_kernel(0, tt, seg)
            for ttt, tseg, frac in ([(0, 0, 0.0)] if quick or B == 1 else TAILS):
                if frac > 0 and ttt < tt:
                    continue
                bt.set_tail(ttt or tt, tseg or seg, frac)
                sec = _time_solve(bt, xs, zs, ds, gam, ws, 2)
                cand = (sec, 0, tt, seg, ttt, tseg, frac)
                log(f"  k={k} {dtype} iso={iso} sw={sw} {side}^{k} x{B}: stream TT={tt} SEG={seg} "
                    f"tail={ttt},{tseg},{frac}: {B / sec:.1f} images/s")
                if best is None or sec < best[0]:
                    best = cand
    else:
        # several iterations per launch: float32 volumes without pixel weights
        tts = (1, 2, 3) if tsize == 4 and not sw else (1,)
        for tt in tts:
            for sl in ((16, 32) if quick else (SMALL_3D if small else SLICES_3D)):
                bt.set_kernel(3, tt, sl)
                sec = _time_solve(bt, xs, zs, ds, gam, ws, 2)
                log(f"  k=3 {dtype} iso={iso} sw={sw} {side}^3 x{B}: s3 TT={tt} SLICES={sl}: "
                    f"{B / sec:.1f} images/s")
                if best is None or sec < best[0]:
                    best = (sec, 3, tt, sl, 0, 0, 0.0)
    sec, kind, tt, ry, ttt, tseg, frac = best
    del bt, xs, zs, ds, ws
    return (k, tsize, int(iso), int(sw), n, kind, tt, ry, ttt, tseg, frac, B / sec, B if small else 0)


# (k, dtype bytes, iso, spatial weights)
CLASSES = [(2, 4, True, False), (2, 4, True, True), (2, 8, True, False), (2, 4, False, False), (2, 8, False, False),
           (3, 4, True, False), (3, 4, False, False)]


def _read_table(path):
    """{class key: [line, ...]} of an existing table."""
    out = {}
    try:
        lines = pathlib.Path(path).read_text().splitlines()
    except OSError:
        return out
    for ln in lines:
        v = ln.split()
        if not v or ln.startswith("#"):
            continue
        out.setdefault((int(v[0]), int(v[1]), bool(int(v[2])), bool(int(v[3]))), []).append(ln)
    return out


def main(argv=None):
    import torch
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(HERE / "tuning.txt"))
    ap.add_argument("--quick", action="store_true", help="a few candidates per class (a smoke run)")
    ap.add_argument("--classes", default=None, help="indices into CLASSES to measure (the rest are kept)")
    args = ap.parse_args(argv)

    def log(m):
        print(m, file=sys.stderr, flush=True)
    pick = range(len(CLASSES)) if args.classes is None else [int(v) for v in args.classes.split(",")]
    kept = _read_table(args.out) if args.classes is not None else {}
    lines = {}
    for i in pick:
        k, tb, iso, sw = CLASSES[i]
        dt = torch.float32 if tb == 4 else torch.float64
        rows = tune_class(k, dt, iso, sw, SIZES_2D if k == 2 else SIZES_3D, log, args.quick)
        lines[CLASSES[i]] = [f"{r[0]} {r[1]} {r[2]} {r[3]} {r[4]} {r[5]} {r[6]} {r[7]} {r[8]} {r[9]} {r[10]:g}"
                             + (f" {r[12]}" if r[12] else "") + f"   # {r[11]:.1f}" for r in rows]
    name = torch.cuda.get_device_name(0)
    with open(args.out, "w") as f:
        f.write(f"# libfastprox_rocm tuning table, measured by fastprox_amd/autotune.py on {name}\n")
        f.write("# k dtype_bytes iso spatial max_pixels kind TT RY tail_TT tail_RY tail_frac [max_images]   "
                "(images/s)\n")
        for c in CLASSES:
            for ln in lines.get(c, kept.get(c, [])):
                f.write(ln + "\n")
    log(f"wrote {args.out}")


if __name__ == "__main__":
    main()

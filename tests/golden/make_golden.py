"""Writes tests/golden/golden_tv.npz: inputs and the reference library's
outputs for the image-TV path, tests/golden/golden_deriv.npz: the reference
library's isotropic JVP and VJP at kept solves, and
tests/golden/cvxpy_images.npz: the
reference's own CVXPY+Clarabel image fixtures (tv_iso and tv_aniso cases of
tests/fixtures/reference_images.npz, copied as data).

Run in the build container, where the reference library is built from its
sources into oracle/_ref (oracle/Makefile.ref):

    python tests/golden/make_golden.py
"""

from __future__ import annotations

import pathlib
import sys
import zlib

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))
sys.path.insert(0, str(HERE.parent))

import refbind  # noqa: E402

REF_FIXTURES = pathlib.Path("/root/reference/tests/fixtures/reference_images.npz")


def image(shape, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return rng.normal(scale=scale, size=shape) + (np.indices(shape[-2:]).sum(0) // 4 % 2)


def cases():
    """(name, iso, x, gamma or steps, w, spatial, tol, max_iter, warm_tol)."""
    c = []
    c.append(("iso_16x18", True, image((16, 18), 1), 0.3, [1.0, 1.0], None, 1e-6, 100000, None))
    c.append(("iso_12x13_w", True, image((12, 13), 2), 0.5, [0.7, 1.4], None, 1e-9, 100000, None))
    c.append(("iso_20x20_tight", True, image((20, 20), 3), 0.1, [1.0, 1.0], None, 1e-12, 100000, None))
    c.append(("iso_33x47_g1", True, image((33, 47), 4), 1.0, [1.0, 1.0], None, 1e-8, 100000, None))
    c.append(("iso_9x1", True, image((9, 1), 5), 0.4, [1.0, 1.0], None, 1e-6, 1000, None))
    c.append(("iso_1x9", True, image((1, 9), 6), 0.4, [1.0, 1.0], None, 1e-6, 1000, None))
    c.append(("iso_4x5_oneaxis", True, image((4, 5), 7), 0.4, [0.0, 1.0], None, 1e-6, 1000, None))
    c.append(("iso_7x7_gamma0", True, image((7, 7), 8), 0.0, [1.0, 1.0], None, 1e-6, 1000, None))
    c.append(("iso_batch3", True, image((3, 40, 70), 9), 0.2, [1.0, 1.0], None, 1e-7, 100000, None))
    c.append(("iso_huge_scale", True, image((10, 11), 10) * 2.0 ** 80, 0.3 * 2.0 ** 80, [1.0, 1.0], None, 1e-8,
              100000, None))
    c.append(("iso_tiny_scale", True, image((10, 11), 11) * 2.0 ** -90, 0.3 * 2.0 ** -90, [1.0, 1.0], None, 1e-8,
              100000, None))
    c.append(("iso_huge_gamma", True, image((8, 9), 12), 1e6, [1.0, 1.0], None, 1e-8, 100000, None))
    c.append(("iso_3d", True, image((6, 7, 8), 13), 0.25, [1.0, 0.5, 1.2], None, 1e-8, 100000, None))
    c.append(("iso_3d_inactive", True, image((6, 7, 5), 14), 0.3, [1.0, 0.0, 0.8], None, 1e-7, 100000, None))
    rng = np.random.default_rng(15)
    al = rng.lognormal(0.0, 0.3, size=(10, 12))
    al[2, 3] = 0.0
    c.append(("iso_pixel", True, image((10, 12), 16), 0.3, [1.0, 1.0], al, 1e-8, 100000, None))
    al3 = rng.lognormal(0.0, 0.3, size=(5, 6, 7))
    c.append(("iso_pixel_3d", True, image((5, 6, 7), 17), 0.2, [1.0, 1.0, 1.0], al3, 1e-7, 100000, None))
    c.append(("iso_warm", True, image((24, 30), 18), 0.2, [1.0, 1.0], None, 1e-9, 100000, 1e-4))
    c.append(("iso_steps", True, image((3, 20, 22), 19), np.array([0.1, 0.3, 0.0]), [1.0, 1.0], None, 1e-8,
              100000, None))
    xn = image((2, 9, 10), 20)
    xn[1, 3, 4] = np.nan
    c.append(("iso_nan", True, xn, 0.2, [1.0, 1.0], None, 1e-8, 100000, None))
    c.append(("iso_f32", True, image((2, 30, 34), 21).astype(np.float32), 0.15, [1.0, 1.0], None, 1e-6, 100000,
              None))
    c.append(("iso_maxiter", True, image((16, 16), 22), 0.5, [1.0, 1.0], None, 1e-14, 37, None))
    # anisotropic: the reference's ADBPG solver, tight, for certificate comparisons
    c.append(("aniso_16x18", False, image((16, 18), 1), 0.3, [1.0, 1.0], None, 1e-13, 100000, None))
    c.append(("aniso_12x13_w", False, image((12, 13), 2), 0.5, [0.7, 1.4], None, 1e-13, 100000, None))
    c.append(("aniso_3d", False, image((6, 7, 8), 13), 0.25, [1.0, 0.5, 1.2], None, 1e-13, 100000, None))
    be = rng.uniform(0.2, 2.0, size=(2, 11, 13))
    be[0, 4, 5] = 0.0
    c.append(("aniso_edge", False, image((11, 13), 23), 0.4, [1.0, 0.7], be, 1e-13, 100000, None))
    return c


def deriv_cases():
    """(name, x, gamma, w, pixel weights, tol): isotropic derivatives."""
    rng = np.random.default_rng(40)
    yy, xx = np.mgrid[0:64, 0:64] / 64
    audit = ((xx - 0.5) ** 2 + (yy - 0.5) ** 2 < 0.1) + 0.5 * (xx > 0.7)
    audit = audit + 0.1 * np.random.default_rng(0).standard_normal((64, 64))
    return [
        ("d_audit", audit, 0.2, [1.0, 1.0], None, 1e-9),
        ("d_audit_w", audit, 0.2, [1.0, 0.6], np.random.default_rng(2).uniform(0.5, 1.5, audit.shape), 1e-9),
        ("d_72x70", image((72, 70), 41), 0.15, [1.0, 1.0], None, 1e-8),
        ("d_3d", image((5, 16, 14), 42), 0.3, [1.0, 0.7, 1.2], None, 1e-8),
        ("d_pixel", image((40, 44), 43), 0.1, [0.8, 1.0], rng.uniform(0.5, 1.5, (40, 44)), 1e-7),
        ("d_loose", image((48, 40), 44), 0.1, [1.0, 1.0], None, 1e-5),
        ("d_scaled", image((30, 31), 45) * 2.0 ** 80, 0.2 * 2.0 ** 80, [1.0, 1.0], None, 1e-9),
        ("d_oneaxis", image((17, 23), 46), 0.3, [1.0, 0.0], None, 1e-9),
    ]


def write_deriv():
    out = {}
    for name, x, gamma, w, al, tol in deriv_cases():
        k = x.ndim
        op = refbind.Op(True, k, w, spatial=al)
        z, dual, gaps, its, zerrs, status, ws = op.solve(x, gamma, tol=tol, max_iter=10 ** 6, keep=True)
        rng = np.random.default_rng(zlib.crc32(name.encode()))
        P = k + (x.size if al is not None else 0)
        g, dx, dp = rng.standard_normal(x.shape), rng.standard_normal(x.shape), rng.standard_normal(P)
        gx, gg, gp = op.vjp(ws, g)
        jv = op.jvp(ws, dx, 0.3, dp)
        refbind.Op.free_ws(ws)
        for key, val in (("x", x), ("gamma", np.array(gamma)), ("w", np.asarray(w, np.float64)), ("tol", np.array(tol)),
                         ("z", z), ("dual", dual), ("gap", np.array(gaps[0])), ("iters", np.array(its[0])),
                         ("g", g), ("dx", dx), ("dgamma", np.array(0.3)), ("dparams", dp), ("gx", gx),
                         ("ggamma", np.array(gg)), ("gparams", gp), ("jvp", jv)):
            out[f"{name}/{key}"] = val
        if al is not None:
            out[f"{name}/alpha"] = al
        print(f"{name}: iters {its[0]} gap {gaps[0]:.3e}")
    np.savez_compressed(HERE / "golden_deriv.npz", **out)


def main():
    write_deriv()
    out = {}
    for name, iso, x, gamma, w, sp, tol, max_iter, warm_tol in cases():
        k = len(w)
        op = refbind.Op(iso, k, w, spatial=sp)
        warm = None
        if warm_tol is not None:
            _, warm, *_ = op.solve(x, gamma, tol=warm_tol, max_iter=max_iter)
            out[f"{name}/warm"] = warm
        steps = gamma if isinstance(gamma, np.ndarray) else None
        z, dual, gaps, its, zerrs, status, _ = op.solve(x, 0.0 if steps is not None else gamma, tol=tol,
                                                      max_iter=max_iter, warm=warm, steps=steps)
        out[f"{name}/iso"] = np.array(iso)
        out[f"{name}/x"] = x
        out[f"{name}/gamma"] = np.asarray(gamma, np.float64)
        out[f"{name}/w"] = np.asarray(w, np.float64)
        if sp is not None:
            out[f"{name}/spatial"] = sp
        out[f"{name}/tol"] = np.array(tol)
        out[f"{name}/max_iter"] = np.array(max_iter)
        out[f"{name}/z"] = z
        out[f"{name}/dual"] = dual
        out[f"{name}/gaps"] = gaps
        out[f"{name}/iters"] = its
        out[f"{name}/zerrs"] = zerrs
        out[f"{name}/status"] = np.array(status)
        print(f"{name}: iters {list(its)} gaps {list(gaps)} status {status}")
    np.savez_compressed(HERE / "golden_tv.npz", **out)
    if REF_FIXTURES.exists():
        d = np.load(REF_FIXTURES)
        keep = {}
        for key in d.files:
            p = key.split("/", 1)[0]
            case = str(d[f"{p}/case"])
            if case.startswith("tv_iso") or case.startswith("tv_aniso"):
                keep[key] = d[key]
        np.savez_compressed(HERE / "cvxpy_images.npz", **keep)
        print(f"cvxpy fixtures: {len({k.split('/')[0] for k in keep})} problems")


if __name__ == "__main__":
    main()

"""fastprox_amd: fastprox's image total-variation operators (isotropic and
anisotropic, 2D and 3D, dual FISTA with a duality-gap certificate) as HIP
kernels for AMD Instinct MI355X (gfx950), behind fastprox's C ABI
(include/fastprox_rocm.h) and its Python entry points.

    from fastprox_amd.torch import tv_iso, tv_aniso

The compute path is libfastprox_rocm.so, built in-tree by
__graft_entry__.build() (or `make -C fastprox_amd/csrc`). Importing the
package does not load it; the first call does, and raises when it is missing.
"""

from __future__ import annotations

import dataclasses

__version__ = "1.0.6"


@dataclasses.dataclass
class IterInfo:
    """The result record of an iterative solve (the reference's IterInfo,
    bindings/python/fastprox/__init__.py:97-120): the gap summed over the
    images, the largest iteration count, whether every image converged, the
    dual in the C form (ndim, *x.shape), and the certified relative error of
    each image's z."""

    gap: float
    iters: int
    ties: int
    converged: bool
    dual: object = None
    zerr: object = None
    gaps: object = None
    image_iters: object = None


def library_path():
    from . import _capi
    return str(_capi.LIB_PATH)


from .core import FastproxError, Op, Workspace  # noqa: E402

__all__ = ["IterInfo", "FastproxError", "Op", "Workspace", "library_path"]

"""The C ABI of libfastprox_rocm without a GPU: the library loads, exports
every entry point include/fastprox_rocm.h declares, and its operator
constructors and host-only queries follow the reference's rules
(src/api/api.cpp, src/ops/tv_iso.cpp:1072-1095, tv_aniso.cpp:1432-1459)."""

from __future__ import annotations

import ctypes as C
import pathlib
import re

import numpy as np
import pytest

from fastprox_amd import _capi as A

ROOT = pathlib.Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "fastprox_rocm.h"


def declared():
    text = "\n".join(l for l in HEADER.read_text().splitlines() if not l.lstrip().startswith("#"))
    return sorted(set(re.findall(r"FP_API[^;(]*?\b(\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    return A.load()


def test_every_declared_symbol_is_exported(lib):
    names = declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_only_public_symbols_are_exported():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", str(A.LIB_PATH)], capture_output=True, text=True,
                         check=True).stdout
    syms = [l.split()[-1] for l in out.splitlines() if " T " in l]
    ours = [s for s in syms if s.startswith(("fp_", "fpr_", "fastprox_"))]
    assert set(declared()) <= set(ours)


def _w(v):
    a = np.ascontiguousarray(np.atleast_1d(np.asarray(v, np.float64)))
    return a, A.numpy_array(a)


def test_constructor_rules(lib):
    op = C.c_void_p()
    assert lib.fp_op_tv_iso(4, None, C.byref(op)) == A.FP_EPARAM
    assert lib.fp_op_tv_iso(1, None, C.byref(op)) == A.FP_EPARAM
    k, a = _w([1.0, 2.0, 3.0])
    assert lib.fp_op_tv_iso(2, a.ref, C.byref(op)) == A.FP_ESHAPE
    k, a = _w([1.0, -1.0])
    assert lib.fp_op_tv_iso(2, a.ref, C.byref(op)) == A.FP_EPARAM
    k, a = _w([1.0, np.nan])
    assert lib.fp_op_tv_aniso(2, a.ref, C.byref(op)) == A.FP_EPARAM
    k, a = _w([0.5])
    assert lib.fp_op_tv_iso(3, a.ref, C.byref(op)) == A.FP_OK
    assert lib.fp_op_name(op).decode() == "tv_iso"
    assert lib.fp_op_nparams(op) == 3
    lib.fp_op_free(op)
    assert lib.fp_op_tv_aniso(2, None, C.byref(op)) == A.FP_OK
    assert lib.fp_op_name(op).decode() == "tv_aniso"
    lib.fp_op_free(op)


def test_weighted_constructors(lib):
    op = C.c_void_p()
    al = np.ones((4, 5))
    al[1, 1] = 0.0
    arr = A.numpy_array(al)
    assert lib.fp_op_tv_iso_weighted(2, None, arr.ref, C.byref(op)) == A.FP_OK
    assert lib.fp_op_nparams(op) == 2 + 20
    lib.fp_op_free(op)
    bad = al.copy()
    bad[0, 0] = -1.0
    assert lib.fp_op_tv_iso_weighted(2, None, A.numpy_array(bad).ref, C.byref(op)) == A.FP_EPARAM
    assert lib.fp_op_tv_iso_weighted(2, None, None, C.byref(op)) == A.FP_EINVAL
    be = np.ones((2, 4, 5))
    assert lib.fp_op_tv_aniso_weighted(2, None, A.numpy_array(be).ref, C.byref(op)) == A.FP_OK
    assert lib.fp_op_nparams(op) == 2 + 40
    lib.fp_op_free(op)
    assert lib.fp_op_tv_aniso_weighted(2, None, A.numpy_array(np.ones((3, 4, 5))).ref, C.byref(op)) == A.FP_ESHAPE


def test_host_queries(lib):
    o = A.fp_options()
    lib.fp_options_init(C.byref(o))
    assert (o.axis, o.threads, o.tol, o.max_iter, o.tie, o.keep_structure) == (-1, 0, 1e-6, 1000, 0, 0)
    assert lib.fp_abi_version() == 1
    assert lib.fp_version_string().decode() == "1.0.6"
    assert lib.fp_strerror(A.FP_ENOTCONV).decode().startswith("iteration stopped")
    assert lib.fp_device_available(A.FP_DEVICE_CPU) == 1
    assert lib.fp_device_available(A.FP_DEVICE_CUDA) == 0
    op = C.c_void_p()
    assert lib.fp_op_tv_iso(2, None, C.byref(op)) == A.FP_OK
    rpo = C.c_int(7)
    assert lib.fp_op_device_check(op, A.FP_DEVICE_CPU, C.byref(rpo)) == A.FP_OK and rpo.value == 0
    assert lib.fp_op_device_check(op, A.FP_DEVICE_CUDA, None) == A.FP_EINVAL
    assert lib.fp_op_device_check(None, A.FP_DEVICE_CPU, None) == A.FP_EINVAL
    lib.fp_op_free(op)
    assert lib.fp_prox(None, 0.1, None, None, None, None, None) == A.FP_EINVAL
    assert lib.fp_workspace_info(None, None) == A.FP_EINVAL


class _Plugin(C.Structure):
    _fields_ = [("abi", C.c_int), ("build", C.c_char_p), ("device_count", C.c_void_p), ("op_check", C.c_void_p)]


def test_plugin_table(lib):
    lib.fastprox_rocm_plugin.restype = C.POINTER(_Plugin)
    t = lib.fastprox_rocm_plugin().contents
    assert t.abi == 1 and t.build.decode() == "1.0.6" and t.device_count and t.op_check


def test_oracle_is_not_linked_by_the_product():
    """The product library and package never reference the oracle."""
    import subprocess
    out = subprocess.run(["ldd", str(A.LIB_PATH)], capture_output=True, text=True).stdout
    assert "oracle" not in out and "fastprox_ref" not in out
    for f in (ROOT / "fastprox_amd").rglob("*.py"):
        src = f.read_text()
        assert "import oracle" not in src and "from oracle" not in src, f

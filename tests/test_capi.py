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


def _descriptor_cases(buf):
    """(name, fp_array Array, is_valid) for descriptors of x: the boundary's
    rules, each broken once, and valid ones beside them."""
    p = buf.ctypes.data
    big, huge = 1 << 40, 1 << 62
    return [
        ("int32 dtype", A.Array(p, A.FP_I32, (4, 5)), False),
        ("unknown dtype", A.Array(p, 99, (4, 5)), False),
        ("cuda device", A.Array(p, A.FP_F64, (4, 5), device=A.FP_DEVICE_CUDA), False),
        ("unknown device", A.Array(p, A.FP_F64, (4, 5), device=7), False),
        ("negative extent", A.Array(p, A.FP_F64, (-1, 5)), False),
        ("too many entries", A.Array(p, A.FP_F64, (big, big)), False),
        ("stride beyond the range", A.Array(p, A.FP_F64, (4, 5), (huge, 1)), False),
        ("negative stride beyond the range", A.Array(p, A.FP_F64, (4, 5), (-huge, 1)), False),
        ("offset span beyond the range", A.Array(p, A.FP_F64, (1 << 20, 1 << 20), (1 << 41, 1)), False),
        ("no data", A.Array(0, A.FP_F64, (4, 5)), False),
        ("one axis for a 2D operator", A.Array(p, A.FP_F64, (20,)), False),
        ("contiguous", A.Array(p, A.FP_F64, (4, 5)), True),
        ("strided", A.Array(p, A.FP_F64, (2, 5), (10, 1)), True),
        ("negative stride", A.Array(p + 8 * 15, A.FP_F64, (4, 5), (-5, 1)), True),
        ("a huge stride on an extent of 1", A.Array(p, A.FP_F64, (1, 4, 5), (huge, 5, 1)), True),
        ("huge extents beside a zero one", A.Array(0, A.FP_F64, (0, big, big)), False),  # extents above 1 count
        ("no entries and no data", A.Array(0, A.FP_F64, (0, 5)), True),
    ]


def _ndim_case(buf, ndim, shape_null=False):
    a = A.Array(buf.ctypes.data, A.FP_F64, (4, 5))
    a.s.ndim = ndim
    if shape_null:
        a.s.shape = None
    return a


def test_descriptor_rules_match_the_reference(lib):
    """fp_workspace_create checks x's descriptor before any device work: each
    malformed descriptor gets the status the reference library gives it
    (oracle/_ref, built from its sources), and the valid ones pass the checks
    (here, without a GPU, they stop at FP_EDEVICE)."""
    import refbind
    buf = np.zeros(64)
    cases = _descriptor_cases(buf)
    cases += [("ndim -1", _ndim_case(buf, -1), False), ("ndim 33", _ndim_case(buf, 33), False),
              ("null shape", _ndim_case(buf, 2, shape_null=True), False)]
    ref = refbind.lib() if refbind.available() else None
    ops = []
    for L in [lib] + ([ref] if ref else []):
        op = C.c_void_p()
        assert L.fp_op_tv_iso(2, None, C.byref(op)) == A.FP_OK
        ops.append(op)
    expect = {"int32 dtype": A.FP_EDTYPE, "unknown dtype": A.FP_EDTYPE, "cuda device": A.FP_EINVAL,
              "unknown device": A.FP_EINVAL, "no data": A.FP_EINVAL, "null shape": A.FP_EINVAL}
    for name, arr, valid in cases:
        ws = C.c_void_p()
        ours = lib.fp_workspace_create(ops[0], arr.ref, None, C.byref(ws))
        if ours == A.FP_OK:
            lib.fp_workspace_free(ws)
        if valid:
            assert ours in (A.FP_OK, A.FP_EDEVICE), (name, A.STATUS.get(ours))
        else:
            assert ours == expect.get(name, A.FP_ESHAPE), (name, A.STATUS.get(ours))
        if ref:
            theirs = ref.fp_workspace_create(ops[1], arr.ref, None, C.byref(ws))
            if theirs == A.FP_OK:
                ref.fp_workspace_free(ws)
            assert (theirs == A.FP_OK) == valid, (name, A.STATUS.get(theirs))
            if not valid:
                assert ours == theirs, (name, A.STATUS.get(ours), A.STATUS.get(theirs))
    lib.fp_op_free(ops[0])
    if ref:
        ref.fp_op_free(ops[1])


def test_weight_descriptor_rules(lib):
    """The axis-weight array of the constructors follows the same rules."""
    op = C.c_void_p()
    w = np.ones(2, np.int32)
    assert lib.fp_op_tv_iso(2, A.numpy_array(w).ref, C.byref(op)) == A.FP_EDTYPE
    w2 = np.ones((2, 1))
    assert lib.fp_op_tv_iso(2, A.numpy_array(w2).ref, C.byref(op)) == A.FP_ESHAPE
    assert lib.fp_op_tv_iso(2, A.Array(0, A.FP_F64, (2,)).ref, C.byref(op)) == A.FP_EINVAL

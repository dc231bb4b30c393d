"""ctypes view of the C ABI in include/fastprox_rocm.h (the fastprox.h
subset for image TV, include/fastprox.h in the reference): the structs, the
enums and a loader for the HIP library built in-tree.

The structs mirror the reference header field for field (fp_array
include/fastprox.h:61-68, fp_options :93-103, fp_info :105-109), so the same
definitions drive both the reference library and libfastprox_rocm."""

from __future__ import annotations

import ctypes as C
import os
import pathlib

FP_F32, FP_F64, FP_I8, FP_I32, FP_I64 = 1, 2, 3, 4, 5
FP_DEVICE_CPU, FP_DEVICE_CUDA, FP_DEVICE_ROCM = 0, 2, 10
FP_ENTRY_FREE, FP_ENTRY_ZERO, FP_ENTRY_LOWER, FP_ENTRY_UPPER = 0, 1, 2, 3
FP_STRUCT_NONE, FP_STRUCT_PROJECTOR, FP_STRUCT_WEIGHTED, FP_STRUCT_CURVED = 0, 1, 2, 3

(FP_OK, FP_EINVAL, FP_EDTYPE, FP_ESHAPE, FP_ENOMEM, FP_ENOTCONV, FP_EPARAM, FP_EINTERNAL,
 FP_EDEVICE) = range(9)
STATUS = {FP_OK: "FP_OK", FP_EINVAL: "FP_EINVAL", FP_EDTYPE: "FP_EDTYPE", FP_ESHAPE: "FP_ESHAPE",
          FP_ENOMEM: "FP_ENOMEM", FP_ENOTCONV: "FP_ENOTCONV", FP_EPARAM: "FP_EPARAM",
          FP_EINTERNAL: "FP_EINTERNAL", FP_EDEVICE: "FP_EDEVICE"}


class fp_array(C.Structure):
    _fields_ = [("data", C.c_void_p), ("dtype", C.c_int), ("ndim", C.c_int),
                ("shape", C.POINTER(C.c_int64)), ("strides", C.POINTER(C.c_int64)),
                ("device", C.c_int)]


class fp_options(C.Structure):
    _fields_ = [("axis", C.c_int), ("threads", C.c_int), ("tol", C.c_double), ("max_iter", C.c_int),
                ("tie", C.c_int), ("keep_structure", C.c_int)]


class fp_info(C.Structure):
    _fields_ = [("gap", C.c_double), ("iters", C.c_int), ("ties", C.c_int64)]


class Array:
    """An fp_array over a buffer, keeping its shape and strides alive.
    strides in elements (None: C-contiguous)."""

    def __init__(self, ptr, dtype, shape, strides=None, device=FP_DEVICE_CPU):
        self.shape_buf = (C.c_int64 * max(1, len(shape)))(*shape)
        self.strides_buf = None if strides is None else (C.c_int64 * max(1, len(strides)))(*strides)
        self.s = fp_array(C.c_void_p(ptr), dtype, len(shape),
                          C.cast(self.shape_buf, C.POINTER(C.c_int64)),
                          None if strides is None else C.cast(self.strides_buf, C.POINTER(C.c_int64)),
                          device)

    @property
    def ref(self):
        return C.byref(self.s)


def numpy_array(a):
    """An Array over a numpy array (float32/64, int8/32/64; any strides)."""
    import numpy as np
    dt = {np.dtype(np.float32): FP_F32, np.dtype(np.float64): FP_F64, np.dtype(np.int8): FP_I8,
          np.dtype(np.int32): FP_I32, np.dtype(np.int64): FP_I64}[a.dtype]
    strides = [s // a.itemsize for s in a.strides]
    return Array(a.ctypes.data, dt, a.shape, strides, FP_DEVICE_CPU)


FPR_METHOD_ADBPG, FPR_METHOD_FISTA = 0, 1
METHODS = {"adbpg": FPR_METHOD_ADBPG, "fista": FPR_METHOD_FISTA}


def declare(lib):
    """Argument and result types of the fp_* entry points present in lib."""
    P, I, D, I64 = C.c_void_p, C.c_int, C.c_double, C.c_int64
    A = C.POINTER(fp_array)
    sigs = {
        "fp_options_init": (None, [C.POINTER(fp_options)]),
        "fp_op_tv_iso": (I, [I, A, C.POINTER(P)]),
        "fp_op_tv_iso_weighted": (I, [I, A, A, C.POINTER(P)]),
        "fp_op_tv_aniso": (I, [I, A, C.POINTER(P)]),
        "fp_op_tv_aniso_weighted": (I, [I, A, A, C.POINTER(P)]),
        "fp_op_free": (None, [P]),
        "fp_op_nparams": (I64, [P]),
        "fp_op_name": (C.c_char_p, [P]),
        "fp_workspace_create": (I, [P, A, C.POINTER(fp_options), C.POINTER(P)]),
        "fp_workspace_free": (None, [P]),
        "fp_prox": (I, [P, D, A, A, C.POINTER(fp_options), P, C.POINTER(fp_info)]),
        "fp_prox_steps": (I, [P, A, A, A, C.POINTER(fp_options), P]),
        "fp_workspace_get_dual": (I, [P, A]),
        "fp_workspace_set_dual": (I, [P, A]),
        "fp_workspace_set_warm": (I, [P, I]),
        "fp_workspace_set_deriv_tol": (I, [P, D, I]),
        "fp_workspace_deriv_info": (I, [P, C.POINTER(D), C.POINTER(I)]),
        "fp_workspace_info": (I, [P, C.POINTER(fp_info)]),
        "fp_workspace_image_info": (I, [P, C.POINTER(D), C.POINTER(I), C.POINTER(D), I64]),
        "fp_workspace_refine": (I, [P, P, D, C.c_int, A, C.POINTER(fp_info)]),
        "fp_workspace_deriv_exact": (I, [P, C.POINTER(I), C.POINTER(D)]),
        "fp_jvp": (I, [P, P, A, D, A, A]),
        "fp_vjp": (I, [P, P, A, A, C.POINTER(D), A]),
        "fp_vjp_device": (I, [P, P, A, A, A, A]),
        "fp_op_struct_kind": (I, [P, C.POINTER(I)]),
        "fp_workspace_units": (I, [P, C.POINTER(I64), C.POINTER(I64)]),
        "fp_workspace_states": (I, [P, A, C.POINTER(I64)]),
        "fp_workspace_regions": (I, [P, A, A, C.POINTER(I64), I64]),
        "fp_workspace_constraint": (I, [P, A, C.POINTER(C.c_int8), I64]),
        "fp_workspace_kinks": (I, [P, D, A, C.POINTER(I64)]),
        "fp_workspace_set_branch": (I, [P, A, D]),
        "fp_device_available": (I, [I]),
        "fp_op_device_check": (I, [P, I, C.POINTER(I)]),
        "fp_workspace_set_stream": (I, [P, P]),
        "fp_workspace_get_stream": (I, [P, C.POINTER(P)]),
        "fp_strerror": (C.c_char_p, [I]),
        "fp_abi_version": (I, []),
        "fp_version_string": (C.c_char_p, []),
        "fpr_alloc_count": (I64, []),
        "fpr_last_error": (C.c_char_p, []),
        "fpr_prox_static": (I, [P, D, A, A, C.POINTER(fp_options), P, C.POINTER(fp_info)]),
        "fpr_workspace_set_method": (I, [P, I]),
        "fpr_workspace_get_method": (I, [P, C.POINTER(I)]),
        "fpr_workspace_polished": (I, [P, C.POINTER(I), I64, C.POINTER(I)]),
        "fpr_batch_set_method": (I, [P, I]),
        "fpr_batch_get_method": (I, [P, C.POINTER(I)]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(lib, name, None)
        if f is None:
            continue
        f.restype = res
        f.argtypes = args
    return lib


LIB_NAME = "libfastprox_rocm.so"
LIB_PATH = pathlib.Path(__file__).resolve().parent / LIB_NAME


def load(path=None):
    """The in-tree HIP library, declared. Raises OSError when it is not built."""
    p = pathlib.Path(path or os.environ.get("FASTPROX_ROCM_LIB", LIB_PATH))
    if not p.exists():
        raise OSError(f"{p} is not built: run __graft_entry__.build() (or make -C fastprox_amd/csrc)")
    return declare(C.CDLL(str(p), mode=C.RTLD_GLOBAL))

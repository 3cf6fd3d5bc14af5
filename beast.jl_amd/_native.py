"""ctypes binding of libbeast_hip.so (the C ABI declared in include/beast_hip.h).

The library is built in-tree (``csrc/libbeast_hip.so``); there is no fallback:
if it is missing or cannot be loaded, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(_HERE, "csrc", "libbeast_hip.so")

BH_OK = 0
BH_ERR_INVALID = -1
BH_ERR_OOM = -2
BH_ERR_HIP = -3
BH_ERR_ASSERT = -4
BH_ERR_UNSUPPORTED = -5
BH_ERR_SINGULAR = -6
BH_ERR_NOCONV = -7

OP_SINGLELAYER, OP_DOUBLELAYER, OP_NCROSS, OP_DOUBLELAYER_ROTATED, OP_IDENTITY = 0, 1, 2, 3, 4
TRACE_TANGENTIAL, TRACE_CROSS = 0, 1
QS_WILTON_SAUTER, QS_SAUTER, QS_DOUBLENUM, QS_SINGLENUM = 0, 1, 2, 3
FP64, FP32, FP32_SPLIT = 0, 1, 2


class BHError(RuntimeError):
    """A failure reported by the native library."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class UnsupportedOnGPU(BHError):
    """Valid input the GPU path does not cover (the reference's CPU path would)."""


class SingularException(BHError):
    """LU hit an exactly zero pivot (LinearAlgebra.SingularException)."""


class NoConvergence(BHError):
    """Iterative refinement did not reach its tolerance."""


class bh_operator(C.Structure):
    _fields_ = [("kind", C.c_int32), ("gamma", C.c_double * 2), ("alpha", C.c_double * 2),
                ("beta", C.c_double * 2)]


class bh_quadstrat(C.Structure):
    _fields_ = [("kind", C.c_int32), ("params", C.c_int32 * 8)]


class bh_solve_stats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("inner", C.c_int32), ("residual", C.c_double),
                ("seconds_factor", C.c_double), ("seconds_solve", C.c_double)]


# (name, restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_I32 = C.POINTER(C.c_int32)
_I64 = C.POINTER(C.c_int64)
_SIGS = [
    ("bh_version", C.c_int, []),
    ("bh_last_error", C.c_char_p, []),
    ("bh_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("bh_ctx_destroy", C.c_int, [_P]),
    ("bh_ctx_synchronize", C.c_int, [_P]),
    ("bh_ctx_pool", C.c_int, [_P, C.c_int64, _I64]),
    ("bh_ctx_profile", C.c_int, [_P, C.c_int]),
    ("bh_ctx_kernel_stats", C.c_int, [_P, C.c_char_p, _I64, _D, _D]),
    ("bh_ctx_profile_only", C.c_int, [_P, C.c_char_p]),
    ("bh_mesh_create", C.c_int, [_P, _D, C.c_int64, _I32, C.c_int64, C.POINTER(_P)]),
    ("bh_mesh_destroy", C.c_int, [_P]),
    ("bh_space_create", C.c_int, [_P, C.c_int64, _I32, _I32, _I32, _D, C.POINTER(_P)]),
    ("bh_space_destroy", C.c_int, [_P]),
    ("bh_space_e2d", C.c_int, [_P, _I32, _I32, _D, C.c_int64]),
    ("bh_assemble", C.c_int, [_P, C.POINTER(bh_operator), C.POINTER(bh_quadstrat), C.c_int, _P, _P, _D,
                              C.c_int, _P, C.c_int64, C.c_int64, C.c_int64]),
    ("bh_check_strategy", C.c_int, [C.POINTER(bh_operator), C.POINTER(bh_quadstrat)]),
    ("bh_classify", C.c_int, [_P, C.POINTER(bh_operator), C.POINTER(bh_quadstrat), _P, _P,
                              C.POINTER(C.c_int8)]),
    ("bh_local_blocks", C.c_int, [_P, C.POINTER(bh_operator), C.POINTER(bh_quadstrat), _P, _P, C.c_int64,
                                  _I32, _I32, _D]),
    ("bh_assemble_planewave", C.c_int, [_P, _P, _D, _D, _D, _D, _D, C.c_int, C.c_int, _D]),
    ("bh_assemble_excitation", C.c_int, [_P, _P, C.c_int, _D, _D, _D, _D, _D, C.c_int, C.c_int, _D]),
    ("bh_farfield", C.c_int, [_P, C.c_int, _D, _D, _P, _D, C.c_int64, _D, C.c_int, _D]),
    ("bh_matrix_assemble", C.c_int, [_P, C.POINTER(bh_operator), C.POINTER(bh_quadstrat), _P, C.POINTER(_P)]),
    ("bh_matrix_assemble_batch", C.c_int, [_P, C.c_int64, C.POINTER(bh_operator), C.POINTER(bh_quadstrat),
                                           C.POINTER(_P), C.POINTER(_P)]),
    ("bh_matrix_accumulate", C.c_int, [_P, C.POINTER(bh_operator), C.POINTER(bh_quadstrat), _P, _D]),
    ("bh_matrix_from_host", C.c_int, [_P, C.c_int64, _D, C.c_int64, C.POINTER(_P)]),
    ("bh_matrix_destroy", C.c_int, [_P]),
    ("bh_matrix_size", C.c_int, [_P, _I64]),
    ("bh_matrix_download", C.c_int, [_P, _D, C.c_int64]),
    ("bh_matrix_factor", C.c_int, [_P, C.c_int]),
    ("bh_matrix_factor_batch", C.c_int, [_P, C.c_int64, C.POINTER(_P), C.c_int]),
    ("bh_matrix_solve", C.c_int, [_P, C.c_int64, _D, _D, C.c_double, C.c_int, C.POINTER(bh_solve_stats)]),
    ("bh_matrix_pivots", C.c_int, [_P, _I64]),
    ("bh_matrix_solve_noalloc", C.c_int, [_P, C.c_int64, _D, _D, C.c_double, C.c_int, C.POINTER(bh_solve_stats)]),
]

EXPORTED = [s[0] for s in _SIGS]

_lib = None


def load(path=None):
    """Load the library (once) and declare every entry point."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIBPATH
    if not os.path.exists(p):
        raise BHError(BH_ERR_HIP, f"native library not built: {p} (run __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, res, args in _SIGS:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    if path is None:
        _lib = lib
    return lib


def check(rc):
    if rc == BH_OK:
        return
    msg = (load().bh_last_error() or b"").decode("utf-8", "replace")
    if rc == BH_ERR_ASSERT:
        raise AssertionError(msg)
    if rc == BH_ERR_UNSUPPORTED:
        raise UnsupportedOnGPU(rc, msg)
    if rc == BH_ERR_SINGULAR:
        raise SingularException(rc, msg)
    if rc == BH_ERR_NOCONV:
        raise NoConvergence(rc, msg)
    if rc == BH_ERR_INVALID:
        raise ValueError(msg)
    raise BHError(rc, msg)


def dptr(a):
    return a.ctypes.data_as(_D)


def i32ptr(a):
    return a.ctypes.data_as(_I32)


def i64ptr(a):
    return a.ctypes.data_as(_I64)


def carr2(z):
    z = complex(z)
    return (C.c_double * 2)(z.real, z.imag)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)

"""ctypes binding of libmpol.so (the C ABI declared in include/mpol.h).

Loading fails loudly: there is no CPU fallback for the operator or the solve.
"""
from __future__ import annotations

import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmpol.so")

MPOL_OK, MPOL_EINVAL, MPOL_ERANGE, MPOL_EHIP, MPOL_ENOMEM, MPOL_ESTATE = range(6)
MPOL_FP64, MPOL_FP32 = 0, 1


class SolveOpts(ctypes.Structure):
    _fields_ = [("tol", ctypes.c_double), ("restart", ctypes.c_int), ("maxit", ctypes.c_int),
                ("mixed", ctypes.c_int), ("max_refine", ctypes.c_int), ("slab_factor", ctypes.c_double)]


class SolveInfo(ctypes.Structure):
    _fields_ = [("matvecs_fp64", ctypes.c_int64), ("matvecs_fp32", ctypes.c_int64),
                ("converged", ctypes.c_int32), ("refinements", ctypes.c_int32),
                ("max_relres", ctypes.c_double), ("matvecs_inner", ctypes.c_int64)]


class GeometrySummary(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("npairs", ctypes.c_int64), ("ngrid", ctypes.c_int32 * 3),
                ("P", ctypes.c_int32), ("ncell", ctypes.c_int32 * 3), ("tile_rows", ctypes.c_int32),
                ("fft_e", ctypes.c_int32 * 3)]


# (name, restype, argtypes) of every exported symbol
_VP = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)
_I32P = ctypes.POINTER(ctypes.c_int32)
SIGNATURES = [
    ("mpol_version", ctypes.c_int, []),
    ("mpol_last_error", ctypes.c_char_p, []),
    ("mpol_device_count", ctypes.c_int, [ctypes.POINTER(ctypes.c_int)]),
    ("mpol_create", ctypes.c_int, [ctypes.POINTER(_VP), ctypes.c_int, _VP]),
    ("mpol_destroy", ctypes.c_int, [_VP]),
    ("mpol_set_stream", ctypes.c_int, [_VP, _VP]),
    ("mpol_set_inner_operator", ctypes.c_int, [_VP, _VP, ctypes.c_double]),
    ("mpol_synchronize", ctypes.c_int, [_VP]),
    ("mpol_set_tables", ctypes.c_int, [_VP, _DP, _DP, _DP, ctypes.c_int64]),
    ("mpol_set_geometry", ctypes.c_int, [_VP, _VP, ctypes.c_int64, _DP, _I32P, ctypes.c_int32, ctypes.c_double,
                                         _DP, ctypes.c_double]),
    ("mpol_set_permittivity", ctypes.c_int, [_VP, _VP, ctypes.c_int32]),
    ("mpol_matvec", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_double]),
    ("mpol_matvec_timed", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_double,
                                         ctypes.c_int32, _DP]),
    ("mpol_debug_spectrum", ctypes.c_int, [_VP, _VP, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, _VP]),
    ("mpol_matvec_components", ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, ctypes.c_int32, ctypes.c_int32,
                                              ctypes.c_double]),
    ("mpol_solve", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int32, ctypes.POINTER(SolveOpts),
                                  ctypes.POINTER(SolveInfo)]),
    ("mpol_set_tuning", ctypes.c_int, [_VP, _I32P, ctypes.c_int32]),
    ("mpol_set_tuning_for", ctypes.c_int, [_VP, ctypes.c_int32, _I32P, ctypes.c_int32]),
    ("mpol_geometry_info", ctypes.c_int, [_VP, ctypes.POINTER(GeometrySummary)]),
    ("mpol_real_space_table", ctypes.c_int, [ctypes.c_double, _DP, ctypes.c_int64, _DP]),
    ("mpol_real_space_table_field", ctypes.c_int, [ctypes.c_double, _DP, ctypes.c_int64, _DP]),
    ("mpol_compute_forces", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int64, _DP, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int32, ctypes.c_int32, _VP, _VP]),
    ("mpol_compute_forces_noalloc", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int64, _DP, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_int32, ctypes.c_int32, _VP, _VP]),
    ("mpol_grid_kernel", ctypes.c_int, [_VP, _VP, ctypes.c_int64, _I32P, _DP, ctypes.c_double, _DP,
                                        ctypes.c_int32, _VP, _VP, _VP, _VP, _VP]),
    ("mpol_spread_precomp", ctypes.c_int, [_VP, _VP, ctypes.c_int64, ctypes.c_int32, _VP, _VP, _I32P, _VP]),
    ("mpol_contract_precomp", ctypes.c_int, [_VP, _VP, _I32P, _VP, _VP, ctypes.c_int64, ctypes.c_int32, _VP]),
    ("mpol_scale_kgrid", ctypes.c_int, [_VP, _VP, _VP, _I32P, ctypes.c_double, _DP, _VP]),
    ("mpol_kspace_table", ctypes.c_int, [_VP, _I32P, _DP, ctypes.c_double, _DP, _VP, _VP]),
    ("mpol_scale_precomp", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int64, _VP]),
    ("mpol_real_space_pairs", ctypes.c_int, [_VP, _VP, _DP, _VP, _VP, ctypes.c_int64, ctypes.c_double, _VP, _VP,
                                             _VP, ctypes.c_int64, _VP, _VP, _VP, _VP, _VP,
                                             ctypes.POINTER(ctypes.c_int64)]),
    ("mpol_real_space_precomp", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int64, _VP, _VP, _VP, _VP, _VP,
                                               ctypes.c_int64, ctypes.c_double, _VP]),
    ("mpol_real_space_direct", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int64, _DP, _VP, _VP, ctypes.c_int64,
                                              ctypes.c_double, _VP, _VP, _VP, ctypes.c_int64, _VP]),
    ("mpol_set_permittivity_complex", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32]),
    ("mpol_set_quad_tables", ctypes.c_int, [_VP, _DP, ctypes.c_int64]),
    ("mpol_quad_matvec", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, ctypes.c_int32]),
    ("mpol_quad_matvec_components", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int32, ctypes.c_int32]),
    ("mpol_quad_solve", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int32, ctypes.POINTER(SolveOpts),
                                       ctypes.POINTER(SolveInfo)]),
    ("mpol_real_space_table_quad", ctypes.c_int, [ctypes.c_double, _DP, ctypes.c_int64, _DP]),
    ("mpol_set_sizes", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, _DP, _DP]),
    ("mpol_sizes_matvec", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, ctypes.c_int32]),
    ("mpol_sizes_matvec_components", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int32, ctypes.c_int32]),
    ("mpol_sizes_solve", ctypes.c_int, [_VP, _VP, _VP, _VP, ctypes.c_int32, ctypes.POINTER(SolveOpts),
                                        ctypes.POINTER(SolveInfo)]),
    ("mpol_matvec_vjp", ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int32, _VP, _VP]),
    ("mpol_ewald_forces", ctypes.c_int, [_VP, _VP, ctypes.c_int32, _DP, _DP, _VP]),
    ("mpol_real_space_table_sizes", ctypes.c_int, [ctypes.c_double, _DP, ctypes.c_int64, _DP, ctypes.c_int32, _DP,
                                                   _DP]),
    ("mpol_real_space_table_sizes_field", ctypes.c_int, [ctypes.c_double, _DP, ctypes.c_int64, _DP,
                                                         ctypes.c_int32, _DP, _DP]),
    ("mpol_real_space_table_quad_field", ctypes.c_int, [ctypes.c_double, _DP, ctypes.c_int64, _DP]),
]

_lib = None


def load():
    """Load libmpol.so once; raise ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mpol_hip.build` "
                          "(there is no CPU fallback for the MI355X operator)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class MpolError(RuntimeError):
    pass


def check(status: int) -> None:
    """Map a C status to the exception types of the reference."""
    if status == MPOL_OK:
        return
    msg = load().mpol_last_error().decode(errors="replace")
    if status in (MPOL_EINVAL, MPOL_ERANGE):
        raise ValueError(msg)
    if status == MPOL_ENOMEM:
        raise MemoryError(msg)
    raise MpolError(msg)


def dptr(a):
    """ctypes double* of a contiguous float64 numpy array (host)."""
    return a.ctypes.data_as(_DP)


def i32ptr(a):
    return a.ctypes.data_as(_I32P)

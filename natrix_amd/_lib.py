"""ctypes binding of ``libnatrix_hip.so`` (the C-ABI in include/natrix.h).

The library is built in-tree by ``python -m natrix_amd.build``.  There is no
fallback: if the shared library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnatrix_hip.so")

NX_OK = 0
NX_ERR_ARG = -1
NX_ERR_HIP = -2
NX_ERR_STATE = -3

FIELD_VELOCITY = 0
FIELD_PRESSURE = 1
FIELD_DIVERGENCE = 2
FIELD_CURL = 3
FIELD_MASK = 4
FIELD_DYE = 5
FIELD_PRESSURE32 = 6

STORAGE_F32 = 0
STORAGE_BF16 = 1
STORAGE_F16 = 2

BC_REFERENCE = 0
BC_FREE_SLIP = 1
BC_INFLOW = 2

TIME_TAGS = ("advect", "vort_div", "jacobi", "project", "dye", "splat", "chebyshev")


class Params(C.Structure):
    _fields_ = [
        ("speed", C.c_float),
        ("dissipation", C.c_float),
        ("vorticity", C.c_float),
        ("viscosity", C.c_float),
        ("iterations", C.c_int),
        ("has_borders", C.c_int),
        ("bc_mode", C.c_int),
        ("inflow", C.c_float),
        ("dye_speed", C.c_float),
        ("dye_dissipation", C.c_float),
        ("dye_in_step", C.c_int),
        ("fused", C.c_int),
        ("jacobi_k", C.c_int),
        ("jacobi_rows", C.c_int),
        ("jacobi_waves", C.c_int),
        ("jacobi_cols", C.c_int),
        ("tol", C.c_double),
        ("max_iters", C.c_int),
        ("check_every", C.c_int),
        ("correction_sweeps", C.c_int),
        ("jacobi_seg", C.c_int),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("sweeps", C.c_int),
        ("converged", C.c_int),
        ("residual", C.c_double),
        ("residual_sweeps", C.c_int),
    ]


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double
_SZ = C.c_size_t

# name -> argtypes (every function returns int unless listed in _RESTYPE)
SIGNATURES = {
    "nx_create": [_I, _I, C.POINTER(_P)],
    "nx_create_ex": [_I, _I, _I, C.POINTER(_P)],
    "nx_destroy": [_P],
    "nx_default_params": [C.POINTER(Params)],
    "nx_get_params": [_P, C.POINTER(Params)],
    "nx_set_params": [_P, C.POINTER(Params)],
    "nx_check_params": [C.POINTER(Params)],
    "nx_enable_dye": [_P, _I, _I],
    "nx_add_velocity": [_P, _F, _F, _F, _F, _F],
    "nx_add_circle_obstacle": [_P, _F, _F, _F, _I],
    "nx_add_triangle_obstacle": [_P, _F, _F, _F, _F, _F, _F, _I],
    "nx_add_dye": [_P, _F, _F, _F, _F],
    "nx_dye_rgba8": [_P, _P],
    "nx_quiver_rgba8": [_P, _I, _I, _F, _P],
    "nx_add_gaussian_splats": [_P, C.POINTER(_F), _I],
    "nx_set_obstacle_map": [_P, C.POINTER(C.c_uint8)],
    "nx_update": [_P, _F],
    "nx_set_graph": [_P, _I],
    "nx_stage_boundaries": [_P],
    "nx_stage_advect": [_P, _F],
    "nx_stage_curl": [_P],
    "nx_stage_confine": [_P, _F],
    "nx_stage_viscosity": [_P],
    "nx_stage_divergence": [_P],
    "nx_stage_clear_pressure": [_P],
    "nx_stage_jacobi": [_P, _I],
    "nx_stage_subtract_gradient": [_P],
    "nx_stage_clear_obstacles": [_P],
    "nx_stage_advect_dye": [_P, _F],
    "nx_stage_fused_vort_div": [_P, _F],
    "nx_jacobi_blocked": [_P, _I, _I, _I, _I, _I, _I],
    "nx_upload": [_P, _I, _P, _SZ],
    "nx_download": [_P, _I, _P, _SZ],
    "nx_field_ptr": [_P, _I, C.POINTER(_P), C.POINTER(_SZ)],
    "nx_residual": [_P, C.POINTER(_D)],
    "nx_get_stats": [_P, C.POINTER(Stats)],
    "nx_synchronize": [_P],
    "nx_get_stream": [_P, C.POINTER(_P)],
    "nx_set_timing": [_P, _I],
    "nx_collect_timing": [_P, C.POINTER(_D), C.POINTER(_I)],
    "nx_batch_create": [_I, C.POINTER(_I), C.POINTER(_I), _I, C.POINTER(_P)],
    "nx_batch_destroy": [_P],
    "nx_batch_set_params": [_P, C.POINTER(Params)],
    "nx_batch_set_scene_params": [_P, _I, C.POINTER(Params)],
    "nx_batch_set_obstacle_map": [_P, _I, C.POINTER(C.c_uint8)],
    "nx_batch_add_circle_obstacle": [_P, _I, _F, _F, _F, _I],
    "nx_batch_add_triangle_obstacle": [_P, _I, _F, _F, _F, _F, _F, _F, _I],
    "nx_batch_set_step_splats": [_P, C.POINTER(_F), _I],
    "nx_batch_update": [_P, _F],
    "nx_batch_set_graph": [_P, _I],
    "nx_batch_jacobi_blocked": [_P, _I, _I, _I, _I, _I],
    "nx_batch_upload": [_P, _I, _I, _P, _SZ],
    "nx_batch_download": [_P, _I, _I, _P, _SZ],
    "nx_batch_cells": [_P, C.POINTER(C.c_longlong)],
    "nx_batch_residuals": [_P, C.POINTER(_D)],
    "nx_batch_synchronize": [_P],
    "nx_batch_set_timing": [_P, _I],
    "nx_batch_collect_timing": [_P, C.POINTER(_D), C.POINTER(_I)],
    "nx_step_begin": [_P, _F],
    "nx_solve": [_P, _D, _I, C.POINTER(Stats)],
    "nx_restrict_residual": [_P, _I, _P, _P],
    "nx_restrict_residual_faces": [_P, _I, _P, _P, _P],
    "nx_prolong_add": [_P, _I, _P, _P, _F],
    "nx_prolong_line_search": [_P, _I, _P, _P, _P],
    "nx_line_search": [_P, _P, _P],
    "nx_damped_sweeps": [_P, _I, _F],
    "nx_smooth": [_P, _I, _F, _P],
    "nx_smooth_chebyshev": [_P, _I, _F, _P, _P],
    "nx_set_smoother_tile": [_P, C.POINTER(_I), _I],
    "nx_get_smoother_tile": [_P, C.POINTER(_I)],
    "nx_coarse_smooth": [_P, _I, _I, _P, _P, _P, _I, _F],
    "nx_coarse_smooth_faces": [_P, _I, _I, _P, _P, _P, _P, _I, _F],
    "nx_refine_begin": [_P],
    "nx_refine_residual": [_P, C.POINTER(_D)],
    "nx_refine_accumulate": [_P],
    "nx_refine_end": [_P],
    "nx_step_finish": [_P],
    "nx_conv3x3_fwd": [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _P],
    "nx_conv3x3_fwd_split": [_P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _I, _P],
    "nx_unet_create": [_I, _I, _I, _P, _P, _P, _P, _P],
    "nx_unet_destroy": [_P],
    "nx_unet_forward": [_P, _P, _P, _P],
    "nx_learned_correct": [_P, _P, _I, _I, _F, _P],
    # typed pointers: ctypes turns a Python int given for a c_void_p into an
    # address, so a misplaced integer would reach the library as a pointer
    "nx_learned_cycles": [_P, _P, _I, _I, _F, _I, _F, C.POINTER(_I), _D, _D, _I,
                          C.POINTER(_D), _I, C.POINTER(_I)],
    "nx_unet_coarse": [_P, _P, _P, _P],
    "nx_unet_set_bypass": [_P, _I],
    "nx_unet_set_levels": [_P, _I, _I, _F],
    "nx_get_size": [_P, _P, _P],
    "nx_conv3x3_dgrad": [_P, _P, _P, _P, _I, _I, _I, _I, _I, _P],
    "nx_conv3x3_wgrad": [_P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _P],
    "nx_relu_bwd": [_P, _P, _P, C.c_longlong, _P],
    "nx_colsum": [_P, _P, _P, C.c_longlong, _I, _P],
    "nx_avgpool2": [_P, _P, _I, _I, _I, _I, _I, _P],
    "nx_upsample2": [_P, _P, _I, _I, _I, _I, _I, _P],
    "nx_concat": [_P, _P, _P, C.c_longlong, _I, _I, _I, _P],
    "nx_head_fwd": [_P, _P, _P, _P, C.c_longlong, _I, _P],
    "nx_head_bwd": [_P, _P, _P, _P, _P, _P, C.c_longlong, _I, _P],
    "nx_fp64_solve": [_P, _P, _I, _I, _P, _D, _I, C.POINTER(_D), _P],
    "nx_fp64_solve_noalloc": [_P, _P, _I, _I, _P, _D, _I, C.POINTER(_D), _P],
    "nx_fp64_compare": [_P, _P, _P, _I, _I, C.POINTER(_D), _P],
    "nx_fp64_solve_batch": [_P, _P, _I, _I, _P, _P, _D, _I, _P, _P, _P],
    "nx_copy_bandwidth": [_SZ, _I, C.POINTER(_D)],
    "nx_copy_bandwidth_variant": [_SZ, _I, _I, C.POINTER(_D)],
    "nx_last_error": [],
}
_RESTYPE = {"nx_last_error": C.c_char_p}

_lock = threading.Lock()
_lib = None


class NatrixError(RuntimeError):
    """A HIP or state failure reported by the native library."""


def lib():
    """Load the native library once; raise if it has not been built."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise NatrixError(
                    f"{LIB_PATH} is missing: build it with "
                    "`python -m natrix_amd.build` (there is no CPU fallback)")
            h = C.CDLL(LIB_PATH)
            for name, args in SIGNATURES.items():
                fn = getattr(h, name)
                fn.argtypes = args
                fn.restype = _RESTYPE.get(name, C.c_int)
            _lib = h
        return _lib


def last_error():
    return lib().nx_last_error().decode("utf-8", "replace")


def check(rc):
    """Map a status code to the exception the reference would raise."""
    if rc == NX_OK:
        return
    msg = last_error()
    if rc == NX_ERR_ARG:
        raise ValueError(msg)
    raise NatrixError(f"natrix error {rc}: {msg}")


def call(name, *args):
    check(getattr(lib(), name)(*args))

"""ctypes binding of libtreams_gpu.so (C ABI in include/treams_gpu.h).

The library must be built (treams-rs_amd/build.py); there is no fallback: every
compute call goes to the HIP library and fails loudly if it is missing.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtreams_gpu.so")

TG_OK = 0
TG_INVALID_INPUT = 1
TG_SPECIAL_FUNCTION = 2
TG_NON_FINITE = 3
TG_NOT_CONVERGED = 4
TG_SINGULAR = 5
TG_OUT_OF_MEMORY = 6
TG_DEVICE_ERROR = 7

TG_STORE_NONE, TG_STORE_C64, TG_STORE_C128 = 0, 1, 2
TG_OP_MATRIX_FREE_F64, TG_OP_STORED, TG_OP_MATRIX_FREE_C64, TG_OP_MATRIX_FREE_C64_HI = 0, 1, 2, 3
TG_PC_NONE, TG_PC_BLOCK_JACOBI, TG_PC_LEARNED = 0, 1, 2

#: every symbol include/treams_gpu.h declares
EXPORTED = [
    "tg_last_error", "tg_modes_per_sphere", "tg_plan_terms", "tg_wigner3j", "tg_rotation_matrix", "tg_cluster_create", "tg_cluster_create_particles",
    "tg_cluster_destroy", "tg_cluster_dimension", "tg_cluster_mie_blocks", "tg_pair_matrix",
    "tg_translation_blocks", "tg_translation_blocks_noalloc", "tg_cluster_assemble", "tg_cluster_release", "tg_cluster_reset", "tg_cluster_apply", "tg_cluster_coupling",
    "tg_cluster_local", "tg_cluster_overlap", "tg_cluster_plane_wave", "tg_cluster_solve", "tg_cluster_cross_sections",
    "tg_cluster_pullback", "tg_cluster_overlap_pullback", "tg_cluster_timing_enable",
    "tg_cluster_timing_get", "tg_cluster_timing_reset", "tg_cluster_set_group_size",
    "tg_cluster_set_table_limit", "tg_cluster_set_inner_gemm", "tg_cluster_set_inner_relaxation",
    "tg_cluster_near_pairs",
    "tg_cluster_pair_list", "tg_cluster_pair_features", "tg_cluster_set_pair_output", "tg_cluster_learned_apply",
    "tg_cluster_learned_loss", "tg_cluster_set_mlp", "tg_mlp_create", "tg_mlp_destroy", "tg_mlp_param_count",
    "tg_mlp_get_params", "tg_mlp_set_params", "tg_mlp_get_grads", "tg_mlp_zero_grad", "tg_mlp_forward",
    "tg_mlp_backward", "tg_mlp_adam_step", "tg_factor_create", "tg_factor_destroy", "tg_factor_solve",
    "tg_factor_pullback", "tg_factor_equilibrated", "tg_factor_global_tmatrix", "tg_cluster_field",
    "tg_plane_wave_field", "tg_expansion_field", "tg_translation_blocks_rect", "tg_translation_blocks_tangent", "tg_cluster_pushforward",
    "tg_factor_pushforward", "tg_factor_info", "tg_dense_factor_create", "tg_dense_factor_solve",
    "tg_cluster_pushforward_local", "tg_factor_pushforward_local", "tg_cluster_pullback_local",
    "tg_factor_pullback_local",
]


class GmresOptions(ctypes.Structure):
    """linalg/gmres.rs:22-43 plus precision / preconditioner (include/treams_gpu.h)."""

    _fields_ = [
        ("rtol", ctypes.c_double),
        ("atol", ctypes.c_double),
        ("restart", ctypes.c_int),
        ("max_iterations", ctypes.c_int),
        ("mixed_precision", ctypes.c_int),
        ("preconditioner", ctypes.c_int),
    ]


class Convergence(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int),
        ("refinements", ctypes.c_int),
        ("residual_norm", ctypes.c_double),
        ("rhs_norm", ctypes.c_double),
    ]


class NotConverged(ValueError):
    """GMRES did not reach its tolerance (Error::NotConverged, lib.rs:379-411)."""


_lib = None

_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_i64 = ctypes.c_int64


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libtreams_gpu.so is not built ({LIB_PATH}); run treams-rs_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "tg_last_error": (ctypes.c_char_p, []),
        "tg_modes_per_sphere": (_i, [_i]),
        "tg_plan_terms": (_i, [_i, _vp, _vp, _vp, _vp, ctypes.POINTER(_i64)]),
        "tg_wigner3j": (_d, [_i] * 6),
        "tg_rotation_matrix": (_i, [_i, _d, _d, _d, _vp]),
        "tg_cluster_create": (_i, [_i, _d, _i, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
        "tg_cluster_create_particles": (_i, [_i, _d, _i, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
        "tg_cluster_destroy": (_i, [_vp]),
        "tg_cluster_dimension": (_i64, [_vp]),
        "tg_cluster_mie_blocks": (_i, [_vp, _vp]),
        "tg_pair_matrix": (_i, [_vp, _i, _vp]),
        "tg_translation_blocks": (_i, [_i, _d, _i64, _vp, _i, _vp, _vp]),
        "tg_translation_blocks_noalloc": (_i, [_i, _d, _i64, _vp, _i, _vp, _vp]),
        "tg_cluster_assemble": (_i, [_vp, _i]),
        "tg_cluster_release": (_i, [_vp]),
        "tg_cluster_reset": (_i, [_vp]),
        "tg_cluster_set_table_limit": (_i, [_vp, _i64]),
        "tg_cluster_set_inner_gemm": (_i, [_vp, _i]),
        "tg_cluster_set_inner_relaxation": (_i, [_vp, _d]),
        "tg_cluster_apply": (_i, [_vp, _vp, _vp, _i, _i, _i]),
        "tg_cluster_coupling": (_i, [_vp, _vp, _vp, _i, _i, _i]),
        "tg_cluster_local": (_i, [_vp, _vp, _vp, _i, _i]),
        "tg_cluster_overlap": (_i, [_vp, _vp, _vp, _i, _i]),
        "tg_cluster_plane_wave": (_i, [_vp, _vp, _i, _vp]),
        "tg_cluster_solve": (_i, [_vp, _vp, _i, ctypes.POINTER(GmresOptions), _i, _vp, _vp]),
        "tg_cluster_cross_sections": (_i, [_vp, _vp, _vp, _d, ctypes.POINTER(_d), ctypes.POINTER(_d)]),
        "tg_cluster_pullback": (_i, [_vp, _vp, _vp, _vp, _i, ctypes.POINTER(GmresOptions), _vp, _vp, _vp, _vp,
                                     _vp, _vp]),
        "tg_cluster_overlap_pullback": (_i, [_vp, _vp, _vp, _i, _vp]),
        "tg_cluster_timing_enable": (_i, [_vp, _i]),
        "tg_cluster_timing_get": (_i, [_vp, ctypes.c_char_p, ctypes.POINTER(_d), ctypes.POINTER(_i64)]),
        "tg_cluster_timing_reset": (_i, [_vp]),
        "tg_cluster_set_group_size": (_i, [_vp, _i]),
        "tg_cluster_near_pairs": (_i, [_vp, _i, ctypes.POINTER(_i64)]),
        "tg_cluster_pair_list": (_i, [_vp, _vp]),
        "tg_cluster_pair_features": (_i, [_vp, _vp]),
        "tg_cluster_set_pair_output": (_i, [_vp, _vp]),
        "tg_cluster_learned_apply": (_i, [_vp, _vp, _vp, _i, _i]),
        "tg_cluster_learned_loss": (_i, [_vp, _vp, _i, ctypes.POINTER(_d), _vp]),
        "tg_cluster_set_mlp": (_i, [_vp, _vp]),
        "tg_mlp_create": (_i, [_i, _i, _i, _i, ctypes.c_uint64, _vp, ctypes.POINTER(_vp)]),
        "tg_mlp_destroy": (_i, [_vp]),
        "tg_mlp_param_count": (_i64, [_vp]),
        "tg_mlp_get_params": (_i, [_vp, _vp]),
        "tg_mlp_set_params": (_i, [_vp, _vp]),
        "tg_mlp_get_grads": (_i, [_vp, _vp]),
        "tg_mlp_zero_grad": (_i, [_vp]),
        "tg_mlp_forward": (_i, [_vp, _vp, _i64, _vp]),
        "tg_mlp_backward": (_i, [_vp, _vp]),
        "tg_mlp_adam_step": (_i, [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _i]),
        "tg_factor_create": (_i, [_vp, _i, ctypes.POINTER(_vp)]),
        "tg_factor_destroy": (_i, [_vp]),
        "tg_factor_solve": (_i, [_vp, _vp, _i, _d, _i, _vp, _vp]),
        "tg_factor_equilibrated": (_i, [_vp]),
        "tg_factor_global_tmatrix": (_i, [_vp, _i, _d, _vp, _vp]),
        "tg_translation_blocks_rect": (_i, [_i, _i, _d, _i64, _vp, _i, _vp, _vp]),
        "tg_cluster_field": (_i, [_vp, _vp, _i, _i64, _vp, _i, _vp]),
        "tg_plane_wave_field": (_i, [_d, _vp, _i, _i64, _vp, _vp, _vp]),
        "tg_expansion_field": (_i, [_i, _d, _vp, _vp, _i, _i64, _vp, _i, _vp, _vp]),
        "tg_factor_pullback": (_i, [_vp, _vp, _vp, _vp, _i, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
        "tg_translation_blocks_tangent": (_i, [_i, _d, _i64, _vp, _vp, _d, _i, _vp, _vp]),
        "tg_cluster_pushforward": (_i, [_vp, _vp, _vp, _i, ctypes.POINTER(GmresOptions), _d, _vp, _vp, _vp, _vp,
                                        _vp, _vp]),
        "tg_factor_pushforward": (_i, [_vp, _vp, _vp, _i, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
        "tg_factor_info": (_i, [_vp, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i)]),
        "tg_dense_factor_create": (_i, [_i64, _vp, _i, _vp, ctypes.POINTER(_vp)]),
        "tg_dense_factor_solve": (_i, [_vp, _vp, _i, _d, _i, _vp, _vp]),
        "tg_cluster_pushforward_local": (_i, [_vp, _vp, _vp, _i, ctypes.POINTER(GmresOptions), _d, _vp, _vp, _vp,
                                              _vp, _vp]),
        "tg_factor_pushforward_local": (_i, [_vp, _vp, _vp, _i, _d, _d, _vp, _vp, _vp, _vp, _vp]),
        "tg_cluster_pullback_local": (_i, [_vp, _vp, _vp, _vp, _i, ctypes.POINTER(GmresOptions), _vp, _vp, _vp,
                                           _vp, _vp]),
        "tg_factor_pullback_local": (_i, [_vp, _vp, _vp, _vp, _i, _d, _vp, _vp, _vp, _vp, _vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(status: int) -> None:
    """Map a tg_status to the reference's Python exceptions (treams-py context.rs:69-74)."""
    if status == TG_OK:
        return
    msg = lib().tg_last_error().decode(errors="replace")
    if status == TG_OUT_OF_MEMORY:
        raise MemoryError(msg)
    if status == TG_NOT_CONVERGED:
        raise NotConverged(msg)
    if status == TG_DEVICE_ERROR:
        raise RuntimeError(msg)
    raise ValueError(msg)

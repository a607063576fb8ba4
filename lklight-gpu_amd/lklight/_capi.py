"""ctypes binding of liblklight_gpu.so (include/lklight_gpu.h).

Loading fails loudly: there is no silent substitute for the HIP library.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblklight_gpu.so")

F_FAR, F_ELEC, F_CLASH, F_DESOLV = 1, 2, 4, 16
FLAGS_DNA, FLAGS_PYDOCK, FLAGS_VDW = 7, 3, 0

_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_u8 = C.POINTER(C.c_ubyte)
_i = C.POINTER(C.c_int)
_u16 = C.POINTER(C.c_ushort)
_u32 = C.POINTER(C.c_uint)


SCORE_FALLBACK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.c_int,
                             C.POINTER(C.c_double))


class ComplexDesc(C.Structure):
    """lkl_complex_desc."""
    _fields_ = [
        ("phi", _f), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("ox", C.c_double), ("oy", C.c_double), ("oz", C.c_double), ("sp", C.c_double),
        ("nr", C.c_int), ("r_xyz", _f), ("r_ele", _f), ("r_svdw", _f), ("r_vdwr", _f),
        ("r_heavy", _u8),
        ("nl", C.c_int), ("l_xyz", _f), ("l_ele", _f), ("l_svdw", _f), ("l_vdwr", _f),
        ("l_heavy", _u8),
        ("r_flag", _u8), ("r_des", _f), ("r_asa", _f),
        ("l_flag", _u8), ("l_des", _f), ("l_asa", _f),
        ("cell_size", C.c_float),
    ]


class LibraryMissing(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded library (raises LibraryMissing if it was never built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LibraryMissing(
            f"{LIB_PATH} is missing: build it with `python lklight-gpu_amd/build.py` "
            "(there is no CPU substitute for the HIP path)")
    # PyTorch (device memory and streams for callers) ships its own HIP runtime
    # with the same soname as /opt/rocm's.  Loading torch first lets this
    # library bind to that runtime, so the process holds ONE HIP runtime; the
    # other order leaves two, and the second cannot open the device.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    sig = {
        "lkl_device_available": (C.c_int, []),
        "lkl_last_error": (C.c_char_p, []),
        "lkl_cache_clear": (None, []),
        "lkl_ctx_create": (C.c_void_p, [C.POINTER(ComplexDesc)]),
        "lkl_ctx_destroy": (None, [C.c_void_p]),
        "lkl_ctx_score": (C.c_int, [C.c_void_p, _d, C.c_int, C.c_uint, _d, _d]),
        "lkl_ctx_score_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint,
                                          C.c_void_p, C.c_void_p, C.c_void_p]),
        "lkl_ctx_score_async_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_uint,
                                                  C.c_void_p, C.c_void_p, C.c_void_p]),
        "lkl_ctx_transform": (C.c_int, [C.c_void_p, _d, C.c_int, _d]),
        "lkl_ctx_set_tiles": (C.c_int, [C.c_void_p, C.c_float, C.c_int, C.c_int]),
        "lkl_ctx_autotune": (C.c_int, [C.c_void_p, _d, C.c_int, C.c_uint, _f, _f, _i, _i, _f, _i]),
        "lkl_ctx_set_cell_x": (C.c_int, [C.c_void_p, C.c_float]),
        "lkl_ctx_set_work_sort": (C.c_int, [C.c_void_p, C.c_int]),
        "lkl_ctx_set_list_spacing": (C.c_int, [C.c_void_p, C.c_float]),
        "lkl_ctx_set_phi_packed": (C.c_int, [C.c_void_p, C.c_int]),
        "lkl_ctx_set_interface": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_int, C.c_void_p, C.c_int]),
        "lkl_ctx_has_interface": (C.c_int, [C.c_void_p]),
        "lkl_ctx_interface_sizes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "lkl_ctx_interface": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
        "lkl_ctx_interface_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
        "lkl_ctx_phi_packed": (C.c_int, [C.c_void_p]),
        "lkl_ctx_set_param_types": (C.c_int, [C.c_void_p, C.c_int]),
        "lkl_ctx_param_types": (C.c_int, [C.c_void_p]),
        "lkl_ctx_list_info": (C.c_longlong, [C.c_void_p, C.c_void_p]),
        "lkl_batch_score": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                      C.c_float, C.c_float, _f, _f, _f, _f, _u8, C.c_int,
                                      _i, _i, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                      C.c_float, C.c_float, _f, _d, _f, _f, _f, _u8, C.c_int,
                                      C.c_int, C.c_uint, _d]),
        "lkl_cpydock_solv": (C.c_int, [_f, _u8, C.c_int, _i, _i, C.c_int, C.c_int, C.c_int,
                                       C.c_float, C.c_float, C.c_float, C.c_float, _f, _d, _u8,
                                       _f, C.c_int, C.c_int, _f, _f, _f, _f, _d]),
        "lkl_full_score": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                     C.c_float, C.c_float, _f, _f, _f, _f, _u8, C.c_int, _i,
                                     _i, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                     C.c_float, C.c_float, _f, _f, _f, _f, _u8, C.c_int, _f,
                                     _f]),
        "lkl_far_field": (C.c_int, [_f, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                                    C.c_float, C.c_float, _f, _f, C.c_int, _d]),
        "lkl_field_dims": (C.c_longlong, [_f, C.c_int, _d, _i]),
        "lkl_field_build": (C.c_int, [_f, _f, C.c_int, _f]),
        "lkl_lr_sasa": (C.c_int, [_d, _d, C.c_int, C.c_double, C.c_int, _d]),
        "lkl_ctx_score_coords": (C.c_int, [C.c_void_p, _f, C.c_int, C.c_uint, _d, _d]),
        "lkl_dataset_poses": (C.c_int, [C.c_uint64, C.c_int, _d]),
        "lkl_ctx_set_pose_order": (C.c_int, [C.c_void_p, C.c_int]),
        "lkl_ctx_pose_order": (C.c_int, [C.c_void_p]),
        "lkl_ctx_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
        "lkl_ctx_kernel_ms": (C.c_int, [C.c_void_p, _d, _d]),
        "lkl_synth_rsa": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_uint64, C.c_int64, _d, C.POINTER(C.c_int64)]),
        "lkl_ctx_build_field": (C.c_int, [C.c_void_p]),
        "lkl_ctx_field_info": (C.c_int, [C.c_void_p, _d, _i]),
        "lkl_multi_create": (C.c_void_p, [C.POINTER(C.c_void_p), C.POINTER(C.c_uint), _i, C.c_int]),
        "lkl_multi_destroy": (None, [C.c_void_p]),
        "lkl_multi_poses": (C.c_int, [C.c_void_p]),
        "lkl_multi_set_tiles": (C.c_int, [C.c_void_p, C.c_int, C.c_int]),
        "lkl_multi_score": (C.c_int, [C.c_void_p, _d, _d]),
        "lkl_multi_score_async": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
        "lkl_multi_use_graph": (C.c_int, [C.c_void_p, C.c_int]),
        "lkl_multi_set_trace": (C.c_int, [C.c_void_p, C.c_int]),
        "lkl_multi_trace": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.c_int, _i]),
        "lkl_multi_autotune": (C.c_int, [C.c_void_p, _d, _i, _i, _f, _i]),
        "lkl_gso_create": (C.c_void_p, [C.c_void_p, C.c_int, C.c_int, _i, _d, C.c_uint64,
                                        SCORE_FALLBACK, C.c_void_p]),
        "lkl_gso_destroy": (None, [C.c_void_p]),
        "lkl_gso_update_luciferin": (C.c_int, [C.c_void_p]),
        "lkl_gso_movement_phase": (C.c_int, [C.c_void_p]),
        "lkl_gso_run": (C.c_int, [C.c_void_p, C.c_int]),
        "lkl_gso_device_run": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p]),
        "lkl_gso_size": (C.c_int, [C.c_void_p]),
        "lkl_gso_state": (C.c_int, [C.c_void_p, _d, _d, _d, _d, _i, _u8]),
        "lkl_gso_set_poses": (C.c_int, [C.c_void_p, _d, _u8]),
        "lkl_gso_stats": (C.c_int, [C.c_void_p, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong),
                                    C.POINTER(C.c_longlong)]),
        "lkl_gso_save": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_char_p]),
        "lkl_gso_plan": (C.c_int, [C.c_void_p, _d]),
        "lkl_gso_commit": (C.c_int, [C.c_void_p, _d, _d, _u8]),
        "lkl_feat_create": (C.c_void_p, [_f, _u8, C.c_int, _f, _u8, C.c_int]),
        "lkl_feat_create_grid": (C.c_void_p, [_f, _u8, C.c_int, _f, _u8, C.c_int, C.c_double]),
        "lkl_feat_grid_info": (C.c_longlong, [C.c_void_p, _d, _i, _d]),
        "lkl_feat_grid_read": (C.c_int, [C.c_void_p, C.c_longlong, C.c_longlong, _u8]),
        "lkl_feat_destroy": (None, [C.c_void_p]),
        "lkl_feat_compute": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
        "lkl_feat_compute_host": (C.c_int, [C.c_void_p, _d, C.c_int, _u16, _u32]),
        "lkl_mlp_param_count": (C.c_longlong, []),
        "lkl_mlp_create": (C.c_void_p, [C.c_uint64]),
        "lkl_mlp_destroy": (None, [C.c_void_p]),
        "lkl_mlp_get_params": (C.c_int, [C.c_void_p, _f]),
        "lkl_mlp_set_params": (C.c_int, [C.c_void_p, _f]),
        "lkl_mlp_get_grads": (C.c_int, [C.c_void_p, _f]),
        "lkl_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
        "lkl_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, _f]),
        "lkl_mlp_adam": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float]),
        "lkl_mlp_sync": (C.c_int, [C.c_void_p]),
        "lkl_mlp_forward_host": (C.c_int, [C.c_void_p, _u16, C.c_int, _f]),
        "lkl_mlp_grad_host": (C.c_int, [C.c_void_p, _u16, _f, C.c_int, _f]),
        "lkl_gemm_bf16_host": (C.c_int, [_u16, C.c_long, C.c_long, C.c_long, _u16, C.c_long,
                                         C.c_long, C.c_long, C.c_int, C.c_int, C.c_int, _f]),
        "lkl_screen_candidates": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.c_double,
                                            C.c_uint64, C.c_void_p, C.c_void_p]),
        "lkl_screen_select": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                        C.c_void_p, C.c_void_p, C.c_void_p]),
        "lkl_rng_f64s": (None, [C.c_uint64, C.c_int, _d]),
        "lkl_quat_slerp": (None, [_d, _d, C.c_double, _d]),
        "lkl_quat_slerp_cr": (None, [_d, _d, C.c_double, _d]),
        "lkl_sin_cr": (C.c_double, [C.c_double]),
        "lkl_acos_cr": (C.c_double, [C.c_double]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def ptr(a, ctype):
    """Pointer to a contiguous numpy array (None passes NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))


def last_error() -> str:
    return lib().lkl_last_error().decode(errors="replace")


class GpuError(RuntimeError):
    pass


def check(rc: int, what: str):
    if rc != 0:
        raise GpuError(f"{what} failed: {last_error()}")


def as_f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def as_f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)

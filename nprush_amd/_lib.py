"""ctypes binding of libnprush_hip.so (include/nprush_hip.h).

The library is built in-tree by ``__graft_entry__.build()`` (nprush_amd/csrc/Makefile).
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

MODE_REF = 0
MODE_EXT = 1

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnprush_hip.so")


class NpsConfig(C.Structure):
    _fields_ = [
        ("max_tile_splats", C.c_int32),
        ("adc_period", C.c_int32),
        ("adc_enabled", C.c_int32),
        ("device", C.c_int32),
        ("mode", C.c_int32),
        ("sh_degree", C.c_int32),
        ("ext_tile", C.c_int32),
        ("ext_batch", C.c_int32),
        ("reserved", C.c_int32 * 8),
    ]


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_I = C.POINTER(C.c_int32)
_U = C.POINTER(C.c_uint32)
_L = C.POINTER(C.c_int64)
_D = C.POINTER(C.c_double)
i32 = C.c_int32

# name -> (restype, argtypes)
SIGNATURES = {
    "nps_config_default": (i32, [C.POINTER(NpsConfig)]),
    "nps_create": (i32, [C.POINTER(NpsConfig), i32, C.POINTER(_P)]),
    "nps_destroy": (None, [_P]),
    "nps_last_error": (C.c_char_p, [_P]),
    "nps_num_splats": (i32, [_P]),
    "nps_write_splats": (i32, [_P, _F, i32]),
    "nps_read_splats": (i32, [_P, _F, i32]),
    "nps_write_grads": (i32, [_P, _I, i32]),
    "nps_read_grads": (i32, [_P, _I, i32]),
    "nps_write_adam": (i32, [_P, _F, _F, C.c_float, i32]),
    "nps_read_adam": (i32, [_P, _F, _F, _F, i32]),
    "nps_write_accum": (i32, [_P, _F, i32]),
    "nps_read_accum": (i32, [_P, _F, i32]),
    "nps_write_sh": (i32, [_P, _F, i32]),
    "nps_read_sh": (i32, [_P, _F, i32]),
    "nps_write_grads_f32": (i32, [_P, _F, i32]),
    "nps_read_grads_f32": (i32, [_P, _F, i32]),
    "nps_set_camera": (i32, [_P, _F, _F, i32]),
    "nps_set_target": (i32, [_P, _F, _F, i32, i32]),
    "nps_set_target_dev": (i32, [_P, _P, _P, i32, i32]),
    "nps_backward": (i32, [_P]),
    "nps_forward": (i32, [_P]),
    "nps_autotune": (i32, [_P, _I, _I, _F]),
    "nps_step": (i32, [_P]),
    "nps_adc": (i32, [_P]),
    "nps_set_adc_ext": (i32, [_P, C.c_float, C.c_float, C.c_float, C.c_float]),
    "nps_dispatch": (i32, [_P]),
    "nps_sync": (i32, [_P]),
    "nps_read_image": (i32, [_P, _F, _F]),
    "nps_read_loss": (i32, [_P, _D]),
    "nps_read_tiles": (i32, [_P, _U, _U, _U, _I, _L]),
    "nps_read_stats": (i32, [_P, _L]),
    "nps_read_tile_work": (i32, [_P, _U]),
    "nps_enable_timing": (i32, [_P, i32]),
    "nps_read_timing": (i32, [_P, _F]),
    "npf_create": (i32, [_P, C.POINTER(_P)]),
    "npf_destroy": (None, [_P]),
    "npf_write_vp": (i32, [_P, _F]),
    "npf_set_target": (i32, [_P, i32, i32]),
    "npf_dispatch": (i32, [_P, i32, i32]),
    "npf_read_target": (i32, [_P, _F, _F]),
    "npf_write_target": (i32, [_P, _F, _F]),
    "npt_create": (i32, [C.POINTER(_P), i32, i32, C.POINTER(_P)]),
    "npt_destroy": (None, [_P]),
    "npt_last_error": (C.c_char_p, [_P]),
    "npt_add_view": (i32, [_P, i32, _F, _F, _P, i32, i32]),
    "npt_step": (i32, [_P, i32]),
    "npt_autotune": (i32, [_P]),
    "npt_read_tune": (i32, [_P, _I]),
    "npt_read_losses": (i32, [_P, _D]),
    "npt_graph_info": (i32, [_P, C.POINTER(C.c_int64), _F]),
    "npp_create": (i32, [i32, C.POINTER(_P)]),
    "npp_destroy": (None, [_P]),
    "npp_last_error": (C.c_char_p, [_P]),
    "npp_set_env": (i32, [_P, _F, i32, i32]),
    "npp_set_meshes": (i32, [_P, _F, i32, _U, i32]),
    "npp_set_output_size": (i32, [_P, i32, i32]),
    "npp_reset": (i32, [_P]),
    "npp_write_inv_vp": (i32, [_P, _F]),
    "npp_dispatch": (i32, [_P]),
    "npp_read_output": (i32, [_P, _F]),
    "npp_read_accum": (i32, [_P, _F]),
    "npp_output_dev": (i32, [_P, C.POINTER(_P)]),
    "npp_read_bvh": (i32, [_P, _U, _U, _I]),
    "npp_env_background": (i32, [_P, _F, _F, i32, i32]),
    "npp_raster": (i32, [_P, _F, _F, _F, C.c_float, _F, _F, i32, i32]),
    "npb_create": (i32, [i32, C.POINTER(_P)]),
    "npb_destroy": (None, [_P]),
    "npb_last_error": (C.c_char_p, [_P]),
    "npb_blur": (i32, [_P, _F, _F, i32, i32, i32, C.c_float, i32]),
    "npb_blur_dev": (i32, [_P, _P, _P, i32, i32, i32, C.c_float, i32]),
    "npb_blur_dev_noalloc": (i32, [_P, _P, _P, i32, i32, i32, C.c_float, i32]),
    "npd_compose": (i32, [i32, C.POINTER(_F), _I, _F, _F, i32, i32]),
    "nps_dev_alloc": (i32, [_P, C.c_int64, C.POINTER(_P)]),
    "nps_dev_free": (i32, [_P, _P]),
    "nps_dev_upload": (i32, [_P, _P, _P, C.c_int64]),
    "nps_debug_math": (i32, [i32, i32, _F, _F, _F, i32]),
    "nps_debug_depth_sort": (i32, [i32, C.POINTER(C.c_uint64), C.c_uint32, i32, C.POINTER(C.c_uint64), _U]),
    "nps_adc_features": (i32, [_P, _F, i32, _I]),
    "nps_adc_children": (i32, [_P, _I, _I, i32, _I]),
    "nps_adc_apply_deltas": (i32, [_P, _F, i32]),
    "nps_snapshot": (i32, [_P, i32]),
    "nps_adc_features_dev": (i32, [_P, C.POINTER(_P), _I]),
    "nps_adc_apply_deltas_dev": (i32, [_P, _P, i32]),
    "nml_create": (i32, [i32, i32, C.POINTER(_P)]),
    "nml_destroy": (None, [_P]),
    "nml_last_error": (C.c_char_p, [_P]),
    "nml_num_params": (i32, []),
    "nml_write_params": (i32, [_P, _F]),
    "nml_read_params": (i32, [_P, _F]),
    "nml_set_norm": (i32, [_P, _F, _F, _F, _F]),
    "nml_forward": (i32, [_P, _F, i32, _F]),
    "nml_forward_dev": (i32, [_P, _P, i32, _P]),
    "nml_grads": (i32, [_P, _F, _F, i32, _F, _D]),
    "nml_train_dev": (i32, [_P, _P, _P, i32, C.c_float, _D]),
    "nml_train": (i32, [_P, _F, _F, i32, C.c_float, _D]),
    "nml_enable_timing": (i32, [_P, i32]),
    "nml_read_timing": (i32, [_P, _F]),
    "nml_dev_alloc": (i32, [_P, C.c_int64, C.POINTER(_P)]),
    "nml_dev_free": (i32, [_P, _P]),
    "nml_dev_upload": (i32, [_P, _P, _P, C.c_int64]),
    "nbz_create": (i32, [i32, i32, C.POINTER(_P)]),
    "nbz_destroy": (None, [_P]),
    "nbz_last_error": (C.c_char_p, [_P]),
    "nbz_write_beziers": (i32, [_P, _F, i32]),
    "nbz_read_beziers": (i32, [_P, _F, i32]),
    "nbz_write_grads": (i32, [_P, _I, i32]),
    "nbz_read_grads": (i32, [_P, _I, i32]),
    "nbz_write_adam": (i32, [_P, _F, _F, C.c_float, i32]),
    "nbz_read_adam": (i32, [_P, _F, _F, _F, i32]),
    "nbz_write_adc": (i32, [_P, _F, _L, i32]),
    "nbz_read_adc": (i32, [_P, _F, _L, i32]),
    "nbz_write_pixel_loss": (i32, [_P, _I]),
    "nbz_read_pixel_loss": (i32, [_P, _I]),
    "nbz_write_vp_matrix": (i32, [_P, _F]),
    "nbz_write_vp_inv_matrix": (i32, [_P, _F]),
    "nbz_write_mode": (i32, [_P, C.c_float]),
    "nbz_write_bg_penalty": (i32, [_P, C.c_float]),
    "nbz_write_max_width": (i32, [_P, C.c_float]),
    "nbz_write_kill_thresholds": (i32, [_P, C.c_float, C.c_float]),
    "nbz_set_adc_period": (i32, [_P, i32]),
    "nbz_write_no_kill": (i32, [_P, i32]),
    "nbz_set_backward_target": (i32, [_P, _F, _F, _F, _F, _F, i32, i32]),
    "nbz_set_backward_target_dev": (i32, [_P, _P, _P, _P, _P, _P, i32, i32]),
    "nbz_dispatch": (i32, [_P]),
    "nbz_backward": (i32, [_P]),
    "nbz_step": (i32, [_P]),
    "nbz_adc": (i32, [_P]),
    "nbz_forward": (i32, [_P]),
    "nbz_sync": (i32, [_P]),
    "nbz_read_image": (i32, [_P, _F, _F, _F]),
    "nbz_read_loss": (i32, [_P, _D]),
    "nbz_read_tiles": (i32, [_P, _U, _U, _I, _L]),
    "nbz_read_stats": (i32, [_P, _L]),
    "nbz_dev_alloc": (i32, [_P, C.c_int64, C.POINTER(_P)]),
    "nbz_dev_free": (i32, [_P, _P]),
    "nbz_dev_upload": (i32, [_P, _P, _P, C.c_int64]),
    "nbf_create": (i32, [_P, _P, i32, i32, C.POINTER(_P)]),
    "nbf_destroy": (None, [_P]),
    "nbf_write_vp": (i32, [_P, _F]),
    "nbf_set_target": (i32, [_P, i32, i32]),
    "nbf_dispatch": (i32, [_P, i32, i32]),
    "nbf_read_target": (i32, [_P, _F]),
    "nbf_write_target": (i32, [_P, _F]),
    "nbz_enable_timing": (i32, [_P, i32]),
    "nbz_read_timing": (i32, [_P, _F]),
}

_lib = None


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() (no CPU fallback exists)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_F)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_I)


def uptr(a: np.ndarray):
    assert a.dtype == np.uint32 and a.flags.c_contiguous
    return a.ctypes.data_as(_U)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


def lptr(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_L)


class NpsError(RuntimeError):
    pass


def check(rc: int, handle=None, what: str = ""):
    if rc != 0:
        msg = ""
        if handle is not None:
            m = load().nps_last_error(handle)
            msg = m.decode() if m else ""
        raise NpsError(f"{what} failed ({rc}): {msg}")

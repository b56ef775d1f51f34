"""ctypes binding of libksj_gpu.so (the C ABI declared in include/ksj_gpu.h).

The library is required: importing an op whose library is missing raises, so no
call can silently fall back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
GPU_LIB_PATH = os.path.join(PKG, "libksj_gpu.so")

OK = 0
LAYOUT_COIL_LAST = 0
LAYOUT_COIL_MAJOR = 1
KERNEL_LINEAR = 0
KERNEL_KAISER_BESSEL = 1
DENSITY_NONE = 0
DENSITY_RAMP = 1
DENSITY_GIVEN = 2
DENSITY_RAMP_SQUARED = 3
STATIC_SCRATCH_BYTES = 128 * 1024 * 1024


class KsjGpuError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"ksj_gpu error {code}: {message}")
        self.code = code


class Plan2Desc(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("image_rows", C.c_uint32), ("image_cols", C.c_uint32),
        ("grid_rows", C.c_uint32), ("grid_cols", C.c_uint32),
        ("kernel", C.c_uint32), ("kernel_width", C.c_uint32),
        ("density", C.c_uint32), ("centred", C.c_uint32),
        ("samples", C.c_uint32), ("item_visits", C.c_uint32), ("reserved0", C.c_uint32),
        ("beta", C.c_double), ("row_origin", C.c_double), ("col_origin", C.c_double),
    ]


class Plan2Info(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("items", C.c_uint32), ("visits", C.c_uint64),
                ("device_bytes", C.c_uint64), ("beta", C.c_double)]


class Plan3Desc(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("image_z", C.c_uint32), ("image_y", C.c_uint32), ("image_x", C.c_uint32),
        ("grid_z", C.c_uint32), ("grid_y", C.c_uint32), ("grid_x", C.c_uint32),
        ("kernel_width", C.c_uint32), ("density", C.c_uint32), ("samples", C.c_uint32),
        ("item_visits", C.c_uint32), ("reserved0", C.c_uint32),
        ("beta", C.c_double), ("origin_z", C.c_double), ("origin_y", C.c_double), ("origin_x", C.c_double),
    ]


class Plan3Info(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("items", C.c_uint32), ("visits", C.c_uint64),
                ("device_bytes", C.c_uint64), ("beta", C.c_double)]


class ServeConfig(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("device", C.c_int32), ("hbm_budget_bytes", C.c_uint64),
                ("lanes", C.c_uint32), ("max_batch_scans", C.c_uint32), ("plan_cache_entries", C.c_uint32),
                ("host_staging_mib", C.c_uint32)]


class ServeRequest(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("kind", C.c_uint32), ("image_rows", C.c_uint32),
                ("image_cols", C.c_uint32), ("channels", C.c_uint32), ("samples", C.c_uint32),
                ("kernel_width", C.c_uint32), ("density", C.c_uint32), ("centred", C.c_uint32),
                ("image_depth", C.c_uint32), ("trajectory_key", C.c_uint64), ("d_traj", C.c_void_p),
                ("d_kspace", C.c_void_p), ("d_images", C.c_void_p), ("d_rss", C.c_void_p),
                ("memory", C.c_uint32), ("reserved1", C.c_uint32)]


class ServeStats(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("reserved0", C.c_uint32)] + [
        (n, C.c_uint64) for n in ("submitted", "completed", "failed", "batches", "plan_builds", "plan_hits",
                                  "plan_evictions", "peak_device_bytes", "plan_bytes", "max_batch_seen")]


SERVE_RADIAL2D = 0
SERVE_CARTESIAN2D = 1
SERVE_RADIAL3D = 2
SERVE_DEVICE_MEMORY = 0
SERVE_HOST_MEMORY = 1

_vp = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_sz = C.c_size_t
_int = C.c_int
_dbl = C.c_double

# name -> (restype, argtypes); every symbol include/ksj_gpu.h declares.
SIGNATURES = {
    "ksj_gpu_init": (_int, [_int]),
    "ksj_gpu_last_error": (C.c_char_p, []),
    "ksj_gpu_stream_synchronize": (_int, [_vp]),
    "ksj_gpu_ifft2_workspace_bytes": (_sz, [_u32, _u32, _u32]),
    "ksj_gpu_ifft2_coil_images": (_int, [_vp, _vp, _u32, _u32, _u32, _int, _int, _vp, _sz, _vp]),
    "ksj_gpu_rss": (_int, [_vp, _vp, _u64, _u32, _vp]),
    "ksj_gpu_sense_combine": (_int, [_vp, _vp, _vp, _u64, _u32, _vp]),
    "ksj_gpu_ramp_dcf2": (_int, [_vp, _vp, _u64, _vp]),
    "ksj_gpu_nudft2_adjoint_workspace_bytes": (_sz, [_u32, _u32, _u32]),
    "ksj_gpu_nudft2_adjoint": (_int, [_vp, _vp, _vp, _u32, _u32, _u32, _u32, _dbl, _dbl, _int, _vp, _sz, _vp]),
    "ksj_gpu_nudft2_forward_f64": (_int, [_vp, _vp, _vp, _u32, _u32, _u32, _u32, _dbl, _dbl, _vp]),
    "ksj_gpu_golden_angle2": (_int, [_vp, _u32, _u32, _vp]),
    "ksj_gpu_plan2_create": (_int, [C.POINTER(Plan2Desc), C.POINTER(_vp)]),
    "ksj_gpu_plan2_set_trajectory": (_int, [_vp, _vp, _vp, _vp]),
    "ksj_gpu_plan2_get_info": (_int, [_vp, C.POINTER(Plan2Info)]),
    "ksj_gpu_plan2_export_bins": (_int, [_vp, _vp, _vp, _vp]),
    "ksj_gpu_plan2_destroy": (None, [_vp]),
    "ksj_gpu_adjoint2_workspace_bytes": (_sz, [_vp, _u32]),
    "ksj_gpu_adjoint2_execute": (_int, [_vp, _vp, _u32, _vp, _int, _vp, _vp, _sz, _vp]),
    "ksj_gpu_adjoint2_execute_static": (_int, [_vp, _vp, _u32, _vp, _int, _vp, _vp]),
    "ksj_gpu_calibration_workspace_bytes": (_sz, [_u32, _u64]),
    "ksj_gpu_channel_covariance": (_int, [_vp, _u32, _u64, _int, _vp, _vp, _sz, _vp]),
    "ksj_gpu_hermitian_eigen": (_int, [_vp, _u32, _vp, _vp, _vp, _vp]),
    "ksj_gpu_noise_model_estimate": (_int, [_vp, _u32, _u64, C.c_float, _vp, _vp, _sz, _vp]),
    "ksj_gpu_coil_compression_basis_estimate": (_int, [_vp, _u32, _u32, _u64, _vp, _vp, _sz, _vp]),
    "ksj_gpu_phase_correction_estimate": (_int, [_vp, _u32, _u32, _u32, _vp, _vp]),
    "ksj_gpu_sensitivity_maps_walsh": (_int, [_vp, _u32, _u32, _u32, _u32, _u32, _vp, _vp]),
    "ksj_gpu_plan3_create": (_int, [C.POINTER(Plan3Desc), C.POINTER(_vp)]),
    "ksj_gpu_plan3_set_trajectory": (_int, [_vp, _vp, _vp, _vp]),
    "ksj_gpu_plan3_get_info": (_int, [_vp, C.POINTER(Plan3Info)]),
    "ksj_gpu_plan3_export_bins": (_int, [_vp, _vp, _vp, _vp]),
    "ksj_gpu_plan3_destroy": (None, [_vp]),
    "ksj_gpu_adjoint3_workspace_bytes": (_sz, [_vp, _u32, _u32]),
    "ksj_gpu_adjoint3_execute": (_int, [_vp, _vp, _u32, _u32, _vp, _vp, _vp, _sz, _vp]),
    "ksj_gpu_golden_means3": (_int, [_vp, _u32, _u32, _vp]),
    "ksj_gpu_nufft3_forward_f64_workspace_bytes": (_sz, [C.POINTER(_u32)]),
    "ksj_gpu_nufft3_forward_f64": (_int, [_vp, C.POINTER(_u32), C.POINTER(_dbl), _vp, _u64, _u32, _vp, _vp, _sz,
                                          _vp]),
    "ksj_gpu_normalize_trajectory2": (_int, [_vp, _u64, _int, _u32, _u32, _vp, _vp]),
    "ksj_gpu_assemble_workspace_bytes": (_sz, [_u32]),
    "ksj_gpu_assemble_acquisitions": (_int, [_vp, _vp, _u32, _u32, _vp, _vp, _sz, _vp]),
    "ksj_gpu_trajectory_check": (_int, [_vp, _vp, _u64, C.c_float, _vp, _vp]),
    "ksj_gpu_channel_mix": (_int, [_vp, _vp, _vp, _u32, _u32, _u64, _vp]),
    "ksj_gpu_phase_correct": (_int, [_vp, _vp, _vp, _u32, _u32, _u32, _vp]),
    "ksj_gpu_readout_crop": (_int, [_vp, _vp, _u64, _u32, _u32, _u32, _vp]),
    "ksj_gpu_adjoint2_batch_workspace_bytes": (_sz, [_vp, _u32, _u32]),
    "ksj_gpu_adjoint2_execute_batch": (_int, [_vp, _u32, C.POINTER(_vp), _u32, C.POINTER(_vp), _int,
                                               C.POINTER(_vp), _vp, _sz, _vp]),
    "ksj_gpu_adjoint2_multi_workspace_bytes": (_sz, [_u32, C.POINTER(_vp), C.POINTER(_u32)]),
    "ksj_gpu_adjoint2_execute_multi": (_int, [_u32, C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_u32), C.POINTER(_vp),
                                              _int, C.POINTER(_vp), _vp, _sz, _vp]),
    "ksj_gpu_chain_faults": (_int, [C.POINTER(_u64)]),
    "ksj_gpu_graph_begin": (_int, [_vp]),
    "ksj_gpu_graph_end": (_int, [_vp, C.POINTER(_vp)]),
    "ksj_gpu_graph_launch": (_int, [_vp, _vp]),
    "ksj_gpu_graph_destroy": (None, [_vp]),
    "ksj_gpu_serve_create": (_int, [C.POINTER(ServeConfig), C.POINTER(_vp)]),
    "ksj_gpu_serve_submit": (_int, [_vp, C.POINTER(ServeRequest), C.POINTER(_u64)]),
    "ksj_gpu_serve_submit_batch": (_int, [_vp, _u32, C.POINTER(ServeRequest), C.POINTER(_u64)]),
    "ksj_gpu_serve_wait": (_int, [_vp, _u64, C.POINTER(_dbl)]),
    "ksj_gpu_serve_drain": (_int, [_vp]),
    "ksj_gpu_serve_hold": (_int, [_vp, _int]),
    "ksj_gpu_serve_get_stats": (_int, [_vp, C.POINTER(ServeStats)]),
    "ksj_gpu_serve_destroy": (None, [_vp]),
    "ksj_gpu_tune_set": (_int, [C.c_char_p, C.c_int64]),
    "ksj_gpu_tune_get": (_int, [C.c_char_p, C.POINTER(C.c_int64)]),
    "ksj_gpu_profile_enable": (_int, [_int]),
    "ksj_gpu_profile_reset": (_int, []),
    "ksj_gpu_profile_read": (_int, [C.c_char_p, C.POINTER(_dbl), C.POINTER(_u64)]),
}

_lib = None


def load():
    """Loads libksj_gpu.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(GPU_LIB_PATH):
            raise ImportError(f"{GPU_LIB_PATH} is missing; run `python -m kspacejet_amd.build`")
        lib = C.CDLL(GPU_LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code: int) -> None:
    if code != OK:
        raise KsjGpuError(code, load().ksj_gpu_last_error().decode(errors="replace"))


def call(name: str, *args) -> None:
    check(getattr(load(), name)(*args))

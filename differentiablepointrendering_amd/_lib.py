"""Loader of the HIP renderer library (libpale_hip.so) with ctypes signatures.

The library is built in-tree by ``differentiablepointrendering_amd/build.py``
(or ``__graft_entry__.build()``).  There is no CPU fallback: if the library
is missing or no GPU is present, loading or creating a renderer raises.
"""
from __future__ import annotations

import ctypes as C
import os

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpale_hip.so")

_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_i32p = C.POINTER(C.c_int32)
_lib = None

EXPORTED = (
    "pale_last_error", "pale_default_settings", "pale_create", "pale_destroy", "pale_camera_count",
    "pale_camera_info", "pale_render_forward", "pale_render_backward", "pale_point_count", "pale_get_points",
    "pale_apply_points", "pale_add_points", "pale_remove_points", "pale_rebuild_bvh", "pale_set_point_properties",
    "pale_set_point_opacity", "pale_set_point_beta", "pale_trace_rays", "pale_gather_irradiance", "pale_get_surfel_bvh", "pale_get_photons",
    "pale_get_stats", "pale_set_count_work", "pale_render_forward_device", "pale_render_backward_device",
    "pale_set_opacity_device", "pale_build_photon_map", "pale_render_view", "pale_render_view_device",
    "pale_render_view_backward", "pale_device_bytes", "pale_view_bytes", "pale_views_bytes", "pale_view_buffer_bytes",
    "pale_release_view_buffers", "pale_set_batch_tuning", "pale_get_batch_tuning", "pale_render_views",
    "pale_render_views_backward", "pale_render_view_backward_device", "pale_apply_points_device",
    "pale_get_debug_image", "pale_render_view_backward_noalloc",
)


class PaleError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load the library and declare every exported signature."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise PaleError(f"HIP renderer library not built: {path} (run differentiablepointrendering_amd/build.py)")
    # PyTorch ships its own HIP runtime under the same soname
    # (libamdhip64.so.7).  Importing torch first makes the library bind to
    # that runtime, so one process holds one HIP runtime and device tensors
    # can be exchanged; loading ours first would leave torch a second
    # runtime that finds no GPU.  Without torch the system ROCm runtime is
    # used.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(path)
    H = C.c_void_p
    L.pale_last_error.restype = C.c_char_p
    L.pale_last_error.argtypes = []
    L.pale_default_settings.restype = None
    L.pale_default_settings.argtypes = [C.POINTER(_abi.Settings)]
    L.pale_create.argtypes = [C.POINTER(_abi.SceneDesc), C.POINTER(_abi.Settings), C.POINTER(H)]
    L.pale_destroy.argtypes = [H]
    L.pale_camera_count.argtypes = [H, _u32p]
    L.pale_camera_info.argtypes = [H, C.c_uint32, C.c_char_p, _u32p, _u32p, _i32p]
    L.pale_render_forward.argtypes = [H, C.c_char_p, C.POINTER(_fp), C.POINTER(_fp), _u32p]
    L.pale_render_backward.argtypes = [H, C.c_uint32, C.POINTER(C.c_char_p), C.POINTER(_fp), C.POINTER(_fp),
                                       C.POINTER(_fp)]
    L.pale_render_forward_device.argtypes = [H, C.c_char_p, C.POINTER(_fp), C.POINTER(_fp), _u32p]
    L.pale_render_backward_device.argtypes = [H, C.c_uint32, C.POINTER(C.c_char_p), C.POINTER(_fp), C.POINTER(_fp)]
    L.pale_set_opacity_device.argtypes = [H, _fp, C.c_uint32]
    L.pale_apply_points_device.argtypes = [H, C.c_uint32] + [C.c_void_p] * 7
    L.pale_get_debug_image.argtypes = [H, C.c_char_p, _fp]
    L.pale_render_view_backward_noalloc.argtypes = [H, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]
    L.pale_build_photon_map.argtypes = [H]
    L.pale_device_bytes.argtypes = [C.POINTER(C.c_uint64)]
    L.pale_view_bytes.argtypes = [H, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int32, C.POINTER(C.c_uint64)]
    L.pale_release_view_buffers.argtypes = [H]
    L.pale_view_buffer_bytes.argtypes = [H, C.POINTER(C.c_uint64)]
    L.pale_views_bytes.argtypes = [H, C.c_uint32, _u32p, _u32p, _u32p, _i32p, C.POINTER(C.c_uint64)]
    L.pale_set_batch_tuning.argtypes = [H, C.POINTER(_abi.BatchTuning)]
    L.pale_get_batch_tuning.argtypes = [H, C.POINTER(_abi.BatchTuning)]
    L.pale_render_views.argtypes = [H, C.c_uint32, C.POINTER(_abi.CameraDesc), _u32p, _u32p, C.POINTER(_fp),
                                    C.POINTER(_fp), _u32p]
    L.pale_render_view_backward_device.argtypes = [H, C.POINTER(_abi.CameraDesc), C.c_uint32, _fp, C.POINTER(_fp)]
    L.pale_render_views_backward.argtypes = [H, C.c_uint32, C.POINTER(_abi.CameraDesc), _u32p, C.POINTER(_fp),
                                             C.POINTER(_fp), _u32p]
    _camp = C.POINTER(_abi.CameraDesc)
    L.pale_render_view.argtypes = [H, _camp, C.c_uint32, C.c_uint32, _fp, _fp]
    L.pale_render_view_device.argtypes = [H, _camp, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
    L.pale_render_view_backward.argtypes = [H, _camp, C.c_uint32, _fp, C.POINTER(_fp)]
    L.pale_point_count.argtypes = [H, _u32p]
    L.pale_get_points.argtypes = [H] + [_fp] * 8
    L.pale_apply_points.argtypes = [H, C.POINTER(_abi.PointsDesc)]
    L.pale_add_points.argtypes = [H, C.POINTER(_abi.PointsDesc)]
    L.pale_remove_points.argtypes = [H, C.POINTER(C.c_int64), C.c_size_t]
    L.pale_rebuild_bvh.argtypes = [H]
    L.pale_set_point_properties.argtypes = [H, _fp, _fp, _fp, _fp, C.c_float, C.c_float, C.c_int32]
    L.pale_set_point_opacity.argtypes = [H, C.c_float, C.c_int32]
    L.pale_set_point_beta.argtypes = [H, C.c_float, C.c_int32]
    L.pale_trace_rays.argtypes = [H, C.c_uint32, _fp, _fp, _i32p, _i32p, _fp, _fp, _fp]
    L.pale_gather_irradiance.argtypes = [H, C.c_uint32, _fp, _fp, _fp]
    L.pale_get_surfel_bvh.argtypes = [H, _u32p, _fp, _u32p, _u32p, _u32p]
    L.pale_get_photons.argtypes = [H, _u32p, C.POINTER(C.c_uint64), _fp, _fp, _fp]
    L.pale_get_stats.argtypes = [H, C.POINTER(_abi.Stats)]
    L.pale_set_count_work.argtypes = [H, C.c_int32]
    for name in EXPORTED:
        if name != "pale_last_error" and name != "pale_default_settings":
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(rc: int):
    """Raise the Python exception the reference's pybind11 layer would."""
    if rc == _abi.PALE_OK:
        return
    msg = load().pale_last_error().decode(errors="replace")
    if rc == _abi.PALE_ERR_RANGE:
        raise IndexError(msg)
    raise RuntimeError(msg)

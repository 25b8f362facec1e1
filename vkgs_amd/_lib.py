"""ctypes binding of libvkgs_hip.so (declared in include/vkgs_hip.h).

The library is the only compute path: if it is missing, every entry point that
needs it raises ``LibraryNotBuilt``; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# VKGS_HIP_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("VKGS_HIP_LIB") or os.path.join(_HERE, "lib", "libvkgs_hip.so")

VKGS_OK = 0
STATUS_NAMES = {
    0: "VKGS_OK",
    1: "VKGS_ERR_INVALID_ARG",
    2: "VKGS_ERR_CAPACITY",
    3: "VKGS_ERR_HIP",
    4: "VKGS_ERR_IO",
    5: "VKGS_ERR_FORMAT",
    6: "VKGS_ERR_OVERFLOW",
    7: "VKGS_ERR_NO_DEVICE",
}


class LibraryNotBuilt(RuntimeError):
    pass


class VkgsError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {STATUS_NAMES.get(status, status)}")


class Camera(C.Structure):
    """Byte layout of vk::shader::Camera (uniforms.h:10-15), 192 bytes."""

    _fields_ = [
        ("projection", C.c_float * 16),
        ("view", C.c_float * 16),
        ("camera_position", C.c_float * 3),
        ("pad0", C.c_float),
        ("screen_size", C.c_uint32 * 2),
        ("pad1", C.c_uint32 * 10),
    ]


class PlyInfo(C.Structure):
    _fields_ = [
        ("point_count", C.c_uint64),
        ("stride", C.c_uint32),
        ("offsets", C.c_uint32 * 60),
        ("header_bytes", C.c_uint64),
    ]


class RasterConfig(C.Structure):
    _fields_ = [
        ("tile_w", C.c_uint32),
        ("tile_h", C.c_uint32),
        ("t_min", C.c_float),
        ("variant", C.c_uint32),
        ("flags", C.c_uint32),
        ("lds_batch", C.c_uint32),
        ("waves_per_cu", C.c_uint32),
        ("reserved", C.c_uint32 * 1),
    ]


_P = C.c_void_p
_U32 = C.c_uint32
_U64 = C.c_uint64
_SZ = C.c_size_t
_I = C.c_int
_F = C.c_float

# name -> (restype, argtypes)
SIGNATURES = {
    "vkgs_status_string": (C.c_char_p, [_I]),
    "vkgs_abi_version": (_I, []),
    "vkgs_camera_orbit": (_I, [_F, _U32, _U32, _F, _F, _P, _F, _F, _F, _P]),
    "vkgs_mat4_inverse": (_I, [_P, _P]),
    "vkgs_ply_read_header": (_I, [C.c_char_p, _P]),
    "vkgs_ply_buffer_bytes": (_U64, [_P]),
    "vkgs_ply_read": (_I, [C.c_char_p, _P, _P, _U64]),
    "vkgs_parse_ply": (_I, [_P, _U32, _P, _P, _P, _P, _P]),
    "vkgs_rank_storage_bytes": (_SZ, [_U32]),
    "vkgs_rank": (_I, [_P, _P, _P, _U32, _P, _P, _P, _P, _SZ, _P]),
    "vkgs_sort_kv_storage_bytes": (_SZ, [_U32]),
    "vkgs_sort_kv": (_I, [_U32, _P, _P, _P, _U32, _P, _SZ, _P]),
    "vkgs_sort_kv_noalloc": (_I, [_U32, _P, _P, _P, _U32, _P]),
    "vkgs_inverse_index": (_I, [_P, _P, _U32, _P, _P]),
    "vkgs_projection": (_I, [_P, _P, _P, _P, _P, _P, _P, _U32, _P, _P, _P, _P]),
    "vkgs_raster_config_default": (None, [_P]),
    "vkgs_draw_storage_bytes": (_SZ, [_U32, _U32, _U32, _U64, _P]),
    "vkgs_draw": (_I, [_P, _P, _U32, _U32, _U32, _U64, _P, _P, _P, _SZ, _P]),
    "vkgs_draw_status": (_I, [_P, _SZ, _P, _P]),
    "vkgs_draw_lists": (_I, [_P, _SZ, _P, _P, _P, _P]),
    "vkgs_renderer_create": (_I, [_U32, _P]),
    "vkgs_renderer_destroy": (_I, [_P]),
    "vkgs_renderer_load_ply": (_I, [_P, C.c_char_p]),
    "vkgs_renderer_load_ply_buffer": (_I, [_P, _P, _U32]),
    "vkgs_renderer_load_ply_async": (_I, [_P, C.c_char_p]),
    "vkgs_renderer_cancel_load": (_I, [_P]),
    "vkgs_renderer_load_progress": (_I, [_P, _P, _P, _P]),
    "vkgs_renderer_splat_count": (_U32, [_P]),
    "vkgs_renderer_storage": (_I, [_P, _P, _P, _P, _P]),
    "vkgs_renderer_set_config": (_I, [_P, _P]),
    "vkgs_renderer_render": (_I, [_P, _P, _P, _P, _P]),
    "vkgs_renderer_frame_info": (_I, [_P, _P, _P, _P]),
    "vkgs_renderer_sorted": (_I, [_P, _P, _P]),
    "vkgs_renderer_tile_lists": (_I, [_P, _P, _P, _P, _P]),
    "vkgs_renderer_set_timing": (_I, [_P, _I]),
    "vkgs_renderer_stage_ms": (_I, [_P, _P]),
    "vkgs_renderer_raster_work": (_I, [_P, _P]),
    "vkgs_renderer_autotune": (_I, [_P, _P, _P, _P, _U32, _P, _P, _P]),
    "vkgs_diff_create": (_I, [_P]),
    "vkgs_diff_destroy": (_I, [_P]),
    "vkgs_diff_forward": (_I, [_P, _P, _U32, _P, _U32, _P, _P]),
    "vkgs_diff_backward": (_I, [_P, _P, _U32, _P, _U32, _P]),
    "vkgs_diff_info": (_I, [_P, _P, _P, _P]),
    "vkgs_diff_check": (_I, [_P, _P, _P]),
    "vkgs_diff_set_timing": (_I, [_P, _I]),
    "vkgs_diff_stage_ms": (_I, [_P, _P]),
    "vkgs_diff_raster_work": (_I, [_P, _P, _P]),
    "vkgs_diff_autotune": (_I, [_P, _P, _U32, _P, _U32, _P, _U32, _P, _P, _P]),
}

_lib = None


def lib() -> C.CDLL:
    """Loads the library once; raises LibraryNotBuilt when it is absent."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(
                f"{LIB_PATH} is missing: build it with `python -m vkgs_amd.build` "
                "(there is no CPU fallback)")
        handle = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int, where: str) -> None:
    if status != VKGS_OK:
        raise VkgsError(status, where)


def ptr(x) -> int | None:
    """Device or host address of a torch tensor / numpy array / ctypes object."""
    if x is None:
        return None
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    return C.addressof(x)


def default_raster_config() -> RasterConfig:
    cfg = RasterConfig()
    lib().vkgs_raster_config_default(C.byref(cfg))
    return cfg

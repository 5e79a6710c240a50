"""ctypes binding of librm_hip.so (include/rm.h).

The library is built in-tree by ``make -C renderer_ray_marching_amd/csrc`` (``__graft_entry__.build``
does it). Loading fails loudly when it is missing: there is no fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librm_hip.so")

RM_OK = 0
RM_ERR_INVALID = -1
RM_ERR_HIP = -2
RM_ERR_UNSUPPORTED = -3
RM_ERR_NOMEM = -4
RM_ERR_INTERNAL = -5
RM_MAX_K = 32
RM_MARCH_K = 10
RM_NOALLOC_MAX_VIEWS = 1024
RM_MAX_STACK_DEPTH = 64
RM_MAX_EXIT_DEPTH = 10


class rm_view(C.Structure):
    _fields_ = [("pose", C.c_float * 16), ("fov_radians", C.c_float), ("width", C.c_int32),
                ("height", C.c_int32), ("reserved", C.c_int32), ("offset", C.c_int64)]


class rm_tune(C.Structure):
    _fields_ = [("tile_w", C.c_int32), ("tile_h", C.c_int32), ("mode", C.c_int32),
                ("warm", C.c_int32), ("waves_per_block", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("refill", C.c_int32),
                ("stack_depth", C.c_int32), ("exit_depth", C.c_int32)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_f32 = C.c_float
_pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every symbol include/rm.h declares
SIGNATURES = {
    "rm_ctx_create": (_i32, [_i32, _pp]),
    "rm_ctx_destroy": (_i32, [_vp]),
    "rm_ctx_set_stream": (_i32, [_vp, _vp]),
    "rm_ctx_get_stream": (_i32, [_vp, C.POINTER(_vp)]),
    "rm_ctx_synchronize": (_i32, [_vp]),
    "rm_last_error": (C.c_char_p, [_vp]),
    "rm_ctx_enable_timing": (_i32, [_vp, _i32]),
    "rm_ctx_last_kernel_ms": (_f32, [_vp]),
    "rm_tree_build": (_i32, [_vp, _vp, _i64, _i32, _pp]),
    "rm_tree_destroy": (_i32, [_vp]),
    "rm_tree_info": (_i32, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i32),
                            C.POINTER(_i32)]),
    "rm_knn": (_i32, [_vp, _vp, _vp, _i64, _i32, _vp, _vp]),
    "rm_radius_search": (_i32, [_vp, _vp, _vp, _i64, _f32, _i32, _vp, _vp, _vp]),
    "rm_radii": (_i32, [_vp, _vp, _f32, _vp, _vp]),
    "rm_cloud_create": (_i32, [_vp, _vp, _vp, _vp, _vp, _i64, _pp]),
    "rm_cloud_destroy": (_i32, [_vp]),
    "rm_cloud_set_radii": (_i32, [_vp, _vp, _vp]),
    "rm_cloud_set_colors": (_i32, [_vp, _vp, _vp]),
    "rm_cloud_set_culling": (_i32, [_vp, _vp, _i32]),
    "rm_cloud_refit": (_i32, [_vp, _vp, _vp]),
    "rm_render": (_i32, [_vp, _vp, C.POINTER(_f32), _f32, _i32, _i32, _vp, _vp]),
    "rm_batch_create": (_i32, [_vp, _vp, C.POINTER(rm_view), _i32, C.POINTER(rm_tune), _vp,
                               _vp, _pp]),
    "rm_batch_autotune": (_i32, [_vp, _i32, C.POINTER(rm_tune)]),
    "rm_batch_get_tune": (_i32, [_vp, C.POINTER(rm_tune)]),
    "rm_batch_launch": (_i32, [_vp]),
    "rm_batch_render_diff": (_i32, [_vp, _vp, _vp]),
    "rm_batch_destroy": (_i32, [_vp]),
    "rm_render_batch": (_i32, [_vp, _vp, C.POINTER(rm_view), _i32, _vp, _vp]),
    "rm_render_batch_noalloc": (_i32, [_vp, _vp, C.POINTER(rm_view), _i32, _vp, _vp]),
    "rm_render_diff": (_i32, [_vp, _vp, C.POINTER(_f32), _f32, _i32, _i32, _vp, _vp, _vp]),
    "rm_render_adjoint": (_i32, [_vp, _vp, C.POINTER(_f32), _f32, _i32, _i32, _vp, _vp, _vp,
                                 _vp, _vp, _vp]),
    "rm_batch_render_adjoint": (_i32, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


def lib() -> C.CDLL:
    """Load librm_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C renderer_ray_marching_amd/csrc` "
            "(or __graft_entry__.build()); there is no fallback path")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


class RMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"rm error {code}: {msg}")
        self.code = code


def check(code: int, ctx=None) -> None:
    if code != RM_OK:
        msg = lib().rm_last_error(ctx).decode() if ctx else ""
        raise RMError(code, msg)

"""ctypes binding of librm_host.so (include/rm_host.h): the host-side I/O of the drop-in CLI.

Built in-tree by ``make -C renderer_ray_marching_amd/host`` (``__graft_entry__.build`` does it).
No function here touches the GPU. ``MARCHER`` is the path of the CLI built from the same sources.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librm_host.so")
MARCHER = os.path.join(_HERE, "marcher.bin")

RMH_OK = 0
RMH_ERR_INVALID = -1
RMH_ERR_IO = -2


class rmh_view_spec(C.Structure):
    _fields_ = [("key", C.c_char_p), ("split", C.c_int32), ("has_ply_path", C.c_int32),
                ("ply_path", C.c_char_p), ("pose", C.c_float * 16), ("calib", C.c_float * 16),
                ("width", C.c_int32), ("height", C.c_int32), ("fov_radians", C.c_float)]


_vp = C.c_void_p
_i64 = C.c_int64
_i32 = C.c_int
_cp = C.c_char_p
_pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); every symbol include/rm_host.h declares
SIGNATURES = {
    "rmh_last_error": (_cp, []),
    "rmh_free": (None, [_vp]),
    "rmh_ply_load": (_i32, [_cp, C.POINTER(_i64), _pp, _pp]),
    "rmh_ply_write": (_i32, [_cp, _vp, _vp, _i64]),
    "rmh_radii_write": (_i32, [_cp, _vp, _i64]),
    "rmh_radii_read": (_i32, [_cp, C.POINTER(_i64), _pp]),
    "rmh_save_png": (_i32, [_cp, _vp, _i32, _i32]),
    "rmh_save_depth": (_i32, [_cp, _vp, _i32, _i32]),
    "rmh_matrices_load": (_i32, [_cp, _pp]),
    "rmh_matrices_count": (_i64, [_vp]),
    "rmh_matrices_key": (_cp, [_vp, _i64]),
    "rmh_matrices_get": (_i32, [_vp, _i64, C.POINTER(rmh_view_spec)]),
    "rmh_matrices_free": (None, [_vp]),
    "rmh_view_paths": (_i32, [_cp, _cp, _cp, _cp, _cp, C.c_size_t]),
    "rmh_lock_try": (_i32, [_cp, C.POINTER(_i32)]),
    "rmh_lock_release": (_i32, [_i32, _cp]),
    "rmh_camera_poses": (_i32, [_cp, _i32, C.c_float, _vp, _vp]),
}

_lib = None


class RMHError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RMHError(f"{LIB_PATH} is not built: run make -C renderer_ray_marching_amd/host")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _b(s) -> bytes:
    return os.fsencode(s)


def _check(rc):
    if rc != RMH_OK:
        raise RMHError(lib().rmh_last_error().decode(errors="replace"))


def _take(ptr, count, dtype, shape):
    """Copy a library-allocated buffer into numpy and free it."""
    try:
        if count == 0:
            return np.zeros(shape, dtype)
        buf = (C.c_char * (count * np.dtype(dtype).itemsize)).from_address(ptr.value)
        return np.frombuffer(bytes(buf), dtype=dtype).reshape(shape).copy()
    finally:
        lib().rmh_free(ptr)


def load_ply(path):
    """(xyzw (n, 4) float32 with w = 1, rgba (n, 4) float32 = rgb / 255 with a = 1)."""
    n = _i64()
    a, b = _vp(), _vp()
    _check(lib().rmh_ply_load(_b(path), C.byref(n), C.byref(a), C.byref(b)))
    return _take(a, 4 * n.value, np.float32, (n.value, 4)), _take(b, 4 * n.value, np.float32, (n.value, 4))


def write_ply(path, xyz, rgb):
    xyz = np.ascontiguousarray(xyz, np.float32)
    rgb = np.ascontiguousarray(rgb, np.uint8)
    assert xyz.ndim == 2 and xyz.shape[1] == 3 and rgb.shape == xyz.shape
    _check(lib().rmh_ply_write(_b(path), xyz.ctypes.data, rgb.ctypes.data, xyz.shape[0]))


def write_radii(path, radii):
    r = np.ascontiguousarray(radii, np.float32)
    _check(lib().rmh_radii_write(_b(path), r.ctypes.data, r.shape[0]))


def read_radii(path):
    n = _i64()
    p = _vp()
    _check(lib().rmh_radii_read(_b(path), C.byref(n), C.byref(p)))
    return _take(p, n.value, np.float32, (n.value,))


def save_png(path, rgba_texture):
    t = np.ascontiguousarray(rgba_texture, np.float32)
    H, W, _ = t.shape
    _check(lib().rmh_save_png(_b(path), t.ctypes.data, W, H))


def save_depth(path, rgba_texture):
    t = np.ascontiguousarray(rgba_texture, np.float32)
    H, W, _ = t.shape
    _check(lib().rmh_save_depth(_b(path), t.ctypes.data, W, H))


class Matrices:
    """The views of a matrices file, train then val, each in sorted key order."""

    def __init__(self, path):
        self._h = None
        h = _vp()
        _check(lib().rmh_matrices_load(_b(path), C.byref(h)))
        self._h = h

    def __len__(self):
        return int(lib().rmh_matrices_count(self._h))

    def key(self, i):
        return lib().rmh_matrices_key(self._h, i).decode()

    def get(self, i):
        s = rmh_view_spec()
        _check(lib().rmh_matrices_get(self._h, i, C.byref(s)))
        return {"key": s.key.decode(), "split": s.split,
                "ply_path": s.ply_path.decode() if s.has_ply_path else None,
                "pose": np.array(s.pose[:], np.float32), "calib": np.array(s.calib[:], np.float32),
                "width": s.width, "height": s.height, "fov_radians": float(s.fov_radians)}

    def close(self):
        if getattr(self, "_h", None):
            lib().rmh_matrices_free(self._h)
            self._h = None

    __del__ = close


def view_paths(output_dir, key):
    cap = 8192
    bufs = [C.create_string_buffer(cap) for _ in range(3)]
    _check(lib().rmh_view_paths(_b(output_dir), _b(key), bufs[0], bufs[1], bufs[2], cap))
    return tuple(os.fsdecode(b.value) for b in bufs)


def lock_try(lock_path):
    """File descriptor of the taken lock, or None when another process holds it."""
    fd = _i32(-1)
    r = lib().rmh_lock_try(_b(lock_path), C.byref(fd))
    if r < 0:
        _check(r)
    return fd.value if r == 1 else None


def lock_release(fd, lock_path):
    _check(lib().rmh_lock_release(fd, _b(lock_path)))


def camera_poses(frames: int, dt: float, script: str | None = None):
    """Render poses (frames, 16) column-major and fields of view (frames,) of the viewer's FPS
    camera under a scripted input (None: the built-in tour), fixed frame time dt."""
    poses = np.empty((frames, 16), np.float32)
    fovs = np.empty(frames, np.float32)
    _check(lib().rmh_camera_poses(None if script is None else script.encode(), frames, float(dt),
                                  poses.ctypes.data_as(C.c_void_p),
                                  fovs.ctypes.data_as(C.c_void_p)))
    return poses, fovs

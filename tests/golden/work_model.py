"""ctypes binding of work_model.cpp (measurement infrastructure: what one lane of the library's
radii and march kernels reads when it searches alone). Built on first use with g++ into
tests/golden/_build/ (git-ignored)."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
SRC = [os.path.join(HERE, "work_model.cpp"),
       os.path.join(ROOT, "renderer_ray_marching_amd", "csrc", "rm_tree_host.cpp")]
LIB = os.path.join(HERE, "_build", "libwork_model.so")

# bytes per counted item (see work_model.cpp)
BYTES = {"nodes": 64, "retests": 64, "points": 16, "seeds": 12, "sdf": 4}


class Counts(C.Structure):
    _fields_ = [(f, C.c_uint64) for f in
                ("nodes", "retests", "points", "seeds", "sdf", "steps", "rays", "hits")]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}

    def traversal_bytes(self):
        return sum(BYTES[f] * int(getattr(self, f)) for f in BYTES)


class PacketCounts(C.Structure):
    """Per-wavefront-step work of the batched kernel's 64-ray packets (wm_render_packet)."""
    _fields_ = [(f, C.c_uint64) for f in
                ("wave_steps", "lane_steps", "expansions", "retests", "scans", "candidates",
                 "wave_inserts", "lane_inserts", "lane_child_passes", "lane_candidates", "rays",
                 "hits", "tiles", "containing_depth", "containing_steps", "insert_slots",
                 "inserts_no_move", "seed_moves", "seed_steps")]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    deps = SRC + [os.path.join(ROOT, "renderer_ray_marching_amd", "csrc", "rm_tree_host.h")]
    if not os.path.exists(LIB) or any(os.path.getmtime(s) > os.path.getmtime(LIB) for s in deps):
        os.makedirs(os.path.dirname(LIB), exist_ok=True)
        subprocess.run(["g++", "-O3", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-pthread",
                        "-ffp-contract=off", "-mfma", "-o", LIB + ".tmp", *SRC], check=True)
        os.replace(LIB + ".tmp", LIB)
    L = C.CDLL(LIB)
    vp = C.c_void_p
    L.wm_tree_build.restype = vp
    L.wm_tree_build.argtypes = [vp, C.c_int64, C.c_int]
    L.wm_tree_free.argtypes = [vp]
    L.wm_tree_depth.argtypes = [vp]
    L.wm_set_radii.argtypes = [vp, vp]
    L.wm_radii.argtypes = [vp, vp, C.POINTER(Counts), C.c_int]
    L.wm_render.argtypes = [vp, vp, C.c_float, C.c_int, C.c_int, vp, C.c_int64, vp, vp,
                            C.POINTER(Counts), C.c_int]
    L.wm_render_packet.argtypes = [vp, vp, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, vp,
                                   C.POINTER(PacketCounts), C.c_int]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class WorkTree:
    """The library's search tree (host build, leaf_max points per leaf), for work counting."""

    def __init__(self, xyzw: np.ndarray, leaf_max: int = 16):
        self._xyzw = np.ascontiguousarray(xyzw, np.float32)
        self.n = self._xyzw.shape[0]
        self.h = lib().wm_tree_build(_p(self._xyzw), self.n, leaf_max)
        if not self.h:
            raise ValueError("wm_tree_build failed")

    def set_radii(self, radii):
        r = np.ascontiguousarray(radii, np.float32)
        lib().wm_set_radii(self.h, _p(r))

    def radii(self, nthreads=0):
        out = np.empty(self.n, np.float32)
        c = Counts()
        lib().wm_radii(self.h, _p(out), C.byref(c), nthreads)
        return out, c

    def render(self, pose_colmajor, pane, W, H, pix=None, nthreads=0):
        pose = np.ascontiguousarray(pose_colmajor, np.float32).ravel()
        if pix is not None:
            pix = np.ascontiguousarray(pix, np.int32)
        m = W * H if pix is None else pix.shape[0]
        hit = np.empty(m, np.int32)
        depth = np.empty(m, np.float32)
        c = Counts()
        lib().wm_render(self.h, _p(pose), float(pane), W, H, _p(pix) if pix is not None else None,
                        m, _p(hit), _p(depth), C.byref(c), nthreads)
        return hit, depth, c

    def render_packet(self, pose_colmajor, pane, W, H, tile_stride=1, bound_only=False,
                      nthreads=0):
        """Hit of every pixel of the counted tiles (-2 elsewhere) and the packet counts."""
        pose = np.ascontiguousarray(pose_colmajor, np.float32).ravel()
        hit = np.full(W * H, -2, np.int32)
        c = PacketCounts()
        lib().wm_render_packet(self.h, _p(pose), float(pane), W, H, int(tile_stride),
                               int(bool(bound_only)), _p(hit), C.byref(c), nthreads)
        return hit, c

    def __del__(self):
        try:
            if self.h:
                lib().wm_tree_free(self.h)
                self.h = None
        except Exception:
            pass

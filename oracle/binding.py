"""ctypes binding of the CPU oracle (liboracle.so) and of the FLANN CPU cross-check
(_ref/libflannref.so). TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg, never by the product path."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_LIB = os.path.join(HERE, "_ref", "libflannref.so")


class Counters(C.Structure):
    _fields_ = [("c", C.c_uint64), ("v", C.c_uint64), ("a", C.c_uint64), ("q", C.c_uint64),
                ("s", C.c_uint64)]

    def as_dict(self):
        return {f: int(getattr(self, f)) for f, _ in self._fields_}

    def traversal_bytes(self):
        """12 B per candidate, 16 B per node record, 24 B per far-child box test."""
        return 12 * self.c + 16 * self.v + 24 * self.a


_lib = None
_ref = None

_f = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
_i = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_d = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            subprocess.run(["make", "-C", HERE, "liboracle.so"], check=True,
                           stdout=subprocess.DEVNULL)
        L = C.CDLL(LIB)
        L.orc_tree_build.restype = C.c_void_p
        L.orc_tree_build.argtypes = [_f, C.c_int64, C.c_int]
        L.orc_tree_free.argtypes = [C.c_void_p]
        L.orc_tree_num_nodes.restype = C.c_int64
        L.orc_tree_num_nodes.argtypes = [C.c_void_p]
        L.orc_tree_depth.restype = C.c_int
        L.orc_tree_depth.argtypes = [C.c_void_p]
        L.orc_tree_export.argtypes = [C.c_void_p, _i, _i, _i, _i, _f, _f, _i]
        L.orc_knn.argtypes = [C.c_void_p, _f, C.c_int64, C.c_int, _i, _f, C.POINTER(Counters),
                              C.c_int]
        L.orc_knn_brute.argtypes = [_f, C.c_int64, _f, C.c_int64, C.c_int, _i, _f, C.c_int]
        L.orc_radii.argtypes = [C.c_void_p, C.c_float, _f, _f, C.POINTER(Counters), C.c_int]
        L.orc_subsample_indices.argtypes = [C.c_int64, _i]
        L.orc_subsample_indices.restype = None
        L.orc_mt19937_10000th.restype = C.c_uint32
        L.orc_mt19937_10000th.argtypes = []
        L.orc_pane_distance.restype = C.c_float
        L.orc_pane_distance.argtypes = [C.c_int, C.c_float]
        L.orc_render.argtypes = [C.c_void_p, _f, _f, _f, C.c_float, C.c_int, C.c_int,
                                 C.c_void_p, C.c_int64, _f, _i, C.POINTER(Counters), C.c_int]
        L.orc_render_steps.argtypes = [C.c_void_p, _f, _f, _f, C.c_float, C.c_int, C.c_int,
                                       C.c_void_p, C.c_int64, _f, _i, _i,
                                       C.POINTER(Counters), C.c_int]
        L.orc_frustum_inside.argtypes = [_f, C.c_float, C.c_int, C.c_int, _f, C.c_int64,
                                         np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")]
        L.orc_loss_grad_fixed.restype = C.c_double
        L.orc_loss_grad_fixed.argtypes = [_d, _d, _d, C.c_int64, _f, C.c_float, C.c_int,
                                          C.c_int, _i, _f, _f, C.c_double, C.c_double, _d, _d,
                                          _d]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_LIB)


def ref_lib():
    global _ref
    if _ref is None:
        L = C.CDLL(REF_LIB)
        L.flannref_knn.restype = C.c_int
        L.flannref_knn.argtypes = [_f, C.c_int64, _f, C.c_int64, C.c_int, _i, _f]
        _ref = L
    return _ref


def _xyzw(p):
    p = np.asarray(p, dtype=np.float32)
    if p.shape[1] == 3:
        q = np.zeros((p.shape[0], 4), np.float32)
        q[:, :3] = p
        p = q
    return np.ascontiguousarray(p)


class RefTree:
    """The reference tree rule (leaf <= leaf_max, midpoint of the longest axis)."""

    def __init__(self, pts, leaf_max: int = 64):
        self.p = _xyzw(pts)
        self.n = self.p.shape[0]
        self.h = lib().orc_tree_build(self.p, self.n, leaf_max)
        if not self.h:
            raise ValueError("orc_tree_build failed")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tree_free(self.h)
            self.h = None

    @property
    def num_nodes(self):
        return int(lib().orc_tree_num_nodes(self.h))

    @property
    def depth(self):
        return int(lib().orc_tree_depth(self.h))

    def export(self):
        nn = self.num_nodes
        child1 = np.empty(nn, np.int32)
        parent = np.empty(nn, np.int32)
        begin = np.empty(nn, np.int32)
        end = np.empty(nn, np.int32)
        lo = np.empty((nn, 3), np.float32)
        hi = np.empty((nn, 3), np.float32)
        order = np.empty(self.n, np.int32)
        lib().orc_tree_export(self.h, child1, parent, begin, end, lo, hi, order)
        return dict(child1=child1, parent=parent, begin=begin, end=end, lo=lo, hi=hi,
                    order=order)

    def knn(self, q, k, nthreads=0):
        q = _xyzw(q)
        m = q.shape[0]
        idx = np.empty((m, k), np.int32)
        d2 = np.empty((m, k), np.float32)
        cnt = Counters()
        lib().orc_knn(self.h, q, m, k, idx, d2, C.byref(cnt), nthreads)
        return idx, d2, cnt

    def radii(self, max_radius=0.1, nthreads=0):
        ru = np.empty(self.n, np.float32)
        rc = np.empty(self.n, np.float32)
        cnt = Counters()
        lib().orc_radii(self.h, max_radius, ru, rc, C.byref(cnt), nthreads)
        return ru, rc, cnt

    def render(self, radii, colors4, pose_colmajor, fov, W, H, pix=None, nthreads=0,
               steps=False):
        pane = lib().orc_pane_distance(W, fov)
        radii = np.ascontiguousarray(radii, np.float32)
        colors4 = np.ascontiguousarray(colors4, np.float32)
        pose = np.ascontiguousarray(pose_colmajor, np.float32)
        if pix is None:
            npix = W * H
            pptr = None
        else:
            pix = np.ascontiguousarray(pix, np.int32)
            npix = pix.shape[0]
            pptr = pix.ctypes.data_as(C.c_void_p)
        rgba = np.empty((npix, 4), np.float32)
        hit = np.empty(npix, np.int32)
        cnt = Counters()
        if steps:
            st = np.empty(npix, np.int32)
            lib().orc_render_steps(self.h, radii, colors4, pose, pane, W, H, pptr, npix, rgba,
                                   hit, st, C.byref(cnt), nthreads)
            return rgba, hit, cnt, st
        lib().orc_render(self.h, radii, colors4, pose, pane, W, H, pptr, npix, rgba, hit,
                         C.byref(cnt), nthreads)
        return rgba, hit, cnt


def knn_brute(pts, q, k, nthreads=0):
    p = _xyzw(pts)
    q = _xyzw(q)
    idx = np.empty((q.shape[0], k), np.int32)
    d2 = np.empty((q.shape[0], k), np.float32)
    lib().orc_knn_brute(p, p.shape[0], q, q.shape[0], k, idx, d2, nthreads)
    return idx, d2


def frustum_inside(pose_colmajor, fov, W, H, pts):
    """1 where a point (n x 3 or n x 4) is inside the view frustum (src/ray_marching.cu:231-258)."""
    p = _xyzw(pts)
    out = np.empty(p.shape[0], np.uint8)
    pose = np.ascontiguousarray(pose_colmajor, np.float32)
    lib().orc_frustum_inside(pose, lib().orc_pane_distance(W, fov), W, H, p, p.shape[0], out)
    return out


def render_culled(pts, radii, colors4, pose_colmajor, fov, W, H, pix=None, nthreads=0):
    """The march over the culled cloud: the points inside the view frustum (frustum_inside), in
    their original order, with their radii and colours; hit indices refer to the full cloud.
    Returns (rgba, hit, inside)."""
    p = _xyzw(pts)
    inside = frustum_inside(pose_colmajor, fov, W, H, p)
    sub = np.nonzero(inside)[0].astype(np.int32)
    npix = W * H if pix is None else len(pix)
    if sub.size == 0:
        return np.zeros((npix, 4), np.float32), np.full(npix, -1, np.int32), inside
    t = RefTree(p[sub], 64)
    rgba, hit, _ = t.render(np.asarray(radii, np.float32)[sub],
                            np.asarray(colors4, np.float32)[sub], pose_colmajor, fov, W, H,
                            pix=pix, nthreads=nthreads)
    return rgba, np.where(hit >= 0, sub[np.maximum(hit, 0)], -1).astype(np.int32), inside


def subsample_indices(n):
    """Order of main.cu:75-87's -s subsample (its first mp entries are the subsample)."""
    out = np.empty(n, np.int32)
    lib().orc_subsample_indices(n, out)
    return out


def mt19937_10000th():
    return int(lib().orc_mt19937_10000th())


def pane_distance(W, fov):
    return float(lib().orc_pane_distance(W, fov))


def loss_grad_fixed(xyz, r, rgb, pose_colmajor, fov, W, H, hit, target_rgb, target_depth,
                    w_rgb=1.0, w_depth=1.0):
    xyz = np.ascontiguousarray(xyz, np.float64)
    r = np.ascontiguousarray(r, np.float64)
    rgb = np.ascontiguousarray(rgb, np.float64)
    n = xyz.shape[0]
    gx = np.zeros((n, 3), np.float64)
    gr = np.zeros(n, np.float64)
    gc = np.zeros((n, 3), np.float64)
    pane = lib().orc_pane_distance(W, fov)
    loss = lib().orc_loss_grad_fixed(
        xyz, r, rgb, n, np.ascontiguousarray(pose_colmajor, np.float32), pane, W, H,
        np.ascontiguousarray(hit, np.int32), np.ascontiguousarray(target_rgb, np.float32),
        np.ascontiguousarray(target_depth, np.float32), w_rgb, w_depth, gx, gr, gc)
    return loss, gx, gr, gc


def flann_knn(pts, q, k):
    """FLANN 1.9.1 CPU kd-tree, exact mode (oracle/_ref; built from the reference headers)."""
    p = _xyzw(pts)
    q = _xyzw(q)
    idx = np.empty((q.shape[0], k), np.int32)
    d2 = np.empty((q.shape[0], k), np.float32)
    ref_lib().flannref_knn(p, p.shape[0], q, q.shape[0], k, idx, d2)
    return idx, d2

"""Python mirror of the reference's hot-path interface over librm_hip.so.

``KDTreeFlann`` follows kdtree::KDTreeFlann (src/kdtree/kdtree_flann.h:15-43): Build,
Search / SearchKNN / SearchRadius with KDTreeSearchParams. ``PointcloudRayMarcher`` follows
src/ray_marching.h:11-44 (get_instance, render_to_texture). Device memory comes from torch
tensors (plumbing only); every computation runs in the HIP library.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _abi
from ._abi import check, lib, rm_tune, rm_view
from .scenes import fov_radians


def _ptr(t: torch.Tensor | None):
    if t is None:
        return None
    assert t.is_cuda and t.is_contiguous(), "device tensors must be contiguous CUDA tensors"
    return C.c_void_p(t.data_ptr())


def _f32(t: torch.Tensor, name: str, numel: int | None = None, cols: int | None = None):
    """Contiguous float32 CUDA view of t, checked before its pointer reaches the library (which
    reads exactly numel values, or rows of `cols` values)."""
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{name}: expected a CUDA tensor")
    if t.dtype != torch.float32:
        raise TypeError(f"{name}: expected float32, got {t.dtype}")
    if cols is not None and t.numel() % cols:
        raise ValueError(f"{name}: {t.numel()} values are not rows of {cols}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name}: expected {numel} values, got {t.numel()}")
    return t.contiguous()


class Context:
    """One rm_ctx: a device and a stream (the current torch stream, so torch ops and library
    kernels are ordered)."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        self.device = device
        h = C.c_void_p()
        check(lib().rm_ctx_create(device, C.byref(h)))
        self.h = h
        if use_torch_stream:
            s = torch.cuda.current_stream(device)
            check(lib().rm_ctx_set_stream(self.h, C.c_void_p(s.cuda_stream)), self.h)

    def sync(self):
        check(lib().rm_ctx_synchronize(self.h), self.h)

    def timing(self, on: bool):
        check(lib().rm_ctx_enable_timing(self.h, 1 if on else 0), self.h)

    def last_kernel_ms(self) -> float:
        return float(lib().rm_ctx_last_kernel_ms(self.h))

    def close(self):
        if self.h:
            lib().rm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx: Context | None = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(torch.cuda.current_device())
    return _default_ctx


# ------------------------------------------------------------------------------------------
# kd-tree
# ------------------------------------------------------------------------------------------

@dataclass
class KDTreeSearchParams:
    """src/kdtree/kdtree_search_param.h:5-23: KNN with knn, or Radius with radius/max_nn."""
    knn: int = 30
    radius: float = 0.0
    max_nn: int = 30
    search_type: str = "knn"

    @classmethod
    def KNN(cls, knn: int):
        return cls(knn=knn, search_type="knn")

    @classmethod
    def Radius(cls, radius: float, max_nn: int):
        return cls(radius=radius, max_nn=max_nn, search_type="radius")


class KDTreeFlann:
    """Drop-in for kdtree::KDTreeFlann over (N, 4) float32 CUDA tensors (stride 16 B)."""

    def __init__(self, ctx: Context | None = None, leaf_max: int = 0):
        self.ctx = ctx or default_context()
        self.leaf_max = leaf_max
        self.h = None
        self.n = 0

    def Build(self, pcd: torch.Tensor) -> bool:
        pcd = _f32(pcd, "Build: points", cols=4).reshape(-1, 4)
        self._free()
        h = C.c_void_p()
        check(lib().rm_tree_build(self.ctx.h, _ptr(pcd), pcd.shape[0], self.leaf_max,
                                  C.byref(h)), self.ctx.h)
        self.h = h
        self.n = int(pcd.shape[0])
        return True

    def info(self):
        n, nn, d, lm = C.c_int64(), C.c_int64(), C.c_int(), C.c_int()
        check(lib().rm_tree_info(self.h, C.byref(n), C.byref(nn), C.byref(d), C.byref(lm)))
        return {"n": n.value, "nodes": nn.value, "depth": d.value, "leaf_max": lm.value}

    def Search(self, query: torch.Tensor, param: KDTreeSearchParams):
        if param.search_type == "knn":
            return self.SearchKNN(query, param.knn)
        if param.search_type == "radius":
            return self.SearchRadius(query, param.radius, param.max_nn)
        return -1, None, None

    def SearchKNN(self, query: torch.Tensor, knn: int):
        """Returns (k * M, idx (M, k) int32, d2 (M, k) float32), as kdtree_flann.cu:65-80."""
        q = _f32(query, "SearchKNN: query", cols=4).reshape(-1, 4)
        m = q.shape[0]
        idx = torch.empty((m, knn), dtype=torch.int32, device=q.device)
        d2 = torch.empty((m, knn), dtype=torch.float32, device=q.device)
        check(lib().rm_knn(self.ctx.h, self.h, _ptr(q), m, knn, _ptr(idx), _ptr(d2)),
              self.ctx.h)
        return knn * m, idx, d2

    def SearchRadius(self, query: torch.Tensor, radius: float, max_nn: int):
        """Returns (total found, idx (M, max_nn), d2 (M, max_nn), per-query counts (M,)).

        Neighbours have d2 < radius**2 (kdtree_flann.cu:82-99, result_set.h:372-476). With
        max_nn == 0 only the uncapped per-query counts are computed (idx and d2 are (M, 0)),
        the counting mode of radiusSearchGpu (kdtree_cuda_3d_index.cu:436-452), whose return
        value is their sum."""
        q = _f32(query, "SearchRadius: query", cols=4).reshape(-1, 4)
        m = q.shape[0]
        idx = torch.empty((m, max_nn), dtype=torch.int32, device=q.device)
        d2 = torch.empty((m, max_nn), dtype=torch.float32, device=q.device)
        cnt = torch.empty((m,), dtype=torch.int32, device=q.device)
        check(lib().rm_radius_search(self.ctx.h, self.h, _ptr(q), m, float(radius), max_nn,
                                     _ptr(idx) if max_nn else None, _ptr(d2) if max_nn else None,
                                     _ptr(cnt)), self.ctx.h)
        return int(cnt.sum(dtype=torch.int64).item()), idx, d2, cnt

    def radii(self, max_radius: float = 0.1):
        """Per-point radii of main.cu:99-149 (unclamped, clamped), original order."""
        dev = torch.device("cuda", self.ctx.device)
        ru = torch.empty((self.n,), dtype=torch.float32, device=dev)
        rc = torch.empty((self.n,), dtype=torch.float32, device=dev)
        check(lib().rm_radii(self.ctx.h, self.h, float(max_radius), _ptr(ru), _ptr(rc)),
              self.ctx.h)
        return ru, rc

    def _free(self):
        if self.h:
            lib().rm_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._free()
        except Exception:
            pass


# ------------------------------------------------------------------------------------------
# Marcher
# ------------------------------------------------------------------------------------------

def fov_from_intrinsics(width: int, fx: float) -> float:
    """fov = 2 atanf(W / (2 f)), float32, main.cu:206."""
    return fov_radians(width, fx)


class PointcloudRayMarcher:
    """src/ray_marching.h:11-44 without the singleton: one instance per cloud."""

    def __init__(self, vertices: torch.Tensor, colors: torch.Tensor, radii: torch.Tensor,
                 tree: KDTreeFlann):
        self.ctx = tree.ctx
        self.tree = tree
        v = _f32(vertices, "vertices", numel=4 * tree.n)
        c = _f32(colors, "colors", numel=4 * tree.n)
        r = _f32(radii, "radii", numel=tree.n)
        self.n = tree.n
        h = C.c_void_p()
        check(lib().rm_cloud_create(self.ctx.h, tree.h, _ptr(v), _ptr(c), _ptr(r), self.n,
                                    C.byref(h)), self.ctx.h)
        self.h = h

    @classmethod
    def get_instance(cls, vertices, colors, radii, tree):
        return cls(vertices, colors, radii, tree)

    def render_to_texture(self, camera_pose_colmajor, fov_radians: float, width: int,
                          height: int, rgba: torch.Tensor | None = None,
                          hit: torch.Tensor | None = None):
        """Render one view into an (H, W, 4) float32 'texture' (row 0 = bottom, a = depth)
        and an (H, W) int32 hit map; returns (rgba, hit)."""
        dev = torch.device("cuda", self.ctx.device)
        if rgba is None:
            rgba = torch.empty((height, width, 4), dtype=torch.float32, device=dev)
        else:
            rgba = _f32(rgba, "rgba", numel=4 * width * height)
        if hit is None:
            hit = torch.empty((height, width), dtype=torch.int32, device=dev)
        elif hit.dtype != torch.int32 or hit.numel() != width * height or not hit.is_cuda:
            raise ValueError("hit: expected an int32 CUDA tensor of width * height values")
        pose = (C.c_float * 16)(*np.asarray(camera_pose_colmajor, dtype=np.float32).ravel())
        check(lib().rm_render(self.ctx.h, self.h, pose, float(fov_radians), width, height,
                              _ptr(rgba), _ptr(hit)), self.ctx.h)
        return rgba, hit

    def set_radii(self, radii: torch.Tensor):
        check(lib().rm_cloud_set_radii(self.ctx.h, self.h,
                                       _ptr(_f32(radii, "set_radii", numel=self.n))), self.ctx.h)

    def set_colors(self, colors: torch.Tensor):
        check(lib().rm_cloud_set_colors(self.ctx.h, self.h,
                                        _ptr(_f32(colors, "set_colors", numel=4 * self.n))),
              self.ctx.h)

    def refit(self, xyz: torch.Tensor):
        check(lib().rm_cloud_refit(self.ctx.h, self.h, _ptr(_f32(xyz, "refit", numel=3 * self.n))),
              self.ctx.h)

    def set_culling(self, enable: bool):
        """March each view over the points inside its frustum (is_inside of
        src/ray_marching.cu:231-258, commented out in the reference, so off by default).
        Batches keep the setting they were created with."""
        check(lib().rm_cloud_set_culling(self.ctx.h, self.h, 1 if enable else 0), self.ctx.h)

    def close(self):
        if self.h:
            lib().rm_cloud_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_views(views, offsets=None):
    """rm_view array from objects with pose (4x4 row-major camera-to-world), width, height,
    fov (radians)."""
    arr = (rm_view * len(views))()
    off = 0
    for i, v in enumerate(views):
        arr[i].pose[:] = [float(x) for x in np.asarray(v.pose, np.float32).T.reshape(-1)]
        arr[i].fov_radians = float(v.fov)
        arr[i].width = int(v.width)
        arr[i].height = int(v.height)
        arr[i].offset = off if offsets is None else int(offsets[i])
        off += int(v.width) * int(v.height)
    return arr, off


class Batch:
    """Persistent, graph-launched march over many views (rm_batch_*)."""

    def __init__(self, marcher: PointcloudRayMarcher, views, tune: dict | None = None,
                 with_hit: bool = True):
        self.m = marcher
        self.ctx = marcher.ctx
        self.views_c, self.total = make_views(views)
        dev = torch.device("cuda", self.ctx.device)
        self.rgba = torch.empty((self.total, 4), dtype=torch.float32, device=dev)
        self.hit = (torch.empty((self.total,), dtype=torch.int32, device=dev)
                    if with_hit else None)
        t = None
        if tune is not None:
            t = rm_tune(**tune)
        h = C.c_void_p()
        check(lib().rm_batch_create(self.ctx.h, marcher.h, self.views_c, len(views),
                                    C.byref(t) if t is not None else None, _ptr(self.rgba),
                                    _ptr(self.hit), C.byref(h)), self.ctx.h)
        self.h = h

    def autotune(self, sample_stride: int = 16) -> dict:
        t = rm_tune()
        check(lib().rm_batch_autotune(self.h, sample_stride, C.byref(t)), self.ctx.h)
        return t.as_dict()

    def tune(self) -> dict:
        t = rm_tune()
        check(lib().rm_batch_get_tune(self.h, C.byref(t)), self.ctx.h)
        return t.as_dict()

    def launch(self):
        check(lib().rm_batch_launch(self.h), self.ctx.h)

    def close(self):
        if self.h:
            lib().rm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


"""Seeded synthetic point clouds and camera views for tests and benchmarks.

The reference ships no data. These generators produce the workloads the renderer is measured
on: a box scene (points on the faces of random axis-aligned boxes) and a room scene (walls,
floor, ceiling and furniture boxes with position noise, 4x density variation and uniform
outliers). All randomness comes from ``numpy.random.default_rng(seed)`` (PCG64), so a seed
fixes the cloud bit for bit on any machine with the same numpy.

Conventions follow the reference:
  * points are float32 xyz, colours uint8 RGB (src/pointcloud.cpp:43-58 loads them so);
  * a camera pose is a 4x4 camera-to-world matrix in the OpenGL convention (camera looks down
    -z, +y up), passed to the renderer column-major (glm layout, src/ray_marching.cu:75-77);
  * intrinsics give W = 2 cx, H = 2 cy and fov = 2 atan(W / (2 fx)) (main.cu:202-206).
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass

import numpy as np

_libm = ctypes.CDLL("libm.so.6")
_libm.atanf.restype = ctypes.c_float
_libm.atanf.argtypes = [ctypes.c_float]


def fov_radians(width: int, fx: float) -> float:
    """fov = 2 * atanf(W / (2 f)) in float32 with the C library's atanf, as main.cu:206."""
    w = np.float32(width)
    f = np.float32(fx)
    a = np.float32(_libm.atanf(float(w / (np.float32(2.0) * f))))
    return float(np.float32(2.0) * a)


@dataclass
class Cloud:
    xyz: np.ndarray            # (N, 3) float32
    rgb: np.ndarray            # (N, 3) uint8

    @property
    def n(self) -> int:
        return int(self.xyz.shape[0])

    def xyzw(self) -> np.ndarray:
        """(N, 4) float32 with w = 1, the layout of main.cu:88-91."""
        out = np.ones((self.n, 4), dtype=np.float32)
        out[:, :3] = self.xyz
        return out

    def colors4(self) -> np.ndarray:
        """(N, 4) float32 (r, g, b)/255 with a = 1, as src/pointcloud.cpp:55 builds them."""
        out = np.ones((self.n, 4), dtype=np.float32)
        out[:, :3] = self.rgb.astype(np.float32) / np.float32(255.0)
        return out


@dataclass
class View:
    pose: np.ndarray           # (4, 4) float32 camera-to-world, row-major numpy layout
    width: int
    height: int
    fx: float
    name: str = ""

    @property
    def fov(self) -> float:
        """fov = 2 atanf(W / (2 f)) in float32, as main.cu:206 computes it."""
        return fov_radians(self.width, self.fx)

    def pose_colmajor(self) -> np.ndarray:
        """16 float32 in glm's column-major order."""
        return np.ascontiguousarray(self.pose.T.reshape(-1).astype(np.float32))

    @property
    def pixels(self) -> int:
        return self.width * self.height


def look_at(eye, target, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    """Camera-to-world pose with the camera at ``eye`` looking at ``target``."""
    eye = np.asarray(eye, dtype=np.float64)
    f = np.asarray(target, dtype=np.float64) - eye
    f /= np.linalg.norm(f)
    return pose_from_forward(eye, f, up)


def pose_from_forward(eye, forward, up=(0.0, 0.0, 1.0)) -> np.ndarray:
    f = np.asarray(forward, dtype=np.float64)
    f = f / np.linalg.norm(f)
    upv = np.asarray(up, dtype=np.float64)
    r = np.cross(f, upv)
    if np.linalg.norm(r) < 1e-9:
        r = np.cross(f, np.array([0.0, 1.0, 0.0]))
    r /= np.linalg.norm(r)
    u = np.cross(r, f)
    m = np.eye(4)
    m[:3, 0] = r
    m[:3, 1] = u
    m[:3, 2] = -f
    m[:3, 3] = eye
    return m.astype(np.float32)


# --------------------------------------------------------------------------------------------
# Surface sampling helpers
# --------------------------------------------------------------------------------------------

def _sample_rects(rng, rects, weights, n):
    """Uniform samples on a set of axis-aligned rectangles chosen with prob ~ weights.

    rects: (R, 3, 3) array of (origin, edge_u, edge_v) triples. Returns (points, rect_id).
    """
    w = np.asarray(weights, dtype=np.float64)
    p = w / w.sum()
    rid = rng.choice(len(rects), size=n, p=p)
    s = rng.random((n, 2))
    o = rects[rid, 0]
    pts = o + rects[rid, 1] * s[:, :1] + rects[rid, 2] * s[:, 1:2]
    return pts, rid


def _box_faces(lo, hi, bottom=True):
    """Six (or five, without the bottom) faces of the box [lo, hi] as rects."""
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    d = hi - lo
    ex, ey, ez = np.array([d[0], 0, 0]), np.array([0, d[1], 0]), np.array([0, 0, d[2]])
    faces = [
        (lo, ey, ez), (lo + ex, ey, ez),          # x = lo, x = hi
        (lo, ex, ez), (lo + ey, ex, ez),          # y = lo, y = hi
        (lo + ez, ex, ey),                        # z = hi (top)
    ]
    if bottom:
        faces.append((lo, ex, ey))                # z = lo
    return [np.stack(f) for f in faces]


def _rect_area(r):
    return float(np.linalg.norm(np.cross(r[1], r[2])))


def _colours(rng, base, rid, n):
    noise = rng.uniform(-10.0, 10.0, size=(n, 3))
    c = base[rid] + noise
    return np.clip(np.rint(c), 0, 255).astype(np.uint8)


# --------------------------------------------------------------------------------------------
# Box scene
# --------------------------------------------------------------------------------------------

def box_scene(n_points: int = 100_000, seed: int = 1, n_boxes: int = 20,
              extent: float = 4.0) -> Cloud:
    """Points uniform on the faces of ``n_boxes`` random axis-aligned boxes in a cube of side
    ``extent`` metres (area-weighted face choice). RGB is uniform per box plus U(-10, 10)
    per point, clamped."""
    rng = np.random.default_rng(seed)
    rects, weights, owner = [], [], []
    for b in range(n_boxes):
        size = rng.uniform(0.2, 1.2, size=3) * extent / 4.0
        lo = rng.uniform(0.0, extent - size)
        for f in _box_faces(lo, lo + size):
            rects.append(f)
            weights.append(_rect_area(f))
            owner.append(b)
    rects = np.stack(rects)
    pts, rid = _sample_rects(rng, rects, weights, n_points)
    base = rng.uniform(0.0, 255.0, size=(n_boxes, 3))
    rgb = _colours(rng, base, np.asarray(owner)[rid], n_points)
    return Cloud(pts.astype(np.float32), rgb)


def box_scene_view(cloud: Cloud, width=640, height=480, fx=500.0, distance=3.0,
                   seed: int = 1, elevation_deg: float = 25.0) -> View:
    """One view ``distance`` metres from the centroid, looking at it."""
    rng = np.random.default_rng(seed + 1000)
    c = cloud.xyz.astype(np.float64).mean(axis=0)
    az = rng.uniform(-math.pi, math.pi)
    el = math.radians(elevation_deg)
    eye = c + distance * np.array([math.cos(el) * math.cos(az), math.cos(el) * math.sin(az),
                                   math.sin(el)])
    return View(look_at(eye, c), width, height, fx, name=f"box_s{seed}")


# --------------------------------------------------------------------------------------------
# Room scene
# --------------------------------------------------------------------------------------------

@dataclass
class RoomSpec:
    size: tuple = (10.0, 8.0, 3.0)
    n_boxes: int = 200
    noise_sigma: float = 0.005
    outlier_frac: float = 0.005
    tile: float = 1.0          # density tiles on walls / floor / ceiling (metres)
    density_ratio: float = 4.0


def room_scene(n_points: int = 2_000_000, seed: int = 2, spec: RoomSpec | None = None) -> Cloud:
    """Walls, floor and ceiling of a room plus random boxes standing on the floor; Gaussian
    position noise; surface density varied ``density_ratio``x across 1 m tiles and box faces;
    a fraction of uniform outliers inside the room volume."""
    spec = spec or RoomSpec()
    rng = np.random.default_rng(seed)
    X, Y, Z = spec.size
    rects, weights, base_id = [], [], []

    def tiles(origin, eu, ev):
        lu, lv = np.linalg.norm(eu), np.linalg.norm(ev)
        nu, nv = max(1, int(round(lu / spec.tile))), max(1, int(round(lv / spec.tile)))
        du, dv = eu / nu, ev / nv
        for i in range(nu):
            for j in range(nv):
                yield np.stack([origin + du * i + dv * j, du, dv])

    room_faces = [
        (np.array([0, 0, 0.0]), np.array([X, 0, 0.0]), np.array([0, Y, 0.0])),   # floor
        (np.array([0, 0, Z]), np.array([X, 0, 0.0]), np.array([0, Y, 0.0])),     # ceiling
        (np.array([0, 0, 0.0]), np.array([X, 0, 0.0]), np.array([0, 0, Z])),     # y = 0
        (np.array([0, Y, 0.0]), np.array([X, 0, 0.0]), np.array([0, 0, Z])),     # y = Y
        (np.array([0, 0, 0.0]), np.array([0, Y, 0.0]), np.array([0, 0, Z])),     # x = 0
        (np.array([X, 0, 0.0]), np.array([0, Y, 0.0]), np.array([0, 0, Z])),     # x = X
    ]
    for f in room_faces:
        for r in tiles(*f):
            rects.append(r)
    n_room = len(rects)
    for b in range(spec.n_boxes):
        sx, sy = rng.uniform(0.2, 1.5, size=2)
        sz = rng.uniform(0.2, 2.0)
        lo = np.array([rng.uniform(0.0, X - sx), rng.uniform(0.0, Y - sy), 0.0])
        hi = lo + np.array([sx, sy, min(sz, Z * 0.9)])
        for f in _box_faces(lo, hi, bottom=False):
            rects.append(f)
    rects = np.stack(rects)
    dens = 1.0 + (spec.density_ratio - 1.0) * rng.random(len(rects))
    dens[rng.integers(0, len(rects), size=max(1, len(rects) // 8))] = 1.0
    dens[rng.integers(0, len(rects), size=max(1, len(rects) // 8))] = spec.density_ratio
    weights = np.array([_rect_area(r) for r in rects]) * dens
    n_out = int(round(n_points * spec.outlier_frac))
    n_surf = n_points - n_out
    pts, rid = _sample_rects(rng, rects, weights, n_surf)
    pts = pts + rng.normal(0.0, spec.noise_sigma, size=pts.shape)
    # colour: one base colour per room face group of tiles and per box
    group = np.where(np.arange(len(rects)) < n_room, np.arange(len(rects)) % 7,
                     7 + (np.arange(len(rects)) - n_room) // 5)
    base = rng.uniform(0.0, 255.0, size=(int(group.max()) + 1, 3))
    rgb = _colours(rng, base, group[rid], n_surf)
    out = rng.uniform([0, 0, 0], [X, Y, Z], size=(n_out, 3))
    out_rgb = rng.integers(0, 256, size=(n_out, 3)).astype(np.uint8)
    xyz = np.concatenate([pts, out]).astype(np.float32)
    rgb = np.concatenate([rgb, out_rgb])
    perm = rng.permutation(n_points)
    return Cloud(np.ascontiguousarray(xyz[perm]), np.ascontiguousarray(rgb[perm]))


def room_views(n_views: int, spec: RoomSpec | None = None, seed: int = 102,
               sizes=((1067, 690, 803.8),), margin: float = 0.3) -> list[View]:
    """Camera positions uniform inside the room (``margin`` from the walls), yaw U(-pi, pi),
    pitch U(-20, 20) degrees. ``sizes`` is cycled: a list of (W, H, fx) triples with the
    principal point at the image centre."""
    spec = spec or RoomSpec()
    rng = np.random.default_rng(seed)
    X, Y, Z = spec.size
    views = []
    for i in range(n_views):
        eye = rng.uniform([margin, margin, margin], [X - margin, Y - margin, Z - margin])
        yaw = rng.uniform(-math.pi, math.pi)
        pitch = math.radians(rng.uniform(-20.0, 20.0))
        fwd = np.array([math.cos(pitch) * math.cos(yaw), math.cos(pitch) * math.sin(yaw),
                        math.sin(pitch)])
        w, h, fx = sizes[i % len(sizes)]
        views.append(View(pose_from_forward(eye, fwd), int(w), int(h), float(fx),
                          name=f"room_v{i}"))
    return views


# --------------------------------------------------------------------------------------------
# The measured workloads
# --------------------------------------------------------------------------------------------

def workload_box_100k():
    """100k points on 20 boxes in a 4 m cube, one 640x480 view, fx = fy = 500, 3 m away."""
    cloud = box_scene(100_000, seed=1)
    return cloud, [box_scene_view(cloud, 640, 480, 500.0, 3.0, seed=1)], 0.1


def workload_room_2m(n_views: int = 32):
    """2M points in a 10x8x3 m room with 200 boxes; 32 views at 1067x690, fx = 803.8."""
    spec = RoomSpec()
    cloud = room_scene(2_000_000, seed=2, spec=spec)
    return cloud, room_views(n_views, spec, seed=102, sizes=((1067, 690, 803.8),)), 0.05


def workload_box_1m_diff(n_views: int = 16):
    """Differentiable workload: 1M points on the box scene (seed 4), 16 views at 512x512,
    fx = fy = 450, each 3 m from the centroid at its own azimuth."""
    cloud = box_scene(1_000_000, seed=4)
    views = [box_scene_view(cloud, 512, 512, 450.0, distance=3.0, seed=400 + i)
             for i in range(n_views)]
    return cloud, views, 0.1


def room_10m_spec() -> RoomSpec:
    return RoomSpec(size=(20.0, 16.0, 3.0), n_boxes=800)


def mixed_sizes_256():
    """96 views at 640x480 (fx 500), 96 at 1067x690 and 64 at 1920x1080 (fx 803.8)."""
    return [(640, 480, 500.0)] * 96 + [(1067, 690, 803.8)] * 96 + [(1920, 1080, 803.8)] * 64


def workload_room_10m(n_points: int = 10_000_000):
    """10M points in a 20x16x3 m room; 256 views of mixed resolution."""
    spec = room_10m_spec()
    cloud = room_scene(n_points, seed=3, spec=spec)
    sizes = mixed_sizes_256()
    rng = np.random.default_rng(1103)
    order = rng.permutation(len(sizes))
    sizes = [sizes[i] for i in order]
    views = room_views(len(sizes), spec, seed=103, sizes=sizes)
    return cloud, views, 0.05


def scale_50m_spec() -> RoomSpec:
    return RoomSpec(size=(40.0, 30.0, 4.0), n_boxes=3000, outlier_frac=0.01)


def workload_scale_50m(n_points: int = 50_000_000, n_views: int = 1024):
    """50M points in a 40x30x4 m room (3000 boxes, 1 % outliers, radii clamped at 0.05 m);
    1024 views at 1920x1080, fx = fy = 803.8 (2.12G rays)."""
    spec = scale_50m_spec()
    cloud = room_scene(n_points, seed=5, spec=spec)
    views = room_views(n_views, spec, seed=105, sizes=((1920, 1080, 803.8),))
    return cloud, views, 0.05


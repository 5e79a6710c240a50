"""Differentiable march over torch tensors.

``DiffRenderer.render(xyz, radii, rgb)`` returns (rgb images (V, H, W, 3), depth (V, H, W),
hit (V, H, W) int32). It is a torch::autograd::Function written in libtorch C++
(torch_ext/rm_torch.cpp) over librm_hip.so: the forward marches every view (hit selection as in
the reference march) and reports the hit point's colour and the analytic entry depth of the ray
into its sphere; the backward holds the hit map fixed and returns d/d(centres, radii, rgb).
The search tree is built once over the initial centres and refitted to every new set of
centres, which keeps kNN exact.
"""
from __future__ import annotations

import os

import numpy as np
import torch

from . import api

_HERE = os.path.dirname(os.path.abspath(__file__))
_EXT = os.path.join(_HERE, "_rm_torch.so")


_loaded = False


def _ext():
    global _loaded
    if not os.path.exists(_EXT):
        raise RuntimeError(f"{_EXT} is missing: build it with "
                           "`python renderer_ray_marching_amd/torch_ext/build.py`")
    if not _loaded:
        torch.ops.load_library(_EXT)
        _loaded = True
    return torch.ops.rm_diff


class DiffRenderer:
    def __init__(self, xyz0: torch.Tensor, radii0: torch.Tensor, rgb0: torch.Tensor, views,
                 ctx: api.Context | None = None, leaf_max: int = 0):
        ext = _ext()
        self.ctx = ctx or api.default_context()
        n = int(xyz0.shape[0])
        self.n = n
        xyzw = torch.cat([xyz0.detach().float(), torch.ones((n, 1), device=xyz0.device)], 1)
        xyzw = xyzw.contiguous()
        col4 = torch.cat([rgb0.detach().float(), torch.ones((n, 1), device=xyz0.device)], 1)
        self.tree = api.KDTreeFlann(self.ctx, leaf_max=leaf_max)
        self.tree.Build(xyzw)
        self.marcher = api.PointcloudRayMarcher(xyzw, col4.contiguous(),
                                                radii0.detach().float().contiguous(), self.tree)
        self.scene = torch.classes.rm_diff.DiffScene(self.ctx.h.value, self.marcher.h.value, n)
        self.views = list(views)
        for v in self.views:
            self.scene.add_view([float(x) for x in v.pose_colmajor()], float(v.fov),
                                int(v.width), int(v.height))
        self._render = ext.render

    def render(self, xyz: torch.Tensor, radii: torch.Tensor, rgb: torch.Tensor):
        """Views of one size: img (V, H, W, 3), depth (V, H, W), hit (V, H, W). Views of mixed
        sizes: the views concatenated, img (P, 3), depth (P,), hit (P,), with view v at
        ``view_offsets()[v]``."""
        return self._render(xyz, radii, rgb, self.scene)

    def view_offsets(self) -> list[int]:
        """First pixel of each view in the concatenated outputs, then the total."""
        return [int(x) for x in self.scene.view_offsets()]


def refine(renderer: DiffRenderer, xyz, radii, rgb, target_rgb, target_depth, steps=200,
           lr_xyz=1e-3, lr_r=5e-4, lr_rgb=1e-2, w_rgb=1.0, w_depth=1.0, callback=None):
    """Adam on (centres, radii, rgb) against target views: L2 on RGB plus L1 on depth over the
    pixels where both the render and the target hit something."""
    xyz = xyz.clone().requires_grad_(True)
    radii = radii.clone().requires_grad_(True)
    rgb = rgb.clone().requires_grad_(True)
    opt = torch.optim.Adam([{"params": [xyz], "lr": lr_xyz}, {"params": [radii], "lr": lr_r},
                            {"params": [rgb], "lr": lr_rgb}])
    tmask = target_depth > 0
    losses = []
    for it in range(steps):
        opt.zero_grad(set_to_none=True)
        img, depth, hit = renderer.render(xyz, radii, rgb)
        m = (hit >= 0) & tmask
        mf = m.float()
        loss = (w_rgb * ((img - target_rgb) ** 2).sum(-1) * mf).sum() + \
               (w_depth * (depth - target_depth).abs() * mf).sum()
        loss.backward()
        opt.step()
        with torch.no_grad():
            radii.clamp_(min=1e-4)
            rgb.clamp_(0.0, 1.0)
        losses.append(float(loss.item()))
        if callback:
            callback(it, losses[-1])
    return xyz.detach(), radii.detach(), rgb.detach(), losses


def perturbed_cloud(xyz: np.ndarray, radii: np.ndarray, rgb_u8: np.ndarray, seed: int = 104):
    """The perturbation of the differentiable workload: centres jittered with sigma = 1 cm,
    radii scaled by U(0.8, 1.2), RGB moved by U(-20, 20) (on the 0..255 scale)."""
    rng = np.random.default_rng(seed)
    x = (xyz + rng.normal(0.0, 0.01, xyz.shape)).astype(np.float32)
    r = (radii * rng.uniform(0.8, 1.2, radii.shape)).astype(np.float32)
    c = np.clip(rgb_u8.astype(np.float32) + rng.uniform(-20, 20, rgb_u8.shape), 0, 255)
    return x, r, (c / np.float32(255.0)).astype(np.float32)

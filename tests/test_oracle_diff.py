"""Pins the oracle's fp64 gradient of the fixed-hit differentiable form (orc_loss_grad_fixed)
against central finite differences of its own loss, and the analytic entry depth against a
closed form."""
import math

import numpy as np

from oracle import binding as ob
from renderer_ray_marching_amd import scenes


def _setup(seed=4):
    cloud = scenes.box_scene(3000, seed=seed)
    p = cloud.xyzw()
    tree = ob.RefTree(p, 64)
    _, rc, _ = tree.radii(0.05)
    rc = np.maximum(rc, np.float32(0.02))
    view = scenes.box_scene_view(cloud, 64, 48, 50.0, distance=5.0, seed=seed)
    rgba, hit, _ = tree.render(rc, cloud.colors4(), view.pose_colmajor(), view.fov, 64, 48)
    return cloud, rc, view, hit


def test_oracle_gradient_matches_central_fd():
    cloud, rc, view, hit = _setup()
    xyz = cloud.xyz.astype(np.float64)
    r = rc.astype(np.float64)
    rgb = cloud.colors4()[:, :3].astype(np.float64)
    rng = np.random.default_rng(0)
    W, H = view.width, view.height
    tgt_rgb = rng.random((W * H, 3)).astype(np.float32)
    tgt_d = rng.uniform(0.5, 9.0, W * H).astype(np.float32)
    pose = view.pose_colmajor()
    loss, gx, gr, gc = ob.loss_grad_fixed(xyz, r, rgb, pose, view.fov, W, H, hit, tgt_rgb,
                                          tgt_d, 0.7, 1.3)
    hits = np.unique(hit[hit >= 0])
    assert hits.size > 20
    h = 1e-6
    checked = 0
    for i in hits[:40]:
        for a in range(3):
            for arr, g in ((xyz, gx), (rgb, gc)):
                old = arr[i, a]
                arr[i, a] = old + h
                lp = ob.loss_grad_fixed(xyz, r, rgb, pose, view.fov, W, H, hit, tgt_rgb, tgt_d,
                                        0.7, 1.3)[0]
                arr[i, a] = old - h
                lm = ob.loss_grad_fixed(xyz, r, rgb, pose, view.fov, W, H, hit, tgt_rgb, tgt_d,
                                        0.7, 1.3)[0]
                arr[i, a] = old
                fd = (lp - lm) / (2 * h)
                assert abs(fd - g[i, a]) <= 1e-4 * max(1.0, abs(g[i, a])), (i, a, fd, g[i, a])
                checked += 1
        old = r[i]
        r[i] = old + h
        lp = ob.loss_grad_fixed(xyz, r, rgb, pose, view.fov, W, H, hit, tgt_rgb, tgt_d, 0.7, 1.3)[0]
        r[i] = old - h
        lm = ob.loss_grad_fixed(xyz, r, rgb, pose, view.fov, W, H, hit, tgt_rgb, tgt_d, 0.7, 1.3)[0]
        r[i] = old
        fd = (lp - lm) / (2 * h)
        assert abs(fd - gr[i]) <= 1e-4 * max(1.0, abs(gr[i]))
    assert checked > 100


def test_entry_depth_closed_form():
    # camera at the origin looking down -z; a sphere on the axis: the centre pixel's ray enters
    # it at distance |c| - r, and the L1 depth gradient w.r.t. r is -1 there (dt/dr = -r/r).
    W, H = 8, 6
    pose = np.eye(4, dtype=np.float32).T.ravel()
    fov = scenes.fov_radians(W, 4.0)
    xyz = np.array([[0.0, 0.0, -3.0]])
    r = np.array([0.5])
    rgb = np.array([[0.1, 0.2, 0.3]])
    hit = np.full(W * H, -1, np.int32)
    lin = (H // 2) * W + W // 2
    hit[lin] = 0
    tgt_rgb = np.zeros((W * H, 3), np.float32)
    tgt_d = np.zeros(W * H, np.float32)
    loss, gx, gr, gc = ob.loss_grad_fixed(xyz, r, rgb, pose, fov, W, H, hit, tgt_rgb, tgt_d,
                                          1.0, 1.0)
    assert math.isclose(loss, 2.5 + (0.01 + 0.04 + 0.09), rel_tol=1e-12)
    assert math.isclose(gr[0], -1.0, rel_tol=1e-12)
    # moving the centre away along -z increases the depth one-for-one
    assert math.isclose(gx[0, 2], -1.0, rel_tol=1e-12) and abs(gx[0, 0]) < 1e-12
    assert np.allclose(gc[0], 2 * rgb[0])

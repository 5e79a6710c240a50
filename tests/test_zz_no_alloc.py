"""The allocation-free batched march (rm_render_batch_noalloc) against rm_render_batch, through the
C ABI: equal outputs on two calls in a row (the second smaller, so stale scratch contents from the
first would show), and its limits rejected with clear errors."""
import ctypes as C

import numpy as np
import pytest
import torch

from renderer_ray_marching_amd import _abi, api, scenes

pytestmark = pytest.mark.gpu


def _scene(ctx, n, seed):
    cloud = scenes.room_scene(n, seed=seed)
    p = torch.from_numpy(cloud.xyzw()).cuda()
    tree = api.KDTreeFlann(ctx)
    tree.Build(p)
    _, rc = tree.radii(0.05)
    m = api.PointcloudRayMarcher(p, torch.from_numpy(cloud.colors4()).cuda(), rc, tree)
    return cloud, tree, m


def _render(fn, ctx, m, views):
    arr, total = api.make_views(views)
    rgba = torch.full((total, 4), -7.0, dtype=torch.float32, device="cuda")
    hit = torch.full((total,), -7, dtype=torch.int32, device="cuda")
    rc = fn(ctx.h, m.h, arr, len(views), api._ptr(rgba), api._ptr(hit))
    _abi.check(rc, ctx.h)
    ctx.sync()
    return rgba.cpu().numpy(), hit.cpu().numpy()


def test_noalloc_equals_render_batch_on_two_calls(rm_ctx):
    L = _abi.lib()
    print("case: noalloc vs rm_render_batch, 150k-point room, 9 then 3 views", flush=True)
    _, _, m = _scene(rm_ctx, 150_000, seed=21)
    big = scenes.room_views(9, seed=210, sizes=((160, 120, 120.0), (96, 64, 80.0),
                                                 (200, 90, 150.0)))
    small = scenes.room_views(3, seed=211, sizes=((64, 48, 50.0), (33, 17, 30.0)))
    for views in (big, small):
        ref_rgba, ref_hit = _render(L.rm_render_batch, rm_ctx, m, views)
        rgba, hit = _render(L.rm_render_batch_noalloc, rm_ctx, m, views)
        assert (hit == ref_hit).all()
        assert np.array_equal(rgba, ref_rgba)
        assert (hit >= 0).mean() > 0.5          # the views look at surfaces
    # a cloud smaller than ten points (no warm start) still matches
    tiny = scenes.box_scene(7, seed=3)
    p = torch.from_numpy(tiny.xyzw()).cuda()
    tree = api.KDTreeFlann(rm_ctx)
    tree.Build(p)
    _, rc = tree.radii(0.1)
    mt = api.PointcloudRayMarcher(p, torch.from_numpy(tiny.colors4()).cuda(), rc, tree)
    v = [scenes.box_scene_view(tiny, 40, 30, 30.0, 3.0, seed=1)]
    a = _render(L.rm_render_batch, rm_ctx, mt, v)
    b = _render(L.rm_render_batch_noalloc, rm_ctx, mt, v)
    assert (a[1] == b[1]).all() and np.array_equal(a[0], b[0])
    # a box scene seen from outside: rays that leave the cloud end at the boxes of the default
    # early-miss frontier, kept in the same array; equal to the plain kernel on every pixel
    print("case: noalloc vs the plain kernel, 30k-point box scene seen from outside", flush=True)
    box = scenes.box_scene(30_000, seed=9)
    p = torch.from_numpy(box.xyzw()).cuda()
    tree = api.KDTreeFlann(rm_ctx)
    tree.Build(p)
    _, rc = tree.radii(0.05)
    mb = api.PointcloudRayMarcher(p, torch.from_numpy(box.colors4()).cuda(), rc, tree)
    vs = [scenes.box_scene_view(box, 96, 72, 80.0, distance=4.0, seed=s) for s in (1, 2)]
    rgba, hit = _render(L.rm_render_batch_noalloc, rm_ctx, mb, vs)
    off = 0
    for view in vs:
        n = view.width * view.height
        prgba, phit = mb.render_to_texture(view.pose_colmajor(), view.fov, view.width,
                                           view.height)
        rm_ctx.sync()
        assert (hit[off:off + n] == phit.cpu().numpy().ravel()).all()
        assert np.array_equal(rgba[off:off + n], prgba.cpu().numpy().reshape(-1, 4))
        off += n
    assert (hit < 0).mean() > 0.2


def test_noalloc_limits(rm_ctx):
    L = _abi.lib()
    _, _, m = _scene(rm_ctx, 20_000, seed=22)
    n = _abi.RM_NOALLOC_MAX_VIEWS + 1
    views = scenes.room_views(n, seed=220, sizes=((8, 8, 8.0),))
    arr, total = api.make_views(views)
    rgba = torch.empty((total, 4), dtype=torch.float32, device="cuda")
    rc = L.rm_render_batch_noalloc(rm_ctx.h, m.h, arr, n, api._ptr(rgba), None)
    assert rc == _abi.RM_ERR_UNSUPPORTED
    assert b"RM_NOALLOC_MAX_VIEWS" in L.rm_last_error(rm_ctx.h)
    # exactly the maximum is accepted
    rc = L.rm_render_batch_noalloc(rm_ctx.h, m.h, arr, n - 1, api._ptr(rgba), None)
    assert rc == _abi.RM_OK
    rm_ctx.sync()
    m.set_culling(True)
    rc = L.rm_render_batch_noalloc(rm_ctx.h, m.h, arr, 2, api._ptr(rgba), None)
    assert rc == _abi.RM_ERR_UNSUPPORTED
    m.set_culling(False)
    assert L.rm_render_batch_noalloc(rm_ctx.h, m.h, arr, 0, api._ptr(rgba), None) == \
        _abi.RM_ERR_INVALID
    bad = (_abi.rm_view * 1)()
    bad[0] = arr[0]
    bad[0].width = 0
    assert L.rm_render_batch_noalloc(rm_ctx.h, m.h, bad, 1, api._ptr(rgba), None) == \
        _abi.RM_ERR_INVALID
    assert L.rm_render_batch_noalloc(rm_ctx.h, m.h, arr, 1, None, None) == _abi.RM_ERR_INVALID

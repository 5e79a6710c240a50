"""Parity of the HIP path (through the C ABI) with the CPU oracle on the same seeded inputs.

Integer/index outputs must be bit-exact; float outputs (d2, radii, depth, colour) are compared
exactly too, since both sides evaluate the same float expressions: the squared distance as
fma(dz, dz, fma(dy, dy, dx * dx)), everything else without contraction.
"""
import numpy as np
import pytest
import torch

import fp32
from oracle import binding as ob
from renderer_ray_marching_amd import _abi, api, scenes
from renderer_ray_marching_amd._abi import RMError

pytestmark = pytest.mark.gpu


def _dev(a, dtype=torch.float32):
    return torch.from_numpy(np.ascontiguousarray(a)).to(device="cuda", dtype=dtype)


def _tree(ctx, p, leaf=0):
    t = api.KDTreeFlann(ctx, leaf_max=leaf)
    t.Build(_dev(p))
    return t


def _queries(p, seed, m):
    rng = np.random.default_rng(seed)
    lo, hi = p[:, :3].min(0), p[:, :3].max(0)
    q = np.concatenate([
        p[rng.integers(0, p.shape[0], m // 3)],
        p[rng.integers(0, p.shape[0], m // 3)] + rng.normal(0, 0.03, (m // 3, 4)).astype(np.float32),
        np.concatenate([rng.uniform(lo - 1, hi + 1, (m - 2 * (m // 3), 3)),
                        np.ones((m - 2 * (m // 3), 1))], 1).astype(np.float32)])
    return np.ascontiguousarray(q, np.float32)


# ------------------------------------------------------------------------------------------
# kNN
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("k", [1, 2, 3, 5, 10, 16, 32])
@pytest.mark.parametrize("leaf", [0, 1, 8, 64])
def test_knn_bit_exact_vs_oracle(rm_ctx, k, leaf):
    p = scenes.box_scene(20000, seed=1).xyzw()
    q = _queries(p, 3, 3000)
    t = _tree(rm_ctx, p, leaf)
    _, idx, d2 = t.SearchKNN(_dev(q), k)
    rm_ctx.sync()
    oi, od, _ = ob.RefTree(p, 64).knn(q, k)
    assert (idx.cpu().numpy() == oi).all()
    assert (d2.cpu().numpy() == od).all()


def test_knn_room_scene_and_info(rm_ctx):
    c = scenes.room_scene(200_000, seed=2)
    p = c.xyzw()
    q = _queries(p, 4, 20000)
    t = _tree(rm_ctx, p)
    info = t.info()
    assert info["n"] == p.shape[0] and info["depth"] < 60
    _, idx, d2 = t.SearchKNN(_dev(q), 10)
    rm_ctx.sync()
    oi, od, _ = ob.RefTree(p, 64).knn(q, 10)
    assert (idx.cpu().numpy() == oi).all() and (d2.cpu().numpy() == od).all()


def test_knn_ties_duplicates_and_padding(rm_ctx):
    h = np.float32(0.25)
    g = np.arange(6, dtype=np.float32) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.ones(X.size, np.float32)], 1)
    p = np.concatenate([p, p[:20]])                       # exact duplicates
    p = p[np.random.default_rng(0).permutation(p.shape[0])].astype(np.float32)
    q = np.concatenate([p[:60], p[:60] + np.float32(h / 2)]).astype(np.float32)
    # k = 10 runs the exec-masked insertion chain (every other k the select network); lattice
    # queries tie on d2 everywhere, so the order rests on the index half of the keys
    for k in (10, 12):
        for leaf in (1, 4, 16):
            t = _tree(rm_ctx, p, leaf)
            _, idx, d2 = t.SearchKNN(_dev(q), k)
            rm_ctx.sync()
            oi, od = ob.knn_brute(p, q, k)
            assert (idx.cpu().numpy() == oi).all() and (d2.cpu().numpy() == od).all(), (k, leaf)
    # fewer points than k: -1 / +INF padding
    small = p[:5]
    t = _tree(rm_ctx, small)
    for k in (8, 10):
        _, idx, d2 = t.SearchKNN(_dev(q[:10]), k)
        rm_ctx.sync()
        idx, d2 = idx.cpu().numpy(), d2.cpu().numpy()
        assert (idx[:, 5:] == -1).all() and np.isinf(d2[:, 5:]).all()
        oi, od = ob.knn_brute(small, q[:10], k)
        assert (idx == oi).all() and (d2 == od).all(), k


def test_tree_over_the_node_limit_is_refused(rm_ctx):
    """Trees of more than 2^26 node records are refused before any layout work (the packet
    walks address node records by 32-bit byte offsets): 2^26 + 2 points at one per leaf make
    2^26 + 1 inner nodes."""
    p = torch.zeros(((1 << 26) + 2, 4), dtype=torch.float32, device="cuda")
    with pytest.raises(RMError, match=r"2\^26"):
        api.KDTreeFlann(rm_ctx, leaf_max=1).Build(p)
    del p
    torch.cuda.empty_cache()


def test_empty_queries_and_empty_cloud(rm_ctx):
    p = scenes.box_scene(2000, seed=3).xyzw()
    t = _tree(rm_ctx, p)
    q0 = torch.zeros((0, 4), dtype=torch.float32, device="cuda")
    total, idx, d2 = t.SearchKNN(q0, 10)
    assert total == 0 and idx.shape == (0, 10) and d2.shape == (0, 10)
    total, idx, d2, cnt = t.SearchRadius(q0, 0.1, 8)
    assert total == 0 and idx.shape == (0, 8) and cnt.shape == (0,)
    total, _, _, cnt = t.SearchRadius(q0, 0.1, 0)
    assert total == 0 and cnt.shape == (0,)
    rm_ctx.sync()
    with pytest.raises(RMError):
        api.KDTreeFlann(rm_ctx).Build(q0)
    q4, cnt = _dev(p[:4]), torch.zeros(4, dtype=torch.int32, device="cuda")
    f = _abi.lib().rm_radius_search
    assert f(rm_ctx.h, t.h, api._ptr(q4), 4, 0.1, -1, None, None, api._ptr(cnt)) == \
        _abi.RM_ERR_INVALID                                   # negative max_nn
    assert f(rm_ctx.h, t.h, api._ptr(q4), 4, 0.1, 0, None, None, None) == \
        _abi.RM_ERR_INVALID                                   # counting mode without counts
    assert f(rm_ctx.h, t.h, api._ptr(q4), 4, -1.0, 0, None, None, api._ptr(cnt)) == \
        _abi.RM_ERR_INVALID                                   # negative radius


@pytest.mark.parametrize("n,leaf", [(1, 16), (16, 16), (17, 16), (1000, 1), (50_000, 16),
                                    (300_000, 32)])
def test_device_and_host_builds_agree(rm_ctx, n, leaf, monkeypatch):
    """The device build and the host build make the same partitions (same rule, same ties), so
    every search result is identical; both equal the oracle."""
    p = np.ascontiguousarray(scenes.room_scene(max(n, 2), seed=9).xyzw()[:n])
    q = _queries(p, 5, 900) if n > 3 else p
    t_dev = _tree(rm_ctx, p, leaf)
    monkeypatch.setenv("RM_TREE_BUILD", "host")
    t_host = _tree(rm_ctx, p, leaf)
    monkeypatch.delenv("RM_TREE_BUILD")
    assert t_dev.info() == t_host.info()
    k = min(10, n)
    _, i1, d1 = t_dev.SearchKNN(_dev(q), k)
    _, i2, d2 = t_host.SearchKNN(_dev(q), k)
    ru1, rc1 = t_dev.radii(0.05)
    ru2, rc2 = t_host.radii(0.05)
    rm_ctx.sync()
    assert torch.equal(i1, i2) and torch.equal(d1, d2)
    assert torch.equal(ru1, ru2) and torch.equal(rc1, rc2)
    oi, od = ob.knn_brute(p, q, k) if n <= 50_000 else ob.RefTree(p, 64).knn(q, k)[:2]
    assert (i1.cpu().numpy() == oi).all() and (d1.cpu().numpy() == od).all()


def test_knn_rejects_bad_k(rm_ctx):
    p = scenes.box_scene(1000, seed=1).xyzw()
    t = _tree(rm_ctx, p)
    from renderer_ray_marching_amd._abi import RMError
    with pytest.raises(RMError):
        t.SearchKNN(_dev(p[:4]), 33)
    with pytest.raises(RMError):
        t.SearchKNN(_dev(p[:4]), 0)


def test_radius_search_vs_brute(rm_ctx):
    p = scenes.box_scene(20000, seed=5).xyzw()
    q = _queries(p, 6, 2000)
    t = _tree(rm_ctx, p)
    radius, max_nn = 0.05, 16
    _, idx, d2, cnt = t.SearchRadius(_dev(q), radius, max_nn)
    rm_ctx.sync()
    oi, od = ob.knn_brute(p, q, max_nn)
    r2 = np.float32(radius) * np.float32(radius)
    inside = od < r2
    exp_i = np.where(inside, oi, -1)
    exp_d = np.where(inside, od, np.float32(np.inf))
    assert (idx.cpu().numpy() == exp_i).all()
    assert (d2.cpu().numpy() == exp_d).all()
    assert (cnt.cpu().numpy() == inside.sum(1)).all()


@pytest.mark.parametrize("radius", [0.0, 0.05, 0.25])
def test_radius_count_mode_vs_brute(rm_ctx, radius):
    """max_nn == 0: uncapped per-query counts of d2 < r^2 (CountingRadiusResultSet with no cap),
    checked by brute force over exact float32 fma distances; counts go well past RM_MAX_K."""
    p = scenes.box_scene(20000, seed=5).xyzw()
    q = _queries(p, 7, 900)
    t = _tree(rm_ctx, p)
    total, idx, d2, cnt = t.SearchRadius(_dev(q), radius, 0)
    rm_ctx.sync()
    assert idx.shape == (q.shape[0], 0) and d2.shape == (q.shape[0], 0)
    r2 = np.float32(radius) * np.float32(radius)
    exp = np.zeros(q.shape[0], np.int64)
    for a in range(0, q.shape[0], 100):
        exp[a:a + 100] = (fp32.dist2(p[None, :, :], q[a:a + 100, None, :]) < r2).sum(1)
    got = cnt.cpu().numpy()
    assert (got == exp).all()
    assert total == int(exp.sum())
    if radius == 0.25:
        assert exp.max() > 32
    # the capped mode agrees wherever the cap is not reached
    _, _, _, cap = t.SearchRadius(_dev(q), radius, 32)
    cap = cap.cpu().numpy()
    assert (cap == np.minimum(exp, 32)).all()


def test_radius_search_edge_radii(rm_ctx):
    """Radius 0 finds nothing, not even a query's own duplicate (d2 < 0 never holds); a radius
    whose float square overflows to +inf admits every point, so the capped mode returns the
    max_nn nearest and counting returns n; negative and NaN radii are errors."""
    p = scenes.box_scene(3000, seed=8).xyzw()
    p[100:110] = p[:10]                               # exact duplicates
    q = np.ascontiguousarray(p[:40])
    t = _tree(rm_ctx, p)
    total, idx, d2, cnt = t.SearchRadius(_dev(q), 0.0, 8)
    rm_ctx.sync()
    assert total == 0 and (cnt.cpu().numpy() == 0).all()
    assert (idx.cpu().numpy() == -1).all() and np.isinf(d2.cpu().numpy()).all()
    with np.errstate(over="ignore"):
        assert np.isinf(np.float32(1e30) * np.float32(1e30))
    total, idx, d2, cnt = t.SearchRadius(_dev(q), 1e30, 32)
    rm_ctx.sync()
    oi, od = ob.knn_brute(p, q, 32)
    assert (idx.cpu().numpy() == oi).all() and (d2.cpu().numpy() == od).all()
    assert (cnt.cpu().numpy() == 32).all() and total == 32 * q.shape[0]
    total, _, _, cnt = t.SearchRadius(_dev(q), 1e30, 0)
    rm_ctx.sync()
    assert (cnt.cpu().numpy() == p.shape[0]).all() and total == p.shape[0] * q.shape[0]
    for bad in (-0.5, float("nan")):
        with pytest.raises(RMError):
            t.SearchRadius(_dev(q), bad, 8)


# ------------------------------------------------------------------------------------------
# radii
# ------------------------------------------------------------------------------------------

@pytest.mark.parametrize("leaf", [0, 32])
def test_radii_bit_exact_box_100k(rm_ctx, leaf):
    c = scenes.box_scene(100_000, seed=1)
    p = c.xyzw()
    t = _tree(rm_ctx, p, leaf)
    ru, rc = t.radii(0.1)
    rm_ctx.sync()
    oru, orc, _ = ob.RefTree(p, 64).radii(0.1)
    assert (ru.cpu().numpy() == oru).all()
    assert (rc.cpu().numpy() == orc).all()


@pytest.mark.parametrize("leaf,n", [(1, 3001), (3, 5003), (8, 4099), (16, 7000), (24, 6001),
                                    (32, 70), (64, 50)])
def test_radii_leaf_sizes_and_ragged_tails(rm_ctx, leaf, n):
    # leaf sizes that do and do not divide the 64-point wavefront group, tails of every length
    c = scenes.box_scene(n, seed=leaf + 3)
    p = c.xyzw()
    t = _tree(rm_ctx, p, leaf)
    ru, rc = t.radii(0.05)
    rm_ctx.sync()
    oru, orc, _ = ob.RefTree(p, 64).radii(0.05)
    assert (ru.cpu().numpy() == oru).all() and (rc.cpu().numpy() == orc).all()


def test_radii_room_2m_full_size(rm_ctx):
    c = scenes.room_scene(2_000_000, seed=2)
    p = c.xyzw()
    t = _tree(rm_ctx, p)
    ru, rc = t.radii(0.05)
    rm_ctx.sync()
    oru, orc, _ = ob.RefTree(p, 64).radii(0.05)
    assert (ru.cpu().numpy() == oru).all() and (rc.cpu().numpy() == orc).all()


def test_radii_duplicates_lattice_single(rm_ctx):
    h = np.float32(0.125)
    g = np.arange(10, dtype=np.float32) * h
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    p = np.stack([X.ravel(), Y.ravel(), Z.ravel(), np.ones(X.size, np.float32)], 1)
    p = np.concatenate([p, p[:7]]).astype(np.float32)
    t = _tree(rm_ctx, p)
    ru, rc = t.radii(0.0)
    rm_ctx.sync()
    ru = ru.cpu().numpy()
    dup = np.zeros(p.shape[0], bool)
    dup[:7] = True
    dup[-7:] = True
    assert (ru[dup] == 0).all() and (ru[~dup] == h).all()
    t1 = _tree(rm_ctx, p[:1])
    ru1, rc1 = t1.radii(0.1)
    rm_ctx.sync()
    assert np.isinf(ru1.cpu().numpy()[0]) and rc1.cpu().numpy()[0] == np.float32(0.1)


def test_radii_single_output_and_after_refit(rm_ctx):
    """rm_radii with one output only (the other NULL), and on a tree whose points were moved by
    a refit (same leaf order, so the leaf-to-original mapping still holds)."""
    c = scenes.box_scene(30_000, seed=11)
    p = c.xyzw()
    n = p.shape[0]
    t = _tree(rm_ctx, p)
    L = _abi.lib()
    ru = torch.full((n,), -1.0, device="cuda")
    rc = torch.full((n,), -1.0, device="cuda")
    _abi.check(L.rm_radii(rm_ctx.h, t.h, 0.02, api._ptr(ru), None), rm_ctx.h)
    _abi.check(L.rm_radii(rm_ctx.h, t.h, 0.02, None, api._ptr(rc)), rm_ctx.h)
    rm_ctx.sync()
    oru, orc, _ = ob.RefTree(p, 64).radii(0.02)
    assert (ru.cpu().numpy() == oru).all() and (rc.cpu().numpy() == orc).all()
    # move every point, refit through a cloud, recompute the radii of the moved cloud
    moved = p.copy()
    moved[:, :3] += np.random.default_rng(5).normal(0, 0.01, (n, 3)).astype(np.float32)
    m = api.PointcloudRayMarcher.get_instance(_dev(p), _dev(c.colors4()), rc, t)
    m.refit(_dev(np.ascontiguousarray(moved[:, :3])))
    ru2, rc2 = t.radii(0.02)
    rm_ctx.sync()
    oru2, orc2, _ = ob.RefTree(moved, 64).radii(0.02)
    assert (ru2.cpu().numpy() == oru2).all() and (rc2.cpu().numpy() == orc2).all()


# ------------------------------------------------------------------------------------------
# plain march
# ------------------------------------------------------------------------------------------

def _marcher(ctx, cloud, max_radius, leaf=0):
    p = cloud.xyzw()
    t = _tree(ctx, p, leaf)
    _, rc = t.radii(max_radius)
    m = api.PointcloudRayMarcher.get_instance(_dev(p), _dev(cloud.colors4()), rc, t)
    return t, m, rc


def _oracle_render(cloud, radii, view, W, H, pix=None):
    rt = ob.RefTree(cloud.xyzw(), 64)
    return rt.render(radii, cloud.colors4(), view.pose_colmajor(), view.fov, W, H, pix=pix)


def test_plain_march_box_full_view_small(rm_ctx):
    cloud = scenes.box_scene(100_000, seed=1)
    view = scenes.box_scene_view(cloud, 160, 120, 125.0)
    t, m, rc = _marcher(rm_ctx, cloud, 0.1)
    rgba, hit = m.render_to_texture(view.pose_colmajor(), view.fov, 160, 120)
    rm_ctx.sync()
    orgba, ohit, cnt = _oracle_render(cloud, rc.cpu().numpy(), view, 160, 120)
    hit = hit.cpu().numpy().ravel()
    rgba = rgba.cpu().numpy().reshape(-1, 4)
    assert (hit >= 0).mean() > 0.3
    assert (hit == ohit).all()
    assert (rgba == orgba).all()


def test_plain_march_box_640x480_subsample(rm_ctx):
    """The full box workload view (640x480, fx 500), checked on every 7th pixel."""
    cloud, views, max_r = scenes.workload_box_100k()
    view = views[0]
    t, m, rc = _marcher(rm_ctx, cloud, max_r)
    rgba, hit = m.render_to_texture(view.pose_colmajor(), view.fov, 640, 480)
    rm_ctx.sync()
    pix = np.arange(0, 640 * 480, 7, dtype=np.int32)
    orgba, ohit, _ = _oracle_render(cloud, rc.cpu().numpy(), view, 640, 480, pix)
    hit = hit.cpu().numpy().ravel()[pix]
    rgba = rgba.cpu().numpy().reshape(-1, 4)[pix]
    assert (hit == ohit).all()
    assert (rgba == orgba).all()


def test_plain_march_room_views(rm_ctx):
    c = scenes.room_scene(300_000, seed=2)
    views = scenes.room_views(3, seed=102, sizes=((200, 130, 150.0),))
    t, m, rc = _marcher(rm_ctx, c, 0.05)
    for v in views:
        rgba, hit = m.render_to_texture(v.pose_colmajor(), v.fov, v.width, v.height)
        rm_ctx.sync()
        orgba, ohit, _ = _oracle_render(c, rc.cpu().numpy(), v, v.width, v.height)
        assert (hit.cpu().numpy().ravel() == ohit).all()
        assert (rgba.cpu().numpy().reshape(-1, 4) == orgba).all()


def test_plain_march_tiny_cloud_and_misses(rm_ctx):
    # fewer than 10 points: kNN-10 padding must be skipped; most rays miss
    rng = np.random.default_rng(3)
    xyz = rng.uniform([-1, -1, -5], [1, 1, -3], (6, 3)).astype(np.float32)
    cloud = scenes.Cloud(xyz, rng.integers(0, 256, (6, 3)).astype(np.uint8))
    view = scenes.View(np.eye(4, dtype=np.float32), 64, 48, 40.0)
    t, m, rc = _marcher(rm_ctx, cloud, 0.0)
    radii = torch.full((6,), 0.3, device="cuda")
    m.set_radii(radii)
    rgba, hit = m.render_to_texture(view.pose_colmajor(), view.fov, 64, 48)
    rm_ctx.sync()
    orgba, ohit, _ = _oracle_render(cloud, np.full(6, 0.3, np.float32), view, 64, 48)
    assert (hit.cpu().numpy().ravel() == ohit).all()
    assert (rgba.cpu().numpy().reshape(-1, 4) == orgba).all()
    assert (ohit == -1).any() and (ohit >= 0).any()


def test_mirror_rejects_bad_tensors(rm_ctx):
    """The Python mirror checks dtype and length before a pointer reaches the library, which
    reads exactly n values: a short or float64 tensor is an error, not an out-of-bounds read."""
    cloud = scenes.box_scene(2_000, seed=5)
    t, m, rc = _marcher(rm_ctx, cloud, 0.1)
    n = cloud.n
    with pytest.raises(ValueError):
        m.set_radii(rc[:-1])
    with pytest.raises(TypeError):
        m.set_radii(rc.double())
    with pytest.raises(ValueError):
        m.set_colors(torch.ones((n - 1, 4), device="cuda"))
    with pytest.raises(ValueError):
        m.refit(torch.zeros((n, 4), device="cuda"))
    with pytest.raises(TypeError):
        api.KDTreeFlann(rm_ctx).Build(_dev(cloud.xyzw(), torch.float64))
    with pytest.raises(ValueError):
        t.SearchKNN(torch.zeros((5, 3), device="cuda"), 2)
    with pytest.raises(ValueError):
        api.PointcloudRayMarcher(_dev(cloud.xyzw()), _dev(cloud.colors4()[:-1]), rc, t)
    with pytest.raises(ValueError):
        m.render_to_texture(np.eye(4), 1.0, 32, 24, rgba=torch.empty((24, 31, 4), device="cuda"))
    # still usable afterwards
    rgba, hit = m.render_to_texture(np.eye(4), 1.0, 32, 24)
    rm_ctx.sync()
    assert hit.shape == (24, 32)


def test_early_miss_edge_cases(rm_ctx):
    """Rays are ended as misses past their exit from the cloud's dilated box. Check the plain
    and batched marches against the oracle (which marches every step) on cameras inside the box,
    outside looking at it, outside looking away, grazing a face, with a huge radius in the cloud,
    after the radii change (the bound's largest radius is recomputed) and after the points move
    (the box is refitted)."""
    rng = np.random.default_rng(17)
    n = 3000
    xyz = np.concatenate([rng.uniform([-1, -1, -1], [1, 1, -0.98], (n // 2, 3)),     # a floor
                          rng.uniform([-1, 0.98, -1], [1, 1, 1], (n - n // 2, 3))])  # a wall
    xyz = xyz.astype(np.float32)
    cloud = scenes.Cloud(xyz, rng.integers(0, 256, (n, 3)).astype(np.uint8))
    t, m, rc = _marcher(rm_ctx, cloud, 0.0)
    W, H, fx = 48, 36, 30.0
    views = [scenes.View(scenes.look_at(e, c), W, H, fx) for e, c in [
        ((0.0, 0.0, 0.0), (0.0, 1.0, -0.5)),        # inside the box
        ((0.0, -4.0, 0.5), (0.0, 0.0, -0.5)),       # outside, looking at it
        ((0.0, -4.0, 0.5), (0.0, -8.0, 0.5)),       # outside, looking away
        ((-3.0, -0.5, -0.97), (3.0, -0.5, -0.97)),  # grazing the floor
    ]]
    pts = cloud.xyzw()

    def check(radii_np):
        rt = ob.RefTree(pts, 64)
        b = api.Batch(m, views, tune=TUNES[1])
        b.launch()
        rm_ctx.sync()
        off = 0
        for v in views:
            rgba, hit = m.render_to_texture(v.pose_colmajor(), v.fov, W, H)
            rm_ctx.sync()
            orgba, ohit, _ = rt.render(radii_np, cloud.colors4(), v.pose_colmajor(), v.fov, W, H)
            assert (hit.cpu().numpy().ravel() == ohit).all()
            assert (rgba.cpu().numpy().reshape(-1, 4) == orgba).all()
            assert (b.hit[off:off + W * H].cpu().numpy() == ohit).all()
            off += W * H
        b.close()
        return off

    check(rc.cpu().numpy())
    big = rc.clone()
    big[7] = 10.0                                    # one huge sphere
    m.set_radii(big)
    check(big.cpu().numpy())
    small = torch.full_like(rc, 0.02)
    m.set_radii(small)                               # the largest radius shrinks again
    check(small.cpu().numpy())
    moved = xyz + np.float32(0.5)                    # the box moves with the points
    m.refit(_dev(moved))
    pts = np.concatenate([moved, np.ones((n, 1), np.float32)], 1)
    check(small.cpu().numpy())


# ------------------------------------------------------------------------------------------
# batched march == plain march, for every traversal mode / tile / warm start
# ------------------------------------------------------------------------------------------

TUNES = [
    dict(tile_w=8, tile_h=8, mode=0, warm=0, waves_per_block=2, blocks_per_cu=0, refill=1),
    dict(tile_w=8, tile_h=8, mode=0, warm=1, waves_per_block=1, blocks_per_cu=0, refill=1),
    dict(tile_w=16, tile_h=4, mode=0, warm=1, waves_per_block=4, blocks_per_cu=2, refill=1),
    dict(tile_w=32, tile_h=2, mode=1, warm=0, waves_per_block=2, blocks_per_cu=0, refill=1),
    dict(tile_w=8, tile_h=8, mode=1, warm=1, waves_per_block=4, blocks_per_cu=1, refill=1),
    dict(tile_w=16, tile_h=16, mode=0, warm=1, waves_per_block=2, blocks_per_cu=0, refill=1),
    dict(tile_w=32, tile_h=8, mode=0, warm=0, waves_per_block=1, blocks_per_cu=3, refill=16),
    dict(tile_w=16, tile_h=8, mode=1, warm=1, waves_per_block=2, blocks_per_cu=0, refill=4),
    dict(tile_w=64, tile_h=64, mode=1, warm=0, waves_per_block=4, blocks_per_cu=16, refill=64),
    dict(tile_w=24, tile_h=8, mode=0, warm=1, waves_per_block=2, blocks_per_cu=0, refill=8),
    # per-lane stacks split between LDS and the global overflow area (1: every entry but the
    # first overflows)
    dict(tile_w=8, tile_h=8, mode=1, warm=1, waves_per_block=2, blocks_per_cu=0, refill=1,
         stack_depth=1),
    dict(tile_w=16, tile_h=4, mode=1, warm=0, waves_per_block=4, blocks_per_cu=0, refill=1,
         stack_depth=8),
    dict(tile_w=8, tile_h=8, mode=1, warm=1, waves_per_block=1, blocks_per_cu=2, refill=8,
         stack_depth=16),
    # early miss over a frontier of the tree instead of the whole cloud's box
    dict(tile_w=8, tile_h=8, mode=0, warm=1, waves_per_block=2, blocks_per_cu=0, refill=1,
         exit_depth=6),
    dict(tile_w=16, tile_h=8, mode=1, warm=0, waves_per_block=4, blocks_per_cu=0, refill=4,
         exit_depth=10),
    dict(tile_w=8, tile_h=8, mode=0, warm=0, waves_per_block=1, blocks_per_cu=0, refill=1,
         exit_depth=1),
]


@pytest.fixture(scope="module")
def room_small(rm_ctx):
    c = scenes.room_scene(400_000, seed=2)
    sizes = [(160, 120, 125.0), (213, 138, 160.0), (96, 54, 80.0), (7, 5, 5.0)]
    views = scenes.room_views(12, seed=102, sizes=sizes)
    t, m, rc = _marcher(rm_ctx, c, 0.05)
    plain = []
    for v in views:
        rgba, hit = m.render_to_texture(v.pose_colmajor(), v.fov, v.width, v.height)
        plain.append((rgba.reshape(-1, 4).clone(), hit.reshape(-1).clone()))
    rm_ctx.sync()
    return c, views, t, m, rc, plain


@pytest.mark.parametrize("ti", range(len(TUNES)))
def test_batch_equals_plain(rm_ctx, room_small, ti):
    c, views, t, m, rc, plain = room_small
    b = api.Batch(m, views, tune=TUNES[ti])
    for rep in range(3):           # graph replays: every replay equals the eager result
        b.rgba.fill_(7.0)
        b.hit.fill_(-7)
        b.launch()
        rm_ctx.sync()
        off = 0
        for v, (prgba, phit) in zip(views, plain):
            n = v.width * v.height
            assert torch.equal(b.hit[off:off + n], phit)
            assert torch.equal(b.rgba[off:off + n], prgba)
            off += n
    b.close()


def test_batch_stack_depth_bounds(rm_ctx, room_small):
    """stack_depth outside 0..RM_MAX_STACK_DEPTH is refused; values at or above the tree's
    need keep the whole stack in LDS, and the tune reads back as given."""
    c, views, t, m, rc, plain = room_small
    for bad in (-1, _abi.RM_MAX_STACK_DEPTH + 1):
        tune = dict(TUNES[3], stack_depth=bad)
        with pytest.raises(_abi.RMError) as e:
            api.Batch(m, views, tune=tune)
        assert e.value.code == _abi.RM_ERR_INVALID
    b = api.Batch(m, views, tune=dict(TUNES[3], stack_depth=_abi.RM_MAX_STACK_DEPTH))
    assert b.tune()["stack_depth"] == _abi.RM_MAX_STACK_DEPTH
    b.launch()
    rm_ctx.sync()
    assert torch.equal(b.hit, torch.cat([p[1] for p in plain]))
    b.close()


def test_early_miss_frontier_keeps_results(rm_ctx):
    """The early miss over a frontier of the tree (exit_depth 1..RM_MAX_EXIT_DEPTH) gives the
    plain march's outputs on every pixel: box scenes seen from outside (many rays pass between
    the boxes), an axis-aligned camera (rays with a direction component of exactly 0 take the
    cloud box's bound), a cloud smaller than one leaf, and after a refit that moves every point
    (the frontier is read from the tree at each launch). Depths out of range are refused."""
    def views_of(cloud, n=3):
        vs = [scenes.box_scene_view(cloud, 96, 72, 80.0, distance=4.0, seed=s)
              for s in range(1, n + 1)]
        pose = np.eye(4, dtype=np.float32)
        pose[:3, 3] = (cloud.xyz.astype(np.float64).mean(0) + [0.0, 0.0, 5.0]).astype(np.float32)
        vs.append(scenes.View(pose, 64, 48, 40.0, name="axis_aligned"))
        return vs

    def check(m, views, depths):
        plain = []
        for v in views:
            rgba, hit = m.render_to_texture(v.pose_colmajor(), v.fov, v.width, v.height)
            plain.append((rgba.reshape(-1, 4).clone(), hit.reshape(-1).clone()))
        rm_ctx.sync()
        ref_hit = torch.cat([p[1] for p in plain])
        ref_rgba = torch.cat([p[0] for p in plain])
        for ed in depths:
            b = api.Batch(m, views, tune=dict(TUNES[1], exit_depth=ed))
            for rep in range(2):
                b.hit.fill_(-7)
                b.launch()
                rm_ctx.sync()
                assert torch.equal(b.hit, ref_hit), ed
                assert torch.equal(b.rgba, ref_rgba), ed
            b.close()
        return ref_hit

    cloud = scenes.box_scene(30_000, seed=9)
    t, m, rc = _marcher(rm_ctx, cloud, 0.05)
    views = views_of(cloud)
    hit = check(m, views, (0, 1, 3, 6, _abi.RM_MAX_EXIT_DEPTH))
    assert (hit < 0).float().mean() > 0.2          # many rays leave the cloud
    for bad in (-1, _abi.RM_MAX_EXIT_DEPTH + 1):
        with pytest.raises(_abi.RMError) as e:
            api.Batch(m, views, tune=dict(TUNES[1], exit_depth=bad))
        assert e.value.code == _abi.RM_ERR_INVALID
    # the points moved by 0.2 m along x: a frontier kept from before the refit would cut rays
    # that now hit the shifted boxes
    moved = torch.from_numpy(cloud.xyz + np.float32([0.2, 0.0, 0.0])).cuda()
    m.refit(moved)
    check(m, views, (4, 8))
    # a cloud smaller than one leaf: the frontier is that leaf alone
    tiny = scenes.box_scene(7, seed=3)
    t2, m2, rc2 = _marcher(rm_ctx, tiny, 0.5)
    check(m2, views_of(tiny, 1), (2, 6))


def test_batch_plain_vs_oracle_subsample(rm_ctx, room_small):
    c, views, t, m, rc, plain = room_small
    v = views[1]
    orgba, ohit, _ = _oracle_render(c, rc.cpu().numpy(), v, v.width, v.height)
    assert (plain[1][1].cpu().numpy() == ohit).all()
    assert (plain[1][0].cpu().numpy() == orgba).all()


def test_batch_autotune_keeps_results(rm_ctx, room_small):
    c, views, t, m, rc, plain = room_small
    b = api.Batch(m, views)
    tune = b.autotune(sample_stride=4)
    assert tune["tile_w"] * tune["tile_h"] == 64
    b.launch()
    rm_ctx.sync()
    ref_hit = torch.cat([p[1] for p in plain])
    ref_rgba = torch.cat([p[0] for p in plain])
    assert torch.equal(b.hit, ref_hit) and torch.equal(b.rgba, ref_rgba)
    b.close()


# ------------------------------------------------------------------------------------------
# frustum culling (opt-in): the march over the points inside each view's frustum
# ------------------------------------------------------------------------------------------

@pytest.fixture(scope="module")
def room_culled(rm_ctx):
    c = scenes.room_scene(400_000, seed=2)
    sizes = [(160, 120, 125.0), (213, 138, 160.0), (96, 54, 80.0), (7, 5, 5.0)]
    views = scenes.room_views(8, seed=102, sizes=sizes)
    t, m, rc = _marcher(rm_ctx, c, 0.05)
    m.set_culling(True)
    plain = []
    for v in views:
        rgba, hit = m.render_to_texture(v.pose_colmajor(), v.fov, v.width, v.height)
        plain.append((rgba.reshape(-1, 4).clone(), hit.reshape(-1).clone()))
    rm_ctx.sync()
    return c, views, t, m, rc, plain


def test_culled_plain_vs_oracle(rm_ctx, room_culled):
    c, views, t, m, rc, plain = room_culled
    pts, col, r = c.xyzw(), c.colors4(), rc.cpu().numpy()
    changed = 0
    for v, (rgba, hit) in zip(views, plain):
        orgba, ohit, _ = ob.render_culled(pts, r, col, v.pose_colmajor(), v.fov, v.width,
                                          v.height)
        assert (hit.cpu().numpy() == ohit).all()
        assert (rgba.cpu().numpy() == orgba).all()
        _, uhit, _ = ob.RefTree(pts, 64).render(r, col, v.pose_colmajor(), v.fov, v.width,
                                                 v.height)
        changed += int((uhit != ohit).sum())
    assert changed > 0          # culling changes pixels on these views


@pytest.mark.parametrize("ti", [0, 1, 2, 3, 4, 10, 11, 13, 14])
def test_culled_batch_equals_plain(rm_ctx, room_culled, ti):
    c, views, t, m, rc, plain = room_culled
    b = api.Batch(m, views, tune=TUNES[ti])
    for rep in range(2):
        b.rgba.fill_(7.0)
        b.hit.fill_(-7)
        b.launch()
        rm_ctx.sync()
        off = 0
        for v, (prgba, phit) in zip(views, plain):
            n = v.width * v.height
            assert torch.equal(b.hit[off:off + n], phit)
            assert torch.equal(b.rgba[off:off + n], prgba)
            off += n
    b.close()


def test_culled_render_diff_uses_the_culled_march(rm_ctx, room_culled):
    """rm_render_diff on a culling cloud selects its hits by the culled march."""
    import ctypes as C
    c, views, t, m, rc, plain = room_culled
    L = _abi.lib()
    for v, (prgba, phit) in list(zip(views, plain))[:3]:
        W, H = v.width, v.height
        img = torch.empty((H, W, 3), device="cuda")
        dep = torch.empty((H, W), device="cuda")
        hit = torch.empty((H, W), dtype=torch.int32, device="cuda")
        pose = (C.c_float * 16)(*[float(x) for x in v.pose_colmajor()])
        assert L.rm_render_diff(rm_ctx.h, m.h, pose, C.c_float(v.fov), W, H,
                                C.c_void_p(img.data_ptr()), C.c_void_p(dep.data_ptr()),
                                C.c_void_p(hit.data_ptr())) == 0
        rm_ctx.sync()
        assert torch.equal(hit.reshape(-1), phit)
        # colour of the hit point, as the plain march wrote it
        h = phit >= 0
        assert torch.equal(img.reshape(-1, 3)[h], prgba[h][:, :3])


def test_culled_few_and_no_visible_points(rm_ctx):
    """Fewer than 10 points inside the frustum (padding in the warm-started list) and none."""
    fx, w, h = 40.0, 64, 48
    fov = scenes.fov_radians(w, fx)
    pane = ob.pane_distance(w, fov)
    rng = np.random.default_rng(11)
    M = scenes.look_at((0.0, 0.0, 0.0), (0.0, -1.0, 0.0))
    # 6 points in view, 300 behind the camera
    u = rng.uniform(-0.8, 0.8, 6) * w / 2
    v = rng.uniform(-0.8, 0.8, 6) * h / 2
    dep = rng.uniform(3.0, 6.0, 6)
    cam = np.stack([u, v, -np.full(6, pane)], 1) * (dep / pane)[:, None]
    front = cam @ M[:3, :3].T.astype(np.float64)
    back = rng.normal(0, 1, (300, 3)) + np.array([0.0, 5.0, 0.0])
    xyz = np.concatenate([front, back]).astype(np.float32)
    pts = np.concatenate([xyz, np.ones((xyz.shape[0], 1), np.float32)], 1)
    radii = np.full(pts.shape[0], 0.3, np.float32)
    col = rng.uniform(0, 1, (pts.shape[0], 4)).astype(np.float32)
    col[:, 3] = 1
    tree = _tree(rm_ctx, pts)
    mm = api.PointcloudRayMarcher(_dev(pts), _dev(col), _dev(radii), tree)
    mm.set_culling(True)
    # the second camera looks away from every point
    views = [scenes.View(M, w, h, fx),
             scenes.View(scenes.look_at((0.0, -20.0, 0.0), (0.0, -21.0, 0.0)), w, h, fx)]
    for vw in views:
        rgba, hit = mm.render_to_texture(vw.pose_colmajor(), vw.fov, w, h)
        rm_ctx.sync()
        orgba, ohit, ins = ob.render_culled(pts, radii, col, vw.pose_colmajor(), vw.fov, w, h)
        assert (hit.cpu().numpy().ravel() == ohit).all()
        assert (rgba.cpu().numpy().reshape(-1, 4) == orgba).all()
    assert ob.frustum_inside(views[0].pose_colmajor(), fov, w, h, pts).sum() == 6
    assert ob.frustum_inside(views[1].pose_colmajor(), fov, w, h, pts).sum() == 0
    for tune in (TUNES[0], TUNES[1]):
        b = api.Batch(mm, views, tune=tune)
        b.launch()
        rm_ctx.sync()
        for k, vw in enumerate(views):
            _, ohit, _ = ob.render_culled(pts, radii, col, vw.pose_colmajor(), vw.fov, w, h)
            assert (b.hit[k * w * h:(k + 1) * w * h].cpu().numpy() == ohit).all()
        b.close()
    mm.close()


# ------------------------------------------------------------------------------------------
# full-size configurations: the batched march equals the plain path on every pixel and the
# oracle on a deterministic pixel subsample
# ------------------------------------------------------------------------------------------

def _subsample(W, H, step):
    lin = np.arange(W * H, dtype=np.int64)
    x, y = lin % W, lin // W
    return lin[(x + 7 * y) % step == 0].astype(np.int32)


def _full_size_check(rm_ctx, cloud, views, max_r, oracle_views, step):
    p = cloud.xyzw()
    t = _tree(rm_ctx, p)
    ru, rc = t.radii(max_r)
    m = api.PointcloudRayMarcher(_dev(p), _dev(cloud.colors4()), rc, t)
    b = api.Batch(m, views)
    b.launch()
    rm_ctx.sync()
    rt = ob.RefTree(p, 64)
    oru, orc, _ = rt.radii(max_r)
    assert (ru.cpu().numpy() == oru).all() and (rc.cpu().numpy() == orc).all()
    offs = np.cumsum([0] + [v.width * v.height for v in views])
    hit = b.hit.cpu().numpy()
    rgba = b.rgba.cpu().numpy()
    checked = 0
    for i in oracle_views:
        v = views[i]
        pix = _subsample(v.width, v.height, step)
        orgba, ohit, _ = rt.render(orc, cloud.colors4(), v.pose_colmajor(), v.fov, v.width,
                                   v.height, pix=pix)
        assert (hit[offs[i] + pix] == ohit).all(), i
        assert (rgba[offs[i] + pix] == orgba).all(), i
        checked += pix.size
    # every pixel of a few views also against the plain per-view launch
    for i in oracle_views[:2]:
        v = views[i]
        prgba, phit = m.render_to_texture(v.pose_colmajor(), v.fov, v.width, v.height)
        rm_ctx.sync()
        n = v.width * v.height
        assert (phit.cpu().numpy().ravel() == hit[offs[i]:offs[i] + n]).all()
        assert (prgba.cpu().numpy().reshape(-1, 4) == rgba[offs[i]:offs[i] + n]).all()
    b.close()
    return checked


@pytest.mark.slow
def test_full_size_room_2m_32_views(rm_ctx):
    cloud, views, max_r = scenes.workload_room_2m()
    n = _full_size_check(rm_ctx, cloud, views, max_r, list(range(0, 32, 4)), 16)
    assert n > 300_000


@pytest.mark.slow
def test_full_size_room_10m_256_views(rm_ctx):
    cloud, views, max_r = scenes.workload_room_10m()
    n = _full_size_check(rm_ctx, cloud, views, max_r, [0, 37, 101, 190, 255], 16)
    assert n > 100_000


@pytest.fixture(scope="module")
def scale50m():
    """50M points in a 40 x 30 x 4 m room (1 % outliers), 1024 views at 1920x1080."""
    return scenes.workload_scale_50m()


@pytest.mark.slow
@pytest.mark.timeout(600)
def test_full_size_scale_50m_view_subset(rm_ctx, scale50m):
    """The largest configuration's cloud (50M points, radii against the oracle on all of them)
    with every 64th of its 1024 views at 1920x1080 (the batch output of all 1024 views would be
    42 GB on the host)."""
    cloud, views, max_r = scale50m
    n = _full_size_check(rm_ctx, cloud, views[::64], max_r, [0, 5, 10, 15], 64)
    assert n > 100_000


@pytest.mark.slow
@pytest.mark.timeout(900)
def test_full_size_scale_50m_culled_1024_views_one_launch(rm_ctx, scale50m):
    """The largest configuration as stated: all 1024 views at 1920x1080 (2.12G rays) in ONE
    persistent launch, each view marching over its frustum-culled share of the 50M points. The
    42 GB output stays on the device: four views (0, 341, 682, 1023) are compared with the
    oracle's culled march (a tree over each view's inside points) on the (x + 7 y) mod 64 pixels,
    and two views with the culled plain kernel on every pixel."""
    cloud, views, max_r = scale50m
    print("case: scale50m, 1024 culled views in one batched launch", flush=True)
    p = cloud.xyzw()
    t = _tree(rm_ctx, p)
    _, rc = t.radii(max_r)
    m = api.PointcloudRayMarcher(_dev(p), _dev(cloud.colors4()), rc, t)
    m.set_culling(True)
    b = api.Batch(m, views)
    assert b.total == sum(v.width * v.height for v in views) == 2_123_366_400
    b.launch()
    rm_ctx.sync()
    offs = np.cumsum([0] + [v.width * v.height for v in views])
    rch, col = rc.cpu().numpy(), cloud.colors4()
    checked = 0
    for i in (0, 341, 682, 1023):
        v = views[i]
        pix = _subsample(v.width, v.height, 64)
        sel = torch.from_numpy(offs[i] + pix.astype(np.int64)).cuda()
        ghit, grgba = b.hit[sel].cpu().numpy(), b.rgba[sel].cpu().numpy()
        orgba, ohit, inside = ob.render_culled(p, rch, col, v.pose_colmajor(), v.fov, v.width,
                                               v.height, pix=pix)
        assert 0 < inside.sum() < p.shape[0], i
        assert (ghit == ohit).all(), i
        assert (grgba == orgba).all(), i
        checked += pix.size
    for i in (0, 1023):
        v = views[i]
        n = v.width * v.height
        prgba, phit = m.render_to_texture(v.pose_colmajor(), v.fov, v.width, v.height)
        rm_ctx.sync()
        assert torch.equal(phit.reshape(-1), b.hit[offs[i]:offs[i] + n]), i
        assert torch.equal(prgba.reshape(-1, 4), b.rgba[offs[i]:offs[i] + n]), i
    assert checked > 100_000
    assert float((b.hit >= 0).float().mean().item()) > 0.5
    b.close()

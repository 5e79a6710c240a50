"""The differentiable march (libtorch autograd Function over rm_render_diff /
rm_render_adjoint) against the fp64 oracle gradient, central finite differences of the GPU
forward, and a short Adam refinement."""
import numpy as np
import pytest
import torch

from oracle import binding as ob
from renderer_ray_marching_amd import scenes

pytestmark = pytest.mark.gpu


def _diff():
    from renderer_ray_marching_amd import diff
    return diff


@pytest.fixture(scope="module")
def small_scene(rm_ctx):
    diff = _diff()
    cloud = scenes.box_scene(30000, seed=4)
    p = cloud.xyzw()
    tree = ob.RefTree(p, 64)
    _, rc, _ = tree.radii(0.05)
    rc = np.maximum(rc, np.float32(0.01))
    views = [scenes.box_scene_view(cloud, 96, 72, 80.0, distance=5.0, seed=s) for s in (4, 5)]
    xyz = torch.from_numpy(cloud.xyz).cuda()
    r = torch.from_numpy(rc).cuda()
    rgb = torch.from_numpy(cloud.colors4()[:, :3].copy()).cuda()
    R = diff.DiffRenderer(xyz, r, rgb, views, ctx=rm_ctx)
    return cloud, rc, views, R, xyz, r, rgb


def _ray_sphere64(view, xyz64, r64, hit, lin):
    """fp64 ray/sphere geometry of each pixel's ray and its hit sphere: (d.m, h, r) with
    m = o - c and h = m - (d.m) d the centre-to-ray offset."""
    W, H = view.width, view.height
    pane = ob.pane_distance(W, view.fov)
    M = view.pose.astype(np.float64)
    x, y = lin % W, lin // W
    d = np.stack([x - W / 2.0, y - H / 2.0, -np.full(x.shape, pane)], 1) @ M[:3, :3].T
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = M[:3, 3]
    m = o - xyz64[hit]
    b = (d * m).sum(1)
    hv = m - b[:, None] * d
    return b, hv, r64[hit]


def _disc_ok(view, xyz64, r64, hit, lin):
    """fp64 discriminant / r^2 of each hit pixel (for selecting well-conditioned pixels)."""
    _, hv, rh = _ray_sphere64(view, xyz64, r64, hit, lin)
    return (rh ** 2 - (hv ** 2).sum(1)) / rh ** 2


def _entry_depth64(view, xyz64, r64, hit, lin):
    """fp64 entry distance -(d.m) - sqrt(r^2 - |h|^2) of each pixel's ray into its hit sphere."""
    b, hv, rh = _ray_sphere64(view, xyz64, r64, hit, lin)
    return -b - np.sqrt(rh ** 2 - (hv ** 2).sum(1))


def test_forward_matches_oracle(rm_ctx, small_scene):
    cloud, rc, views, R, xyz, r, rgb = small_scene
    img, depth, hit = R.render(xyz, r, rgb)
    rm_ctx.sync()
    tree = ob.RefTree(cloud.xyzw(), 64)
    for v, view in enumerate(views):
        orgba, ohit, _ = tree.render(rc, cloud.colors4(), view.pose_colmajor(), view.fov,
                                     view.width, view.height)
        h = hit[v].cpu().numpy().ravel()
        assert (h == ohit).all()
        im = img[v].cpu().numpy().reshape(-1, 3)
        assert (im[h >= 0] == orgba[h >= 0, :3]).all() and (im[h < 0] == 0).all()
        lin = np.nonzero(h >= 0)[0]
        q = _disc_ok(view, cloud.xyz.astype(np.float64), rc.astype(np.float64), h[lin], lin)
        good = lin[q > 0.05]
        dep = depth[v].cpu().numpy().ravel()
        # analytic entry depth vs the marched depth: within a few MIN_DIST where well posed
        assert np.abs(dep[good] - orgba[good, 3]).max() < 0.01
        assert (dep[h < 0] == 0).all()


def test_adjoint_matches_fp64_oracle(rm_ctx, small_scene):
    cloud, rc, views, R, xyz, r, rgb = small_scene
    rng = np.random.default_rng(1)
    x = xyz.clone().requires_grad_(True)
    rr = r.clone().requires_grad_(True)
    cc = rgb.clone().requires_grad_(True)
    img, depth, hit = R.render(x, rr, cc)
    V, H, W = hit.shape
    tgt_rgb = torch.from_numpy(rng.random((V, H, W, 3)).astype(np.float32)).cuda()
    # depth targets clearly away from the rendered depth so |e| has a stable sign
    off = rng.uniform(0.01, 0.05, (V, H, W)) * rng.choice([-1, 1], (V, H, W))
    tgt_d = (depth.detach() + torch.from_numpy(off.astype(np.float32)).cuda())
    m = (hit >= 0).float()
    loss = (0.7 * ((img - tgt_rgb) ** 2).sum(-1) * m).sum() + (1.3 * (depth - tgt_d).abs() * m).sum()
    loss.backward()
    gx, gr, gc = x.grad.cpu().numpy(), rr.grad.cpu().numpy(), cc.grad.cpu().numpy()
    ex = np.zeros((cloud.n, 3))
    er = np.zeros(cloud.n)
    ec = np.zeros((cloud.n, 3))
    bad = np.zeros(cloud.n, bool)
    xyz64, r64, rgb64 = cloud.xyz.astype(np.float64), rc.astype(np.float64), \
        cloud.colors4()[:, :3].astype(np.float64)
    for v, view in enumerate(views):
        h = hit[v].cpu().numpy().ravel()
        _, a, b, c = ob.loss_grad_fixed(xyz64, r64, rgb64, view.pose_colmajor(), view.fov, W, H,
                                        h, tgt_rgb[v].cpu().numpy().reshape(-1, 3),
                                        tgt_d[v].cpu().numpy().ravel(), 0.7, 1.3)
        ex += a
        er += b
        ec += c
        lin = np.nonzero(h >= 0)[0]
        q = _disc_ok(view, xyz64, r64, h[lin], lin)
        bad[h[lin][q < 0.05]] = True           # near-grazing pixels: 1/sqrt(disc) blows up
    hitpts = np.unique(np.concatenate([hit[v].cpu().numpy().ravel() for v in range(V)]))
    hitpts = hitpts[hitpts >= 0]
    good = hitpts[~bad[hitpts]]
    assert good.size > 200
    assert np.allclose(gc[hitpts], ec[hitpts], rtol=1e-4, atol=1e-4)
    assert np.allclose(gx[good], ex[good], rtol=2e-3, atol=2e-3)
    assert np.allclose(gr[good], er[good], rtol=2e-3, atol=2e-3)
    miss = np.setdiff1d(np.arange(cloud.n), hitpts)
    assert (gx[miss] == 0).all() and (gr[miss] == 0).all() and (gc[miss] == 0).all()


def test_central_fd_of_gpu_forward(rm_ctx, small_scene):
    """d(loss)/d(centre) and d/d(radius) by central differences (step 1e-4 m) of the GPU forward,
    re-marched, on points away from silhouettes (every pixel of the point well inside its
    sphere's disc and the hit map unchanged by the step). The loss covers the pixels whose depth
    is the analytic entry depth; pixels that only graze their sphere report the marched depth,
    which the adjoint treats as constant, so they are masked out (mask fixed before the step)."""
    cloud, rc, views, R, xyz, r, rgb = small_scene
    rng = np.random.default_rng(2)
    img, depth, hit = R.render(xyz, r, rgb)
    tgt_d = depth + torch.from_numpy(rng.uniform(0.01, 0.05, depth.shape).astype(np.float32)).cuda()
    hv = hit.cpu().numpy()
    xyz64, r64 = cloud.xyz.astype(np.float64), rc.astype(np.float64)
    mask = np.zeros(hv.shape, bool)
    bad = np.zeros(cloud.n, bool)
    for v, view in enumerate(views):
        hh = hv[v].ravel()
        lin = np.nonzero(hh >= 0)[0]
        q = _disc_ok(view, xyz64, r64, hh[lin], lin)
        mv = mask[v].reshape(-1)
        mv[lin[q > 0.01]] = True
        bad[hh[lin][q < 0.2]] = True
    m = torch.from_numpy(mask.astype(np.float32)).cuda()

    def loss_of(x, rr):
        _, d, h = R.render(x, rr, rgb)
        return float(((d.double() - tgt_d.double()).abs() * m.double()).sum().item()), h

    x = xyz.clone().requires_grad_(True)
    rr = r.clone().requires_grad_(True)
    _, d, h0 = R.render(x, rr, rgb)
    ((d - tgt_d).abs() * m).sum().backward()
    gx, gr = x.grad.cpu().numpy(), rr.grad.cpu().numpy()
    counts = np.bincount(hv[hv >= 0].ravel(), minlength=cloud.n)
    cand = np.nonzero((counts >= 4) & ~bad)[0]
    rng.shuffle(cand)
    step = 1e-4
    checked = 0
    for i in cand[:60]:
        for a in range(4):
            xp, xm = xyz.clone(), xyz.clone()
            rp, rm_ = r.clone(), r.clone()
            if a < 3:
                xp[i, a] += step
                xm[i, a] -= step
            else:
                rp[i] += step
                rm_[i] -= step
            lp, hp = loss_of(xp, rp)
            lm, hm = loss_of(xm, rm_)
            if not (torch.equal(hp, h0) and torch.equal(hm, h0)):
                continue
            fd = (lp - lm) / (2 * step)
            g = gx[i, a] if a < 3 else gr[i]
            assert abs(fd - g) <= 0.02 * max(1.0, abs(g)), (i, a, fd, g)
            checked += 1
    R.render(xyz, r, rgb)          # restore the scene's tree to the unperturbed centres
    assert checked >= 40


def test_adam_refinement_reduces_loss(rm_ctx):
    diff = _diff()
    cloud = scenes.box_scene(60000, seed=4)
    tree = ob.RefTree(cloud.xyzw(), 64)
    _, rc, _ = tree.radii(0.05)
    views = [scenes.box_scene_view(cloud, 96, 96, 84.0, distance=5.0, seed=s) for s in range(4)]
    xyz_c = torch.from_numpy(cloud.xyz).cuda()
    r_c = torch.from_numpy(rc).cuda()
    rgb_c = torch.from_numpy(cloud.colors4()[:, :3].copy()).cuda()
    Rt = diff.DiffRenderer(xyz_c, r_c, rgb_c, views, ctx=rm_ctx)
    timg, tdep, thit = Rt.render(xyz_c, r_c, rgb_c)
    tdep = torch.where(thit >= 0, tdep, torch.zeros_like(tdep))
    x0, r0, c0 = diff.perturbed_cloud(cloud.xyz, rc, cloud.rgb, seed=104)
    x0, r0, c0 = (torch.from_numpy(a).cuda() for a in (x0, r0, c0))
    R = diff.DiffRenderer(x0, r0, c0, views, ctx=rm_ctx)
    _, _, _, losses = diff.refine(R, x0, r0, c0, timg, tdep, steps=40)
    assert np.isfinite(losses).all()
    assert losses[-1] < 0.7 * losses[0], losses[::8]


def test_central_fd_1000_points_full_workload(rm_ctx):
    """The differentiable workload (1M box-scene points, 16 views at 512x512, fx 450): central
    differences with step 1e-4 m on 1000 random non-silhouette points, for every centre
    coordinate and the radius. Each pixel's depth depends only on its hit point, so one
    perturbed render of the whole cloud gives every point's own finite difference: its loss
    is summed over the pixels it hits, and a point is used only where all of its pixels keep
    their hit in both perturbed renders and sit well inside their sphere's disc.

    Every one of the 1000 points must pass three checks, with the perturbation actually
    applied (x + h and x - h rounded to fp32) as the divisor:
      * forward: in every perturbed render, each of the point's pixels has the GPU depth (fp32)
        of the fp64 closed form at the same fixed hit and perturbed coordinates within 2e-5 m;
      * adjoint: the GPU gradient equals the fp64 closed form's central difference with step
        1e-7 m (truncation and rounding far below the tolerance) within 2e-3 of
        max(1, |gradient|);
      * step 1e-4: the GPU gradient equals the GPU forward's difference within 2 % plus the
        fp32 rounding of the GPU depths (their deviations from the closed form, summed over the
        point's pixels in both renders) plus the difference's own truncation error, measured
        as the gap between the
        fp64 closed form's differences at 1e-4 and at 1e-7 (the step is not small against the
        half-chords of spheres of about 1 cm, so a central difference is off by its
        third-derivative term)."""
    diff = _diff()
    cloud, views, max_r = scenes.workload_box_1m_diff()
    rt = ob.RefTree(cloud.xyzw(), 64)
    _, rc, _ = rt.radii(max_r)
    xyz = torch.from_numpy(cloud.xyz).cuda()
    r = torch.from_numpy(rc).cuda()
    rgb = torch.from_numpy(cloud.colors4()[:, :3].copy()).cuda()
    R = diff.DiffRenderer(xyz, r, rgb, views, ctx=rm_ctx)
    rng = np.random.default_rng(7)
    img, depth, hit = R.render(xyz, r, rgb)
    V, H, W = hit.shape
    hv = hit.cpu().numpy().reshape(V, -1)
    tgt_d = depth + torch.from_numpy(rng.uniform(0.01, 0.05, depth.shape).astype(np.float32)).cuda()
    xyz64, r64 = cloud.xyz.astype(np.float64), rc.astype(np.float64)
    mask = np.zeros(hv.shape, bool)
    bad = np.zeros(cloud.n, bool)
    for v, view in enumerate(views):
        lin = np.nonzero(hv[v] >= 0)[0]
        q = _disc_ok(view, xyz64, r64, hv[v][lin], lin)
        mask[v, lin[q > 0.01]] = True
        bad[hv[v][lin][q < 0.2]] = True
    m = torch.from_numpy(mask.reshape(V, H, W).astype(np.float32)).cuda()
    x = xyz.clone().requires_grad_(True)
    rr = r.clone().requires_grad_(True)
    _, d, _ = R.render(x, rr, rgb)
    ((d - tgt_d).abs() * m).sum().backward()
    g = np.concatenate([x.grad.cpu().numpy(), rr.grad.cpu().numpy()[:, None]], 1)
    pix_pt = hv[mask]                                # hit point of every masked pixel
    pix_lin = [np.nonzero(mask[v])[0] for v in range(V)]
    tgt64 = tgt_d.double().cpu().numpy().reshape(V, -1)[mask]
    npix = np.bincount(pix_pt, minlength=cloud.n)

    def loss64(xp64, rp64):
        """Per-point loss of the fp64 closed form at the fixed hits, and the depths. A pixel
        near its sphere's rim (a point marked bad, never picked) can lose its intersection
        under a perturbation: its depth is NaN and so are that point's sums."""
        with np.errstate(invalid="ignore"):
            t64 = np.concatenate([_entry_depth64(view, xp64, rp64, hv[v][pix_lin[v]],
                                                 pix_lin[v]) for v, view in enumerate(views)])
        return np.bincount(pix_pt, weights=np.abs(t64 - tgt64), minlength=cloud.n), t64

    step = 1e-4
    ok = (npix >= 4) & ~bad
    fd = np.zeros((cloud.n, 4))          # GPU forward, step 1e-4
    fd64 = np.zeros((cloud.n, 4))        # fp64 closed form, same perturbation
    fd64_tiny = np.zeros((cloud.n, 4))   # fp64 closed form, step 1e-7
    noise = np.zeros((cloud.n, 4))       # fp32 rounding of the GPU depths, per point
    maxdev = np.zeros(cloud.n)           # largest |GPU depth - fp64 depth| of a point's pixels
    for a in range(4):
        lo, lo64, coord = [], [], []
        rnd = np.zeros(cloud.n)
        for sgn in (1.0, -1.0):
            xp, rp = xyz.clone(), r.clone()
            if a < 3:
                xp[:, a] += sgn * step
            else:
                rp += sgn * step
            _, dp, hp = R.render(xp, rp, rgb)
            hp = hp.cpu().numpy().reshape(V, -1)
            dpm = dp.double().cpu().numpy().reshape(V, -1)[mask]
            lo.append(np.bincount(pix_pt, weights=np.abs(dpm - tgt64), minlength=cloud.n))
            # a point whose pixels change hit in either render is not comparable
            changed = np.bincount(pix_pt, weights=(hp[mask] != pix_pt), minlength=cloud.n)
            ok &= changed == 0
            xp64 = xp.double().cpu().numpy()
            rp64 = rp.double().cpu().numpy()
            l64, t64 = loss64(xp64, rp64)
            lo64.append(l64)
            dev = np.abs(dpm - t64)
            rnd += np.bincount(pix_pt, weights=dev, minlength=cloud.n)
            with np.errstate(invalid="ignore"):
                np.maximum.at(maxdev, pix_pt, dev)
            coord.append(xp64[:, a] if a < 3 else rp64)
        dx = coord[0] - coord[1]                     # the perturbation actually applied
        fd[:, a] = (lo[0] - lo[1]) / dx
        fd64[:, a] = (lo64[0] - lo64[1]) / dx
        noise[:, a] = rnd / dx
        tiny = []
        for sgn in (1.0, -1.0):
            xp64, rp64 = xyz64.copy(), r64.copy()
            if a < 3:
                xp64[:, a] += sgn * 1e-7
            else:
                rp64 += sgn * 1e-7
            tiny.append(loss64(xp64, rp64)[0])
        fd64_tiny[:, a] = (tiny[0] - tiny[1]) / 2e-7
    R.render(xyz, r, rgb)          # restore the tree to the unperturbed centres
    cand = np.nonzero(ok)[0]
    assert cand.size >= 1000, cand.size
    pick = rng.choice(cand, 1000, replace=False)
    g, fd, fd64, fd64_tiny, noise = g[pick], fd[pick], fd64[pick], fd64_tiny[pick], noise[pick]
    assert np.isfinite(fd64).all() and np.isfinite(fd64_tiny).all() and np.isfinite(noise).all()
    scale = np.maximum(1.0, np.abs(g))
    trunc = np.abs(fd64 - fd64_tiny)
    e_adj = np.abs(g - fd64_tiny)
    e_fd = np.abs(fd - g)
    print(f"forward: max |depth - fp64| {maxdev[pick].max():.3e} m; adjoint vs fp64 step "
          f"1e-7: max {(e_adj / scale).max():.3e}, median {np.median(e_adj / scale):.3e}; "
          f"step 1e-4: max err {(e_fd / scale).max():.3e}, truncation max "
          f"{(trunc / scale).max():.3e}, rounding max {(noise / scale).max():.3e}")
    assert maxdev[pick].max() < 2e-5, maxdev[pick].max()
    assert (e_adj <= 2e-3 * scale).all(), (e_adj / scale).max()
    assert (e_fd <= 0.02 * scale + noise + trunc).all(), ((e_fd - noise - trunc) / scale).max()
    assert np.median(e_adj / scale) < 2e-4


def test_batch_render_diff_mixed_sizes(rm_ctx, small_scene):
    """rm_batch_render_diff over views of different (odd) sizes writes each view's rgb / depth
    at that view's offset, equal to rm_render_diff of the view alone on every pixel."""
    import ctypes as C
    from renderer_ray_marching_amd import _abi, api
    cloud, rc, views, R, xyz, r, rgb = small_scene
    vs = [scenes.box_scene_view(cloud, w, h, f, distance=5.0, seed=s)
          for (w, h, f, s) in ((96, 72, 80.0, 4), (33, 17, 30.0, 5), (64, 130, 70.0, 6))]
    b = api.Batch(R.marcher, vs)
    total = sum(v.width * v.height for v in vs)
    rgb_b = torch.full((total, 3), -1.0, device="cuda")
    dep_b = torch.full((total,), -1.0, device="cuda")
    L = _abi.lib()
    assert L.rm_batch_render_diff(b.h, C.c_void_p(rgb_b.data_ptr()),
                                  C.c_void_p(dep_b.data_ptr())) == 0
    rm_ctx.sync()
    off = 0
    for view in vs:
        W, H = view.width, view.height
        im1 = torch.empty((H, W, 3), device="cuda")
        d1 = torch.empty((H, W), device="cuda")
        h1 = torch.empty((H, W), dtype=torch.int32, device="cuda")
        pose = (C.c_float * 16)(*[float(x) for x in view.pose_colmajor()])
        assert L.rm_render_diff(rm_ctx.h, R.marcher.h, pose, C.c_float(view.fov), W, H,
                                C.c_void_p(im1.data_ptr()), C.c_void_p(d1.data_ptr()),
                                C.c_void_p(h1.data_ptr())) == 0
        rm_ctx.sync()
        n = W * H
        assert (h1 >= 0).any()
        assert torch.equal(b.hit[off:off + n], h1.reshape(-1))
        assert torch.equal(rgb_b[off:off + n], im1.reshape(-1, 3))
        assert torch.equal(dep_b[off:off + n], d1.reshape(-1))
        off += n
    b.close()


def test_op_runs_on_the_current_stream(rm_ctx, small_scene):
    """Forward and backward called on a side stream (torch.cuda.stream) give the default
    stream's outputs and gradients bit for bit, and the context queues on the stream it had
    before the calls afterwards."""
    import ctypes as C
    from renderer_ray_marching_amd import _abi
    cloud, rc, views, R, xyz, r, rgb = small_scene

    def run():
        x = xyz.clone().requires_grad_(True)
        rr = r.clone().requires_grad_(True)
        c = rgb.clone().requires_grad_(True)
        img, d, h = R.render(x, rr, c)
        (img.sum() + d.sum()).backward()
        return [img.detach().clone(), d.detach().clone(), h.clone(), x.grad.clone(),
                rr.grad.clone(), c.grad.clone()]

    before = C.c_void_p()
    assert _abi.lib().rm_ctx_get_stream(rm_ctx.h, C.byref(before)) == 0
    ref = run()
    torch.cuda.synchronize()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        got = run()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    after = C.c_void_p()
    assert _abi.lib().rm_ctx_get_stream(rm_ctx.h, C.byref(after)) == 0
    assert after.value == before.value
    assert (ref[2] >= 0).any()
    for a, b in zip(ref, got):
        assert torch.equal(a, b)


def test_batched_forward_equals_per_view(rm_ctx, small_scene):
    """The op's forward (one batched march over all views, rm_batch_render_diff) equals
    rm_render_diff called view by view, on every pixel."""
    import ctypes as C
    from renderer_ray_marching_amd import _abi
    cloud, rc, views, R, xyz, r, rgb = small_scene
    img, depth, hit = R.render(xyz, r, rgb)
    L = _abi.lib()
    for v, view in enumerate(views):
        W, H = view.width, view.height
        im1 = torch.empty((H, W, 3), device="cuda")
        d1 = torch.empty((H, W), device="cuda")
        h1 = torch.empty((H, W), dtype=torch.int32, device="cuda")
        pose = (C.c_float * 16)(*[float(x) for x in view.pose_colmajor()])
        rc_ = L.rm_render_diff(rm_ctx.h, R.marcher.h, pose, C.c_float(view.fov), W, H,
                               C.c_void_p(im1.data_ptr()), C.c_void_p(d1.data_ptr()),
                               C.c_void_p(h1.data_ptr()))
        assert rc_ == 0
        rm_ctx.sync()
        assert torch.equal(h1, hit[v]) and torch.equal(im1, img[v]) and torch.equal(d1, depth[v])


def _per_view_forward_and_adjoint(rm_ctx, marcher, vs, g_img, g_dep, offs):
    """rm_render_diff and rm_render_adjoint view by view (at the cloud's current parameters)."""
    import ctypes as C
    from renderer_ray_marching_amd import _abi
    L = _abi.lib()
    n = marcher.n
    g = [torch.zeros((n, 3), device="cuda"), torch.zeros((n,), device="cuda"),
         torch.zeros((n, 3), device="cuda")]
    outs = []
    for i, view in enumerate(vs):
        W, H = view.width, view.height
        im1 = torch.empty((H, W, 3), device="cuda")
        d1 = torch.empty((H, W), device="cuda")
        h1 = torch.empty((H, W), dtype=torch.int32, device="cuda")
        pose = (C.c_float * 16)(*[float(x) for x in view.pose_colmajor()])
        assert L.rm_render_diff(rm_ctx.h, marcher.h, pose, C.c_float(view.fov), W, H,
                                C.c_void_p(im1.data_ptr()), C.c_void_p(d1.data_ptr()),
                                C.c_void_p(h1.data_ptr())) == 0
        a, b = offs[i], offs[i + 1]
        gi = g_img[a:b].contiguous()
        gd = g_dep[a:b].contiguous()
        assert L.rm_render_adjoint(rm_ctx.h, marcher.h, pose, C.c_float(view.fov), W, H,
                                   C.c_void_p(h1.data_ptr()), C.c_void_p(gi.data_ptr()),
                                   C.c_void_p(gd.data_ptr()), *[C.c_void_p(t.data_ptr())
                                                                for t in g]) == 0
        outs.append((im1.reshape(-1, 3), d1.reshape(-1), h1.reshape(-1)))
    rm_ctx.sync()
    return outs, g


def test_op_mixed_view_sizes_forward_and_backward(rm_ctx):
    """The autograd op over views of three sizes: outputs concatenated view after view, each
    equal to rm_render_diff of the view alone; the backward (one adjoint over all views) equals
    the per-view adjoints summed, up to the order of the float sums."""
    diff = _diff()
    cloud = scenes.box_scene(30000, seed=14)
    tree = ob.RefTree(cloud.xyzw(), 64)
    _, rc, _ = tree.radii(0.05)
    rc = np.maximum(rc, np.float32(0.01))
    vs = [scenes.box_scene_view(cloud, w, h, f, distance=5.0, seed=s)
          for (w, h, f, s) in ((96, 72, 80.0, 4), (33, 17, 30.0, 5), (64, 130, 70.0, 6))]
    xyz = torch.from_numpy(cloud.xyz).cuda()
    r = torch.from_numpy(rc).cuda()
    rgb = torch.from_numpy(cloud.colors4()[:, :3].copy()).cuda()
    R = diff.DiffRenderer(xyz, r, rgb, vs, ctx=rm_ctx)
    offs = R.view_offsets()
    assert offs[-1] == sum(v.width * v.height for v in vs)
    x = xyz.clone().requires_grad_(True)
    rr = r.clone().requires_grad_(True)
    cc = rgb.clone().requires_grad_(True)
    img, depth, hit = R.render(x, rr, cc)
    assert img.shape == (offs[-1], 3) and depth.shape == (offs[-1],) and hit.shape == (offs[-1],)
    rng = np.random.default_rng(3)
    g_img = torch.from_numpy(rng.normal(size=(offs[-1], 3)).astype(np.float32)).cuda()
    g_dep = torch.from_numpy(rng.normal(size=offs[-1]).astype(np.float32)).cuda()
    ((img * g_img).sum() + (depth * g_dep).sum()).backward()
    outs, g = _per_view_forward_and_adjoint(rm_ctx, R.marcher, vs, g_img, g_dep, offs)
    for i, (im1, d1, h1) in enumerate(outs):
        a, b = offs[i], offs[i + 1]
        assert (h1 >= 0).any()
        assert torch.equal(hit[a:b], h1) and torch.equal(img[a:b], im1)
        assert torch.equal(depth[a:b], d1)
    for got, ref in ((x.grad, g[0]), (rr.grad, g[1]), (cc.grad, g[2])):
        assert torch.allclose(got, ref, rtol=1e-5, atol=1e-5 * float(ref.abs().max())), \
            float((got - ref).abs().max())
        assert float(ref.abs().max()) > 0


def test_backward_of_an_earlier_forward(rm_ctx, small_scene):
    """Two forwards at different centres, then the backward of the first: equal to a lone
    forward + backward at the first centres (the backward uses the centres and radii its own
    forward marched, not the cloud's latest)."""
    cloud, rc, views, R, xyz, r, rgb = small_scene
    rng = np.random.default_rng(7)
    g_img = torch.from_numpy(rng.normal(size=(len(views), 72, 96, 3)).astype(np.float32)).cuda()
    g_dep = torch.from_numpy(rng.normal(size=(len(views), 72, 96)).astype(np.float32)).cuda()

    def fwd(x0, r0):
        x = x0.clone().requires_grad_(True)
        rr = r0.clone().requires_grad_(True)
        cc = rgb.clone().requires_grad_(True)
        img, depth, _ = R.render(x, rr, cc)
        return (img * g_img).sum() + (depth * g_dep).sum(), (x, rr, cc)

    moved = xyz + 0.003 * torch.randn(xyz.shape, device="cuda",
                                      generator=torch.Generator("cuda").manual_seed(1))
    loss_a, pa = fwd(xyz, r)
    loss_b, pb = fwd(moved, r * 1.1)
    loss_a.backward()
    loss_b.backward()
    loss_c, pc = fwd(xyz, r)
    loss_c.backward()
    for t_a, t_c in zip(pa, pc):
        assert torch.equal(t_a.grad, t_c.grad)
    assert not torch.equal(pa[0].grad, pb[0].grad)
    # editing a saved input in place after its forward makes autograd refuse the backward
    loss_d, pd = fwd(xyz, r)
    with torch.no_grad():
        pd[0].add_(0.001)
    with pytest.raises(RuntimeError):
        loss_d.backward()


def test_refit_through_another_cloud_is_refused(rm_ctx):
    """Two clouds on one tree: a refit through one leaves the other's centres stale, so the
    other's march calls fail until it is refit too."""
    from renderer_ray_marching_amd import _abi, api
    from renderer_ray_marching_amd._abi import RMError
    cloud = scenes.box_scene(5000, seed=2)
    p = torch.from_numpy(cloud.xyzw()).cuda()
    col = torch.from_numpy(cloud.colors4()).cuda()
    t = api.KDTreeFlann(rm_ctx)
    t.Build(p)
    _, rc = t.radii(0.1)
    a = api.PointcloudRayMarcher(p, col, rc, t)
    b = api.PointcloudRayMarcher(p, col, rc, t)
    v = scenes.box_scene_view(cloud, 40, 30, 30.0, 3.0, seed=1)
    a.render_to_texture(v.pose_colmajor(), v.fov, 40, 30)
    moved = (p[:, :3] + 0.001).contiguous()
    a.refit(moved)
    a.render_to_texture(v.pose_colmajor(), v.fov, 40, 30)
    with pytest.raises(RMError) as e:
        b.render_to_texture(v.pose_colmajor(), v.fov, 40, 30)
    assert e.value.code == _abi.RM_ERR_INVALID and "refit" in str(e.value)
    with pytest.raises(RMError):
        api.Batch(b, [v])
    b.refit(moved)
    rgba_b, hit_b = b.render_to_texture(v.pose_colmajor(), v.fov, 40, 30)
    with pytest.raises(RMError):                 # now a is the stale one
        a.render_to_texture(v.pose_colmajor(), v.fov, 40, 30)
    a.refit(moved)
    rgba_a, hit_a = a.render_to_texture(v.pose_colmajor(), v.fov, 40, 30)
    rm_ctx.sync()
    assert torch.equal(hit_a, hit_b) and torch.equal(rgba_a, rgba_b)

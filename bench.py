"""Benchmark: batched vs per-view sphere ray marching on one MI355X.

Workload (default): 10M points from the seeded room generator (20 x 16 x 3 m, 800 boxes, 5 mm
noise, 4x density variation, 0.5 % outliers, radii clamped at 0.05 m) and 256 views of mixed
resolution (96 x 640x480, 96 x 1067x690, 64 x 1920x1080) = 232,879,680 rays per step.

A step is one pass of the batched march over all 256 views (one persistent kernel launched as a
HIP graph). Beside it, on the same inputs:
  * the plain path: one launch per view, the reference's launch shape (256 sequential launches);
  * the same batched kernel launched once per view (256 sequential one-view graphs), which
    separates what the kernel gains from what batching the views gains;
  * the kNN radii kernels, the kNN / radius-search API, and the differentiable form;
  * the CPU oracle (the reference algorithm restated in C++/OpenMP) on the host cores, which also
    checks the batched output pixel by pixel on a fixed 1-in-64 subsample of every view.
Prints one JSON line (one GPU).

    python bench.py [--steps K] [--warmup W] [--workload room10m|room2m|box100k|scale50m]
"""
import argparse
import hashlib
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
SIMDS = 256 * 4                # 256 CUs x 4 SIMDs
XCDS = 8                       # GRBM_GUI_ACTIVE is summed over the 8 XCDs
CYCLES_PER_WAVE64_VALU = 4     # a wave64 VALU instruction occupies its SIMD for 4 cycles


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="room10m",
                    choices=["room10m", "room2m", "box100k", "scale50m"])
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="use the default launch tune")
    ap.add_argument("--tune", default=None,
                    help="JSON rm_tune to use instead of autotuning (e.g. from a previous run)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy first (HBM counter calibration)")
    ap.add_argument("--no-diff", action="store_true",
                    help="skip the differentiable-form leg (fwd and fwd+bwd ms/view)")
    ap.add_argument("--diff-steps", type=int, default=200,
                    help="Adam steps of the differentiable leg")
    ap.add_argument("--plain-views", type=int, default=0,
                    help="time the plain and one-view-batched paths on the first N views only "
                         "(0 = all)")
    ap.add_argument("--leaf", type=int, default=0,
                    help="points per tree leaf (0 = the library default, 16)")
    ap.add_argument("--cull", action="store_true",
                    help="march over each view's frustum-culled cloud (off in the reference)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(views would be dealt to GPUs with no communication; see the design notes).",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from renderer_ray_marching_amd import api, scenes

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    wl = {"room10m": scenes.workload_room_10m, "room2m": scenes.workload_room_2m,
          "box100k": scenes.workload_box_100k, "scale50m": scenes.workload_scale_50m}[a.workload]
    t0 = time.time()
    cloud, views, max_r = wl()
    log(f"scene {a.workload}: {cloud.n} points, {len(views)} views ({time.time() - t0:.1f}s)")
    rays = sum(v.width * v.height for v in views)

    if a.copy_calibration:
        src = torch.ones(256 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)          # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        del src, dst
    ctx = api.Context(0)
    p = torch.from_numpy(cloud.xyzw()).cuda()
    col = torch.from_numpy(cloud.colors4()).cuda()
    t0 = time.time()
    tree = api.KDTreeFlann(ctx, leaf_max=a.leaf)
    tree.Build(p)
    info = tree.info()
    log(f"tree: {info} ({time.time() - t0:.1f}s)")

    # ---- kNN radii kernels (k = 2 in leaf order, then the original order, clamped) ----
    ctx.timing(True)
    ru, rc = tree.radii(max_r)
    ctx.sync()
    radii_ms = []
    for _ in range(5):
        tree.radii(max_r)
        radii_ms.append(ctx.last_kernel_ms())
    radii_ms = float(np.median(radii_ms))
    log(f"radii kernels: {radii_ms:.3f} ms")

    search = search_leg(api, ctx, tree, p, torch)
    marcher = api.PointcloudRayMarcher.get_instance(p, col, rc, tree)
    if a.cull:
        marcher.set_culling(True)

    # ---- plain per-view path: one launch per view, sequential ----
    nplain = len(views) if a.plain_views <= 0 else min(a.plain_views, len(views))
    plain_out = []
    for v in views[:nplain]:
        plain_out.append((torch.empty((v.height, v.width, 4), device="cuda"),
                          torch.empty((v.height, v.width), dtype=torch.int32, device="cuda")))
    poses = [v.pose_colmajor() for v in views]
    marcher.render_to_texture(poses[0], views[0].fov, views[0].width, views[0].height,
                              *plain_out[0])
    ctx.sync()
    torch.cuda.synchronize()
    tp = time.perf_counter()
    for i in range(nplain):
        v = views[i]
        marcher.render_to_texture(poses[i], v.fov, v.width, v.height, *plain_out[i])
    ctx.sync()
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - tp
    plain_rays = sum(v.width * v.height for v in views[:nplain])
    plain_mrays = plain_rays / plain_s / 1e6
    log(f"plain: {nplain} views, {plain_s:.3f}s, {plain_mrays:.1f} Mrays/s")

    # ---- batched path: persistent kernel, tile queue, HIP graph ----
    fixed = json.loads(a.tune) if a.tune else None
    batch = api.Batch(marcher, views, tune=fixed)
    tune = batch.tune()
    if not a.no_tune and fixed is None:
        tt = time.time()
        tune = batch.autotune(sample_stride=16)
        log(f"autotune: {tune} ({time.time() - tt:.1f}s)")

    for _ in range(a.warmup):
        batch.launch()
        ctx.last_kernel_ms()
    ctx.sync()
    torch.cuda.synchronize()
    kernel_ms = []
    t_start = time.perf_counter()
    for _ in range(a.steps):
        batch.launch()
        kernel_ms.append(ctx.last_kernel_ms())
        log(f"step: {kernel_ms[-1]:.1f} ms")
    ctx.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ms_per_step = wall / a.steps * 1e3
    value = rays * a.steps / wall / 1e6
    avg_kernel_ms = float(np.mean(kernel_ms))

    # ---- the same kernel, one view per launch (256 sequential one-view graphs) ----
    singles = [api.Batch(marcher, [v], tune=tune) for v in views[:nplain]]
    singles[0].launch()
    ctx.sync()
    torch.cuda.synchronize()
    ts = time.perf_counter()
    for sb in singles:
        sb.launch()
    ctx.sync()
    torch.cuda.synchronize()
    single_s = time.perf_counter() - ts
    single_mrays = plain_rays / single_s / 1e6
    log(f"batched kernel one view per launch: {nplain} launches, {single_s:.3f}s, "
        f"{single_mrays:.1f} Mrays/s")

    # every pixel of the batched output must equal the plain path and the one-view launches
    agree = agree_single = 0
    same_rgba = same_rgba_single = True
    off = 0
    for i in range(nplain):
        v = views[i]
        n = v.width * v.height
        ph = plain_out[i][1].reshape(-1)
        agree += int((batch.hit[off:off + n] == ph).sum().item())
        agree_single += int((batch.hit[off:off + n] == singles[i].hit).sum().item())
        same_rgba &= bool(torch.equal(batch.rgba[off:off + n], plain_out[i][0].reshape(-1, 4)))
        same_rgba_single &= bool(torch.equal(batch.rgba[off:off + n], singles[i].rgba))
        off += n
    for sb in singles:
        sb.close()
    del singles
    hit_agreement = agree / plain_rays
    hit_frac = float((batch.hit >= 0).float().mean().item())
    log(f"batched: {value:.1f} Mrays/s, agreement with plain {hit_agreement:.6f}, with one-view "
        f"launches {agree_single / plain_rays:.6f}")
    cull_info = None
    if a.cull:
        # what culling changes: hit agreement with the unculled batched march
        marcher.set_culling(False)
        full = api.Batch(marcher, views, tune=tune)
        full.launch()
        ctx.sync()
        cull_info = {"hit_agreement_vs_unculled":
                     float((full.hit == batch.hit).float().mean().item()),
                     "hit_fraction_unculled": round(float((full.hit >= 0).float().mean()
                                                          .item()), 4)}
        log(f"culled vs unculled hit agreement {cull_info['hit_agreement_vs_unculled']:.6f}")
        full.close()
        del full

    # the batched output on the 1-in-64 pixel subsample, kept for the check against the oracle
    sub_pix = [subsample(v.width, v.height) for v in views]
    offs = np.cumsum([0] + [v.width * v.height for v in views])
    sel = torch.from_numpy(np.concatenate([offs[i] + sp for i, sp in enumerate(sub_pix)])).cuda()
    sub_hit = batch.hit[sel].cpu().numpy()
    sub_rgba = batch.rgba[sel].cpu().numpy()
    gpu_radii = rc.cpu().numpy()

    counters = profile_counters(a.workload, tune)
    roofline = march_roofline(a.workload, rays, avg_kernel_ms, counters) if not a.cull else None
    radii_roof = radii_roofline(a.workload, cloud.n, radii_ms, counters)

    del batch, plain_out, marcher, tree, p, col
    torch.cuda.empty_cache()
    diff = None if a.no_diff else diff_leg(api, scenes, torch, a.diff_steps)

    # ---- CPU baseline: the oracle on the host cores; also the check of the batched output ----
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(cloud, views, max_r, sub_pix, sub_hit, sub_rgba, gpu_radii,
                           culled=a.cull)

    out = {
        "metric": "Mrays/s batched vs per-view, 10M pts, >=99.9% hit match; fwd+bwd ms/view",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded room generator)",
        "config": {"workload": f"{a.workload}: {cloud.n} points, {len(views)} views, "
                               f"{rays} rays per step", "points": cloud.n, "views": len(views),
                   "rays_per_step": rays, "max_radius": max_r, "parallelism": "single GPU",
                   "tree": info, "culling": a.cull},
        "plain_mrays_s": round(plain_mrays, 2),
        "plain_views_timed": nplain,
        "batched_vs_plain": round(value / plain_mrays, 3),
        "per_view_batched_mrays_s": round(single_mrays, 2),
        "batched_vs_per_view_batched": round(value / single_mrays, 3),
        "per_view_batched_vs_plain": round(single_mrays / plain_mrays, 3),
        "hit_agreement_vs_plain": hit_agreement,
        "rgba_identical_to_plain": same_rgba,
        "hit_agreement_vs_per_view_batched": agree_single / plain_rays,
        "rgba_identical_to_per_view_batched": same_rgba_single,
        "hit_agreement_vs_oracle": cpu["hit_agreement_vs_oracle"] if cpu else None,
        "hit_fraction": round(hit_frac, 4),
        "tune": tune,
        "radii_ms": round(radii_ms, 4),
        "radii_roofline": radii_roof,
        "search": search,
        "roofline": roofline,
        "diff": diff,
        "cpu_baseline": cpu,
    }
    if a.cull:
        out["cull"] = cull_info
    print(json.dumps(out), flush=True)


def subsample(W, H):
    """Pixels with (x + 7 y) mod 64 == 0: a fixed 1-in-64 sample spread over the whole view."""
    lin = np.arange(W * H, dtype=np.int64)
    return lin[(lin % W + 7 * (lin // W)) % 64 == 0]


def kernel_source_sha256():
    src = os.path.join(ROOT, "renderer_ray_marching_amd", "csrc", "rm_kernels.hip")
    return hashlib.sha256(open(src, "rb").read()).hexdigest()


def search_leg(api, ctx, tree, p, torch):
    """Batched queries against the workload's tree (the SearchKNN / SearchRadius API): 1M
    queries, points of the cloud moved by N(0, 1 cm); device time per launch (median of 5)."""
    g = torch.Generator(device="cuda").manual_seed(7)
    m = 1_000_000
    sel = torch.randint(0, p.shape[0], (m,), device="cuda", generator=g)
    q = p[sel].clone()
    q[:, :3] += 0.01 * torch.randn((m, 3), device="cuda", generator=g)
    q = q.contiguous()
    out = {"queries": m}

    def timed(fn):
        fn()
        ctx.sync()
        ms = []
        for _ in range(5):
            fn()
            ms.append(ctx.last_kernel_ms())
        ctx.sync()
        return float(np.median(ms))

    ms = timed(lambda: tree.SearchKNN(q, 10))
    out["knn10_ms"] = round(ms, 3)
    out["knn10_mqueries_s"] = round(m / ms / 1e3, 1)
    ms = timed(lambda: tree.SearchRadius(q, 0.05, 16))
    out["radius_r005_nn16_ms"] = round(ms, 3)
    out["radius_mqueries_s"] = round(m / ms / 1e3, 1)
    ms = timed(lambda: tree.SearchRadius(q, 0.05, 0))
    out["radius_count_r005_ms"] = round(ms, 3)
    out["radius_count_mqueries_s"] = round(m / ms / 1e3, 1)
    log(f"search: kNN-10 {out['knn10_mqueries_s']} Mq/s, radius {out['radius_mqueries_s']} Mq/s, "
        f"count {out['radius_count_mqueries_s']} Mq/s")
    return out


def profile_counters(workload, tune):
    """Per-launch counters of the march and radii kernels from the committed rocprofv3 summary
    (profiles/counters_<workload>.txt: FETCH_SIZE and WRITE_SIZE calibrated against a 1 GiB copy,
    L2 hits, instruction counts, busy cycles). Used only when that summary was taken on this
    workload from the same kernel source (SHA-256 of rm_kernels.hip); the launch tune it was
    taken with is reported beside it."""
    path = os.path.join(ROOT, "profiles", f"counters_{workload}.txt")
    if not os.path.exists(path):
        return {}
    rec = {}
    for line in open(path):
        line = line.strip()
        if line and not line.startswith("#") and ":" in line:
            k, v = line.split(":", 1)
            rec[k.strip()] = v.strip()
    if rec.get("WORKLOAD") != workload or rec.get("KERNEL_SRC_SHA256") != kernel_source_sha256():
        return {}
    out = {"tune": json.loads(rec["TUNE"]) if "TUNE" in rec else None}
    for k, v in rec.items():
        if "." in k:
            kern, name = k.split(".", 1)
            try:
                out.setdefault(kern, {})[name] = float(v)
            except ValueError:
                pass
    return out


def _alg(workload):
    path = os.path.join(ROOT, "tests", "golden", "alg_bytes.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get(workload)


def _binding(c):
    """VALU issue use and lane use of one kernel from its counters: wave64 VALU instructions x 4
    cycles over the SIMD-cycles of the launch (GRBM_GUI_ACTIVE / 8 XCDs x 1024 SIMDs), and the
    mean fraction of the 64 lanes active per VALU instruction (SQ_THREAD_CYCLES_VALU over
    SQ_ACTIVE_INST_VALU, scaled by the same ratio of a copy kernel, whose lanes are all active)."""
    if not c or "SQ_INSTS_VALU" not in c or "GRBM_GUI_ACTIVE" not in c:
        return None
    cycles = c["GRBM_GUI_ACTIVE"] / XCDS
    b = {"resource": "VALU issue",
         "valu_issue_frac": round(c["SQ_INSTS_VALU"] * CYCLES_PER_WAVE64_VALU /
                                  (cycles * SIMDS), 4)}
    if "SQ_INSTS_SALU" in c:
        b["salu_issue_frac"] = round(c["SQ_INSTS_SALU"] * CYCLES_PER_WAVE64_VALU /
                                     (cycles * SIMDS), 4)
    if "VALU_LANE_FRAC" in c:
        b["active_lane_frac"] = round(c["VALU_LANE_FRAC"], 4)
        b["useful_valu_frac"] = round(b["valu_issue_frac"] * c["VALU_LANE_FRAC"], 4)
    return b


def _hbm(c, ms):
    if not c or "FETCH_BYTES" not in c or "WRITE_BYTES" not in c:
        return None, None
    traffic = c["FETCH_BYTES"] + c["WRITE_BYTES"]
    return traffic, traffic / (ms * 1e-3) / 1e9


def march_roofline(workload, rays, kernel_ms, counters):
    """HBM roofline of the batched march from measured traffic (the bytes its reads and writes
    actually move past the L2), with the binding resource beside it. The algorithmic byte counts
    (what an exact per-lane search of this tree reads, and what the reference's traversal reads)
    are reported as work figures: both exceed what HBM could deliver in the kernel's time, i.e.
    they are served by the caches."""
    c = counters.get("march")
    traffic, gbs = _hbm(c, kernel_ms)
    alg = _alg(workload)
    r = {"bound": "hbm",
         "achieved": round(gbs, 1) if gbs is not None else None,
         "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs is not None else None,
         "traffic": int(traffic) if traffic is not None else None,
         "achieved_basis": "measured FETCH_SIZE + WRITE_SIZE per launch (rocprofv3, calibrated "
                           "with a 1 GiB copy) over the live kernel time",
         "kernel": "batch_kernel (persistent march)", "kernel_ms": round(kernel_ms, 3),
         "l2_hit_rate": round(c["L2_HIT"], 4) if c and "L2_HIT" in c else None,
         "binding": _binding(c), "counters_tune": counters.get("tune")}
    if alg and alg.get("rays") == rays:
        compulsory = 20.0 * rays          # RGBA32F + hit index per ray
        w = {"compulsory_bytes_per_ray": 20.0,
             "compulsory_frac": round(compulsory / (kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
        if "this_march_bytes_per_ray" in alg:
            b = alg["this_march_bytes_per_ray"] * rays
            w["per_lane_bytes_per_ray"] = round(alg["this_march_bytes_per_ray"], 1)
            w["per_lane_gbs"] = round(b / (kernel_ms * 1e-3) / 1e9, 1)
        if "march_bytes_per_ray" in alg:
            b = alg["march_bytes_per_ray"] * rays
            w["reference_traversal_bytes_per_ray"] = round(alg["march_bytes_per_ray"], 1)
            w["reference_traversal_gbs"] = round(b / (kernel_ms * 1e-3) / 1e9, 1)
        if traffic is not None:
            w["traffic_over_compulsory"] = round(traffic / compulsory, 2)
        r["work"] = w
    return r


def radii_roofline(workload, n, radii_ms, counters):
    """The same for the radii pass (its two kernels: the k = 2 search in leaf order and the
    coalesced reorder)."""
    c = counters.get("radii")
    traffic, gbs = _hbm(c, radii_ms)
    alg = _alg(workload)
    r = {"bound": "hbm",
         "achieved": round(gbs, 1) if gbs is not None else None,
         "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs is not None else None,
         "traffic": int(traffic) if traffic is not None else None,
         "fetch_bytes": int(c["FETCH_BYTES"]) if c and "FETCH_BYTES" in c else None,
         "write_bytes": int(c["WRITE_BYTES"]) if c and "WRITE_BYTES" in c else None,
         "kernel_ms": round(radii_ms, 4),
         "l2_hit_rate": round(c["L2_HIT"], 4) if c and "L2_HIT" in c else None,
         "binding": _binding(c)}
    compulsory = 16.0 * n + 8.0 * n       # xyz read (16 B record) + two 4-byte radii written
    w = {"compulsory_bytes": int(compulsory),
         "compulsory_frac": round(compulsory / (radii_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5)}
    if c and "WRITE_BYTES" in c:
        w["write_over_results"] = round(c["WRITE_BYTES"] / (8.0 * n), 2)
    if alg and alg.get("points") == n:
        if "this_radii_bytes_per_query" in alg:
            w["per_lane_bytes_per_query"] = round(alg["this_radii_bytes_per_query"], 1)
            w["per_lane_gbs"] = round(alg["this_radii_bytes_per_query"] * n /
                                      (radii_ms * 1e-3) / 1e9, 1)
        w["reference_traversal_bytes_per_query"] = round(alg["radii_bytes_per_query"], 1)
        w["reference_traversal_gbs"] = round(alg["radii_bytes_total"] / (radii_ms * 1e-3) / 1e9,
                                             1)
    r["work"] = w
    return r


def diff_leg(api, scenes, torch, steps):
    """Differentiable form on 1M box-scene points, 16 views at 512x512: forward and
    forward+backward device time per view, and an Adam refinement from the perturbed cloud
    toward targets rendered from the clean one."""
    from renderer_ray_marching_amd import diff as D
    cloud, views, max_r = scenes.workload_box_1m_diff()
    ctx = api.Context(0)
    p = torch.from_numpy(cloud.xyzw()).cuda()
    tree = api.KDTreeFlann(ctx)
    tree.Build(p)
    _, rc = tree.radii(max_r)
    xyz_c = torch.from_numpy(cloud.xyz).cuda()
    rgb_c = torch.from_numpy(cloud.colors4()[:, :3].copy()).cuda()
    Rt = D.DiffRenderer(xyz_c, rc, rgb_c, views, ctx=ctx)
    timg, tdep, thit = Rt.render(xyz_c, rc, rgb_c)
    tdep = torch.where(thit >= 0, tdep, torch.zeros_like(tdep))
    x0, r0, c0 = D.perturbed_cloud(cloud.xyz, rc.cpu().numpy(), cloud.rgb, seed=104)
    x0, r0, c0 = (torch.from_numpy(v).cuda() for v in (x0, r0, c0))
    R = D.DiffRenderer(x0, r0, c0, views, ctx=ctx)
    x = x0.clone().requires_grad_(True)
    r = r0.clone().requires_grad_(True)
    c = c0.clone().requires_grad_(True)

    def loss_fn(img, depth, hit):
        m = ((hit >= 0) & (tdep > 0)).float()
        return (((img - timg) ** 2).sum(-1) * m).sum() + ((depth - tdep).abs() * m).sum()

    for _ in range(2):                                   # warm-up
        loss_fn(*R.render(x, r, c)).backward()
    torch.cuda.synchronize()
    reps = 5
    t0 = time.perf_counter()
    for _ in range(reps):
        with torch.no_grad():
            R.render(x, r, c)
    torch.cuda.synchronize()
    fwd = (time.perf_counter() - t0) / reps
    t0 = time.perf_counter()
    for _ in range(reps):
        loss_fn(*R.render(x, r, c)).backward()
    torch.cuda.synchronize()
    fb = (time.perf_counter() - t0) / reps
    nv = len(views)
    log(f"diff: fwd {fwd / nv * 1e3:.2f} ms/view, fwd+bwd {fb / nv * 1e3:.2f} ms/view")
    t0 = time.perf_counter()
    _, _, _, losses = D.refine(R, x0, r0, c0, timg, tdep, steps=steps)
    adam_s = time.perf_counter() - t0
    log(f"diff: Adam {steps} steps in {adam_s:.1f}s, loss {losses[0]:.1f} -> {losses[-1]:.1f}")
    return {"workload": f"box1m: {cloud.n} points, {nv} views 512x512",
            "fwd_ms_per_view": round(fwd / nv * 1e3, 3),
            "fwd_bwd_ms_per_view": round(fb / nv * 1e3, 3),
            "bwd_over_fwd": round(fb / fwd, 3),
            "adam_steps": steps, "adam_s": round(adam_s, 2),
            "loss_first": round(losses[0], 2), "loss_last": round(losses[-1], 2)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cloud, views, max_r, sub_pix, sub_hit, sub_rgba, gpu_radii, culled=False):
    """The oracle (C++/OpenMP restatement of the reference algorithm) on the host cores
    available to this process, as the baseline plan states it: the k = 2 radii of all points and
    the march of 8 whole views (every 32nd view: the mix of sizes), 1 warm-up run and the median
    of 5; the march also on 1 thread over a small sample. It then checks the batched GPU output
    against the oracle's march on the 1-in-64 pixel subsample of every view (with the oracle's
    own radii), and the GPU radii against the oracle's."""
    from oracle import binding as ob
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ncores = max(1, min(affinity, 64))
    t0 = time.time()
    tree = ob.RefTree(cloud.xyzw(), 64)
    log(f"cpu baseline: oracle tree ({time.time() - t0:.1f}s)")
    col = cloud.colors4()

    radii_s = []
    for rep in range(6):
        t1 = time.perf_counter()
        _, orc, _ = tree.radii(max_r, nthreads=ncores)
        if rep:
            radii_s.append(time.perf_counter() - t1)
    radii_s = float(np.median(radii_s))
    log(f"cpu baseline: radii of {cloud.n} points, median {radii_s:.2f}s")

    sample = views[::max(1, len(views) // 8)][:8]
    rays8 = sum(v.width * v.height for v in sample)

    def march(vs, nthreads, stride=1):
        t1 = time.perf_counter()
        for v in vs:
            pix = None if stride == 1 else np.arange(0, v.width * v.height, stride,
                                                      dtype=np.int32)
            tree.render(orc, col, v.pose_colmajor(), v.fov, v.width, v.height, pix=pix,
                        nthreads=nthreads)
        return time.perf_counter() - t1

    times = [march(sample, ncores) for _ in range(6)][1:]
    t_med = float(np.median(times))
    log(f"cpu baseline: {len(sample)} whole views ({rays8} rays), median {t_med:.2f}s on "
        f"{ncores} threads")
    rays1 = sum(len(range(0, v.width * v.height, 256)) for v in views[:16])
    march(views[:1], 1, 256)
    t1thr = march(views[:16], 1, 256)
    log(f"cpu baseline: {rays1} rays in {t1thr:.1f}s on 1 thread")

    agree_hit = agree_rgba = total = 0
    off = 0
    if not culled:
        for v, pix in zip(views, sub_pix):
            orgba, ohit, _ = tree.render(orc, col, v.pose_colmajor(), v.fov, v.width, v.height,
                                         pix=pix.astype(np.int32), nthreads=ncores)
            n = pix.shape[0]
            agree_hit += int((sub_hit[off:off + n] == ohit).sum())
            agree_rgba += int((sub_rgba[off:off + n] == orgba).all(1).sum())
            total += n
            off += n
        log(f"oracle check: {total} pixels, hit agreement {agree_hit / total:.6f}")
    return {"value": round(rays8 / t_med / 1e6, 4), "unit": "Mrays/s", "cores": ncores,
            "kind": "port",
            "sample": f"{len(sample)} whole views (every {max(1, len(views) // 8)}th view, "
                      f"{rays8} rays), oracle restatement of the reference march, OpenMP; "
                      "1 warm-up run, median of 5",
            "host_cores_total": os.cpu_count(), "affinity_cores": affinity,
            "cpu_model": cpu_model(),
            "value_1thread": round(rays1 / t1thr / 1e6, 4),
            "sample_1thread": f"every 256th pixel of the first {min(16, len(views))} views "
                              f"({rays1} rays)",
            "radii_s": round(radii_s, 3),
            "radii_mqueries_s": round(cloud.n / radii_s / 1e6, 2),
            "radii_sample": f"k = 2 radii of all {cloud.n} points, {ncores} threads, "
                            "1 warm-up run, median of 5",
            "radii_identical_to_oracle": bool(np.array_equal(gpu_radii, orc)),
            "hit_agreement_vs_oracle": (agree_hit / total) if total else None,
            "rgba_agreement_vs_oracle": (agree_rgba / total) if total else None,
            "oracle_pixels_checked": total,
            "oracle_check": "pixels with (x + 7 y) mod 64 == 0 of every view" if total else
                            "skipped (culled run)"}


if __name__ == "__main__":
    main()

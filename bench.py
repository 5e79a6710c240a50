"""Benchmark: batched vs per-view sphere ray marching on one MI355X.

Workload (default): 10M points from the seeded room generator (20 x 16 x 3 m, 800 boxes, 5 mm
noise, 4x density variation, 0.5 % outliers, radii clamped at 0.05 m) and 256 views of mixed
resolution (96 x 640x480, 96 x 1067x690, 64 x 1920x1080) = 232,879,680 rays per step.

A step is one pass of the batched march over all 256 views (one persistent kernel launched as a
HIP graph). The plain path (256 sequential per-view launches, the reference's launch shape) is
timed beside it on the same inputs; the two must agree on every pixel. The kNN radii kernel is
timed on the same cloud. Prints one JSON line (rank 0, one GPU).

    python bench.py [--steps K] [--warmup W] [--workload room10m|room2m|box100k]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


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
    ap.add_argument("--diff-steps", type=int, default=20,
                    help="Adam steps of the differentiable leg")
    ap.add_argument("--plain-views", type=int, default=0,
                    help="time the plain path on the first N views only (0 = all)")
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

    # ---- kNN radii kernel (k = 2, fused sqrt + clamp) ----
    ctx.timing(True)
    ru, rc = tree.radii(max_r)
    ctx.sync()
    radii_ms = []
    for _ in range(5):
        tree.radii(max_r)
        radii_ms.append(ctx.last_kernel_ms())
    radii_ms = float(np.median(radii_ms))
    log(f"radii kernel: {radii_ms:.3f} ms")

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

    # every pixel of the batched output must equal the plain path
    agree = 0
    same_rgba = True
    off = 0
    for i in range(nplain):
        v = views[i]
        n = v.width * v.height
        ph = plain_out[i][1].reshape(-1)
        agree += int((batch.hit[off:off + n] == ph).sum().item())
        same_rgba &= bool(torch.equal(batch.rgba[off:off + n], plain_out[i][0].reshape(-1, 4)))
        off += n
    hit_agreement = agree / plain_rays
    hit_frac = float((batch.hit >= 0).float().mean().item())
    log(f"batched: {value:.1f} Mrays/s, agreement with plain {hit_agreement:.6f}")
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

    # ---- roofline (algorithmic bytes counted by the oracle, committed per workload) ----
    alg_path = os.path.join(ROOT, "tests", "golden", "alg_bytes.json")
    roofline = None
    radii_roof = None
    if os.path.exists(alg_path):
        with open(alg_path) as f:
            alg = json.load(f).get(a.workload)
        if alg and alg["points"] == cloud.n and alg["rays"] == rays and not a.cull:
            b_launch = alg["march_bytes_per_ray"] * rays
            ach = b_launch / (avg_kernel_ms * 1e-3) / 1e9
            pmc = measured_traffic(a.workload)
            roofline = {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                        "traffic": pmc.get("TRAFFIC_BYTES_PER_LAUNCH"),
                        "kernel": "batch_kernel (persistent march)",
                        "alg_bytes_per_ray": round(alg["march_bytes_per_ray"], 1),
                        "alg_bytes_basis": "reference traversal (leaf 64, restart per step), "
                                           "counted by the oracle on a 1/64 pixel subsample",
                        "hbm_gbs_measured": (round(pmc["TRAFFIC_BYTES_PER_LAUNCH"] /
                                                   (avg_kernel_ms * 1e-3) / 1e9, 1)
                                             if "TRAFFIC_BYTES_PER_LAUNCH" in pmc else None),
                        "l2_hit_rate": pmc.get("L2_HIT_RATE"),
                        "valu_busy": pmc.get("VALU_BUSY"),
                        "pmc_tune": pmc.get("PMC_TUNE"),
                        "kernel_ms": round(avg_kernel_ms, 3)}
            rb = alg["radii_bytes_total"]
            rach = rb / (radii_ms * 1e-3) / 1e9
            radii_roof = {"achieved": round(rach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(rach / HBM_PEAK_GBS, 4), "kernel_ms": round(radii_ms, 4),
                          "alg_bytes_per_query": round(alg["radii_bytes_per_query"], 1),
                          "compulsory_frac": round(16 * cloud.n / (radii_ms * 1e-3) / 1e9 /
                                                   HBM_PEAK_GBS, 5),
                          "traffic": pmc.get("RADII_TRAFFIC_BYTES"),
                          "l2_hit_rate": pmc.get("RADII_L2_HIT_RATE")}

    del batch, plain_out, marcher, tree, p, col
    torch.cuda.empty_cache()
    diff = None if a.no_diff else diff_leg(api, scenes, torch, a.diff_steps)

    # ---- CPU baseline: the oracle's march on a bounded pixel sample, all host cores ----
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(cloud, views, rc.cpu().numpy(), max_r)

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
                   "tree": info},
        "plain_mrays_s": round(plain_mrays, 2),
        "plain_views_timed": nplain,
        "batched_vs_plain": round(value / plain_mrays, 3),
        "hit_agreement_vs_plain": hit_agreement,
        "rgba_identical_to_plain": same_rgba,
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


def kernel_source_sha256():
    import hashlib
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


def measured_traffic(workload):
    """HBM bytes per launch of the march and radii kernels from the committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE passes, calibrated with a 1 GiB copy), used only when that
    summary was taken on the same workload and from the same kernel source (SHA-256 of
    rm_kernels.hip). The launch tune it was taken with is returned as PMC_TUNE: the autotuner may
    pick an equivalent block shape in another run, and the report names the profiled one."""
    path = os.path.join(ROOT, "profiles", "march_batch_room10m_pmc.txt")
    if not os.path.exists(path):
        return {}
    rec = {}
    for line in open(path):
        if ":" in line and not line.startswith("#"):
            k, v = line.split(":", 1)
            rec[k.strip()] = v.strip()
    try:
        if rec.get("WORKLOAD") != workload:
            return {}
        if rec.get("KERNEL_SRC_SHA256") != kernel_source_sha256():
            return {}
        out = {"TRAFFIC_BYTES_PER_LAUNCH": int(rec["TRAFFIC_BYTES_PER_LAUNCH"]),
               "L2_HIT_RATE": float(rec["L2_HIT_RATE"]),
               "PMC_TUNE": json.loads(rec["TUNE"])}
        if "VALU_BUSY" in rec:
            out["VALU_BUSY"] = float(rec["VALU_BUSY"])
        if "RADII_TRAFFIC_BYTES" in rec:
            out["RADII_TRAFFIC_BYTES"] = int(rec["RADII_TRAFFIC_BYTES"])
        if "RADII_L2_HIT_RATE" in rec:
            out["RADII_L2_HIT_RATE"] = float(rec["RADII_L2_HIT_RATE"])
        return out
    except (KeyError, ValueError):
        return {}


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
    _, _, _, losses = D.refine(R, x0, r0, c0, timg, tdep, steps=steps)
    log(f"diff: Adam {steps} steps, loss {losses[0]:.1f} -> {losses[-1]:.1f}")
    return {"workload": f"box1m: {cloud.n} points, {nv} views 512x512",
            "fwd_ms_per_view": round(fwd / nv * 1e3, 3),
            "fwd_bwd_ms_per_view": round(fb / nv * 1e3, 3),
            "bwd_over_fwd": round(fb / fwd, 3),
            "adam_steps": steps, "loss_first": round(losses[0], 2),
            "loss_last": round(losses[-1], 2)}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(cloud, views, radii, max_r):
    """The oracle (C++/OpenMP restatement of the reference algorithm) on the host cores
    available to this process: the march over every 16th pixel of every view (the headline
    value), the same march on one thread over every 256th pixel of the first 16 views, and the
    k = 2 radii over all points."""
    from oracle import binding as ob
    ncores = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    ncores = max(1, min(ncores, 64))
    t0 = time.time()
    tree = ob.RefTree(cloud.xyzw(), 64)
    log(f"cpu baseline: oracle tree ({time.time() - t0:.1f}s)")
    col = cloud.colors4()

    def march(stride, vs, nthreads):
        rays, tcpu = 0, 0.0
        for v in vs:
            pix = np.arange(0, v.width * v.height, stride, dtype=np.int32)
            t1 = time.perf_counter()
            tree.render(radii, col, v.pose_colmajor(), v.fov, v.width, v.height, pix=pix,
                        nthreads=nthreads)
            tcpu += time.perf_counter() - t1
            rays += pix.shape[0]
        return rays, tcpu

    stride = 16
    rays, tcpu = march(stride, views, ncores)
    log(f"cpu baseline: {rays} rays in {tcpu:.1f}s on {ncores} threads")
    rays1, tcpu1 = march(256, views[:16], 1)
    log(f"cpu baseline: {rays1} rays in {tcpu1:.1f}s on 1 thread")
    t1 = time.perf_counter()
    tree.radii(max_r, nthreads=ncores)
    radii_s = time.perf_counter() - t1
    log(f"cpu baseline: radii of {cloud.n} points in {radii_s:.2f}s")
    return {"value": round(rays / tcpu / 1e6, 4), "unit": "Mrays/s", "cores": ncores,
            "kind": "port",
            "sample": f"every {stride}th pixel of all {len(views)} views ({rays} rays), "
                      "oracle restatement of the reference march, OpenMP",
            "value_1thread": round(rays1 / tcpu1 / 1e6, 4),
            "sample_1thread": f"every 256th pixel of the first {min(16, len(views))} views "
                              f"({rays1} rays)",
            "radii_s": round(radii_s, 3),
            "radii_sample": f"k = 2 radii of all {cloud.n} points, {ncores} threads",
            "cpu_model": cpu_model()}


if __name__ == "__main__":
    main()

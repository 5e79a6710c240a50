"""End-to-end runs of the drop-in CLI (marcher.bin) on the GPU, checked against the oracle.

Each output file must be byte-identical to the same writer applied to the oracle's render of
the same view (the writers' pixel rules are pinned by known answers in test_host_io.py), the
radii cache must hold the oracle's unclamped radii bit for bit, and the skip / lock / cache
semantics of main.cu:99-225 must hold.
"""
import fcntl
import json
import os
import subprocess

import numpy as np
import pytest

from oracle import binding as ob
from renderer_ray_marching_amd import host_io as hio
from renderer_ray_marching_amd import scenes

pytestmark = pytest.mark.gpu

N_POINTS = 30_000
MAX_R = 0.1
VIEWS = [("train", "s/r/v0_color.png", 96, 72, 90.0, 1),
         ("train", "s/r/v1_color.png", 80, 60, 70.0, 2),
         ("val", "s/q/v2_color.png", 64, 48, 60.0, 3)]


def _run(args, cwd):
    print("marcher.bin", " ".join(args), flush=True)
    r = subprocess.run([hio.MARCHER] + args, capture_output=True, text=True, timeout=120, cwd=cwd)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    return r


def _view_json(view):
    rows = view.pose.astype(np.float32).astype(np.float64).tolist()
    W, H = view.width, view.height
    fx = float(np.float32(view.fx))
    calib = [[fx, 0.0, W / 2.0, 0.0], [0.0, fx, H / 2.0, 0.0], [0.0, 0.0, 1.0, 0.0],
             [0.0, 0.0, 0.0, 1.0]]
    return {"camera_pose": rows, "calibration_mat": calib}


@pytest.fixture(scope="module")
def scene(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    cloud = scenes.box_scene(N_POINTS, seed=7)
    ply = d / "cloud.ply"
    hio.write_ply(ply, cloud.xyz, cloud.rgb)
    views = {}
    doc = {"train": {}, "val": {}}
    for split, key, W, H, fx, seed in VIEWS:
        v = scenes.box_scene_view(cloud, W, H, fx, distance=3.0, seed=seed)
        views[key] = v
        doc[split][key] = _view_json(v)
    # a view of another scan: skipped (main.cu:195-201)
    other = d / "other.ply"
    hio.write_ply(other, cloud.xyz[:100], cloud.rgb[:100])
    doc["val"]["s/q/v3_color.png"] = dict(_view_json(views["s/q/v2_color.png"]),
                                          source_scan_ply_path=str(other))
    mats = d / "m.json"
    mats.write_text(json.dumps(doc))
    rt = ob.RefTree(cloud.xyzw(), 64)
    ru, rc, _ = rt.radii(MAX_R)
    return {"dir": d, "cloud": cloud, "ply": ply, "mats": mats, "views": views, "tree": rt,
            "ru": ru, "rc": rc}


def _expected_files(scene, tmp, key, cloud, tree, radii):
    """Writer output for the oracle's render of view `key` (texture order, row 0 = bottom)."""
    v = scene["views"][key]
    rgba, _, _ = tree.render(radii, cloud.colors4(), v.pose_colmajor(), v.fov, v.width, v.height)
    tex = rgba.reshape(v.height, v.width, 4)
    base = os.path.join(tmp, key.replace("/", "_"))
    hio.save_png(base + "_color.png", tex)
    hio.save_depth(base + "_depth.png", tex)
    return {"color": open(base + "_color.png", "rb").read(),
            "depth": open(base + "_depth.png", "rb").read(),
            "npy": open(base + "_depth.png.npy", "rb").read()}


def _outputs(out, key):
    c, d, lk = hio.view_paths(str(out), key)
    return c, d, lk


def _assert_view(out, key, exp):
    c, d, lk = _outputs(out, key)
    assert open(c, "rb").read() == exp["color"], key
    assert open(d, "rb").read() == exp["depth"], key
    assert open(d + ".npy", "rb").read() == exp["npy"], key
    assert not os.path.exists(lk)


def test_cli_render_cache_skip_and_batch(scene, tmp_path):
    d = scene["dir"]
    out = d / "out"
    args = ["-d", "-f", str(scene["ply"]), "-m", str(scene["mats"]), "-o", str(out), "-r", str(MAX_R)]
    r = _run(args, d)
    assert r.returncode == 0, r.stderr
    assert f"Number of points: {N_POINTS}" in r.stdout
    assert "Precomputing radii..." in r.stdout and "Matrices loaded." in r.stdout
    # the cache holds the unclamped radii, bit for bit
    cached = hio.read_radii(str(scene["ply"]) + ".kdtree.radii")
    assert np.array_equal(cached.view(np.uint32), scene["ru"].view(np.uint32))
    cloud, tree = scene["cloud"], scene["tree"]
    exp = {k: _expected_files(scene, tmp_path, k, cloud, tree, scene["rc"]) for k in scene["views"]}
    for k in scene["views"]:
        _assert_view(out, k, exp[k])
    # the view of another scan is skipped and leaves its lock file behind
    c3, _, lk3 = _outputs(out, "s/q/v3_color.png")
    assert not os.path.exists(c3) and os.path.exists(lk3)
    assert "rerun with proper ply" in r.stdout

    # second run: every rendered view already exists
    r = _run(args, d)
    assert r.returncode == 0 and r.stdout.count("ALREADY EXISTS") == 3
    assert "Using precomputed radii..." in r.stdout

    # forced re-render, three views per batched launch, radii from the cache
    for k in scene["views"]:
        os.remove(_outputs(out, k)[0])
    r = _run(args + ["-i", "-b", "3"], d)
    assert r.returncode == 0, r.stderr
    assert "batch of 3 views" in r.stdout
    for k in scene["views"]:
        _assert_view(out, k, exp[k])
    # and with the plain per-view kernel (-b 0): the same files
    for k in scene["views"]:
        os.remove(_outputs(out, k)[0])
    r = _run(args + ["-i", "-b", "0"], d)
    assert r.returncode == 0, r.stderr
    for k in scene["views"]:
        _assert_view(out, k, exp[k])


def test_cli_lock_held_by_another_process(scene):
    d = scene["dir"]
    out = d / "out_lock"
    os.makedirs(out / "r", exist_ok=True)
    _, _, lk = _outputs(out, "s/r/v1_color.png")
    with open(lk, "w") as f:
        fcntl.lockf(f, fcntl.LOCK_EX | fcntl.LOCK_NB)
        r = _run(["-d", "-f", str(scene["ply"]), "-m", str(scene["mats"]), "-o", str(out),
                  "-r", str(MAX_R)], d)
    assert r.returncode == 0, r.stderr
    assert "ALREADY LOCKED" in r.stdout
    assert not os.path.exists(_outputs(out, "s/r/v1_color.png")[0])
    assert os.path.exists(_outputs(out, "s/r/v0_color.png")[0])
    assert os.path.exists(_outputs(out, "s/q/v2_color.png")[0])


def test_cli_subsample_with_full_cache(scene, tmp_path):
    d = scene["dir"]
    # the cache of the full cloud exists (written by the first run or here)
    cache = str(scene["ply"]) + ".kdtree.radii"
    if not os.path.exists(cache):
        hio.write_radii(cache, scene["ru"])
    mp = 20_000
    out = d / "out_sub"
    r = _run(["-d", "-f", str(scene["ply"]), "-m", str(scene["mats"]), "-o", str(out),
              "-r", str(MAX_R), "-s", str(mp)], d)
    assert r.returncode == 0, r.stderr
    assert f"Number of points: {mp}" in r.stdout and "Using precomputed radii..." in r.stdout
    cloud = scene["cloud"]
    idx = ob.subsample_indices(cloud.n)[:mp]
    sub = scenes.Cloud(cloud.xyz[idx], cloud.rgb[idx])
    # remapped (main.cu:128-134), then clamped (main.cu:144-149)
    radii = np.where(scene["ru"][idx] > np.float32(MAX_R), np.float32(MAX_R), scene["ru"][idx])
    tree = ob.RefTree(sub.xyzw(), 64)
    for k in scene["views"]:
        exp = _expected_files(scene, tmp_path, k, sub, tree, radii.astype(np.float32))
        _assert_view(out, k, exp)


def test_cli_cache_errors(scene, tmp_path):
    import shutil
    ply = tmp_path / "copy.ply"
    shutil.copy(scene["ply"], ply)
    hio.write_radii(str(ply) + ".kdtree.radii", scene["ru"][:-1])
    r = _run(["-d", "-f", str(ply), "-o", str(tmp_path / "o")], tmp_path)
    assert r.returncode == 1 and "does not contain the same number of points" in r.stderr
    with open(str(ply) + ".kdtree.radii", "w") as f:
        f.write("22 serialization::archive 17 0 0 3 0 1.0 2.0")
    r = _run(["-d", "-f", str(ply), "-o", str(tmp_path / "o")], tmp_path)
    assert r.returncode == 1 and "serialization archive" in r.stderr
    # -p recomputes and rewrites the cache
    r = _run(["-d", "-p", "-f", str(ply), "-o", str(tmp_path / "o")], tmp_path)
    assert r.returncode == 0 and "Precomputing radii..." in r.stdout
    assert "The matrix file '' was not found." in r.stdout
    cached = hio.read_radii(str(ply) + ".kdtree.radii")
    assert np.array_equal(cached.view(np.uint32), scene["ru"].view(np.uint32))
    # a matrices file without "val" fails after the radii step
    bad = tmp_path / "bad.json"
    bad.write_text(json.dumps({"train": {}}))
    r = _run(["-d", "-f", str(ply), "-m", str(bad), "-o", str(tmp_path / "o")], tmp_path)
    assert r.returncode == 1 and "val" in r.stderr


def test_viewer_renders_the_scripted_path_and_counts_fps(rm_ctx, tmp_path):
    """Without -d: the viewer's compute half (main.cu:235-266 without the window). A box scene
    centred on the origin (the camera starts at (0, 0, 3) looking down -z), a scripted path of
    1500 frames at 1024x768: every dumped frame equals rm_render of the dumped pose and field of
    view on every pixel, the poses are the camera's (rmh_camera_poses of the same script), and
    "<n> FPS" is printed once per second of the run; with -b 0 (the plain kernel) the frames are
    byte-identical."""
    import re
    import torch
    from renderer_ray_marching_amd import api
    cloud = scenes.box_scene(30_000, seed=8)
    xyz = (cloud.xyz - cloud.xyz.mean(0)).astype(np.float32)
    ply = tmp_path / "c.ply"
    hio.write_ply(ply, xyz, cloud.rgb)
    script = "\n".join([f"{f} mouse {3 if (f // 150) % 2 == 0 else -3} 0" for f in range(0, 1500, 10)]
                       + [f"{f} mouse 0 {2 if f < 700 else -2}" for f in range(600, 800, 20)]
                       + [f"{f} scroll 0.5" for f in range(1000, 1040)])
    (tmp_path / "path.txt").write_text(script + "\n")
    dump = tmp_path / "frames"
    r = _run(["-f", str(ply), "-r", "0.1", "--script", str(tmp_path / "path.txt"), "--frames",
              "1500", "--dt", "0.01", "--dump", str(dump), "--dump-every", "250"], tmp_path)
    assert r.returncode == 0
    lines = r.stdout.splitlines()
    fps = [int(x.split()[0]) for x in lines if re.fullmatch(r"\d+ FPS", x)]
    secs = float(re.search(r"Rendered 1500 frames at 1024x768 in ([0-9.]+) s", r.stdout).group(1))
    assert int(secs) - 1 <= len(fps) <= int(secs) and sum(fps) <= 1500
    poses, fovs = hio.camera_poses(1500, 0.01, script)
    # the same cloud and radii through the library
    xyzw, rgba = hio.load_ply(ply)
    radii = np.minimum(hio.read_radii(str(ply) + ".kdtree.radii"), np.float32(0.1))
    ctx = rm_ctx
    t = api.KDTreeFlann(ctx)
    p = torch.from_numpy(xyzw).cuda()
    t.Build(p)
    m = api.PointcloudRayMarcher(p, torch.from_numpy(rgba).cuda(), torch.from_numpy(radii).cuda(),
                                 t)
    hits = 0
    for f in range(0, 1500, 250):
        vals = np.array(open(dump / f"frame_{f:05d}.pose").read().split(), np.float32)
        pose, fov = vals[:16], vals[16]
        assert np.array_equal(pose, poses[f]) and fov == fovs[f], f
        frame = np.load(dump / f"frame_{f:05d}.npy").reshape(768, 1024, 4)
        ref, hit = m.render_to_texture(pose, float(fov), 1024, 768)
        ctx.sync()
        assert np.array_equal(ref.cpu().numpy(), frame), f
        hits += int((hit >= 0).sum().item())
    assert hits > 100_000
    # -b 0: the frames come from the plain kernel instead (the default is the batched kernel
    # through the allocation-free entry point); the same images
    dump0 = tmp_path / "frames_plain"
    r0 = _run(["-f", str(ply), "-r", "0.1", "--script", str(tmp_path / "path.txt"), "--frames",
               "260", "--dt", "0.01", "--dump", str(dump0), "--dump-every", "250", "-b", "0"],
              tmp_path)
    assert r0.returncode == 0
    for f in (0, 250):
        assert (dump0 / f"frame_{f:05d}.npy").read_bytes() == \
            (dump / f"frame_{f:05d}.npy").read_bytes(), f

"""Counts the algorithmic bytes of the measured workloads and writes them to
tests/golden/alg_bytes.json for bench.py's work figures, in two forms:

1. the reference's traversal (kdtree_cuda_3d_index.cu:52-129) on the reference tree rule
   (leaf 64), restated by the oracle;
2. what one lane of this library's kernels reads when it searches alone (work_model.cpp: the
   exact per-lane kNN over the library's leaf-16 tree, the warm-started march with the early
   miss), whose hit indices and depths must equal the oracle's on the counted pixels.

Per kNN-radius query:   B_q = 16 + 12 C_q + 16 V_q + 24 A_q
Per marched ray:        B_ray = sum over steps (12 C + 16 V + 24 A + 10 (4 + 4)) + 16 + 4
where C = candidate distance evaluations (12 B of xyz each), V = node records (16 B),
A = far-child box tests (24 B), 10 (4 + 4) = the index remap and radius gathers of the ten
neighbours per step, 16 + 4 = the RGBA32F and hit-index writes. The march is counted on the
pixels with (x + 7 y) mod 64 == 0 of every view and scaled by the exact pixel ratio.

    python tests/golden/make_alg_bytes.py [--workload room10m|box100k|room2m|scale50m]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import binding as ob  # noqa: E402
from renderer_ray_marching_amd import scenes  # noqa: E402

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import work_model as wm  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "alg_bytes.json")

WORKLOADS = {
    "box100k": scenes.workload_box_100k,
    "room2m": scenes.workload_room_2m,
    "room10m": scenes.workload_room_10m,
    "scale50m": scenes.workload_scale_50m,
}


def subsample(W, H):
    lin = np.arange(W * H, dtype=np.int64)
    x, y = lin % W, lin // W
    return lin[(x + 7 * y) % 64 == 0].astype(np.int32)


def count(name):
    t0 = time.time()
    cloud, views, max_r = WORKLOADS[name]()
    p = cloud.xyzw()
    tree = ob.RefTree(p, 64)
    ru, rc, rcnt = tree.radii(max_r)
    radii_bytes = 16 * rcnt.q + rcnt.traversal_bytes()
    col = cloud.colors4()
    tot = ob.Counters()
    rays_counted = 0
    rays_total = 0
    hits = 0
    for v in views:
        pix = subsample(v.width, v.height)
        rgba, hit, c = tree.render(rc, col, v.pose_colmajor(), v.fov, v.width, v.height, pix=pix)
        for f in ("c", "v", "a", "q", "s"):
            setattr(tot, f, getattr(tot, f) + getattr(c, f))
        rays_counted += pix.shape[0]
        rays_total += v.width * v.height
        hits += int((hit >= 0).sum())
    march_bytes = tot.traversal_bytes() + 80 * tot.s + 20 * rays_counted
    # this library's kernels, one lane at a time
    wt = wm.WorkTree(p, 16)
    wr, wrc = wt.radii()
    assert np.array_equal(wr, ru), "work model radii differ from the oracle's"
    wt.set_radii(rc)
    wtot = wm.Counts()
    for v in views:
        pix = subsample(v.width, v.height)
        orgba, ohit, _ = tree.render(rc, col, v.pose_colmajor(), v.fov, v.width, v.height,
                                     pix=pix)
        whit, wdep, c = wt.render(v.pose_colmajor(), ob.pane_distance(v.width, v.fov), v.width,
                                  v.height, pix=pix)
        assert (whit == ohit).all() and np.array_equal(wdep, orgba[:, 3]), v.name
        for f, _ in wm.Counts._fields_:
            setattr(wtot, f, getattr(wtot, f) + getattr(c, f))
    # per ray: traversal reads, RGBA32F + hit index written, the hit point's colour read
    this_march = wtot.traversal_bytes() + 20 * wtot.rays + 12 * wtot.hits
    # per query: the traversal reads, the query point (16 B) read, two 4-byte radii written
    this_radii = wrc.traversal_bytes() + 24 * cloud.n
    rec = {
        "points": cloud.n,
        "views": len(views),
        "rays": rays_total,
        "radii_queries": rcnt.q,
        "radii_bytes_total": int(radii_bytes),
        "radii_bytes_per_query": radii_bytes / rcnt.q,
        "radii_counters": rcnt.as_dict(),
        "march_rays_counted": rays_counted,
        "march_counters": tot.as_dict(),
        "march_steps_per_ray": tot.s / rays_counted,
        "march_bytes_per_ray": march_bytes / rays_counted,
        "march_hit_fraction": hits / rays_counted,
        "this_march_counters": wtot.as_dict(),
        "this_march_bytes_per_ray": this_march / wtot.rays,
        "this_radii_counters": wrc.as_dict(),
        "this_radii_bytes_per_query": this_radii / cloud.n,
        "this_basis": "per-lane exact search of the leaf-16 tree with warm start and early miss "
                      "(work_model.cpp); nodes and re-tested leaves 64 B, leaf points 16 B, "
                      "warm-start gathers 12 B, radius gathers 4 B",
        "subsample": "pixels with (x + 7 y) mod 64 == 0",
        "seconds": time.time() - t0,
    }
    return rec


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", nargs="*", default=["box100k", "room10m"])
    a = ap.parse_args()
    data = {}
    if os.path.exists(OUT):
        with open(OUT) as f:
            data = json.load(f)
    for w in a.workload:
        data[w] = count(w)
        print(w, json.dumps(data[w]), flush=True)
        with open(OUT, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()

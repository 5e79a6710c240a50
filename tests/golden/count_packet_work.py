"""Counts what a wavefront of the batched march executes per step on a measured workload, with
the packet restatement of work_model.cpp (wm_render_packet: 64-ray packets over 8x8 tiles, the
default tune), and writes tests/golden/packet_work.json. Every tile_stride-th tile of every view
is marched; the marched pixels' hits are checked against the per-lane work model (itself checked
against the oracle by make_alg_bytes.py and the tests). Two variants that were studied and not
built are counted as well: the containing-node depth and the bound-only warm seed.

    python tests/golden/count_packet_work.py [--workload room10m] [--tile-stride 61]
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

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "packet_work.json")
WORKLOADS = {"box100k": scenes.workload_box_100k, "room10m": scenes.workload_room_10m}


def count(name, stride, check_every):
    cloud, views, max_r = WORKLOADS[name]()
    p = np.ascontiguousarray(cloud.xyzw())
    wt = wm.WorkTree(p, 16)
    ru, _ = wt.radii()
    rc = np.where((max_r > 0) & (ru > max_r), np.float32(max_r), ru).astype(np.float32)
    wt.set_radii(rc)
    res = {}
    for variant, bound_only in (("warm_seed", False), ("bound_only_seed", True)):
        tot = {f: 0 for f, _ in wm.PacketCounts._fields_}
        checked = 0
        t0 = time.time()
        for vi, v in enumerate(views):
            pane = ob.pane_distance(v.width, v.fov)
            hit, c = wt.render_packet(v.pose_colmajor(), pane, v.width, v.height, stride,
                                      bound_only)
            for f in tot:
                tot[f] += getattr(c, f)
            if vi % check_every == 0:
                pix = np.nonzero(hit != -2)[0].astype(np.int32)
                lh, _, _ = wt.render(v.pose_colmajor(), pane, v.width, v.height, pix=pix)
                assert (lh == hit[pix]).all(), v.name
                checked += pix.size
        ws = tot["wave_steps"]
        res[variant] = {
            "counts": tot,
            "per_wave_step": {f: round(tot[f] / ws, 2) for f in
                              ("expansions", "retests", "scans", "candidates", "wave_inserts")},
            "active_lanes_per_wave_step": round(tot["lane_steps"] / ws, 2),
            "per_lane_step": {
                "child_passes": round(tot["lane_child_passes"] / tot["lane_steps"], 2),
                "candidates": round(tot["lane_candidates"] / tot["lane_steps"], 2),
                "inserts": round(tot["lane_inserts"] / tot["lane_steps"], 2)},
            "steps_per_ray": round(tot["lane_steps"] / tot["rays"], 2),
            "insertion_chain": {
                "slots_per_wave_insert": round(tot["insert_slots"] / max(1, tot["wave_inserts"]), 2),
                "wave_inserts_moving_nothing": round(tot["inserts_no_move"] /
                                                     max(1, tot["wave_inserts"]), 3),
                "seed_moves_per_warm_wave_step": round(tot["seed_moves"] /
                                                       max(1, tot["seed_steps"]), 2)},
            "containing_depth_mean": round(tot["containing_depth"] /
                                           max(1, tot["containing_steps"]), 2),
            "pixels_checked_against_per_lane_model": checked,
            "seconds": round(time.time() - t0, 1)}
        print(variant, json.dumps(res[variant]["per_wave_step"]), flush=True)
    return {"workload": name, "tile_stride": stride, "tiles": "8x8, every tile_stride-th of every view",
            **res}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--workload", default="room10m", choices=sorted(WORKLOADS))
    ap.add_argument("--tile-stride", type=int, default=61)
    ap.add_argument("--check-every", type=int, default=16, help="check every n-th view's hits")
    a = ap.parse_args()
    data = json.load(open(OUT)) if os.path.exists(OUT) else {}
    data[a.workload] = count(a.workload, a.tile_stride, a.check_every)
    with open(OUT, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
        f.write("\n")


if __name__ == "__main__":
    main()

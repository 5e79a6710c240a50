"""The work model (tests/golden/work_model.cpp) restates one lane of the library's radii and march
kernels to count what they read. Its outputs must equal the oracle's (the search is exact, so
any tree gives the same radii, hits and depths), and its counts must follow the traversal on
trees small enough to count by hand."""
import os
import sys

import numpy as np

from oracle import binding as ob
from renderer_ray_marching_amd import scenes

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
import work_model as wm  # noqa: E402


def test_outputs_equal_the_oracle():
    cloud = scenes.box_scene(20_000, seed=3)
    p = cloud.xyzw()
    rt = ob.RefTree(p, 64)
    ru, rc, _ = rt.radii(0.1)
    wt = wm.WorkTree(p, 16)
    wr, c = wt.radii()
    assert np.array_equal(wr, ru)
    assert c.nodes > 0 and c.points > 0 and c.steps == 0
    wt.set_radii(rc)
    v = scenes.box_scene_view(cloud, 80, 60, 60.0, 3.0, seed=2)
    rgba, hit, _ = rt.render(rc, cloud.colors4(), v.pose_colmajor(), v.fov, 80, 60)
    h, d, c = wt.render(v.pose_colmajor(), ob.pane_distance(80, v.fov), 80, 60)
    assert (h == hit).all() and np.array_equal(d, rgba[:, 3])
    assert (hit >= 0).mean() > 0.3
    assert c.rays == 80 * 60 and c.hits == int((hit >= 0).sum())
    # every step gathers the ten radii; every step after a ray's first re-seeds ten neighbours
    assert c.sdf == 10 * c.steps and c.seeds == 10 * (c.steps - c.rays)


def test_counts_on_a_single_leaf_and_on_two_leaves():
    rng = np.random.default_rng(0)
    p = np.ones((12, 4), np.float32)
    p[:, :3] = rng.uniform(0, 1, (12, 3))
    wt = wm.WorkTree(p, 16)                 # one leaf: the root record holds it in slot 0
    _, c = wt.radii()
    assert (c.nodes, c.retests, c.points) == (12, 12, 144)
    # two leaves of 16 far apart: each query expands the root while its list is still empty
    # (bound +INF), so both leaves are pushed; it scans the near one (16 points), then re-tests
    # the far one against the new bound and skips it
    q = np.ones((32, 4), np.float32)
    q[:16, :3] = rng.uniform(0, 1, (16, 3))
    q[16:, :3] = rng.uniform(0, 1, (16, 3)) + np.float32(100.0)
    wt2 = wm.WorkTree(q, 16)
    r2, c2 = wt2.radii()
    assert (c2.nodes, c2.retests, c2.points) == (32, 64, 512)
    ru, _, _ = ob.RefTree(q, 64).radii(0.0)
    assert np.array_equal(r2, ru)

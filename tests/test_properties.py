"""Property-based tests (hypothesis) on the CPU: file round trips of the host drop-in and the
oracle's exact kNN and radius rule against brute force on generated tiny clouds, including
duplicates, coincident coordinates, and k above the point count.
"""
import os
import tempfile

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings
from hypothesis import strategies as st
from hypothesis.extra import numpy as hnp

import fp32
from oracle import binding as ob
from renderer_ray_marching_amd import host_io as hio

SETTINGS = settings(max_examples=60, deadline=None,
                    suppress_health_check=[HealthCheck.too_slow, HealthCheck.data_too_large])

finite32 = st.floats(width=32, allow_nan=False, allow_infinity=False)
coord = st.floats(width=32, min_value=-4.0, max_value=4.0, allow_nan=False)
host_built = pytest.mark.skipif(not os.path.exists(hio.LIB_PATH), reason="librm_host.so not built")


@host_built
@SETTINGS
@given(hnp.arrays(np.float32, st.integers(0, 300),
                  elements=st.floats(width=32, allow_nan=False)))
def test_radii_cache_round_trip_is_exact(radii):
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "c.kdtree.radii")
        hio.write_radii(p, radii)
        back = hio.read_radii(p)
    assert back.dtype == np.float32 and back.shape == radii.shape
    assert (back.view(np.uint32) == radii.view(np.uint32)).all()


@host_built
@SETTINGS
@given(st.integers(1, 200).flatmap(lambda n: st.tuples(
    hnp.arrays(np.float32, (n, 3), elements=finite32),
    hnp.arrays(np.uint8, (n, 3)))))
def test_ply_round_trip(data):
    xyz, rgb = data
    with tempfile.TemporaryDirectory() as d:
        p = os.path.join(d, "c.ply")
        hio.write_ply(p, xyz, rgb)
        xyzw, rgba = hio.load_ply(p)
    assert (xyzw[:, :3].view(np.uint32) == xyz.view(np.uint32)).all()
    assert (xyzw[:, 3] == 1).all() and (rgba[:, 3] == 1).all()
    # src/pointcloud.cpp:55: colour / 255.0f in float
    assert (rgba[:, :3] == rgb.astype(np.float32) / np.float32(255.0)).all()


def _cloud(pts, dup):
    p = np.concatenate([pts, pts[:dup]]) if dup else pts
    return np.concatenate([p, np.ones((p.shape[0], 1), np.float32)], 1).astype(np.float32)


@SETTINGS
@given(st.integers(1, 60).flatmap(lambda n: hnp.arrays(np.float32, (n, 3), elements=coord)),
       st.integers(0, 5), st.integers(1, 12), st.sampled_from([1, 2, 5, 64]))
def test_oracle_knn_is_the_brute_force_top_k(pts, dup, k, leaf):
    p = _cloud(pts, min(dup, pts.shape[0]))
    q = np.concatenate([p[:5], p[:5] + np.float32(0.01)]).astype(np.float32)
    idx, d2, _ = ob.RefTree(p, leaf).knn(q, k)
    # brute force from the definition: the k smallest (d2, index) pairs, d2 as the float fma chain
    d_all = fp32.dist2(p[None, :, :], q[:, None, :])
    n = p.shape[0]
    for i in range(q.shape[0]):
        order = sorted(range(n), key=lambda j: (float(d_all[i, j]), j))[:k]
        want_i = np.array(order + [-1] * (k - len(order)), np.int32)
        want_d = np.array([d_all[i, j] for j in order] + [np.inf] * (k - len(order)), np.float32)
        assert (idx[i] == want_i).all() and (d2[i] == want_d).all()


@SETTINGS
@given(st.integers(2, 60).flatmap(lambda n: hnp.arrays(np.float32, (n, 3), elements=coord)),
       st.integers(0, 3), st.floats(0.0, 2.0))
def test_oracle_radii_rule(pts, dup, max_r):
    p = _cloud(pts, min(dup, pts.shape[0]))
    max_r = np.float32(max_r)
    ru, rc, _ = ob.RefTree(p, 3).radii(max_r)
    d_all = fp32.dist2(p[None, :, :], p[:, None, :])
    second = np.sort(d_all, axis=1)[:, 1]                 # includes self (d2 = 0) first
    want = np.sqrt(second).astype(np.float32)
    assert (ru == want).all()
    clamped = np.where((max_r > 0) & (want > max_r), max_r, want).astype(np.float32)
    assert (rc == clamped).all()

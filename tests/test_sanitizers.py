"""Host drop-in I/O and the CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer.

The parsers of librm_host (PLY, the .kdtree.radii text archive, the matrices JSON, the viewer's
input script) and its
writers are compiled with g++ -fsanitize=address,undefined together with
tests/sanitize/host_io_fuzz.cpp, which feeds every parser its valid input and thousands of
deterministic corruptions of it (truncations, byte changes). Each must be parsed or rejected with
an error code; any sanitizer report or crash fails the test. The oracle is built the same way
with tests/sanitize/oracle_driver.cpp, which calls each of its entry points on small edge-case
clouds. CPU only; no GPU code is built with a sanitizer.
"""
import json
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST = os.path.join(ROOT, "renderer_ray_marching_amd", "host")
SRC = ["json.cpp", "ply.cpp", "radii_cache.cpp", "image_io.cpp", "views.cpp", "viewer.cpp",
       "capi.cpp"]

pytestmark = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")


@pytest.fixture(scope="module")
def fuzzer(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("asan") / "host_io_fuzz")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", os.path.join(ROOT, "include"), "-I", HOST]
    cmd += [os.path.join(HOST, f) for f in SRC]
    cmd += [os.path.join(ROOT, "tests", "sanitize", "host_io_fuzz.cpp"), "-lz", "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "asan" in r.stderr.lower() and "cannot find" in r.stderr.lower():
        pytest.skip("the sanitizer runtimes are not installed")
    assert r.returncode == 0, r.stderr[-3000:]
    return out


def _run(fuzzer, kind, path, tmp_path):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([fuzzer, kind, str(path), str(tmp_path)], capture_output=True, text=True,
                       env=env, timeout=600)
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]
    return r.stdout


def _ply_binary(path, n, seed):
    rng = np.random.default_rng(seed)
    head = (f"ply\nformat binary_little_endian 1.0\ncomment test\nelement vertex {n}\n"
            "property double x\nproperty double y\nproperty double z\nproperty float nx\n"
            "property uchar red\nproperty uchar green\nproperty uchar blue\n"
            "element face 2\nproperty list uchar int vertex_indices\nend_header\n").encode()
    body = b""
    for i in range(n):
        x, y, z = rng.normal(size=3)
        body += struct.pack("<dddfBBB", x, y, z, 0.5, i % 256, (7 * i) % 256, 255 - i % 256)
    body += struct.pack("<Biii", 3, 0, 1, 2) + struct.pack("<Biiii", 4, 0, 1, 2, 3)
    path.write_bytes(head + body)


def _ply_ascii(path, n):
    lines = ["ply", "format ascii 1.0", f"element vertex {n}", "property float x",
             "property float y", "property float z", "property uchar red", "property uchar green",
             "property uchar blue", "end_header"]
    lines += [f"{0.1 * i} {-0.2 * i} {1e-3 * i} {i % 256} {2 * i % 256} 7" for i in range(n)]
    path.write_text("\n".join(lines) + "\n")


def test_ply_parsers_under_sanitizers(fuzzer, tmp_path):
    _ply_binary(tmp_path / "b.ply", 40, 1)
    _ply_ascii(tmp_path / "a.ply", 30)
    assert "corrupted inputs" in _run(fuzzer, "ply", tmp_path / "b.ply", tmp_path)
    assert "corrupted inputs" in _run(fuzzer, "ply", tmp_path / "a.ply", tmp_path)


def test_radii_reader_under_sanitizers(fuzzer, tmp_path):
    vals = " ".join(f"{v:.9e}" for v in np.linspace(0, 0.05, 50, dtype=np.float32))
    (tmp_path / "r.radii").write_text(f"22 serialization::archive 17 0 0 50 0 {vals}\n")
    assert "corrupted inputs" in _run(fuzzer, "radii", tmp_path / "r.radii", tmp_path)


def test_matrices_parser_under_sanitizers(fuzzer, tmp_path):
    pose = [[1, 0, 0, 0.5], [0, 1, 0, -1], [0, 0, 1, 2.25], [0, 0, 0, 1]]
    calib = [[500.0, 0, 320, 0], [0, 500.0, 240, 0], [0, 0, 1, 0], [0, 0, 0, 1]]
    doc = {"train": {"s/a/x_color.png": {"camera_pose": pose, "calibration_mat": calib,
                                         "source_scan_ply_path": "/p/q \\u00e9.ply"}},
           "val": {"s/b/y_color.png": {"camera_pose": pose, "calibration_mat": calib}}}
    (tmp_path / "m.json").write_text(json.dumps(doc))
    assert "corrupted inputs" in _run(fuzzer, "matrices", tmp_path / "m.json", tmp_path)


def test_viewer_script_parser_under_sanitizers(fuzzer, tmp_path):
    (tmp_path / "s.txt").write_text("# a path\n0 key W\n0 mouse 12.5 -3\n1 key A\n2 scroll 1\n"
                                    "3 key E  # up\n5 mouse 0 900\n7 scroll -2.5\n")
    assert "corrupted inputs" in _run(fuzzer, "script", tmp_path / "s.txt", tmp_path)


def test_writers_and_paths_under_sanitizers(fuzzer, tmp_path):
    assert "io: ok" in _run(fuzzer, "io", "-", tmp_path)


def test_oracle_under_sanitizers(tmp_path):
    out = str(tmp_path / "oracle_driver")
    oracle = os.path.join(ROOT, "oracle")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fopenmp", "-ffp-contract=off", "-mfma",
           "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", oracle, os.path.join(oracle, "oracle.cpp"),
           os.path.join(ROOT, "tests", "sanitize", "oracle_driver.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", OMP_NUM_THREADS="2")
    r = subprocess.run([out], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0 and "oracle: ok" in r.stdout, (r.stdout + r.stderr)[-4000:]
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-4000:]

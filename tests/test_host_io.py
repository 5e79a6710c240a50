"""Host-side I/O of the drop-in CLI (librm_host.so, include/rm_host.h), on the CPU.

Known answers are worked out by hand from the reference's rules:
  PLY       src/pointcloud.cpp:43-58 (double -> float, uchar / 255.0f, w = 1)
  radii     main.cu:115-127 (text archive of std::vector<float>)
  PNG       src/utils.cu:151-163 ((unsigned char)(v * 255.0f), alpha 255, rows flipped)
  depth     src/utils.cu:165-192 (0 where |d - 1| < 1e-5, else (uint8)clamp(2.55f d, 0, 255);
            <path>.npy float32 (H, W) flipped)
  matrices  main.cu:193-206, src/utils.cu:137-149
  paths     main.cu:172-177; locks main.cu:186-192, 217-218
"""
import ctypes
import json
import os
import struct
import subprocess
import sys
import zlib

import numpy as np
import pytest

from renderer_ray_marching_amd import host_io as hio

pytestmark = pytest.mark.skipif(not os.path.exists(hio.LIB_PATH), reason="librm_host.so not built")


# ---- a small PNG decoder (independent of the writer) --------------------------------------

def read_png(path):
    data = open(path, "rb").read()
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, ihdr = 8, b"", None
    while pos < len(data):
        (ln,) = struct.unpack(">I", data[pos:pos + 4])
        typ = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + ln]
        (crc,) = struct.unpack(">I", data[pos + 8 + ln:pos + 12 + ln])
        assert zlib.crc32(typ + body) & 0xFFFFFFFF == crc
        if typ == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif typ == b"IDAT":
            idat += body
        pos += 12 + ln
    W, H, depth, ctype, _, _, interlace = ihdr
    assert depth == 8 and interlace == 0
    ch = {0: 1, 2: 3, 6: 4}[ctype]
    raw = zlib.decompress(idat)
    stride = W * ch
    out = np.zeros((H, stride), np.uint8)
    prev = np.zeros(stride, np.int32)
    for y in range(H):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw[y * (stride + 1) + 1:(y + 1) * (stride + 1)], np.uint8).astype(np.int32)
        cur = np.zeros(stride, np.int32)
        for x in range(stride):
            a = cur[x - ch] if x >= ch else 0
            b = prev[x]
            c = prev[x - ch] if x >= ch else 0
            if f == 0:
                p = 0
            elif f == 1:
                p = a
            elif f == 2:
                p = b
            elif f == 3:
                p = (a + b) // 2
            else:
                pa, pb, pc = abs(b - c), abs(a - c), abs(a + b - 2 * c)
                p = a if pa <= pb and pa <= pc else (b if pb <= pc else c)
            cur[x] = (line[x] + p) & 0xFF
        out[y] = cur
        prev = cur
    return out.reshape(H, W, ch) if ch > 1 else out.reshape(H, W)


# ---- PLY ----------------------------------------------------------------------------------

def _ply_header(fmt, n, props, extra_before="", extra_after=""):
    h = f"ply\nformat {fmt} 1.0\ncomment made by a test\n{extra_before}element vertex {n}\n"
    h += "".join(f"property {t} {name}\n" for t, name in props)
    h += extra_after + "end_header\n"
    return h.encode()


def test_ply_binary_le_float_roundtrip(tmp_path):
    rng = np.random.default_rng(1)
    xyz = rng.normal(size=(1000, 3)).astype(np.float32)
    rgb = rng.integers(0, 256, size=(1000, 3), dtype=np.uint8)
    p = tmp_path / "c.ply"
    hio.write_ply(p, xyz, rgb)
    xyzw, rgba = hio.load_ply(p)
    assert np.array_equal(xyzw[:, :3], xyz) and np.all(xyzw[:, 3] == 1.0)
    assert np.array_equal(rgba[:, :3], rgb.astype(np.float32) / np.float32(255.0))
    assert np.all(rgba[:, 3] == 1.0)
    # known answers of c / 255.0f
    assert hio.load_ply(p)[1].dtype == np.float32
    assert np.float32(255) / np.float32(255.0) == 1.0


@pytest.mark.parametrize("fmt", ["binary_little_endian", "binary_big_endian"])
def test_ply_double_positions_extra_props_and_faces(tmp_path, fmt):
    # double coordinates are rounded to float; other properties and elements are skipped
    xyz = np.array([[0.1, -2.5, 1e-3], [1.0 / 3.0, 7.0, -0.2]], np.float64)
    rgb = np.array([[0, 128, 255], [1, 2, 3]], np.uint8)
    e = ">" if fmt.endswith("big_endian") else "<"
    props = [("float", "nx"), ("double", "x"), ("double", "y"), ("double", "z"), ("uchar", "red"),
             ("uchar", "green"), ("uchar", "blue"), ("uchar", "alpha"), ("int", "label")]
    body = b""
    for i in range(2):
        body += struct.pack(e + "f", 9.0) + struct.pack(e + "ddd", *xyz[i])
        body += bytes(rgb[i]) + bytes([200]) + struct.pack(e + "i", -7)
    faces = b""
    for f in ([0, 1, 1], [1, 0]):
        faces += bytes([len(f)]) + b"".join(struct.pack(e + "i", v) for v in f)
    head = _ply_header(fmt, 2, props,
                       extra_before="element camera 1\nproperty float fx\n",
                       extra_after="element face 2\nproperty list uchar int vertex_indices\n")
    p = tmp_path / "d.ply"
    p.write_bytes(head + struct.pack(e + "f", 500.0) + body + faces)
    xyzw, rgba = hio.load_ply(p)
    assert np.array_equal(xyzw[:, :3], xyz.astype(np.float32))
    assert xyzw[0, 0] == np.float32(0.1) and xyzw[1, 0] == np.float32(1.0 / 3.0)
    assert rgba[0, 0] == 0.0 and rgba[0, 2] == 1.0
    assert rgba[0, 1] == np.float32(128) / np.float32(255)


def test_ply_ascii(tmp_path):
    head = _ply_header("ascii", 3, [("float", "x"), ("float", "y"), ("double", "z"), ("uchar", "red"),
                                    ("uchar", "green"), ("uchar", "blue")],
                       extra_after="element face 1\nproperty list uchar int vertex_indices\n")
    rows = "0.1 0.2 0.30000000000000004 1 2 3\r\n-1e-3 5 6 255 0 17\n1 2 3 4 5 6\n3 0 1 2\n"
    p = tmp_path / "a.ply"
    p.write_bytes(head + rows.encode())
    xyzw, rgba = hio.load_ply(p)
    assert xyzw[0, 0] == np.float32("0.1") and xyzw[1, 0] == np.float32("-1e-3")
    assert xyzw[0, 2] == np.float32(np.float64("0.30000000000000004"))
    assert rgba[1, 0] == 1.0 and rgba[1, 2] == np.float32(17) / np.float32(255)
    assert xyzw.shape == (3, 4)


def test_ply_errors(tmp_path):
    def ply(props, body=b"", fmt="binary_little_endian", n=1):
        p = tmp_path / "e.ply"
        p.write_bytes(_ply_header(fmt, n, props) + body)
        return p

    good = [("float", "x"), ("float", "y"), ("float", "z")]
    # colours must be uchar (happly getProperty<unsigned char>)
    with pytest.raises(hio.RMHError, match="coerced"):
        hio.load_ply(ply(good + [("float", "red"), ("float", "green"), ("float", "blue")], b"\0" * 24))
    # positions must be float or double
    with pytest.raises(hio.RMHError, match="coerced"):
        hio.load_ply(ply([("int", "x"), ("float", "y"), ("float", "z"), ("uchar", "red"),
                          ("uchar", "green"), ("uchar", "blue")], b"\0" * 15))
    with pytest.raises(hio.RMHError, match="red"):
        hio.load_ply(ply(good, b"\0" * 12))
    with pytest.raises(hio.RMHError, match="end of binary"):
        hio.load_ply(ply(good + [("uchar", "red"), ("uchar", "green"), ("uchar", "blue")], b"\0" * 10))
    # a vertex count the file cannot hold is rejected before anything is allocated for it
    # (run in a child process with a 2 GiB address-space limit)
    colours = [("uchar", "red"), ("uchar", "green"), ("uchar", "blue")]
    for fmt, body in (("binary_little_endian", b"\0" * 15), ("ascii", b"1 2 3 4 5 6\n")):
        big = ply(good + colours, body, fmt=fmt, n=4_000_000_000)
        code = ("import resource, sys; resource.setrlimit(resource.RLIMIT_AS, (2 << 30, 2 << 30)); "
                "from renderer_ray_marching_amd import host_io as h\n"
                "try:\n h.load_ply(sys.argv[1])\nexcept h.RMHError as e:\n print(e)")
        r = subprocess.run([sys.executable, "-c", code, str(big)], capture_output=True, text=True,
                           cwd=os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
        assert r.returncode == 0 and "exceeds the file" in r.stdout, r.stdout + r.stderr
    bad = tmp_path / "n.ply"
    bad.write_bytes(b"plx\n")
    with pytest.raises(hio.RMHError):
        hio.load_ply(bad)
    with pytest.raises(hio.RMHError, match="could not open"):
        hio.load_ply(tmp_path / "missing.ply")


# ---- radii cache --------------------------------------------------------------------------

def test_radii_cache_layout_and_exact_roundtrip(tmp_path):
    rng = np.random.default_rng(3)
    r = np.concatenate([rng.random(500).astype(np.float32),
                        np.array([0.0, 0.1, 1e-40, 3.4e38, np.inf], np.float32),
                        (rng.random(50) * 1e-3).astype(np.float32)])
    p = tmp_path / "x.ply.kdtree.radii"
    hio.write_radii(p, r)
    text = p.read_text()
    tok = text.split()
    assert tok[:7] == ["22", "serialization::archive", "17", "0", "0", str(len(r)), "0"]
    assert len(tok) == 7 + len(r) and text.endswith("\n")
    assert tok[7 + 501] == "1.000000015e-01"      # 0.1f printed with 9 fractional digits
    assert tok[7 + 500] == "0.000000000e+00"
    back = hio.read_radii(p)
    assert back.dtype == np.float32 and np.array_equal(back.view(np.uint32), r.view(np.uint32))


@pytest.mark.parametrize("text,expect", [
    ("22 serialization::archive 19 0 0 3 0 1.5 2.5 3.5", [1.5, 2.5, 3.5]),     # newer version
    ("22 serialization::archive 3 0 0 3 1.5 2.5 3.5\n", [1.5, 2.5, 3.5]),      # no item version
    ("22 serialization::archive 17\n0 0 2 0\n4.0e-01\n5.0e-01\n", [0.4, 0.5]),  # line breaks
    ("22 serialization::archive 17 0 0 0 0\n", []),                             # empty vector
    ("22 serialization::archive 17 0 0 4 0 1 2 3 4", [1, 2, 3, 4]),             # integer-looking
])
def test_radii_cache_tolerant_reader(tmp_path, text, expect):
    p = tmp_path / "r.radii"
    p.write_text(text)
    assert np.array_equal(hio.read_radii(p), np.array(expect, np.float32))


@pytest.mark.parametrize("text", [
    "",
    "22 serialization::archiv 17 0 0 1 0 1.0",
    "21 serialization::archive 17 0 0 1 0 1.0",
    "22 serialization::archive 17 0 0 5 0 1.0 2.0",
    "22 serialization::archive 17 0 0 2 0 1.0 abc",
    "22 serialization::archive 17 0 0 3 0 1.0 2.0",     # truncated: one value missing
    "22 serialization::archive 17 0 0 1 0 1.0 2.0",     # extra value
    "22 serialization::archive 17 0 0",
])
def test_radii_cache_corrupt(tmp_path, text):
    p = tmp_path / "bad.radii"
    p.write_text(text)
    with pytest.raises(hio.RMHError):
        hio.read_radii(p)


# ---- colour PNG, depth PNG, npy -----------------------------------------------------------

def test_save_png_known_answers(tmp_path):
    tex = np.zeros((2, 3, 4), np.float32)          # row 0 = bottom
    tex[0, 0] = [1.0, 0.5, 0.999, 0.0]
    tex[0, 1] = [0.0, 0.25, 0.1, 7.0]
    tex[1, 2] = [0.2, 0.4, 0.6, 3.0]
    p = tmp_path / "c.png"
    hio.save_png(p, tex)
    img = read_png(p)
    assert img.shape == (2, 3, 4)
    assert img[1, 0].tolist() == [255, 127, 254, 255]   # bottom row written last
    assert img[1, 1].tolist() == [0, 63, 25, 255]
    assert img[0, 2].tolist() == [51, 102, 153, 255]
    assert np.all(img[..., 3] == 255)


def test_save_depth_known_answers(tmp_path):
    d = np.array([[0.0, 1.0, 1.000001, 0.99995],
                  [50.0, 100.0, 200.0, 0.5],
                  [0.99, 3.3, 99.9, 12.0]], np.float32)
    tex = np.zeros((3, 4, 4), np.float32)
    tex[..., 3] = d
    p = tmp_path / "v_depth.png"
    hio.save_depth(p, tex)
    img = read_png(p)
    expect = np.array([[0, 0, 0, 2],        # |d - 1| < 1e-5 -> 0; 0.99995 is 5e-5 away
                       [127, 255, 255, 1],
                       [2, 8, 254, 30]], np.uint8)
    assert np.array_equal(img, expect[::-1])
    arr = np.load(str(p) + ".npy")
    assert arr.dtype == np.float32 and arr.shape == (3, 4)
    assert np.array_equal(arr, d[::-1])
    raw = open(str(p) + ".npy", "rb").read()
    assert raw[:8] == b"\x93NUMPY\x01\x00"
    (hlen,) = struct.unpack("<H", raw[8:10])
    assert (10 + hlen) % 16 == 0
    assert raw[10:10 + hlen].rstrip(b" \n") == b"{'descr': '<f4', 'fortran_order': False, 'shape': (3, 4), }"


# ---- matrices -----------------------------------------------------------------------------

def _calib(fx, cx, cy):
    return [[fx, 0.0, cx, 0.0], [0.0, fx, cy, 0.0], [0.0, 0.0, 1.0, 0.0], [0.0, 0.0, 0.0, 1.0]]


def _libm_atanf(x):
    libm = ctypes.CDLL("libm.so.6")
    libm.atanf.restype = ctypes.c_float
    libm.atanf.argtypes = [ctypes.c_float]
    return np.float32(libm.atanf(np.float32(x)))


def test_matrices_order_layout_and_intrinsics(tmp_path):
    pose = [[1, 2, 3, 4], [5, 6, 7, 8], [9, 10, 11, 12], [0, 0, 0, 1]]
    doc = {"train": {"s/b/z_color.png": {"camera_pose": pose, "calibration_mat": _calib(803.802316909057, 533.5, 345.0)},
                     "s/a/y_color.png": {"camera_pose": pose, "calibration_mat": _calib(500, 320, 240),
                                         "source_scan_ply_path": "/some/where.ply"}},
           "val": {"s/c/x_color.png": {"camera_pose": pose, "calibration_mat": _calib(500, 100.5, 50.25)}}}
    p = tmp_path / "m.json"
    p.write_text(json.dumps(doc))
    m = hio.Matrices(p)
    assert len(m) == 3
    assert [m.key(i) for i in range(3)] == ["s/a/y_color.png", "s/b/z_color.png", "s/c/x_color.png"]
    v = m.get(1)
    assert v["split"] == 0 and v["ply_path"] is None
    assert v["width"] == 1067 and v["height"] == 690
    fx = np.float32(803.802316909057)
    W = np.float32(2.0) * np.float32(533.5)
    assert v["fov_radians"] == np.float32(2.0) * _libm_atanf(W / (np.float32(2.0) * fx))
    # row-major JSON -> column-major: element (row r, column c) at 4 c + r
    assert v["pose"][1] == 5 and v["pose"][4] == 2 and v["pose"][12] == 4 and v["pose"][3] == 0
    assert m.get(0)["ply_path"] == "/some/where.ply"
    w = m.get(2)
    assert w["split"] == 1 and w["width"] == 201 and w["height"] == 100   # (int)(2 cx) truncates
    m.close()


def test_matrices_errors(tmp_path):
    p = tmp_path / "m.json"
    p.write_text(json.dumps({"train": {}}))
    with pytest.raises(hio.RMHError, match="val"):
        hio.Matrices(p)
    p.write_text('{"train": {"a/b.png": {"camera_pose": [[1,0,0,0],[0,1,0,0],[0,0,1,0]]}, '
                 '"a/c.png": {"calibration_mat": []}, "a/a.png": {}}, "val": {}}')
    m = hio.Matrices(p)            # parameters are parsed per view, lazily
    assert len(m) == 3
    for i in range(3):
        with pytest.raises(hio.RMHError):
            m.get(i)
    p.write_text('{"train": {"x": 1,}, "val": {}}')
    with pytest.raises(hio.RMHError, match="JSON"):
        hio.Matrices(p)


def test_matrices_duplicate_keys_and_escapes(tmp_path):
    pose = [[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]]
    view = json.dumps({"camera_pose": pose, "calibration_mat": _calib(10, 8, 6)})
    view2 = json.dumps({"camera_pose": pose, "calibration_mat": _calib(10, 16, 6)})
    p = tmp_path / "m.json"
    p.write_text('{"val": {}, "train": {"d/\\u00e9_color.png": %s, "d/\\u00e9_color.png": %s}} trailing'
                 % (view, view2))
    m = hio.Matrices(p)
    assert len(m) == 1 and m.key(0) == "d/é_color.png"
    assert m.get(0)["width"] == 32            # the last duplicate wins


# ---- output paths and locks ---------------------------------------------------------------

def test_view_paths():
    c, d, lk = hio.view_paths("out", "scene/rgb/frame_color_color.png")
    assert c == "out/rgb/frame_color_color.png"
    assert d == "out/rgb/frame_depth_depth.png"           # every "_color" is replaced
    assert lk == "out/rgb/.frame_color_color.png.lock"
    c, d, lk = hio.view_paths("/o", "/abs/dir/v.png")
    assert (c, d, lk) == ("/o/dir/v.png", "/o/dir/v.png", "/o/dir/.v.png.lock")   # no _color: same file
    c, _, _ = hio.view_paths("out", "single_color.png")
    assert c == "out/single_color.png"


def test_locks(tmp_path):
    lk = str(tmp_path / ".v.png.lock")
    holder = subprocess.Popen(
        [sys.executable, "-c",
         "import fcntl, sys, time\n"
         f"f = open({lk!r}, 'w')\n"
         "fcntl.lockf(f, fcntl.LOCK_EX | fcntl.LOCK_NB)\n"
         "print('held', flush=True)\n"
         "sys.stdin.readline()\n"],
        stdin=subprocess.PIPE, stdout=subprocess.PIPE, text=True)
    try:
        assert holder.stdout.readline().strip() == "held"
        assert hio.lock_try(lk) is None                   # held by another process
    finally:
        holder.stdin.write("\n")
        holder.stdin.flush()
        holder.wait(timeout=30)
    assert os.path.exists(lk)                             # a stale lock file with no holder
    fd = hio.lock_try(lk)
    assert fd is not None
    hio.lock_release(fd, lk)
    assert not os.path.exists(lk)


# ---- the CLI's argument handling (fails before any GPU work) ------------------------------

def _run(args):
    return subprocess.run([hio.MARCHER] + args, capture_output=True, text=True, timeout=60)


@pytest.mark.skipif(not os.path.exists(hio.MARCHER), reason="marcher.bin not built")
def test_cli_arguments(tmp_path):
    r = _run(["--help"])
    assert r.returncode == 0
    for flag in ["-f,--file", "-m,--matrices", "-o,--output_path", "-s,--max_points",
                 "-r,--max_radius", "-d,--headless", "-p,--precompute-radii", "-i,--ignore_existing"]:
        assert flag in r.stdout
    assert _run(["-x"]).returncode == 109
    assert _run(["positional"]).returncode == 109
    assert _run(["-s", "abc"]).returncode == 104
    assert _run(["-r", "1.5x"]).returncode == 104
    assert _run(["-b", "-1"]).returncode == 104                # -b 0 (plain kernel) and up only
    assert _run(["-f"]).returncode == 114
    assert _run(["-f", "a", "--file", "b"]).returncode == 114
    r = _run(["-d", "-f", str(tmp_path / "missing.ply")])
    assert r.returncode == 1 and "could not open" in r.stderr

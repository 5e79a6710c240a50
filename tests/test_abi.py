"""CPU checks of the C ABI: the library loads and exports every symbol include/rm.h declares,
and the host-side tree builder keeps its invariants (no GPU calls)."""
import os
import re
import subprocess

import pytest

from renderer_ray_marching_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "rm.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int|float)\s+(rm_\w+)\(", src, re.M)))


def test_header_declares_the_binding():
    assert set(_declared()) == set(_abi.SIGNATURES)


def test_library_exports_every_declared_symbol():
    L = _abi.lib()
    for name in _declared():
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (rm_\w+)", out))
    assert set(_declared()) <= exported


def test_struct_layouts_match_header(tmp_path):
    """Size and every field offset of the ctypes mirrors equal what a C compiler makes of
    include/rm.h, and the header's limits equal the binding's constants."""
    import ctypes as C
    structs = {"rm_view": _abi.rm_view, "rm_tune": _abi.rm_tune}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rm.h"', "int main(void) {"]
    for name, cls in structs.items():
        lines.append(f'printf("{name} sizeof %zu\\n", sizeof({name}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{name} {f} %zu\\n", offsetof({name}, {f}));')
    consts = ["RM_MAX_K", "RM_MARCH_K", "RM_NOALLOC_MAX_VIEWS", "RM_MAX_STACK_DEPTH",
              "RM_MAX_EXIT_DEPTH"]
    for c in consts:
        lines.append(f'printf("const {c} %d\\n", (int)({c}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = {}
    for ln in subprocess.run([str(exe)], capture_output=True, text=True,
                             check=True).stdout.splitlines():
        a, b, v = ln.split()
        got[(a, b)] = int(v)
    for name, cls in structs.items():
        assert got[(name, "sizeof")] == C.sizeof(cls), name
        for f, _ in cls._fields_:
            assert got[(name, f)] == getattr(cls, f).offset, (name, f)
    for c in consts:
        assert got[("const", c)] == getattr(_abi, c), c


def test_ctx_create_without_device_fails_cleanly():
    import ctypes as C
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    h = C.c_void_p()
    assert _abi.lib().rm_ctx_create(0, C.byref(h)) == _abi.RM_ERR_INVALID


def _declared_host():
    src = open(os.path.join(ROOT, "include", "rm_host.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const char\*|int|int64_t|void)\s+(rmh_\w+)\(", src, re.M)))


def test_host_header_declares_the_binding_and_library_exports_it():
    from renderer_ray_marching_amd import host_io
    assert set(_declared_host()) == set(host_io.SIGNATURES)
    L = host_io.lib()
    out = subprocess.run(["nm", "-D", "--defined-only", host_io.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (rmh_\w+)", out))
    assert set(_declared_host()) <= exported
    for name in _declared_host():
        assert hasattr(L, name), name


def test_host_view_spec_layout():
    import ctypes as C
    from renderer_ray_marching_amd import host_io
    # const char* key; int32 split, has_ply_path; const char* ply_path; float pose[16], calib[16];
    # int32 width, height; float fov
    assert C.sizeof(host_io.rmh_view_spec) == 8 + 4 + 4 + 8 + 64 + 64 + 4 + 4 + 4 + 4

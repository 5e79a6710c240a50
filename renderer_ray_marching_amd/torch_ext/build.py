"""Builds the libtorch extension _rm_torch.so in-tree with g++ (host C++ only; the kernels live
in librm_hip.so, which it links with an $ORIGIN rpath). Load it with torch.ops.load_library;
it registers torch.classes.rm_diff.DiffScene and torch.ops.rm_diff.render.

    python renderer_ray_marching_amd/torch_ext/build.py
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
ROOT = os.path.dirname(PKG)


def build(verbose=False):
    import torch
    from torch.utils import cpp_extension
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    inc = cpp_extension.include_paths() + [ROOT]
    libdirs = cpp_extension.library_paths()
    out = os.path.join(PKG, "_rm_torch.so")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-w",
           f"-D_GLIBCXX_USE_CXX11_ABI={abi}", "-D__HIP_PLATFORM_AMD__", "-DUSE_ROCM"]
    cmd += [f"-I{d}" for d in inc] + ["-I/opt/rocm/include"]
    cmd += [os.path.join(HERE, "rm_torch.cpp"), "-o", out + ".tmp"]
    cmd += [f"-L{d}" for d in libdirs] + ["-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu"]
    cmd += [f"-L{PKG}", "-l:librm_hip.so", "-Wl,-rpath,$ORIGIN"]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    os.replace(out + ".tmp", out)
    return out


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

"""CPU checks of the tree shape rule (renderer_ray_marching_amd/csrc/rm_tree_host.cpp): the node
count rm_tree_build uses to refuse oversized trees before laying them out equals the size of the
laid-out shape, and the limit lands where the packet walks' 32-bit node offsets need it."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "renderer_ray_marching_amd", "csrc")

HARNESS = r"""
#include "rm_tree_host.h"
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(7);
    for (int it = 0; it < 300; ++it) {
        const int64_t n = it < 64 ? it + 1 : 1 + (int64_t)(g() % 300000);
        const int L = 1 + (int)(g() % 64);
        RmShape s;
        std::string err;
        if (!rm_tree_shape(n, L, s, err)) { std::printf("shape failed %s\n", err.c_str()); return 1; }
        if ((int64_t)s.nodes.size() != rm_tree_node_count(n, L)) {
            std::printf("mismatch n %lld L %d\n", (long long)n, L);
            return 1;
        }
    }
    // leaf 1: n - 1 inner nodes; 2^26 + 1 points is the largest accepted tree at that leaf size
    std::printf("%lld %lld %lld\n", (long long)rm_tree_node_count((1ll << 26) + 1, 1),
                (long long)rm_tree_node_count((1ll << 26) + 2, 1),
                (long long)rm_tree_node_count(10000000, 16));
    return 0;
}
"""


def test_node_count_equals_laid_out_shape(tmp_path):
    src = tmp_path / "count.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "count"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", CSRC, str(src),
                    os.path.join(CSRC, "rm_tree_host.cpp"), "-pthread", "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    at_limit, over, room10m = (int(v) for v in out)
    assert at_limit == 1 << 26 and over == (1 << 26) + 1
    assert room10m == 624999          # the room10m tree's node count (rm_tree_info)

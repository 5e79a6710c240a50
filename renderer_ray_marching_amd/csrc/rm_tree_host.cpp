// Host-side construction of the search tree. See rm_tree_host.h for the layout.
#include "rm_tree_host.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <thread>
#include <unordered_map>

namespace {

int64_t split_left_count(int64_t c, int L) {
    int64_t half = c / 2;
    int64_t nl = ((half + L - 1) / L) * L;
    if (nl >= c) nl = c - 1;
    if (nl < 1) nl = 1;
    return nl;
}

struct Builder {
    const float* xyzw;
    int64_t n;
    int L;
    std::vector<int32_t> idx;
    RmHostTree* out;
    std::unordered_map<int64_t, int64_t> inner_memo;

    // Left part of a split of c > L points: half, rounded up to a multiple of L.
    int64_t left_count(int64_t c) const { return split_left_count(c, L); }

    // Number of inner nodes of a subtree over c points (the shape depends on counts only).
    int64_t inner_count(int64_t c) {
        if (c <= L) return 0;
        auto it = inner_memo.find(c);
        if (it != inner_memo.end()) return it->second;
        int64_t nl = left_count(c);
        int64_t v = 1 + inner_count(nl) + inner_count(c - nl);
        inner_memo[c] = v;
        return v;
    }

    int depth_of(int64_t c) {
        if (c <= L) return 0;
        int64_t nl = left_count(c);
        return 1 + std::max(depth_of(nl), depth_of(c - nl));
    }

    void box_of(int64_t b, int64_t e, float bx[6]) const {
        const float inf = std::numeric_limits<float>::infinity();
        float lo[3] = {inf, inf, inf}, hi[3] = {-inf, -inf, -inf};
        for (int64_t i = b; i < e; ++i) {
            const float* p = xyzw + 4 * (int64_t)idx[i];
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], p[d]);
                hi[d] = std::max(hi[d], p[d]);
            }
        }
        for (int d = 0; d < 3; ++d) {
            bx[d] = lo[d];
            bx[3 + d] = hi[d];
        }
    }

    // Build the inner node `node` over idx[b, e) whose tight box is `bx`.
    void build(int64_t node, int64_t b, int64_t e, const float bx[6], int level, int par_depth) {
        int64_t c = e - b;
        float ext[3] = {bx[3] - bx[0], bx[4] - bx[1], bx[5] - bx[2]};
        int dim = 0;
        if (ext[1] > ext[dim]) dim = 1;
        if (ext[2] > ext[dim]) dim = 2;
        int64_t nl = left_count(c);
        const float* P = xyzw;
        auto less = [P, dim](int32_t a, int32_t bb) {
            float va = P[4 * (int64_t)a + dim], vb = P[4 * (int64_t)bb + dim];
            return va < vb || (va == vb && a < bb);
        };
        std::nth_element(idx.begin() + b, idx.begin() + b + nl, idx.begin() + e, less);
        RmNode& nd = out->nodes[node];
        out->level[node] = level;
        float cb[2][6];
        box_of(b, b + nl, cb[0]);
        box_of(b + nl, e, cb[1]);
        int64_t range[2][2] = {{b, b + nl}, {b + nl, e}};
        int64_t child_node[2] = {node + 1, node + 1 + inner_count(nl)};
        for (int s = 0; s < 2; ++s) {
            std::memcpy(nd.box[s], cb[s], sizeof(cb[s]));
            int64_t cc = range[s][1] - range[s][0];
            if (cc <= L) {
                nd.ref[s] = (int32_t)range[s][0];
                nd.cnt[s] = (int32_t)cc;
            } else {
                nd.ref[s] = (int32_t)child_node[s];
                nd.cnt[s] = 0;
            }
        }
        bool spawn = par_depth > 0 && c > 200000;
        if (spawn) {
            std::thread th;
            bool started = false;
            if (nd.cnt[0] == 0) {
                float b0[6];
                std::memcpy(b0, cb[0], sizeof(b0));
                th = std::thread([this, child_node, range, b0, level, par_depth]() {
                    build(child_node[0], range[0][0], range[0][1], b0, level + 1, par_depth - 1);
                });
                started = true;
            }
            if (nd.cnt[1] == 0)
                build(child_node[1], range[1][0], range[1][1], cb[1], level + 1, par_depth - 1);
            if (started) th.join();
        } else {
            for (int s = 0; s < 2; ++s)
                if (nd.cnt[s] == 0)
                    build(child_node[s], range[s][0], range[s][1], cb[s], level + 1, 0);
        }
    }
};

}  // namespace

bool rm_build_host_tree(const float* xyzw, int64_t n, int leaf_max, RmHostTree& out,
                        std::string& err) {
    if (n <= 0) {
        err = "rm_tree_build: empty point set";
        return false;
    }
    if (n > (int64_t)std::numeric_limits<int32_t>::max() - 1) {
        err = "rm_tree_build: more than 2^31-2 points";
        return false;
    }
    if (leaf_max < 1 || leaf_max > 64) {
        err = "rm_tree_build: leaf_max must be in [1, 64]";
        return false;
    }
    Builder B;
    B.xyzw = xyzw;
    B.n = n;
    B.L = leaf_max;
    B.out = &out;
    B.idx.resize(n);
    for (int64_t i = 0; i < n; ++i) B.idx[i] = (int32_t)i;
    out.n = n;
    out.leaf_max = leaf_max;
    int64_t n_inner = B.inner_count(n);
    if (n_inner == 0) {
        // A single leaf: one root node, slot 0 = all points, slot 1 empty.
        out.nodes.assign(1, RmNode{});
        out.level.assign(1, 0);
        RmNode& r = out.nodes[0];
        B.box_of(0, n, r.box[0]);
        const float inf = std::numeric_limits<float>::infinity();
        for (int d = 0; d < 3; ++d) {
            r.box[1][d] = inf;
            r.box[1][3 + d] = -inf;
        }
        r.ref[0] = 0;
        r.cnt[0] = (int32_t)n;
        r.ref[1] = 0;
        r.cnt[1] = -1;
        out.depth = 1;
    } else {
        out.nodes.assign(n_inner, RmNode{});
        out.level.assign(n_inner, 0);
        float bx[6];
        B.box_of(0, n, bx);
        unsigned hw = std::max(1u, std::thread::hardware_concurrency());
        int par_depth = 0;
        while ((1u << par_depth) < hw && par_depth < 6) ++par_depth;
        B.build(0, 0, n, bx, 0, par_depth);
        out.depth = B.depth_of(n);
    }
    out.pts.resize(4 * n);
    for (int64_t i = 0; i < n; ++i) {
        int32_t o = B.idx[i];
        const float* p = xyzw + 4 * (int64_t)o;
        out.pts[4 * i + 0] = p[0];
        out.pts[4 * i + 1] = p[1];
        out.pts[4 * i + 2] = p[2];
        std::memcpy(&out.pts[4 * i + 3], &o, sizeof(float));
    }
    return true;
}

int64_t rm_tree_node_count(int64_t n, int leaf_max) {
    if (n <= leaf_max) return 1;
    std::unordered_map<int64_t, int64_t> memo;
    auto inner = [&](auto&& self, int64_t c) -> int64_t {
        if (c <= leaf_max) return 0;
        auto it = memo.find(c);
        if (it != memo.end()) return it->second;
        const int64_t nl = split_left_count(c, leaf_max);
        const int64_t v = 1 + self(self, nl) + self(self, c - nl);
        memo[c] = v;
        return v;
    };
    return inner(inner, n);
}

bool rm_tree_shape(int64_t n, int leaf_max, RmShape& out, std::string& err) {
    if (n <= 0) {
        err = "rm_tree_build: empty point set";
        return false;
    }
    if (n > (int64_t)std::numeric_limits<int32_t>::max() - 1) {
        err = "rm_tree_build: more than 2^31-2 points";
        return false;
    }
    if (leaf_max < 1 || leaf_max > 64) {
        err = "rm_tree_build: leaf_max must be in [1, 64]";
        return false;
    }
    out = RmShape();
    out.n = n;
    out.leaf_max = leaf_max;
    const float inf = std::numeric_limits<float>::infinity();
    if (n <= leaf_max) {
        RmNode r{};
        for (int d = 0; d < 3; ++d) {
            r.box[1][d] = inf;
            r.box[1][3 + d] = -inf;
        }
        r.ref[0] = 0;
        r.cnt[0] = (int32_t)n;
        r.ref[1] = 0;
        r.cnt[1] = -1;
        out.nodes.push_back(r);
        out.seg_b.push_back(0);
        out.seg_e.push_back(n);
        out.nl.push_back((int32_t)n);
        out.level_start = {0, 1};
        out.depth = 1;
        return true;
    }
    std::vector<int32_t> level;
    out.nodes.push_back(RmNode{});
    out.seg_b.push_back(0);
    out.seg_e.push_back(n);
    level.push_back(0);
    for (size_t qi = 0; qi < out.nodes.size(); ++qi) {
        const int64_t b = out.seg_b[qi], e = out.seg_e[qi];
        const int64_t nl = split_left_count(e - b, leaf_max);
        out.nl.push_back((int32_t)nl);
        const int64_t cb[2] = {b, b + nl}, ce[2] = {b + nl, e};
        for (int s = 0; s < 2; ++s) {
            const int64_t c = ce[s] - cb[s];
            if (c <= leaf_max) {
                out.nodes[qi].ref[s] = (int32_t)cb[s];
                out.nodes[qi].cnt[s] = (int32_t)c;
            } else {
                out.nodes[qi].ref[s] = (int32_t)out.nodes.size();
                out.nodes[qi].cnt[s] = 0;
                out.nodes.push_back(RmNode{});
                out.seg_b.push_back(cb[s]);
                out.seg_e.push_back(ce[s]);
                level.push_back(level[qi] + 1);
            }
        }
    }
    const int nlev = level.back() + 1;
    out.depth = nlev;
    out.level_start.assign(nlev + 1, 0);
    for (int32_t l : level) out.level_start[l + 1]++;
    for (int l = 0; l < nlev; ++l) out.level_start[l + 1] += out.level_start[l];
    return true;
}

// CPU oracle for the point-cloud ray-marching hot path. TEST INFRASTRUCTURE ONLY:
// loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker;
// never by the product library. See oracle.h for what each function restates and what pins it.
//
// Built with -O2 -ffp-contract=off (oracle/Makefile): every float expression below is
// evaluated exactly as written, one IEEE single-precision rounding per operation.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <random>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

const float kInf = std::numeric_limits<float>::infinity();

// Constants of src/common.h:4-7 and the knn of src/ray_marching.h:17.
const float kZFar = 100.0f;
const float kZNear = 0.1f;
const int kMaxSteps = 128;
const float kMinDist = 0.001f;
const int kMarchK = 10;

// One node of the reference tree (kdtree_cuda_builder.h:102-119 SplitInfo, child1, parent,
// aabb_min/max). Inner nodes use dim/val; leaves use [begin, end) into the leaf order.
struct RefNode {
    int child1;  // -1: leaf; otherwise the left child, the right child is child1 + 1
    int parent;  // -1 for the root
    int dim;
    float val;
    int begin, end;
    float lo[3], hi[3];
};

struct RefTree {
    int64_t n = 0;
    int depth = 0;
    std::vector<RefNode> nodes;
    std::vector<int32_t> order;   // leaf position -> original index
    std::vector<float> p;         // leaf-ordered xyz, 3 floats per point
};

// Tight AABB of order[b, e).
void tight_box(const float* xyzw, const int32_t* ord, int b, int e, float lo[3], float hi[3]) {
    for (int d = 0; d < 3; ++d) {
        lo[d] = kInf;
        hi[d] = -kInf;
    }
    for (int i = b; i < e; ++i) {
        const float* q = xyzw + 4 * (int64_t)ord[i];
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], q[d]);
            hi[d] = std::max(hi[d], q[d]);
        }
    }
}

// Reference split rule (kdtree_cuda_builder.h:297-349): a leaf with more than leaf_max points
// whose points are not all equal is split at the midpoint of the longest tight-AABB extent
// (x wins ties, then y: strict '>'), points with value > split go right (:215-223); the
// partition is stable. Children are allocated next to each other.
// Deviation: when the midpoint rounds to an AABB end so that one side is empty, the reference
// re-splits the same node forever (:336-349); the oracle makes such a node a leaf instead.
RefTree* build_ref_tree(const float* xyzw, int64_t n, int leaf_max) {
    RefTree* t = new RefTree();
    t->n = n;
    t->order.resize(n);
    for (int64_t i = 0; i < n; ++i) t->order[i] = (int32_t)i;
    RefNode root{};
    root.child1 = -1;
    root.parent = -1;
    root.begin = 0;
    root.end = (int)n;
    tight_box(xyzw, t->order.data(), 0, (int)n, root.lo, root.hi);
    t->nodes.push_back(root);
    std::vector<std::pair<int, int>> work;  // (node, depth)
    work.push_back({0, 1});
    std::vector<int32_t> tmp;
    while (!work.empty()) {
        auto [ni, depth] = work.back();
        work.pop_back();
        t->depth = std::max(t->depth, depth);
        RefNode nd = t->nodes[ni];
        int cnt = nd.end - nd.begin;
        bool all_equal = nd.lo[0] == nd.hi[0] && nd.lo[1] == nd.hi[1] && nd.lo[2] == nd.hi[2];
        if (cnt <= leaf_max || all_equal) continue;
        float ext[3] = {nd.hi[0] - nd.lo[0], nd.hi[1] - nd.lo[1], nd.hi[2] - nd.lo[2]};
        float mid[3];
        for (int d = 0; d < 3; ++d) mid[d] = (nd.hi[d] + nd.lo[d]) * 0.5f;
        int dim = 0;
        float len = ext[0];
        for (int d = 1; d <= 2; ++d)
            if (ext[d] > len) {
                dim = d;
                len = ext[d];
            }
        float split = mid[dim];
        tmp.clear();
        int nl = 0;
        for (int i = nd.begin; i < nd.end; ++i)
            if (!(xyzw[4 * (int64_t)t->order[i] + dim] > split)) {
                tmp.push_back(t->order[i]);
                ++nl;
            }
        for (int i = nd.begin; i < nd.end; ++i)
            if (xyzw[4 * (int64_t)t->order[i] + dim] > split) tmp.push_back(t->order[i]);
        if (nl == 0 || nl == cnt) continue;  // degenerate midpoint: keep as a leaf
        std::copy(tmp.begin(), tmp.end(), t->order.begin() + nd.begin);
        int left = (int)t->nodes.size();
        RefNode l{}, r{};
        l.child1 = r.child1 = -1;
        l.parent = r.parent = ni;
        l.begin = nd.begin;
        l.end = nd.begin + nl;
        r.begin = nd.begin + nl;
        r.end = nd.end;
        tight_box(xyzw, t->order.data(), l.begin, l.end, l.lo, l.hi);
        tight_box(xyzw, t->order.data(), r.begin, r.end, r.lo, r.hi);
        t->nodes[ni].child1 = left;
        t->nodes[ni].dim = dim;
        t->nodes[ni].val = split;
        t->nodes.push_back(l);
        t->nodes.push_back(r);
        work.push_back({left, depth + 1});
        work.push_back({left + 1, depth + 1});
    }
    t->p.resize(3 * n);
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) t->p[3 * i + d] = xyzw[4 * (int64_t)t->order[i] + d];
    return t;
}

// Squared distance of kdtree_cuda_3d_index.cu:186-190 (dot of the float4 difference, w = 0),
// accumulated left to right with fused multiply-adds: fma(dz, dz, fma(dy, dy, dx * dx)).
// (std::fma on floats; built with -mfma it is one instruction.)
inline float dist2(float px, float py, float pz, float qx, float qy, float qz) {
    float dx = px - qx, dy = py - qy, dz = pz - qz;
    return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}

inline bool key_less(float da, int32_t ia, float db, int32_t ib) {
    return da < db || (da == db && ia < ib);
}

// Top-k set under the (d2, index) total order; entries kept ascending.
struct TopK {
    int k, count = 0;
    float d[64];
    int32_t i[64];
    explicit TopK(int kk) : k(kk) {}
    float worst() const { return count < k ? kInf : d[k - 1]; }
    void insert(float dd, int32_t ii) {
        int pos;
        if (count < k) {
            pos = count++;
        } else {
            if (!key_less(dd, ii, d[k - 1], i[k - 1])) return;
            pos = k - 1;
        }
        while (pos > 0 && key_less(dd, ii, d[pos - 1], i[pos - 1])) {
            d[pos] = d[pos - 1];
            i[pos] = i[pos - 1];
            --pos;
        }
        d[pos] = dd;
        i[pos] = ii;
    }
};

// Reference traversal (kdtree_cuda_3d_index.cu:52-129): stackless, parent pointers, descend to
// the child on the query's side of the split (diff < 0 -> left), scan leaves, and on the way
// back up visit the far child iff it was not visited yet and its AABB mindist <= worstDist.
void ref_search(const RefTree& t, float qx, float qy, float qz, TopK& rs, orc_counters& c) {
    bool backtrack = false;
    int last = -1, cur = 0;
    const float q[3] = {qx, qy, qz};
    while (cur != -1) {
        const RefNode& nd = t.nodes[cur];
        c.v++;
        int left = nd.child1;
        int best = left, other = left;
        if (left != -1) {
            float diff = q[nd.dim] - nd.val;
            if (diff < 0)
                other++;
            else
                best++;
        }
        if (!backtrack) {
            if (left == -1) {
                for (int i = nd.begin; i < nd.end; ++i) {
                    float dd = dist2(t.p[3 * i], t.p[3 * i + 1], t.p[3 * i + 2], qx, qy, qz);
                    rs.insert(dd, t.order[i]);
                }
                c.c += (uint64_t)(nd.end - nd.begin);
                backtrack = true;
                last = cur;
                cur = nd.parent;
            } else {
                last = cur;
                cur = best;
            }
        } else {
            c.a++;
            const RefNode& o = t.nodes[other];
            float md = 0.0f;
            for (int d = 0; d < 3; ++d) {
                if (q[d] < o.lo[d]) {
                    float e = q[d] - o.lo[d];
                    md += e * e;
                } else if (q[d] > o.hi[d]) {
                    float e = q[d] - o.hi[d];
                    md += e * e;
                }
            }
            if (last == best && md <= rs.worst()) {
                last = cur;
                cur = other;
                backtrack = false;
            } else {
                last = cur;
                cur = nd.parent;
            }
        }
    }
}

int threads_or_default(int nthreads) {
#ifdef _OPENMP
    return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    (void)nthreads;
    return 1;
#endif
}

struct MarchHit {
    float s;
    int32_t idx;
};

// The view frustum of render_to_texture (src/ray_marching.cu:112-124): the edge points
// camera_matrix * v(x, y), v(x, y) = (x * (0.5 W), y * (0.5 H), -W / (2 tanf(0.5 fov)), 0), for
// the corners (-1,-1), (1,-1), (1,1), (-1,1) in that order, with glm's mat4 * vec4 summed as
// (m0 v0 + m1 v1) + (m2 v2 + m3 v3). The side planes of is_inside (:243-250) pair corner i with
// corner i+1 and use the normal (v2 x v1) with v2 = corner i, v1 = corner i+1.
struct Frustum {
    float n[4][3];
    float cam[3], dir[3];
};

Frustum make_frustum(const float* M, float pane, int W, int H) {
    const float xs[4] = {-1.0f, 1.0f, 1.0f, -1.0f};
    const float ys[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
    float e[4][3];
    for (int k = 0; k < 4; ++k) {
        float v[4] = {xs[k] * (0.5f * (float)W), ys[k] * (0.5f * (float)H), -pane, 0.0f};
        for (int i = 0; i < 3; ++i) {
            float m01 = M[0 + i] * v[0] + M[4 + i] * v[1];
            float m23 = M[8 + i] * v[2] + M[12 + i] * v[3];
            e[k][i] = m01 + m23;
        }
    }
    Frustum f;
    for (int k = 0; k < 4; ++k) {
        const float* v2 = e[k];
        const float* v1 = e[(k + 1) % 4];
        f.n[k][0] = v2[1] * v1[2] - v2[2] * v1[1];
        f.n[k][1] = v2[2] * v1[0] - v2[0] * v1[2];
        f.n[k][2] = v2[0] * v1[1] - v2[1] * v1[0];
    }
    for (int i = 0; i < 3; ++i) {
        f.cam[i] = M[12 + i];     // make_float3(CAMERA_MATRIX[3])
        f.dir[i] = -M[8 + i];     // make_float3(-CAMERA_MATRIX[2])
    }
    return f;
}

// is_inside (src/ray_marching.cu:231-258, the commented-out body): four side planes, then the
// near plane (pt - dir*ZNEAR) . dir > 0 and the far plane (pt - dir*(ZFAR-ZNEAR)) . -dir > 0.
bool frustum_inside(const Frustum& f, float x, float y, float z) {
    float pt[3] = {x - f.cam[0], y - f.cam[1], z - f.cam[2]};
    bool in = true;
    for (int k = 0; k < 4; ++k)
        in = in && ((f.n[k][0] * pt[0] + f.n[k][1] * pt[1]) + f.n[k][2] * pt[2]) < 0.0f;
    for (int i = 0; i < 3; ++i) pt[i] = pt[i] - f.dir[i] * kZNear;
    in = in && ((pt[0] * f.dir[0] + pt[1] * f.dir[1]) + pt[2] * f.dir[2]) > 0.0f;
    const float span = kZFar - kZNear;
    for (int i = 0; i < 3; ++i) pt[i] = pt[i] - f.dir[i] * span;
    in = in && ((pt[0] * -f.dir[0] + pt[1] * -f.dir[1]) + pt[2] * -f.dir[2]) > 0.0f;
    return in;
}

// distance_function (src/ray_marching.cu:260-292): kNN-10 of the query (ascending), then the
// strict minimum of sqrtf(d2) - radius; initial best is (+INF, 0).
MarchHit sdf(const RefTree& t, const float* radii, float px, float py, float pz,
             orc_counters& c) {
    TopK rs(kMarchK);
    ref_search(t, px, py, pz, rs, c);
    c.q++;
    MarchHit h{kInf, 0};
    for (int j = 0; j < rs.count; ++j) {
        int32_t id = rs.i[j];
        float s = std::sqrt(rs.d[j]) - radii[id];
        if (s < h.s) {
            h.s = s;
            h.idx = id;
        }
    }
    return h;
}

struct Ray {
    float ox, oy, oz, dx, dy, dz, t_near, t_far;
};

// Ray generation of src/ray_marching.cu:70-79 with glm's mat4*vec4 as
// (m0*v0 + m1*v1) + (m2*v2 + m3*v3) and normalize as v * (1 / sqrt(dot4)),
// dot4 = (x*x + y*y) + (z*z + w*w); float3 dot as (ax*bx + ay*by) + az*bz.
Ray make_ray(const float* M, float pane, int W, int H, int x, int y) {
    float u = (2.0f * (float)x - (float)W) / 2.0f;
    float v = (2.0f * (float)y - (float)H) / 2.0f;
    float zc = -pane;
    float r[4];
    for (int i = 0; i < 4; ++i) {
        float mul0 = M[0 + i] * u;
        float mul1 = M[4 + i] * v;
        float add0 = mul0 + mul1;
        float mul2 = M[8 + i] * zc;
        float mul3 = M[12 + i] * 0.0f;
        float add1 = mul2 + mul3;
        r[i] = add0 + add1;
    }
    float dot4 = (r[0] * r[0] + r[1] * r[1]) + (r[2] * r[2] + r[3] * r[3]);
    float inv = 1.0f / std::sqrt(dot4);
    Ray ray;
    ray.dx = r[0] * inv;
    ray.dy = r[1] * inv;
    ray.dz = r[2] * inv;
    ray.ox = M[12];
    ray.oy = M[13];
    ray.oz = M[14];
    float cx = -M[8], cy = -M[9], cz = -M[10];
    float cosang = (cx * ray.dx + cy * ray.dy) + cz * ray.dz;
    ray.t_far = kZFar / cosang;
    ray.t_near = kZNear / cosang;
    return ray;
}

}  // namespace

extern "C" {

void* orc_tree_build(const float* xyzw, int64_t n, int leaf_max) {
    if (n <= 0 || leaf_max <= 0) return nullptr;
    return build_ref_tree(xyzw, n, leaf_max);
}

void orc_tree_free(void* tree) { delete static_cast<RefTree*>(tree); }

int64_t orc_tree_num_nodes(const void* tree) {
    return (int64_t) static_cast<const RefTree*>(tree)->nodes.size();
}

int orc_tree_depth(const void* tree) { return static_cast<const RefTree*>(tree)->depth; }

void orc_tree_export(const void* tree, int32_t* child1, int32_t* parent, int32_t* begin,
                     int32_t* end, float* lo, float* hi, int32_t* order) {
    const RefTree* t = static_cast<const RefTree*>(tree);
    for (size_t i = 0; i < t->nodes.size(); ++i) {
        const RefNode& nd = t->nodes[i];
        child1[i] = nd.child1;
        parent[i] = nd.parent;
        begin[i] = nd.begin;
        end[i] = nd.end;
        for (int d = 0; d < 3; ++d) {
            lo[3 * i + d] = nd.lo[d];
            hi[3 * i + d] = nd.hi[d];
        }
    }
    std::copy(t->order.begin(), t->order.end(), order);
}

void orc_knn(const void* tree, const float* q, int64_t m, int k, int32_t* idx, float* d2,
             orc_counters* cnt, int nthreads) {
    const RefTree* t = static_cast<const RefTree*>(tree);
    orc_counters total{};
    int nt = threads_or_default(nthreads);
#pragma omp parallel num_threads(nt)
    {
        orc_counters c{};
#pragma omp for schedule(dynamic, 256)
        for (int64_t j = 0; j < m; ++j) {
            TopK rs(k);
            ref_search(*t, q[4 * j], q[4 * j + 1], q[4 * j + 2], rs, c);
            c.q++;
            for (int s = 0; s < k; ++s) {
                idx[j * k + s] = s < rs.count ? rs.i[s] : -1;
                d2[j * k + s] = s < rs.count ? rs.d[s] : kInf;
            }
        }
#pragma omp critical
        {
            total.c += c.c;
            total.v += c.v;
            total.a += c.a;
            total.q += c.q;
        }
    }
    if (cnt) *cnt = total;
}

void orc_knn_brute(const float* p, int64_t n, const float* q, int64_t m, int k, int32_t* idx,
                   float* d2, int nthreads) {
    int nt = threads_or_default(nthreads);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 16)
    for (int64_t j = 0; j < m; ++j) {
        TopK rs(k);
        for (int64_t i = 0; i < n; ++i)
            rs.insert(dist2(p[4 * i], p[4 * i + 1], p[4 * i + 2], q[4 * j], q[4 * j + 1],
                            q[4 * j + 2]),
                      (int32_t)i);
        for (int s = 0; s < k; ++s) {
            idx[j * k + s] = s < rs.count ? rs.i[s] : -1;
            d2[j * k + s] = s < rs.count ? rs.d[s] : kInf;
        }
    }
}

void orc_radii(const void* tree, float max_radius, float* r_unclamped, float* r_clamped,
               orc_counters* cnt, int nthreads) {
    const RefTree* t = static_cast<const RefTree*>(tree);
    int64_t n = t->n;
    orc_counters total{};
    int nt = threads_or_default(nthreads);
#pragma omp parallel num_threads(nt)
    {
        orc_counters c{};
#pragma omp for schedule(dynamic, 256)
        for (int64_t pos = 0; pos < n; ++pos) {
            // Queries are the points themselves (main.cu:91, 109).
            const float* pp = &t->p[3 * pos];
            TopK rs(2);
            ref_search(*t, pp[0], pp[1], pp[2], rs, c);
            c.q++;
            float d = rs.count > 1 ? rs.d[1] : kInf;
            float r = std::sqrt(d);
            int32_t id = t->order[pos];
            r_unclamped[id] = r;
            r_clamped[id] = (max_radius > 0.0f && r > max_radius) ? max_radius : r;
        }
#pragma omp critical
        {
            total.c += c.c;
            total.v += c.v;
            total.a += c.a;
            total.q += c.q;
        }
    }
    if (cnt) *cnt = total;
}

// main.cu:75-87: indices 0..n-1 shuffled by std::shuffle with std::mt19937 seeded 42; the
// subsample is the first mp of them. Same standard-library calls as the reference, so equal to
// its subsample when built against the same libstdc++.
void orc_subsample_indices(int64_t n, int32_t* out) {
    std::vector<int> idx((size_t)n);
    std::iota(idx.begin(), idx.end(), 0);
    std::mt19937 g(42);
    std::shuffle(idx.begin(), idx.end(), g);
    std::memcpy(out, idx.data(), sizeof(int) * (size_t)n);
}

// The standard fixes the 10000th output of a default-constructed std::mt19937 (seed 5489).
uint32_t orc_mt19937_10000th(void) {
    std::mt19937 g;
    g.discard(9999);
    return (uint32_t)g();
}

float orc_pane_distance(int W, float fov_radians) {
    return (float)W / (2.0f * std::tan(0.5f * fov_radians));
}

void orc_render(const void* tree, const float* radii, const float* colors, const float* pose,
                float pane, int W, int H, const int32_t* pix, int64_t npix, float* rgba,
                int32_t* hit, orc_counters* cnt, int nthreads) {
    orc_render_steps(tree, radii, colors, pose, pane, W, H, pix, npix, rgba, hit, nullptr, cnt,
                     nthreads);
}

void orc_frustum_inside(const float* pose, float pane, int W, int H, const float* xyzw,
                        int64_t n, uint8_t* inside) {
    const Frustum f = make_frustum(pose, pane, W, H);
    for (int64_t i = 0; i < n; ++i)
        inside[i] = frustum_inside(f, xyzw[4 * i], xyzw[4 * i + 1], xyzw[4 * i + 2]) ? 1 : 0;
}

void orc_render_steps(const void* tree, const float* radii, const float* colors,
                      const float* pose, float pane, int W, int H, const int32_t* pix,
                      int64_t npix, float* rgba, int32_t* hit, int32_t* steps,
                      orc_counters* cnt, int nthreads) {
    const RefTree* t = static_cast<const RefTree*>(tree);
    int64_t total_pix = pix ? npix : (int64_t)W * H;
    orc_counters total{};
    int nt = threads_or_default(nthreads);
#pragma omp parallel num_threads(nt)
    {
        orc_counters c{};
#pragma omp for schedule(dynamic, 64)
        for (int64_t j = 0; j < total_pix; ++j) {
            int64_t lin = pix ? pix[j] : j;
            int x = (int)(lin % W), y = (int)(lin / W);
            Ray ray = make_ray(pose, pane, W, H, x, y);
            // ray_march (src/ray_marching.cu:202-228)
            float tt = ray.t_near;
            int32_t hid = -1;
            float depth = 0.0f;
            int nsteps = 0;
            for (int i = 0; i < kMaxSteps; ++i) {
                ++nsteps;
                float px = ray.ox + ray.dx * tt;
                float py = ray.oy + ray.dy * tt;
                float pz = ray.oz + ray.dz * tt;
                MarchHit h = sdf(*t, radii, px, py, pz, c);
                c.s++;
                tt = tt + h.s;
                if (h.s < kMinDist) {
                    depth = tt;
                    hid = h.idx;
                    break;
                }
                if (tt > ray.t_far) break;
            }
            hit[j] = hid;
            if (steps) steps[j] = nsteps;
            if (hid >= 0) {
                rgba[4 * j + 0] = colors[4 * (int64_t)hid + 0];
                rgba[4 * j + 1] = colors[4 * (int64_t)hid + 1];
                rgba[4 * j + 2] = colors[4 * (int64_t)hid + 2];
                rgba[4 * j + 3] = depth;
            } else {
                rgba[4 * j + 0] = rgba[4 * j + 1] = rgba[4 * j + 2] = rgba[4 * j + 3] = 0.0f;
            }
        }
#pragma omp critical
        {
            total.c += c.c;
            total.v += c.v;
            total.a += c.a;
            total.q += c.q;
            total.s += c.s;
        }
    }
    if (cnt) *cnt = total;
}

double orc_loss_grad_fixed(const double* xyz, const double* r, const double* rgb, int64_t n,
                           const float* pose, float pane, int W, int H, const int32_t* hit,
                           const float* target_rgb, const float* target_depth, double w_rgb,
                           double w_depth, double* g_xyz, double* g_r, double* g_rgb) {
    (void)n;
    double loss = 0.0;
    for (int64_t lin = 0; lin < (int64_t)W * H; ++lin) {
        int32_t h = hit[lin];
        if (h < 0) continue;
        int x = (int)(lin % W), y = (int)(lin / W);
        Ray ray = make_ray(pose, pane, W, H, x, y);
        double o[3] = {ray.ox, ray.oy, ray.oz}, d[3] = {ray.dx, ray.dy, ray.dz};
        // m = o - c, b = d.m, h = m - b d (centre-to-ray offset), disc = r^2 - |h|^2,
        // entry depth t = -b - sqrt(disc)
        double m[3], b = 0.0;
        for (int a = 0; a < 3; ++a) {
            m[a] = o[a] - xyz[3 * h + a];
            b += d[a] * m[a];
        }
        double hv[3], h2 = 0.0;
        for (int a = 0; a < 3; ++a) {
            hv[a] = m[a] - b * d[a];
            h2 += hv[a] * hv[a];
        }
        double disc = r[h] * r[h] - h2;
        // colour term
        for (int a = 0; a < 3; ++a) {
            double e = rgb[3 * h + a] - (double)target_rgb[3 * lin + a];
            loss += w_rgb * e * e;
            g_rgb[3 * h + a] += w_rgb * 2.0 * e;
        }
        if (!(disc > 0.0)) continue;  // grazing hit: depth is not differentiable here
        double sq = std::sqrt(disc);
        double depth = -b - sq;
        double e = depth - (double)target_depth[lin];
        loss += w_depth * std::fabs(e);
        double gd = w_depth * (e > 0 ? 1.0 : (e < 0 ? -1.0 : 0.0));
        // dt/dc = d - h / sqrt(disc),  dt/dr = -r / sqrt(disc)
        for (int a = 0; a < 3; ++a) g_xyz[3 * h + a] += gd * (d[a] - hv[a] / sq);
        g_r[h] += gd * (-r[h] / sq);
    }
    return loss;
}

}  // extern "C"

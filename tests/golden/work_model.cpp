// Work model of this library's radii and march kernels. MEASUREMENT INFRASTRUCTURE: loaded only
// by tests/golden/make_alg_bytes.py and the tests, never by the library.
//
// It restates on the CPU what one lane of the HIP kernels (renderer_ray_marching_amd/csrc/
// rm_kernels.hip) reads when it searches alone: the exact per-lane kNN `knn_lane` (depth-first
// from the root, near child first, a leaf's box re-tested when it is popped) over the tree the
// library builds (rm_tree_host.cpp; the device build makes the same partitions), the march with
// the warm start (`warm_seed`: the previous step's ten neighbours re-measured and listed first,
// later candidates inserted only if not listed) and the early miss (`ray_exit_bound`). Per lane:
//   nodes    inner-node records expanded        64 B each (two child boxes, refs, counts)
//   retests  leaf boxes re-tested on pop         64 B each (the parent's record again)
//   points   leaf points scanned                 16 B each (x, y, z, original index)
//   seeds    warm-start gathers                  12 B each (x, y, z of a previous neighbour)
//   sdf      radius gathers of the signed distance 4 B each
// The search is exact whatever the tree, so the hit indices and depths produced here equal the
// oracle's: the tests check that equality, which pins this restatement.
//
// wm_render_packet restates instead what a WAVEFRONT of the batched kernel executes in its
// default tune (64-ray packets over 8x8 tiles, `knn_packet_rounds` with the warm start): the
// walk enters a child if any active lane's ball reaches its box, the near child first by lane
// majority, expands inner nodes until two leaves are queued, then re-tests and scans them; a
// candidate is inserted by the wavefront when some lane's distance passes its k-th distance.
// It counts per wavefront step (expansions, re-tests, scans, candidates, wavefront inserts,
// active lanes) and per lane (the lane's own passes), and can evaluate two variants that were
// studied and not built: the depth of the deepest node whose box holds every lane's ball
// (`containing`), and seeding the list with the warm bound alone (`bound_only`).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../../renderer_ray_marching_amd/csrc/rm_tree_host.h"

namespace {

const float kInf = std::numeric_limits<float>::infinity();
const float kZFar = 100.0f, kZNear = 0.1f, kMinDist = 0.001f;
const int kMaxSteps = 128, kMarchK = 10;
const uint64_t kPadKey = (uint64_t(0x7f800000u) << 32) | 0xffffffffu;

struct Counts {
    uint64_t nodes = 0, retests = 0, points = 0, seeds = 0, sdf = 0, steps = 0, rays = 0, hits = 0;
    void add(const Counts& o) {
        nodes += o.nodes; retests += o.retests; points += o.points; seeds += o.seeds;
        sdf += o.sdf; steps += o.steps; rays += o.rays; hits += o.hits;
    }
};

inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }
inline uint64_t make_key(float d2, uint32_t idx) { return (uint64_t(fbits(d2)) << 32) | idx; }
inline float key_d2(uint64_t k) { return bitsf(uint32_t(k >> 32)); }

inline float dist2(float px, float py, float pz, float qx, float qy, float qz) {
    float dx = px - qx, dy = py - qy, dz = pz - qz;
    return std::fma(dz, dz, std::fma(dy, dy, dx * dx));
}

inline float mindist(const float* b, float qx, float qy, float qz) {
    float ex = std::fmax(std::fmax(b[0] - qx, qx - b[3]), 0.0f);
    float ey = std::fmax(std::fmax(b[1] - qy, qy - b[4]), 0.0f);
    float ez = std::fmax(std::fmax(b[2] - qz, qz - b[5]), 0.0f);
    return std::fma(ez, ez, std::fma(ey, ey, ex * ex));
}

struct TopK {
    int k;
    uint64_t key[32];
    int moved = 0;   // slots the last insert wrote (0: not inserted)
    void init(int kk, uint64_t pad) { k = kk; for (int j = 0; j < k; ++j) key[j] = pad; }
    float worst() const { return key_d2(key[k - 1]); }
    // true when the list changed
    bool insert(uint64_t kk, bool nodup) {
        moved = 0;
        if (!(kk < key[k - 1])) return false;
        if (nodup)
            for (int j = 0; j < k; ++j)
                if (uint32_t(key[j]) == uint32_t(kk)) return false;
        int pos = k - 1;
        while (pos > 0 && kk < key[pos - 1]) { key[pos] = key[pos - 1]; --pos; }
        key[pos] = kk;
        moved = k - pos;
        return true;
    }
};

struct Tree {
    RmHostTree t;
    std::vector<float> xyzr;   // original order x, y, z, radius (set by wm_set_radii)
    float rmax = 0.0f;
};

// knn_lane: stack entries as the kernel pushes them (inner child: node << 1; leaf child:
// parent << 2 | slot << 1 | 1).
void knn_lane(const RmHostTree& t, float qx, float qy, float qz, TopK& tk, bool nodup, Counts& c) {
    int stk[256];
    int sp = 0;
    auto push_children = [&](int ni) {
        const RmNode& n = t.nodes[ni];
        c.nodes++;
        const float w = tk.worst();
        const float m0 = mindist(n.box[0], qx, qy, qz), m1 = mindist(n.box[1], qx, qy, qz);
        const bool n0 = n.cnt[0] >= 0 && m0 <= w, n1 = n.cnt[1] >= 0 && m1 <= w;
        auto entry = [&](int s) { return n.cnt[s] > 0 ? ((ni << 2) | (s << 1) | 1) : (n.ref[s] << 1); };
        if (n0 && n1) {
            const bool near0 = m0 <= m1;
            stk[sp++] = near0 ? entry(1) : entry(0);
            stk[sp++] = near0 ? entry(0) : entry(1);
        } else if (n0) {
            stk[sp++] = entry(0);
        } else if (n1) {
            stk[sp++] = entry(1);
        }
    };
    push_children(0);
    while (sp > 0) {
        const int e = stk[--sp];
        if (!(e & 1)) { push_children(e >> 1); continue; }
        const int ni = e >> 2, s = (e >> 1) & 1;
        const RmNode& n = t.nodes[ni];
        c.retests++;
        if (mindist(n.box[s], qx, qy, qz) > tk.worst()) continue;
        const float* p = t.pts.data() + 4 * (int64_t)n.ref[s];
        c.points += (uint64_t)n.cnt[s];
        for (int j = 0; j < n.cnt[s]; ++j)
            tk.insert(make_key(dist2(p[4 * j], p[4 * j + 1], p[4 * j + 2], qx, qy, qz),
                               fbits(p[4 * j + 3])), nodup);
    }
}

struct Ray { float ox, oy, oz, dx, dy, dz, tnear, tfar; };

Ray make_ray(const float* M, float pane, int W, int H, int x, int y) {
    float u = (2.0f * (float)x - (float)W) / 2.0f;
    float v = (2.0f * (float)y - (float)H) / 2.0f;
    float zc = -pane;
    float r[4];
    for (int i = 0; i < 4; ++i) {
        float add0 = M[0 + i] * u + M[4 + i] * v;
        float add1 = M[8 + i] * zc + M[12 + i] * 0.0f;
        r[i] = add0 + add1;
    }
    float dot4 = (r[0] * r[0] + r[1] * r[1]) + (r[2] * r[2] + r[3] * r[3]);
    float inv = 1.0f / std::sqrt(dot4);
    Ray ray;
    ray.dx = r[0] * inv; ray.dy = r[1] * inv; ray.dz = r[2] * inv;
    ray.ox = M[12]; ray.oy = M[13]; ray.oz = M[14];
    float cosang = (-M[8] * ray.dx + -M[9] * ray.dy) + -M[10] * ray.dz;
    ray.tfar = kZFar / cosang;
    ray.tnear = kZNear / cosang;
    return ray;
}

// ray_exit_bound: the ray's exit from the box of all centres dilated by the largest radius,
// 2 MIN_DIST and 1e-5 of the coordinate scale (-INF: misses the box, +INF: indeterminate).
float ray_exit_bound(const Tree& T, const Ray& r) {
    const RmNode& root = T.t.nodes[0];
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
    for (int s = 0; s < 2; ++s)
        if (root.cnt[s] >= 0)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::fmin(lo[a], root.box[s][a]);
                hi[a] = std::fmax(hi[a], root.box[s][3 + a]);
            }
    float ext = 0.0f;
    for (int a = 0; a < 3; ++a) ext = std::fmax(ext, std::fmax(std::fabs(lo[a]), std::fabs(hi[a])));
    const float scale = std::fmax(std::fmax(std::fabs(r.ox), std::fabs(r.oy)), std::fabs(r.oz)) +
                        ext + 1000.0f;
    const float e = T.rmax + 2.0f * kMinDist + 1e-5f * scale;
    const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    float tin = -kInf, tout = kInf;
    bool nan = false;
    for (int a = 0; a < 3; ++a) {
        const float inv = 1.0f / d[a];
        const float t1 = (lo[a] - e - o[a]) * inv, t2 = (hi[a] + e - o[a]) * inv;
        nan |= (t1 != t1) | (t2 != t2);
        tin = std::fmax(tin, std::fmin(t1, t2));
        tout = std::fmin(tout, std::fmax(t1, t2));
    }
    if (nan) return kInf;
    if (tin > tout) return -kInf;
    return tout + std::fabs(tout) * 1e-5f + 1e-3f;
}

int nthreads_or_default(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

}  // namespace

extern "C" {

typedef struct wm_counts {
    uint64_t nodes, retests, points, seeds, sdf, steps, rays, hits;
} wm_counts;

void* wm_tree_build(const float* xyzw, int64_t n, int leaf_max) {
    Tree* T = new Tree();
    std::string err;
    if (!rm_build_host_tree(xyzw, n, leaf_max, T->t, err)) {
        delete T;
        return nullptr;
    }
    T->xyzr.assign((size_t)n * 4, 0.0f);
    for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) T->xyzr[4 * i + d] = xyzw[4 * i + d];
    return T;
}

void wm_tree_free(void* t) { delete static_cast<Tree*>(t); }

int wm_tree_depth(const void* t) { return static_cast<const Tree*>(t)->t.depth; }

void wm_set_radii(void* t, const float* radii) {
    Tree* T = static_cast<Tree*>(t);
    float m = 0.0f;
    for (size_t i = 0; i < T->xyzr.size() / 4; ++i) {
        T->xyzr[4 * i + 3] = radii[i];
        if (radii[i] > m) m = radii[i];     // NaN radii are ignored, as max_radius_kernel does
    }
    T->rmax = m;
}

// Exact per-lane top-2 of every point (original order) and the lane's counts; r = sqrtf(2nd d2).
void wm_radii(const void* t, float* r_unclamped, wm_counts* out, int nthreads) {
    const Tree* T = static_cast<const Tree*>(t);
    const int64_t n = T->t.n;
    Counts total;
#pragma omp parallel num_threads(nthreads_or_default(nthreads))
    {
        Counts c;
#pragma omp for schedule(dynamic, 4096)
        for (int64_t i = 0; i < n; ++i) {
            const float* q = T->t.pts.data() + 4 * i;
            TopK tk;
            tk.init(2, kPadKey);
            knn_lane(T->t, q[0], q[1], q[2], tk, false, c);
            r_unclamped[fbits(q[3])] = std::sqrt(key_d2(tk.key[1]));
        }
#pragma omp critical
        total.add(c);
    }
    std::memcpy(out, &total, sizeof(Counts));
}

// The march of one view over the listed pixels (npix ids y * W + x; all when pix is NULL), as
// the batched kernel runs it lane by lane: hit (-1 on a miss) and depth (0 on a miss).
void wm_render(const void* t, const float* pose, float pane, int W, int H, const int32_t* pix,
               int64_t npix, int32_t* hit, float* depth, wm_counts* out, int nthreads) {
    const Tree* T = static_cast<const Tree*>(t);
    const int64_t total_pix = pix ? npix : (int64_t)W * H;
    const bool warm = T->t.n >= kMarchK;
    Counts total;
#pragma omp parallel num_threads(nthreads_or_default(nthreads))
    {
        Counts c;
#pragma omp for schedule(dynamic, 64)
        for (int64_t j = 0; j < total_pix; ++j) {
            const int64_t lin = pix ? pix[j] : j;
            const int x = (int)(lin % W), y = (int)(lin / W);
            const Ray r = make_ray(pose, pane, W, H, x, y);
            const float tfar = std::fmin(r.tfar, ray_exit_bound(*T, r));
            float tt = r.tnear;
            TopK tk;
            tk.init(kMarchK, kPadKey);
            int h = -1;
            float dep = 0.0f;
            c.rays++;
            for (int steps = 0;;) {
                const float px = r.ox + r.dx * tt, py = r.oy + r.dy * tt, pz = r.oz + r.dz * tt;
                const bool seeded = warm && steps > 0;
                if (seeded) {
                    for (int k = 0; k < kMarchK; ++k) {
                        const uint32_t id = uint32_t(tk.key[k]);
                        const float* q = T->xyzr.data() + 4 * (size_t)id;
                        tk.key[k] = make_key(dist2(q[0], q[1], q[2], px, py, pz), id);
                    }
                    std::sort(tk.key, tk.key + kMarchK);
                    c.seeds += kMarchK;
                } else {
                    tk.init(kMarchK, kPadKey);
                }
                knn_lane(T->t, px, py, pz, tk, seeded, c);
                float best = kInf;
                uint32_t bi = 0;
                for (int k = 0; k < kMarchK; ++k) {
                    const uint32_t id = uint32_t(tk.key[k]);
                    if (id < (uint32_t)T->t.n) {
                        c.sdf++;
                        const float s = std::sqrt(key_d2(tk.key[k])) - T->xyzr[4 * (size_t)id + 3];
                        if (s < best) { best = s; bi = id; }
                    }
                }
                tt = tt + best;
                ++steps;
                c.steps++;
                if (best < kMinDist) { h = (int)bi; dep = tt; c.hits++; break; }
                if (tt > tfar || steps >= kMaxSteps) break;
            }
            hit[j] = h;
            depth[j] = dep;
        }
#pragma omp critical
        total.add(c);
    }
    std::memcpy(out, &total, sizeof(Counts));
}

// insert_slots: over the wavefront inserts, the slots written by the lane that writes most
// (what the kernel's exec-masked insertion chain runs for); inserts_no_move: wavefront inserts
// in which no lane's list changes; seed_moves: over the warm-started wavefront steps, the sum
// over keys j of the places the furthest-moving lane moves key j when the re-measured seed is
// sorted by insertion (the kernel's chain sort), seed_steps: those wavefront steps.
typedef struct wm_packet_counts {
    uint64_t wave_steps, lane_steps, expansions, retests, scans, candidates, wave_inserts,
        lane_inserts, lane_child_passes, lane_candidates, rays, hits, tiles, containing_depth,
        containing_steps, insert_slots, inserts_no_move, seed_moves, seed_steps;
} wm_packet_counts;

// Tiles tile_stride * j of one view (8x8, row-major); hit of each counted pixel (others keep
// their value). bound_only: seed each warm lane's list with pads carrying the largest of its ten
// re-measured distances instead of the ten keys (the result is the same exact top-10).
void wm_render_packet(const void* t, const float* pose, float pane, int W, int H,
                      int tile_stride, int bound_only, int32_t* hit, wm_packet_counts* out,
                      int nthreads) {
    const Tree* T = static_cast<const Tree*>(t);
    const int tx = (W + 7) / 8, ty = (H + 7) / 8, ntiles = tx * ty;
    const int kLeafQueue = 2;
    const bool warm_ok = T->t.n >= kMarchK;
    wm_packet_counts total{};
#pragma omp parallel num_threads(nthreads_or_default(nthreads))
    {
        wm_packet_counts c{};
#pragma omp for schedule(dynamic, 1)
        for (int ti = 0; ti < ntiles; ti += std::max(1, tile_stride)) {
            c.tiles++;
            const int x0 = (ti % tx) * 8, y0 = (ti / tx) * 8;
            Ray ray[64];
            float tt[64], tfar[64], px[64], py[64], pz[64];
            int steps[64], pix[64];
            bool has[64], nodup[64];
            TopK tk[64];
            for (int l = 0; l < 64; ++l) {
                const int x = x0 + l % 8, y = y0 + l / 8;
                has[l] = x < W && y < H;
                if (!has[l]) continue;
                ray[l] = make_ray(pose, pane, W, H, x, y);
                tt[l] = ray[l].tnear;
                tfar[l] = std::fmin(ray[l].tfar, ray_exit_bound(*T, ray[l]));
                steps[l] = 0;
                pix[l] = y * W + x;
                c.rays++;
            }
            while (true) {
                int na = 0;
                for (int l = 0; l < 64; ++l) na += has[l];
                if (na == 0) break;
                c.wave_steps++;
                c.lane_steps += (uint64_t)na;
                bool all_warm = true, any_warm = false;
                int seed_mv[kMarchK] = {};
                for (int l = 0; l < 64; ++l) {
                    if (!has[l]) continue;
                    px[l] = ray[l].ox + ray[l].dx * tt[l];
                    py[l] = ray[l].oy + ray[l].dy * tt[l];
                    pz[l] = ray[l].oz + ray[l].dz * tt[l];
                    nodup[l] = warm_ok && steps[l] > 0;
                    if (nodup[l]) {
                        for (int k = 0; k < kMarchK; ++k) {
                            const uint32_t id = uint32_t(tk[l].key[k]);
                            const float* q = T->xyzr.data() + 4 * (size_t)id;
                            tk[l].key[k] = make_key(dist2(q[0], q[1], q[2], px[l], py[l], pz[l]), id);
                        }
                        any_warm = true;
                        for (int j = 1; j < kMarchK; ++j) {
                            int m = 0;
                            for (int i = 0; i < j; ++i) m += tk[l].key[j] < tk[l].key[i];
                            seed_mv[j] = std::max(seed_mv[j], m);
                        }
                        std::sort(tk[l].key, tk[l].key + kMarchK);
                        if (bound_only) {
                            tk[l].init(kMarchK, tk[l].key[kMarchK - 1] | 0xffffffffull);
                            nodup[l] = false;
                        }
                    } else {
                        tk[l].init(kMarchK, kPadKey);
                        all_warm = false;
                    }
                }
                if (any_warm) {
                    c.seed_steps++;
                    for (int j = 1; j < kMarchK; ++j) c.seed_moves += (uint64_t)seed_mv[j];
                }
                if (all_warm) {
                    // deepest node whose box holds every lane's ball with a float margin
                    int depth = 0, ni = 0;
                    while (true) {
                        const RmNode& n = T->t.nodes[ni];
                        int go = -1;
                        for (int s = 0; s < 2 && go < 0; ++s) {
                            if (n.cnt[s] < 0) continue;
                            bool in = true;
                            for (int l = 0; l < 64 && in; ++l) {
                                if (!has[l]) continue;
                                const float r = std::sqrt(tk[l].worst()) * 1.0001f + 1e-6f;
                                const float* b = n.box[s];
                                in = px[l] - b[0] > r && b[3] - px[l] > r && py[l] - b[1] > r &&
                                     b[4] - py[l] > r && pz[l] - b[2] > r && b[5] - pz[l] > r;
                            }
                            if (in) go = s;
                        }
                        if (go < 0) break;
                        ++depth;
                        if (n.cnt[go] > 0) break;
                        ni = n.ref[go];
                    }
                    c.containing_depth += (uint64_t)depth;
                    c.containing_steps++;
                }
                int stk[512];
                int sp = 0;
                auto push_children = [&](int ni) {
                    const RmNode& n = T->t.nodes[ni];
                    c.expansions++;
                    bool b0 = false, b1 = false;
                    int pref0 = 0;
                    for (int l = 0; l < 64; ++l) {
                        if (!has[l]) continue;
                        const float w = tk[l].worst();
                        const float m0 = mindist(n.box[0], px[l], py[l], pz[l]);
                        const float m1 = mindist(n.box[1], px[l], py[l], pz[l]);
                        if (n.cnt[0] >= 0 && m0 <= w) { b0 = true; c.lane_child_passes++; }
                        if (n.cnt[1] >= 0 && m1 <= w) { b1 = true; c.lane_child_passes++; }
                        pref0 += m0 <= m1;
                    }
                    auto entry = [&](int s) {
                        return n.cnt[s] > 0 ? ((ni << 2) | (s << 1) | 1) : (n.ref[s] << 1);
                    };
                    if (b0 && b1) {
                        const bool near0 = 2 * pref0 >= na;
                        stk[sp++] = near0 ? entry(1) : entry(0);
                        stk[sp++] = near0 ? entry(0) : entry(1);
                    } else if (b0) {
                        stk[sp++] = entry(0);
                    } else if (b1) {
                        stk[sp++] = entry(1);
                    }
                };
                push_children(0);
                while (sp > 0) {
                    int lq[2], nq = 0;
                    while (sp > 0 && nq < kLeafQueue) {
                        const int e = stk[--sp];
                        if (e & 1)
                            lq[nq++] = e;
                        else
                            push_children(e >> 1);
                    }
                    for (int i = 0; i < nq; ++i) {
                        const int ni = lq[i] >> 2, s = (lq[i] >> 1) & 1;
                        const RmNode& n = T->t.nodes[ni];
                        c.retests++;
                        bool any = false;
                        for (int l = 0; l < 64; ++l)
                            if (has[l] && mindist(n.box[s], px[l], py[l], pz[l]) <= tk[l].worst()) {
                                any = true;
                                c.lane_candidates += (uint64_t)n.cnt[s];
                            }
                        if (!any) continue;
                        c.scans++;
                        const float* p = T->t.pts.data() + 4 * (int64_t)n.ref[s];
                        for (int j = 0; j < n.cnt[s]; ++j) {
                            c.candidates++;
                            float d[64];
                            bool pass = false;
                            for (int l = 0; l < 64; ++l) {
                                if (!has[l]) continue;
                                d[l] = dist2(p[4 * j], p[4 * j + 1], p[4 * j + 2], px[l], py[l], pz[l]);
                                pass |= d[l] <= tk[l].worst();
                            }
                            if (!pass) continue;
                            c.wave_inserts++;
                            int mv = 0;
                            for (int l = 0; l < 64; ++l) {
                                if (!has[l]) continue;
                                c.lane_inserts += tk[l].insert(make_key(d[l], fbits(p[4 * j + 3])),
                                                               nodup[l]);
                                mv = std::max(mv, tk[l].moved);
                            }
                            c.insert_slots += (uint64_t)mv;
                            c.inserts_no_move += mv == 0;
                        }
                    }
                }
                for (int l = 0; l < 64; ++l) {
                    if (!has[l]) continue;
                    float best = kInf;
                    uint32_t bi = 0;
                    for (int k = 0; k < kMarchK; ++k) {
                        const uint32_t id = uint32_t(tk[l].key[k]);
                        if (id < (uint32_t)T->t.n) {
                            const float sd = std::sqrt(key_d2(tk[l].key[k])) - T->xyzr[4 * (size_t)id + 3];
                            if (sd < best) { best = sd; bi = id; }
                        }
                    }
                    tt[l] = tt[l] + best;
                    ++steps[l];
                    if (best < kMinDist) {
                        hit[pix[l]] = (int)bi;
                        has[l] = false;
                        c.hits++;
                    } else if (tt[l] > tfar[l] || steps[l] >= kMaxSteps) {
                        hit[pix[l]] = -1;
                        has[l] = false;
                    }
                }
            }
        }
#pragma omp critical
        {
            uint64_t* a = reinterpret_cast<uint64_t*>(&total);
            const uint64_t* b = reinterpret_cast<const uint64_t*>(&c);
            for (size_t i = 0; i < sizeof(wm_packet_counts) / 8; ++i) a[i] += b[i];
        }
    }
    *out = total;
}

}  // extern "C"

// HIP kernels for gfx950: exact kNN over the search tree, sphere radii, the plain and the
// batched sphere march, tree refit, and the differentiable forward/adjoint.
//
// Arithmetic contract (shared with the CPU oracle, which restates the reference):
//   squared distance  ((dx*dx + dy*dy) + dz*dz), dx = p.x - q.x      (no fused multiply-add:
//                     the library is built with -ffp-contract=off)
//   ordering          ascending (d2, original index): the 64-bit key (bits(d2) << 32 | index)
//                     compares exactly like that pair because d2 >= 0.
//   pruning           a child box is visited iff its mindist <= the current k-th d2; mindist
//                     is a float lower bound of every candidate d2 inside the box, so the
//                     result is the exact top-k under the total order, whatever the tree.
#include "rm_internal.h"

namespace rm {
namespace {

typedef const __attribute__((address_space(4))) float CFloat;
typedef const __attribute__((address_space(4))) int32_t CInt;

constexpr uint64_t kPadKey = (uint64_t(0x7f800000u) << 32) | 0xffffffffu;
constexpr float kInf = __builtin_huge_valf();

__device__ __forceinline__ uint64_t make_key(float d2, uint32_t idx) {
    return (uint64_t(__float_as_uint(d2)) << 32) | idx;
}
__device__ __forceinline__ float key_d2(uint64_t k) { return __uint_as_float(uint32_t(k >> 32)); }
__device__ __forceinline__ uint32_t key_idx(uint64_t k) { return uint32_t(k); }

// Squared distance of kdtree_cuda_3d_index.cu:186-190 without contraction.
__device__ __forceinline__ float dist2(float px, float py, float pz, float qx, float qy,
                                       float qz) {
    float dx = px - qx;
    float dy = py - qy;
    float dz = pz - qz;
    return (dx * dx + dy * dy) + dz * dz;
}

// Lower bound of dist2(p, q) over p in the box [l, h] (per axis |p - q| >= the gap, and
// float subtraction, squaring and addition are monotone).
__device__ __forceinline__ float mindist(float lx, float ly, float lz, float hx, float hy,
                                         float hz, float qx, float qy, float qz) {
    float ex = fmaxf(fmaxf(lx - qx, qx - hx), 0.0f);
    float ey = fmaxf(fmaxf(ly - qy, qy - hy), 0.0f);
    float ez = fmaxf(fmaxf(lz - qz, qz - hz), 0.0f);
    return (ex * ex + ey * ey) + ez * ez;
}

// Ascending top-K of 64-bit keys held in registers (fully unrolled).
template <int K>
struct TopK {
    uint64_t key[K];
    __device__ __forceinline__ void init(uint64_t pad) {
#pragma unroll
        for (int j = 0; j < K; ++j) key[j] = pad;
    }
    __device__ __forceinline__ float worst_d2() const { return key_d2(key[K - 1]); }
    __device__ __forceinline__ void insert(uint64_t kk) {
        if (kk < key[K - 1]) {
            key[K - 1] = kk;
#pragma unroll
            for (int j = K - 1; j > 0; --j) {
                uint64_t a = key[j - 1], b = key[j];
                bool sw = b < a;
                key[j - 1] = sw ? b : a;
                key[j] = sw ? a : b;
            }
        }
    }
};

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// ------------------------------------------------------------------------------------------
// Packet traversal: the whole wavefront walks one node sequence. A child is entered when any
// active lane's mindist is within that lane's current k-th distance (ballot). Node records and
// leaf points are read with scalar loads (the node index is wave-uniform) and broadcast to
// every lane. The stack is wave-uniform and lives in LDS (one row of kWaveStack per wave);
// an entry is (parent node << 1 | child slot) so the child box can be re-tested on pop.
// ------------------------------------------------------------------------------------------
template <int K>
__device__ __forceinline__ void packet_push_children(CFloat* nf, int ni, float qx, float qy,
                                                     float qz, bool active, const TopK<K>& tk,
                                                     int* wstk, int& sp, int* err) {
    CFloat* f = nf + (int64_t)ni * 16;
    CInt* fi = (CInt*)f;
    int c0 = fi[14], c1 = fi[15];
    float w = tk.worst_d2();
    float m0 = mindist(f[0], f[1], f[2], f[3], f[4], f[5], qx, qy, qz);
    float m1 = mindist(f[6], f[7], f[8], f[9], f[10], f[11], qx, qy, qz);
    uint64_t b0 = __ballot(active && c0 >= 0 && m0 <= w);
    uint64_t b1 = __ballot(active && c1 >= 0 && m1 <= w);
    if (sp + 2 > kWaveStack) {
        if (b0 | b1) *err = 1;
        return;
    }
    if (b0 && b1) {
        uint64_t pref0 = __ballot(active && m0 <= m1);
        uint64_t act = __ballot(active);
        int near = (2 * __popcll(pref0) >= __popcll(act)) ? 0 : 1;
        wstk[sp++] = 2 * ni + (1 - near);
        wstk[sp++] = 2 * ni + near;
    } else if (b0) {
        wstk[sp++] = 2 * ni;
    } else if (b1) {
        wstk[sp++] = 2 * ni + 1;
    }
}

template <int K>
__device__ __forceinline__ void packet_scan_leaf(CFloat* pf, int b, int cnt, float qx, float qy,
                                                 float qz, TopK<K>& tk) {
    CFloat* p = pf + (int64_t)b * 4;
    int j = 0;
    for (; j + 4 <= cnt; j += 4) {
        float x[4], y[4], z[4];
        uint32_t id[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            x[u] = p[4 * (j + u) + 0];
            y[u] = p[4 * (j + u) + 1];
            z[u] = p[4 * (j + u) + 2];
            id[u] = __float_as_uint(p[4 * (j + u) + 3]);
        }
        float d[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) d[u] = dist2(x[u], y[u], z[u], qx, qy, qz);
#pragma unroll
        for (int u = 0; u < 4; ++u) tk.insert(make_key(d[u], id[u]));
    }
    for (; j < cnt; ++j) {
        float d = dist2(p[4 * j], p[4 * j + 1], p[4 * j + 2], qx, qy, qz);
        tk.insert(make_key(d, __float_as_uint(p[4 * j + 3])));
    }
}

// Exact top-K of every active lane's query. Must be called by the whole wavefront
// (convergent); inactive lanes must hold keys that reject every insert (init(0)).
template <int K>
__device__ void knn_packet(const DevTree& t, float qx, float qy, float qz, bool active,
                           TopK<K>& tk, int* wstk, int* err) {
    CFloat* nf = (CFloat*)t.nodes;
    CFloat* pf = (CFloat*)t.pts;
    int sp = 0;
    packet_push_children<K>(nf, 0, qx, qy, qz, active, tk, wstk, sp, err);
    while (sp > 0) {
        --sp;
        int e = __builtin_amdgcn_readfirstlane(wstk[sp]);
        int ni = e >> 1, s = e & 1;
        CFloat* f = nf + (int64_t)ni * 16;
        CInt* fi = (CInt*)f;
        CFloat* bx = f + 6 * s;
        float md = mindist(bx[0], bx[1], bx[2], bx[3], bx[4], bx[5], qx, qy, qz);
        if (__ballot(active && md <= tk.worst_d2()) == 0) continue;
        int ref = fi[12 + s], cnt = fi[14 + s];
        if (cnt > 0)
            packet_scan_leaf<K>(pf, ref, cnt, qx, qy, qz, tk);
        else
            packet_push_children<K>(nf, ref, qx, qy, qz, active, tk, wstk, sp, err);
    }
}

// ------------------------------------------------------------------------------------------
// Per-lane traversal: every lane walks its own path with its own stack, held in LDS as
// column `lane` of a [depth][threads] array (consecutive lanes hit consecutive banks).
// ------------------------------------------------------------------------------------------
struct NodeRec {
    float4 a, b, c;
    int4 r;
};

__device__ __forceinline__ NodeRec load_node(const RmNode* nodes, int ni) {
    const float4* p = reinterpret_cast<const float4*>(nodes) + (int64_t)ni * 4;
    NodeRec n;
    n.a = p[0];
    n.b = p[1];
    n.c = p[2];
    n.r = reinterpret_cast<const int4*>(p)[3];
    return n;
}

template <int K>
__device__ __forceinline__ void lane_push_children(const RmNode* nodes, int ni, float qx,
                                                   float qy, float qz, const TopK<K>& tk,
                                                   int* lstk, int stride, int& sp, int maxsp,
                                                   int* err) {
    NodeRec n = load_node(nodes, ni);
    float w = tk.worst_d2();
    float m0 = mindist(n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, qx, qy, qz);
    float m1 = mindist(n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, qx, qy, qz);
    bool n0 = n.r.z >= 0 && m0 <= w;
    bool n1 = n.r.w >= 0 && m1 <= w;
    if (sp + 2 > maxsp) {
        if (n0 || n1) *err = 1;
        return;
    }
    if (n0 && n1) {
        int near = m0 <= m1 ? 0 : 1;
        lstk[(sp++) * stride] = 2 * ni + (1 - near);
        lstk[(sp++) * stride] = 2 * ni + near;
    } else if (n0) {
        lstk[(sp++) * stride] = 2 * ni;
    } else if (n1) {
        lstk[(sp++) * stride] = 2 * ni + 1;
    }
}

template <int K>
__device__ void knn_lane(const DevTree& t, float qx, float qy, float qz, TopK<K>& tk, int* lstk,
                         int stride, int maxsp, int* err) {
    int sp = 0;
    lane_push_children<K>(t.nodes, 0, qx, qy, qz, tk, lstk, stride, sp, maxsp, err);
    while (sp > 0) {
        int e = lstk[(--sp) * stride];
        int ni = e >> 1, s = e & 1;
        NodeRec n = load_node(t.nodes, ni);
        float lx = s ? n.b.z : n.a.x, ly = s ? n.b.w : n.a.y, lz = s ? n.c.x : n.a.z;
        float hx = s ? n.c.y : n.a.w, hy = s ? n.c.z : n.b.x, hz = s ? n.c.w : n.b.y;
        int ref = s ? n.r.y : n.r.x, cnt = s ? n.r.w : n.r.z;
        float md = mindist(lx, ly, lz, hx, hy, hz, qx, qy, qz);
        if (md > tk.worst_d2()) continue;
        if (cnt > 0) {
            const float4* p = t.pts + ref;
            int j = 0;
            for (; j + 4 <= cnt; j += 4) {
                float4 v[4];
#pragma unroll
                for (int u = 0; u < 4; ++u) v[u] = p[j + u];
#pragma unroll
                for (int u = 0; u < 4; ++u)
                    tk.insert(make_key(dist2(v[u].x, v[u].y, v[u].z, qx, qy, qz),
                                       __float_as_uint(v[u].w)));
            }
            for (; j < cnt; ++j) {
                float4 v = p[j];
                tk.insert(make_key(dist2(v.x, v.y, v.z, qx, qy, qz), __float_as_uint(v.w)));
            }
        } else {
            lane_push_children<K>(t.nodes, ref, qx, qy, qz, tk, lstk, stride, sp, maxsp, err);
        }
    }
}

// ------------------------------------------------------------------------------------------
// kNN / radius search / radii
// ------------------------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(256) void knn_kernel(DevTree t, const float* __restrict__ q,
                                                  int64_t m, int k, int32_t* __restrict__ idx,
                                                  float* __restrict__ d2, int maxsp, int* err) {
    extern __shared__ int lstk[];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float4 qq = reinterpret_cast<const float4*>(q)[i];
    TopK<K> tk;
    tk.init(kPadKey);
    knn_lane<K>(t, qq.x, qq.y, qq.z, tk, lstk + threadIdx.x, blockDim.x, maxsp, err);
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j < k) {
            bool real = tk.key[j] != kPadKey;
            idx[i * k + j] = real ? (int32_t)key_idx(tk.key[j]) : -1;
            d2[i * k + j] = real ? key_d2(tk.key[j]) : kInf;
        }
    }
}

template <int K>
__global__ __launch_bounds__(256) void radius_kernel(DevTree t, const float* __restrict__ q,
                                                     int64_t m, float r2, int max_nn,
                                                     int32_t* __restrict__ idx,
                                                     float* __restrict__ d2,
                                                     int32_t* __restrict__ counts, int maxsp,
                                                     int* err) {
    extern __shared__ int lstk[];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float4 qq = reinterpret_cast<const float4*>(q)[i];
    // Accept exactly d2 < r2 (KnnRadiusResultSet::insert, util/cuda/result_set.h:372-476):
    // a key is below (r2, 0) iff its d2 < r2.
    const uint64_t pad = make_key(r2, 0u);
    TopK<K> tk;
    tk.init(pad);
    knn_lane<K>(t, qq.x, qq.y, qq.z, tk, lstk + threadIdx.x, blockDim.x, maxsp, err);
    int c = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
        if (j < max_nn) {
            bool real = tk.key[j] < pad;
            c += real ? 1 : 0;
            idx[i * max_nn + j] = real ? (int32_t)key_idx(tk.key[j]) : -1;
            d2[i * max_nn + j] = real ? key_d2(tk.key[j]) : kInf;
        }
    }
    if (counts) counts[i] = c;
}

// Radii: one wavefront per 64 consecutive leaf-ordered points (spatially coherent), packet
// traversal with k = 2, then sqrt + clamp fused, scattered to the original order.
template <int WPB>
__global__ __launch_bounds__(WPB * 64) void radii_kernel(DevTree t, float max_radius,
                                                         float* __restrict__ r_unc,
                                                         float* __restrict__ r_cl, int* err) {
    __shared__ int wstk[WPB][kWaveStack];
    int w = threadIdx.x >> 6;
    int64_t pos = ((int64_t)blockIdx.x * WPB + w) * 64 + lane_id();
    bool active = pos < t.n;
    float4 q = active ? t.pts[pos] : make_float4(0.f, 0.f, 0.f, 0.f);
    TopK<2> tk;
    tk.init(active ? kPadKey : 0ull);
    knn_packet<2>(t, q.x, q.y, q.z, active, tk, wstk[w], err);
    if (active) {
        float r = sqrtf(key_d2(tk.key[1]));
        uint32_t o = __float_as_uint(q.w);
        if (r_unc) r_unc[o] = r;
        if (r_cl) r_cl[o] = (max_radius > 0.0f && r > max_radius) ? max_radius : r;
    }
}

// ------------------------------------------------------------------------------------------
// March
// ------------------------------------------------------------------------------------------
struct Ray {
    float ox, oy, oz, dx, dy, dz, tnear, tfar;
};

// Ray generation of src/ray_marching.cu:70-79: uv = (2 xy - res) / 2, direction
// normalize(M * (u, v, -pane, 0)) with glm's (m0 u + m1 v) + (m2 z + m3 0) and
// v * (1 / sqrt((x*x + y*y) + (z*z + w*w))); near/far planes divided by dot(-M[2], dir).
template <class F>
__device__ __forceinline__ Ray make_ray(F M, float pane, int W, int H, int x, int y) {
    float u = (2.0f * (float)x - (float)W) / 2.0f;
    float v = (2.0f * (float)y - (float)H) / 2.0f;
    float zc = -pane;
    float r[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float mul0 = M(0 + i) * u;
        float mul1 = M(4 + i) * v;
        float add0 = mul0 + mul1;
        float mul2 = M(8 + i) * zc;
        float mul3 = M(12 + i) * 0.0f;
        float add1 = mul2 + mul3;
        r[i] = add0 + add1;
    }
    float dot4 = (r[0] * r[0] + r[1] * r[1]) + (r[2] * r[2] + r[3] * r[3]);
    float inv = 1.0f / sqrtf(dot4);
    Ray ray;
    ray.dx = r[0] * inv;
    ray.dy = r[1] * inv;
    ray.dz = r[2] * inv;
    ray.ox = M(12);
    ray.oy = M(13);
    ray.oz = M(14);
    float cx = -M(8), cy = -M(9), cz = -M(10);
    float cosang = (cx * ray.dx + cy * ray.dy) + cz * ray.dz;
    ray.tfar = kZFar / cosang;
    ray.tnear = kZNear / cosang;
    return ray;
}

// Signed distance of src/ray_marching.cu:280-290 over the sorted kNN-10: strict minimum of
// sqrtf(d2) - radius, first entry wins ties; initial best (+INF, 0).
__device__ __forceinline__ void sdf10(const DevCloud& c, const TopK<kMarchK>& tk, float& s_out,
                                      uint32_t& i_out) {
    float best = kInf;
    uint32_t bi = 0;
#pragma unroll
    for (int j = 0; j < kMarchK; ++j) {
        uint32_t id = key_idx(tk.key[j]);
        if (id < (uint32_t)c.tree.n) {
            float s = sqrtf(key_d2(tk.key[j])) - c.xyzr[id].w;
            if (s < best) {
                best = s;
                bi = id;
            }
        }
    }
    s_out = best;
    i_out = bi;
}

// Warm start: the previous step's 10 neighbours, re-measured from the new query point, bound
// the new 10th key from above (they are 10 valid candidates). Start the list with pads one
// above that bound: only keys <= bound enter and pruning starts tight. Needs n >= 10.
__device__ __forceinline__ void warm_init(const DevCloud& c, TopK<kMarchK>& tk, float qx,
                                          float qy, float qz) {
    uint64_t bound = 0;
#pragma unroll
    for (int j = 0; j < kMarchK; ++j) {
        uint32_t id = key_idx(tk.key[j]);
        float4 p = c.xyzr[id];
        uint64_t kk = make_key(dist2(p.x, p.y, p.z, qx, qy, qz), id);
        bound = kk > bound ? kk : bound;
    }
    tk.init(bound + 1);
}

template <bool WARM>
__device__ void march_lane(const DevCloud& c, const Ray& r, int* lstk, int stride, int maxsp,
                           int* err, int& hit, float& depth) {
    TopK<kMarchK> tk;
    tk.init(kPadKey);
    float t = r.tnear;
    hit = -1;
    depth = 0.0f;
    for (int i = 0; i < kMaxSteps; ++i) {
        float px = r.ox + r.dx * t;
        float py = r.oy + r.dy * t;
        float pz = r.oz + r.dz * t;
        if (WARM && i > 0)
            warm_init(c, tk, px, py, pz);
        else
            tk.init(kPadKey);
        knn_lane<kMarchK>(c.tree, px, py, pz, tk, lstk, stride, maxsp, err);
        float s;
        uint32_t id;
        sdf10(c, tk, s, id);
        t = t + s;
        if (s < kMinDist) {
            hit = (int)id;
            depth = t;
            return;
        }
        if (t > r.tfar) return;
    }
}

template <bool WARM>
__device__ void march_packet(const DevCloud& c, const Ray& r, bool valid, int* wstk, int* err,
                             int& hit, float& depth) {
    TopK<kMarchK> tk;
    tk.init(kPadKey);
    bool active = valid;
    float t = r.tnear;
    int steps = 0;
    hit = -1;
    depth = 0.0f;
    while (__ballot(active)) {
        float px = r.ox + r.dx * t;
        float py = r.oy + r.dy * t;
        float pz = r.oz + r.dz * t;
        if (!active)
            tk.init(0ull);
        else if (WARM && steps > 0)
            warm_init(c, tk, px, py, pz);
        else
            tk.init(kPadKey);
        knn_packet<kMarchK>(c.tree, px, py, pz, active, tk, wstk, err);
        if (active) {
            float s;
            uint32_t id;
            sdf10(c, tk, s, id);
            t = t + s;
            ++steps;
            if (s < kMinDist) {
                hit = (int)id;
                depth = t;
                active = false;
            } else if (t > r.tfar || steps >= kMaxSteps) {
                active = false;
            }
        }
    }
}

__device__ __forceinline__ void write_pixel(const DevCloud& c, MarchOut o, int64_t pix, int hit,
                                            float depth) {
    float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
    if (hit >= 0) {
        float4 col = c.colors[hit];
        v = make_float4(col.x, col.y, col.z, depth);
    }
    o.rgba[pix] = v;
    if (o.hit) o.hit[pix] = hit;
}

// Plain per-view path: the launch shape of src/ray_marching.cu:147-149 (16x16 threads, one
// pixel each), per-lane traversal restarted from the root at every step.
__global__ __launch_bounds__(256) void render_plain_kernel(DevCloud c, DevView v, MarchOut o,
                                                           int maxsp, int* err) {
    extern __shared__ int lstk[];
    int x = blockIdx.x * 16 + threadIdx.x;
    int y = blockIdx.y * 16 + threadIdx.y;
    int tid = threadIdx.y * 16 + threadIdx.x;
    if (x >= v.W || y >= v.H) return;
    Ray r = make_ray([&](int i) { return v.M[i]; }, v.pane, v.W, v.H, x, y);
    int hit;
    float depth;
    march_lane<false>(c, r, &lstk[tid], 256, maxsp, err, hit, depth);
    write_pixel(c, o, v.pix_offset + (int64_t)y * v.W + x, hit, depth);
}

// Batched path: a persistent grid; each wavefront takes 64-pixel tiles from one shared
// queue over all views until the queue is empty, and marches the tile as one packet.
template <int MODE, bool WARM, int WPB>
__global__ __launch_bounds__(WPB * 64) void batch_kernel(BatchParams p, int maxsp) {
    __shared__ int wstk[WPB][kWaveStack];
    extern __shared__ int lstk[];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    CFloat* vf = (CFloat*)p.views;
    constexpr int VSTRIDE = sizeof(DevView) / 4;
    while (true) {
        unsigned long long qv = 0;
        if (lane == 0) qv = atomicAdd(p.queue, 1ull);
        qv = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(qv >> 32)) << 32) |
             (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)qv);
        int64_t tile = (int64_t)qv * p.sample_stride;
        if (tile >= p.total_tiles) break;
        // view lookup: last view whose tile_offset <= tile
        int lo = 0, hi = p.nviews - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            int64_t off = *(const __attribute__((address_space(4))) int64_t*)(
                vf + (int64_t)mid * VSTRIDE + offsetof(DevView, tile_offset) / 4);
            if (off <= tile)
                lo = mid;
            else
                hi = mid - 1;
        }
        CFloat* V = vf + (int64_t)lo * VSTRIDE;
        CInt* VI = (CInt*)V;
        const float pane = V[16];
        const int W = VI[17], H = VI[18], tiles_x = VI[19];
        const int64_t pix_off =
            *(const __attribute__((address_space(4))) int64_t*)(V + offsetof(DevView, pix_offset) / 4);
        const int64_t tile_off =
            *(const __attribute__((address_space(4))) int64_t*)(V + offsetof(DevView, tile_offset) / 4);
        int64_t lt = tile - tile_off;
        int tx = (int)(lt % tiles_x), ty = (int)(lt / tiles_x);
        int x = tx * p.tile_w + lane % p.tile_w;
        int y = ty * p.tile_h + lane / p.tile_w;
        bool valid = x < W && y < H;
        Ray r = make_ray([&](int i) { return V[i]; }, pane, W, H, valid ? x : 0, valid ? y : 0);
        int hit = -1;
        float depth = 0.0f;
        if (MODE == 0) {
            march_packet<WARM>(p.cloud, r, valid, wstk[w], p.err, hit, depth);
        } else {
            if (valid)
                march_lane<WARM>(p.cloud, r, &lstk[threadIdx.x], WPB * 64, maxsp, p.err, hit,
                                 depth);
        }
        if (valid) write_pixel(p.cloud, p.out, pix_off + (int64_t)y * W + x, hit, depth);
    }
}

// Batched path with per-lane refill: a lane whose ray has ended takes the next pixel of the
// shared queue (pixels numbered tile by tile, so lanes refilled together stay close on screen),
// and the wavefront keeps stepping until the queue is drained. One queue claim serves every
// lane that needs work (ballot + prefix count), and claims wait until at least `refill_min`
// lanes are idle so the counter sees few atomics. A wavefront step advances every live ray by
// one march step (one kNN-10 query each), in packet (MODE 0) or per-lane (MODE 1) traversal.
template <int MODE, bool WARM, int WPB>
__global__ __launch_bounds__(WPB * 64) void batch_refill_kernel(BatchParams p, int maxsp,
                                                                int refill_min) {
    __shared__ int wstk[WPB][kWaveStack];
    extern __shared__ int lstk[];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    const int64_t ntiles = (p.total_tiles + p.sample_stride - 1) / p.sample_stride;
    const int64_t nwork = ntiles * 64;
    const uint64_t lt_mask = lane == 0 ? 0ull : (~0ull >> (64 - lane));
    bool has = false;
    bool drained = false;          // wave-uniform
    Ray r = {0.f, 0.f, 0.f, 0.f, 0.f, -1.f, 0.f, 0.f};
    float t = 0.0f;
    int steps = 0;
    int64_t out = 0;
    TopK<kMarchK> tk;
    tk.init(kPadKey);
    while (true) {
        uint64_t needm = __ballot(!has);
        int nneed = __popcll(needm);
        if (!drained && (nneed >= refill_min || nneed == 64)) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(p.queue, (unsigned long long)nneed);
            base = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(base >> 32))
                    << 32) | (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)base);
            if ((int64_t)base + nneed >= nwork) drained = true;
            if (!has) {
                int64_t wi = (int64_t)base + __popcll(needm & lt_mask);
                if (wi < nwork) {
                    int64_t tile = (wi >> 6) * p.sample_stride;
                    int sub = (int)(wi & 63);
                    int lo = 0, hi = p.nviews - 1;
                    while (lo < hi) {
                        int mid = (lo + hi + 1) >> 1;
                        if (p.views[mid].tile_offset <= tile)
                            lo = mid;
                        else
                            hi = mid - 1;
                    }
                    const DevView* V = p.views + lo;
                    const int W = V->W, H = V->H, tiles_x = V->tiles_x;
                    int64_t lt = tile - V->tile_offset;
                    int x = (int)(lt % tiles_x) * p.tile_w + sub % p.tile_w;
                    int y = (int)(lt / tiles_x) * p.tile_h + sub / p.tile_w;
                    if (x < W && y < H) {
                        r = make_ray([&](int i) { return V->M[i]; }, V->pane, W, H, x, y);
                        t = r.tnear;
                        steps = 0;
                        out = V->pix_offset + (int64_t)y * W + x;
                        has = true;
                    }
                }
            }
        }
        if (__ballot(has) == 0) {
            if (drained) break;
            continue;
        }
        float px = r.ox + r.dx * t;
        float py = r.oy + r.dy * t;
        float pz = r.oz + r.dz * t;
        if (!has)
            tk.init(0ull);
        else if (WARM && steps > 0)
            warm_init(p.cloud, tk, px, py, pz);
        else
            tk.init(kPadKey);
        if (MODE == 0) {
            knn_packet<kMarchK>(p.cloud.tree, px, py, pz, has, tk, wstk[w], p.err);
        } else if (has) {
            knn_lane<kMarchK>(p.cloud.tree, px, py, pz, tk, &lstk[threadIdx.x], WPB * 64, maxsp,
                              p.err);
        }
        if (has) {
            float s;
            uint32_t id;
            sdf10(p.cloud, tk, s, id);
            t = t + s;
            ++steps;
            bool hitnow = s < kMinDist;
            if (hitnow || t > r.tfar || steps >= kMaxSteps) {
                write_pixel(p.cloud, p.out, out, hitnow ? (int)id : -1, hitnow ? t : 0.0f);
                has = false;
            }
        }
    }
}

__global__ void reset_queue_kernel(unsigned long long* q) { *q = 0ull; }

// ------------------------------------------------------------------------------------------
// Cloud set-up and refit
// ------------------------------------------------------------------------------------------
__global__ void make_cloud_kernel(const float4* __restrict__ xyzw, const float* __restrict__ radii,
                                  const float4* __restrict__ colors, int64_t n,
                                  float4* __restrict__ xyzr, float4* __restrict__ col) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = xyzw[i];
    xyzr[i] = make_float4(p.x, p.y, p.z, radii[i]);
    float4 c = colors[i];
    col[i] = make_float4(c.x, c.y, c.z, 1.0f);
}

__global__ void set_radii_kernel(const float* __restrict__ radii, int64_t n, float4* xyzr) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) xyzr[i].w = radii[i];
}

__global__ void set_colors_kernel(const float4* __restrict__ colors, int64_t n, float4* col) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float4 c = colors[i];
        col[i] = make_float4(c.x, c.y, c.z, 1.0f);
    }
}

__global__ void refit_points_kernel(float4* pts, float4* xyzr, const float* __restrict__ xyz,
                                    int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 p = pts[i];
    uint32_t o = __float_as_uint(p.w);
    float x = xyz[3 * (int64_t)o], y = xyz[3 * (int64_t)o + 1], z = xyz[3 * (int64_t)o + 2];
    pts[i] = make_float4(x, y, z, p.w);
    float4 q = xyzr[o];
    xyzr[o] = make_float4(x, y, z, q.w);
}

__global__ void refit_level_kernel(RmNode* nodes, const float4* __restrict__ pts,
                                   const int32_t* __restrict__ ids, int count) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    RmNode& nd = nodes[ids[i]];
    for (int s = 0; s < 2; ++s) {
        int c = nd.cnt[s];
        if (c < 0) continue;
        float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
        if (c > 0) {
            for (int j = 0; j < c; ++j) {
                float4 p = pts[nd.ref[s] + j];
                lo[0] = fminf(lo[0], p.x);
                lo[1] = fminf(lo[1], p.y);
                lo[2] = fminf(lo[2], p.z);
                hi[0] = fmaxf(hi[0], p.x);
                hi[1] = fmaxf(hi[1], p.y);
                hi[2] = fmaxf(hi[2], p.z);
            }
        } else {
            const RmNode& ch = nodes[nd.ref[s]];
            for (int u = 0; u < 2; ++u) {
                if (ch.cnt[u] < 0) continue;
                for (int d = 0; d < 3; ++d) {
                    lo[d] = fminf(lo[d], ch.box[u][d]);
                    hi[d] = fmaxf(hi[d], ch.box[u][3 + d]);
                }
            }
        }
        for (int d = 0; d < 3; ++d) {
            nd.box[s][d] = lo[d];
            nd.box[s][3 + d] = hi[d];
        }
    }
}

// ------------------------------------------------------------------------------------------
// Differentiable form
// ------------------------------------------------------------------------------------------
// Entry distance of the ray into sphere (c, r), with h = m - (d.m) d the centre-to-ray offset
// (m = o - c): disc = r^2 - |h|^2, t = -(d.m) - sqrt(disc). Forming disc from |h|^2 avoids the
// cancellation of (d.m)^2 - |m|^2 + r^2.
struct SphereHit {
    float t, sq, b, hx, hy, hz;
    bool ok;
};

__device__ __forceinline__ SphereHit sphere_entry(const Ray& ray, float4 cr) {
    SphereHit s;
    float mx = ray.ox - cr.x, my = ray.oy - cr.y, mz = ray.oz - cr.z;
    float b = (ray.dx * mx + ray.dy * my) + ray.dz * mz;
    s.hx = mx - b * ray.dx;
    s.hy = my - b * ray.dy;
    s.hz = mz - b * ray.dz;
    float h2 = (s.hx * s.hx + s.hy * s.hy) + s.hz * s.hz;
    float disc = cr.w * cr.w - h2;
    s.b = b;
    s.ok = disc > 0.0f;
    s.sq = s.ok ? sqrtf(disc) : 0.0f;
    s.t = -b - s.sq;
    return s;
}

__global__ void diff_forward_kernel(DevCloud c, DevView v, const float4* __restrict__ rgba_m,
                                    const int32_t* __restrict__ hit, float* __restrict__ rgb,
                                    float* __restrict__ depth) {
    int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t npix = (int64_t)v.W * v.H;
    if (pix >= npix) return;
    int h = hit[pix];
    float o0 = 0.f, o1 = 0.f, o2 = 0.f, dd = 0.f;
    if (h >= 0) {
        int x = (int)(pix % v.W), y = (int)(pix / v.W);
        Ray r = make_ray([&](int i) { return v.M[i]; }, v.pane, v.W, v.H, x, y);
        SphereHit s = sphere_entry(r, c.xyzr[h]);
        float4 col = c.colors[h];
        o0 = col.x;
        o1 = col.y;
        o2 = col.z;
        dd = s.ok ? s.t : rgba_m[pix].w;
    }
    rgb[3 * pix] = o0;
    rgb[3 * pix + 1] = o1;
    rgb[3 * pix + 2] = o2;
    depth[pix] = dd;
}

// Keys for the sort: the hit point of each pixel (misses sort last as n), values pixel ids.
__global__ void adjoint_keys_kernel(const int32_t* __restrict__ hit, int64_t npix, int32_t miss,
                                    int32_t* __restrict__ keys, int32_t* __restrict__ vals) {
    int64_t pix = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (pix >= npix) return;
    int h = hit[pix];
    keys[pix] = h >= 0 ? h : miss;
    vals[pix] = (int32_t)pix;
}

// One thread per run of equal keys in the sorted order: it sums the run's pixel gradients
// in pixel order and adds them to its point. Runs are disjoint, so no atomics are needed and
// the result is deterministic.
__global__ void adjoint_reduce_kernel(DevCloud c, DevView v, const int32_t* __restrict__ keys,
                                      const int32_t* __restrict__ vals, int64_t npix,
                                      const float* __restrict__ grad_rgb,
                                      const float* __restrict__ grad_depth, float* g_xyz,
                                      float* g_r, float* g_rgb) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npix) return;
    int key = keys[i];
    if (key >= c.tree.n) return;
    if (i > 0 && keys[i - 1] == key) return;
    float4 cr = c.xyzr[key];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, ar = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
    for (int64_t j = i; j < npix && keys[j] == key; ++j) {
        int64_t pix = vals[j];
        c0 += grad_rgb[3 * pix];
        c1 += grad_rgb[3 * pix + 1];
        c2 += grad_rgb[3 * pix + 2];
        float gd = grad_depth[pix];
        if (gd == 0.0f) continue;
        int x = (int)(pix % v.W), y = (int)(pix / v.W);
        Ray r = make_ray([&](int u) { return v.M[u]; }, v.pane, v.W, v.H, x, y);
        SphereHit s = sphere_entry(r, cr);
        if (!s.ok) continue;
        // dt/dc = d - h / sqrt(disc), dt/dr = -r / sqrt(disc)
        float inv = 1.0f / s.sq;
        a0 += gd * (r.dx - s.hx * inv);
        a1 += gd * (r.dy - s.hy * inv);
        a2 += gd * (r.dz - s.hz * inv);
        ar += gd * (-cr.w * inv);
    }
    g_xyz[3 * (int64_t)key] += a0;
    g_xyz[3 * (int64_t)key + 1] += a1;
    g_xyz[3 * (int64_t)key + 2] += a2;
    g_r[key] += ar;
    g_rgb[3 * (int64_t)key] += c0;
    g_rgb[3 * (int64_t)key + 1] += c1;
    g_rgb[3 * (int64_t)key + 2] += c2;
}

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1) / block); }

}  // namespace

// ------------------------------------------------------------------------------------------
// Launchers
// ------------------------------------------------------------------------------------------
#define RM_KNN_CASE(KK)                                                                     \
    case KK:                                                                                \
        hipLaunchKernelGGL(knn_kernel<KK>, dim3(grid_for(m, 256)), dim3(256), shmem, s, t, \
                           q, m, k, idx, d2, maxsp, err);                                  \
        break;

static int lane_stack_depth(const DevTree& t) { return t.depth + 3; }

hipError_t launch_knn(hipStream_t s, const DevTree& t, const float* q, int64_t m, int k,
                      int32_t* idx, float* d2, int* err) {
    if (m == 0) return hipSuccess;
    int maxsp = lane_stack_depth(t);
    size_t shmem = (size_t)maxsp * 256 * sizeof(int);
    int kk = k <= 1 ? 1 : k <= 2 ? 2 : k <= 4 ? 4 : k <= 8 ? 8 : k <= 10 ? 10 : k <= 16 ? 16 : 32;
    switch (kk) {
        RM_KNN_CASE(1)
        RM_KNN_CASE(2)
        RM_KNN_CASE(4)
        RM_KNN_CASE(8)
        RM_KNN_CASE(10)
        RM_KNN_CASE(16)
        RM_KNN_CASE(32)
    }
    return hipGetLastError();
}

#define RM_RADIUS_CASE(KK)                                                                  \
    case KK:                                                                                \
        hipLaunchKernelGGL(radius_kernel<KK>, dim3(grid_for(m, 256)), dim3(256), shmem, s, \
                           t, q, m, r2, max_nn, idx, d2, counts, maxsp, err);              \
        break;

hipError_t launch_radius(hipStream_t s, const DevTree& t, const float* q, int64_t m, float r2,
                         int max_nn, int32_t* idx, float* d2, int32_t* counts, int* err) {
    if (m == 0) return hipSuccess;
    int maxsp = lane_stack_depth(t);
    size_t shmem = (size_t)maxsp * 256 * sizeof(int);
    int kk = max_nn <= 1 ? 1 : max_nn <= 2 ? 2 : max_nn <= 4 ? 4 : max_nn <= 8 ? 8
             : max_nn <= 10 ? 10 : max_nn <= 16 ? 16 : 32;
    switch (kk) {
        RM_RADIUS_CASE(1)
        RM_RADIUS_CASE(2)
        RM_RADIUS_CASE(4)
        RM_RADIUS_CASE(8)
        RM_RADIUS_CASE(10)
        RM_RADIUS_CASE(16)
        RM_RADIUS_CASE(32)
    }
    return hipGetLastError();
}

hipError_t launch_radii(hipStream_t s, const DevTree& t, float max_radius, float* r_unc,
                        float* r_cl, int* err) {
    constexpr int WPB = 4;
    int64_t waves = (t.n + 63) / 64;
    hipLaunchKernelGGL(radii_kernel<WPB>, dim3(grid_for(waves, WPB)), dim3(WPB * 64), 0, s, t,
                       max_radius, r_unc, r_cl, err);
    return hipGetLastError();
}

hipError_t launch_make_cloud(hipStream_t s, const float* xyzw, const float* radii,
                             const float* colors, int64_t n, float4* xyzr, float4* col) {
    hipLaunchKernelGGL(make_cloud_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(xyzw), radii,
                       reinterpret_cast<const float4*>(colors), n, xyzr, col);
    return hipGetLastError();
}

hipError_t launch_set_radii(hipStream_t s, const float* radii, int64_t n, float4* xyzr) {
    hipLaunchKernelGGL(set_radii_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, radii, n,
                       xyzr);
    return hipGetLastError();
}

hipError_t launch_set_colors(hipStream_t s, const float* colors, int64_t n, float4* col) {
    hipLaunchKernelGGL(set_colors_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(colors), n, col);
    return hipGetLastError();
}

hipError_t launch_render_plain(hipStream_t s, const DevCloud& c, const DevView& v, MarchOut o,
                               int* err) {
    int maxsp = lane_stack_depth(c.tree);
    size_t shmem = (size_t)maxsp * 256 * sizeof(int);
    dim3 grid((v.W + 15) / 16, (v.H + 15) / 16);
    hipLaunchKernelGGL(render_plain_kernel, grid, dim3(16, 16), shmem, s, c, v, o, maxsp, err);
    return hipGetLastError();
}

template <int MODE, bool WARM, int WPB>
static hipError_t launch_batch_t(hipStream_t s, const BatchParams& p, int blocks, int refill) {
    int maxsp = lane_stack_depth(p.cloud.tree);
    size_t shmem = MODE == 1 ? (size_t)maxsp * WPB * 64 * sizeof(int) : 0;
    if (refill > 0)
        hipLaunchKernelGGL((batch_refill_kernel<MODE, WARM, WPB>), dim3(blocks), dim3(WPB * 64),
                           shmem, s, p, maxsp, refill);
    else
        hipLaunchKernelGGL((batch_kernel<MODE, WARM, WPB>), dim3(blocks), dim3(WPB * 64), shmem,
                           s, p, maxsp);
    return hipGetLastError();
}

hipError_t launch_batch(hipStream_t s, const BatchParams& p, int mode, int warm, int wpb,
                        int blocks, int refill) {
#define RM_B(M, W, P) \
    if (mode == M && (warm != 0) == W && wpb == P) \
        return launch_batch_t<M, W, P>(s, p, blocks, refill);
    RM_B(0, false, 1) RM_B(0, false, 2) RM_B(0, false, 4)
    RM_B(0, true, 1) RM_B(0, true, 2) RM_B(0, true, 4)
    RM_B(1, false, 1) RM_B(1, false, 2) RM_B(1, false, 4)
    RM_B(1, true, 1) RM_B(1, true, 2) RM_B(1, true, 4)
#undef RM_B
    return hipErrorInvalidValue;
}

hipError_t launch_reset_queue(hipStream_t s, unsigned long long* q) {
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, s, q);
    return hipGetLastError();
}

hipError_t launch_refit(hipStream_t s, RmNode* nodes, float4* pts, float4* xyzr,
                        const float* xyz, int64_t n, const int32_t* level_order,
                        const int32_t* level_start, int nlevels) {
    hipLaunchKernelGGL(refit_points_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, pts, xyzr,
                       xyz, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    for (int l = nlevels - 1; l >= 0; --l) {
        int b = level_start[l], cnt = level_start[l + 1] - b;
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(refit_level_kernel, dim3(grid_for(cnt, 256)), dim3(256), 0, s, nodes,
                           pts, level_order + b, cnt);
        e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_diff_forward(hipStream_t s, const DevCloud& c, const DevView& v,
                               const float4* rgba_marched, const int32_t* hit, float* rgb,
                               float* depth) {
    int64_t npix = (int64_t)v.W * v.H;
    hipLaunchKernelGGL(diff_forward_kernel, dim3(grid_for(npix, 256)), dim3(256), 0, s, c, v,
                       rgba_marched, hit, rgb, depth);
    return hipGetLastError();
}

hipError_t launch_adjoint_keys(hipStream_t s, const int32_t* hit, int64_t npix, int32_t miss,
                               int32_t* keys, int32_t* vals) {
    hipLaunchKernelGGL(adjoint_keys_kernel, dim3(grid_for(npix, 256)), dim3(256), 0, s, hit,
                       npix, miss, keys, vals);
    return hipGetLastError();
}

hipError_t launch_adjoint_reduce(hipStream_t s, const DevCloud& c, const DevView& v,
                                 const int32_t* keys_sorted, const int32_t* vals_sorted,
                                 int64_t npix, const float* grad_rgb, const float* grad_depth,
                                 float* g_xyz, float* g_r, float* g_rgb) {
    hipLaunchKernelGGL(adjoint_reduce_kernel, dim3(grid_for(npix, 256)), dim3(256), 0, s, c, v,
                       keys_sorted, vals_sorted, npix, grad_rgb, grad_depth, g_xyz, g_r, g_rgb);
    return hipGetLastError();
}

}  // namespace rm

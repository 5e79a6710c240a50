// HIP kernels for gfx950: exact kNN over the search tree, sphere radii, the plain and the
// batched sphere march, tree refit, and the differentiable forward/adjoint.
//
// Arithmetic contract (shared with the CPU oracle, which restates the reference):
//   squared distance  fma(dz, dz, fma(dy, dy, dx * dx)), dx = p.x - q.x (explicit fused
//                     multiply-adds; everything else is built with -ffp-contract=off and
//                     evaluated exactly as written)
//   ordering          ascending (d2, original index): the 64-bit key (bits(d2) << 32 | index)
//                     compares exactly like that pair because d2 >= 0.
//   pruning           a child box is visited iff its mindist <= the current k-th d2; mindist
//                     is a float lower bound of every candidate d2 inside the box, so the
//                     result is the exact top-k under the total order, whatever the tree.
#include <hipcub/hipcub.hpp>

#include "rm_internal.h"

namespace rm {
namespace {

typedef const __attribute__((address_space(4))) float CFloat;
typedef const __attribute__((address_space(4))) int32_t CInt;
typedef const __attribute__((address_space(4))) uint32_t CUint;

constexpr uint64_t kPadKey = (uint64_t(0x7f800000u) << 32) | 0xffffffffu;
constexpr float kInf = __builtin_huge_valf();

__device__ __forceinline__ uint64_t make_key(float d2, uint32_t idx) {
    return (uint64_t(__float_as_uint(d2)) << 32) | idx;
}
__device__ __forceinline__ float key_d2(uint64_t k) { return __uint_as_float(uint32_t(k >> 32)); }
__device__ __forceinline__ uint32_t key_idx(uint64_t k) { return uint32_t(k); }

// Squared distance of kdtree_cuda_3d_index.cu:186-190: the dot product of the difference,
// accumulated left to right with fused multiply-adds (the form a contracting compiler gives
// dot(d, d); the w term is 0 and drops out exactly): fma(dz, dz, fma(dy, dy, dx * dx)).
__device__ __forceinline__ float dist2(float px, float py, float pz, float qx, float qy,
                                       float qz) {
    float dx = px - qx;
    float dy = py - qy;
    float dz = pz - qz;
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

// Lower bound of dist2(p, q) over p in the box [l, h]: per axis |p - q| >= the gap, and the
// rounded subtraction, product and fused multiply-adds are monotone in their operands.
__device__ __forceinline__ float mindist(float lx, float ly, float lz, float hx, float hy,
                                         float hz, float qx, float qy, float qz) {
    float ex = fmaxf(fmaxf(lx - qx, qx - hx), 0.0f);
    float ey = fmaxf(fmaxf(ly - qy, qy - hy), 0.0f);
    float ez = fmaxf(fmaxf(lz - qz, qz - hz), 0.0f);
    return __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex));
}

// Frustum culling masks: bit i of a view's mask is set iff point i of the tree (leaf order) is
// inside the view frustum. 32-bit words; every mask carries spare words so that two consecutive
// words can be read at any point's word.
template <class P>
__device__ __forceinline__ bool mask_bit(P m, int64_t i) {
    return (m[i >> 5] >> (i & 31)) & 1u;
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
    // Insert kk < key[K-1]: with c[j] = kk < key[j] (monotone in j), slot j takes key[j-1]
    // if kk goes before it, kk if kk goes exactly there, else keeps key[j]. K compares and
    // 4 (K-1) + 2 selects, no data-dependent control flow.
    __device__ __forceinline__ void insert_sorted(uint64_t kk) {
        bool c[K];
#pragma unroll
        for (int j = 0; j < K; ++j) c[j] = kk < key[j];
#pragma unroll
        for (int j = K - 1; j > 0; --j) key[j] = c[j - 1] ? key[j - 1] : (c[j] ? kk : key[j]);
        key[0] = c[0] ? kk : key[0];
    }
    __device__ __forceinline__ void insert(uint64_t kk) {
        if (kk < key[K - 1]) insert_sorted(kk);
    }
    // Insert unless the same point is already listed (used when the list was seeded with real
    // points: a seeded point met again in a leaf has the identical key).
    __device__ __forceinline__ void insert_nodup(uint64_t kk) {
        if (kk < key[K - 1]) {
            const uint32_t id = key_idx(kk);
            bool dup = false;
#pragma unroll
            for (int j = 0; j < K; ++j) dup |= key_idx(key[j]) == id;
            if (!dup) insert_sorted(kk);
        }
    }
    // The same insertion as a top-down chain of exec-masked moves (GCN assembly; the list sizes
    // the library instantiates). v_cmpx leaves in exec the lanes whose kk goes below slot j, a
    // subset of those below slot j + 1 (the list is ascending); they move slot j up one place
    // and write kk into slot j, and the chain ends as soon as no lane goes further down. A lane
    // for which kk >= key[K-1], or (NODUP) whose list holds kk's index, takes no part. Three
    // vector instructions per slot moved, against K compares and 4 (K-1) + 2 selects for every
    // call of insert_sorted. exec is restored on every path; vcc is clobbered.
#define RM_CS(hi, lo)                                      \
    "v_cmpx_lt_u64_e32 vcc, %[k], %[l" #lo "]\n"           \
    "s_cbranch_execz 1f\n"                                 \
    "v_mov_b64_e32 %[l" #hi "], %[l" #lo "]\n"             \
    "v_mov_b64_e32 %[l" #lo "], %[k]\n"
#define RM_CS_4 RM_CS(3, 2) RM_CS(2, 1) RM_CS(1, 0)
#define RM_CS_8 RM_CS(7, 6) RM_CS(6, 5) RM_CS(5, 4) RM_CS(4, 3) RM_CS_4
#define RM_CS_10 RM_CS(9, 8) RM_CS(8, 7) RM_CS_8
#define RM_CS_16                                                                               \
    RM_CS(15, 14) RM_CS(14, 13) RM_CS(13, 12) RM_CS(12, 11) RM_CS(11, 10) RM_CS(10, 9) RM_CS_10
#define RM_CS_32                                                                               \
    RM_CS(31, 30) RM_CS(30, 29) RM_CS(29, 28) RM_CS(28, 27) RM_CS(27, 26) RM_CS(26, 25)        \
    RM_CS(25, 24) RM_CS(24, 23) RM_CS(23, 22) RM_CS(22, 21) RM_CS(21, 20) RM_CS(20, 19)        \
    RM_CS(19, 18) RM_CS(18, 17) RM_CS(17, 16) RM_CS(16, 15) RM_CS_16
#define RM_OP(i) [l##i] "+v"(key[i])
#define RM_OPS_2 RM_OP(0), RM_OP(1)
#define RM_OPS_4 RM_OPS_2, RM_OP(2), RM_OP(3)
#define RM_OPS_8 RM_OPS_4, RM_OP(4), RM_OP(5), RM_OP(6), RM_OP(7)
#define RM_OPS_10 RM_OPS_8, RM_OP(8), RM_OP(9)
#define RM_OPS_16 RM_OPS_10, RM_OP(10), RM_OP(11), RM_OP(12), RM_OP(13), RM_OP(14), RM_OP(15)
#define RM_OPS_32                                                                              \
    RM_OPS_16, RM_OP(16), RM_OP(17), RM_OP(18), RM_OP(19), RM_OP(20), RM_OP(21), RM_OP(22),    \
        RM_OP(23), RM_OP(24), RM_OP(25), RM_OP(26), RM_OP(27), RM_OP(28), RM_OP(29), RM_OP(30), \
        RM_OP(31)
// the chain of a list of TOP + 1 slots; DUP: the duplicate test after the first compare
#define RM_CHAIN(TOP, DUP, STEPS, OPS, ...)                                                    \
    asm volatile("s_mov_b64 %[sv], exec\n"                                                     \
                 "v_cmpx_lt_u64_e32 vcc, %[k], %[l" #TOP "]\n" DUP                             \
                 "s_cbranch_execz 1f\n"                                                        \
                 "v_mov_b64_e32 %[l" #TOP "], %[k]\n" STEPS                                    \
                 "1:\n"                                                                        \
                 "s_mov_b64 exec, %[sv]\n"                                                     \
                 : [sv] "=&s"(saved), OPS                                                      \
                 : [k] "v"(kk)__VA_ARGS__                                                      \
                 : "vcc")
    // the index of kk against slots 0..8 (slot 9 cannot hold it once kk < key[9])
#define RM_DUP_10                                                                              \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i0]\n"                                                    \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i1]\n"                                                    \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i2]\n"                                                    \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i3]\n"                                                    \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i4]\n"                                                    \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i5]\n"                                                    \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i6]\n"                                                    \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i7]\n"                                                    \
    "v_cmpx_ne_u32_e32 vcc, %[ki], %[i8]\n"
    static constexpr bool kChain = K == 2 || K == 4 || K == 8 || K == 10 || K == 16 || K == 32;
    template <bool NODUP>
    __device__ __forceinline__ void insert_chain(uint64_t kk) {
        uint64_t saved;
        if constexpr (K == 10 && NODUP)
            RM_CHAIN(9, RM_DUP_10, RM_CS_10, RM_OPS_10, , [ki] "v"(key_idx(kk)),
                     [i0] "v"(key_idx(key[0])), [i1] "v"(key_idx(key[1])),
                     [i2] "v"(key_idx(key[2])), [i3] "v"(key_idx(key[3])),
                     [i4] "v"(key_idx(key[4])), [i5] "v"(key_idx(key[5])),
                     [i6] "v"(key_idx(key[6])), [i7] "v"(key_idx(key[7])),
                     [i8] "v"(key_idx(key[8])));
        else if constexpr (K == 10)
            RM_CHAIN(9, "", RM_CS_10, RM_OPS_10);
        else if constexpr (K == 2)
            RM_CHAIN(1, "", RM_CS(1, 0), RM_OPS_2);
        else if constexpr (K == 4)
            RM_CHAIN(3, "", RM_CS_4, RM_OPS_4);
        else if constexpr (K == 8)
            RM_CHAIN(7, "", RM_CS_8, RM_OPS_8);
        else if constexpr (K == 16)
            RM_CHAIN(15, "", RM_CS_16, RM_OPS_16);
        else if constexpr (K == 32)
            RM_CHAIN(31, "", RM_CS_32, RM_OPS_32);
    }
#undef RM_DUP_10
#undef RM_CHAIN
#undef RM_OPS_32
#undef RM_OPS_16
#undef RM_OPS_10
#undef RM_OPS_8
#undef RM_OPS_4
#undef RM_OPS_2
#undef RM_OP
#undef RM_CS_32
#undef RM_CS_16
#undef RM_CS_10
#undef RM_CS_8
#undef RM_CS_4
#undef RM_CS
    template <bool NODUP>
    __device__ __forceinline__ void add(uint64_t kk) {
        // the chain checks duplicates for K = 10 only (the warm-started march); other lists
        // that need the check take the select network
        if constexpr (kChain && (!NODUP || K == 10))
            insert_chain<NODUP>(kk);
        else if (NODUP)
            insert_nodup(kk);
        else
            insert(kk);
    }
};

// Sorts 10 keys by insertion, each key moved down an exec-masked chain as in
// TopK::insert_chain: key j is held in t, and while some lane's t is below slot i (v_cmpx),
// those lanes move slot i up and write t into slot i. Cost per key: 2 + 3 x (places the
// furthest lane moves it). For the warm start's seed (the previous step's ten nearest,
// re-measured from a point a little further along the ray, so nearly in order) that is about
// 80 vector instructions per wavefront step against 160 for a 32-exchange sorting network.
__device__ __forceinline__ void sort10_nearly_sorted(uint64_t* k) {
    uint64_t t, saved;
#define RM_SORT_STEP(hi, lo)                               \
    "v_cmpx_lt_u64_e32 vcc, %[t], %[l" #lo "]\n"           \
    "s_cbranch_execz 1f\n"                                 \
    "v_mov_b64_e32 %[l" #hi "], %[l" #lo "]\n"             \
    "v_mov_b64_e32 %[l" #lo "], %[t]\n"
#define RM_SORT_END "1:\n" "s_mov_b64 exec, %[sv]\n"
    asm volatile(
        "s_mov_b64 %[sv], exec\n"
        "v_mov_b64_e32 %[t], %[l1]\n" RM_SORT_STEP(1, 0) RM_SORT_END
        "v_mov_b64_e32 %[t], %[l2]\n" RM_SORT_STEP(2, 1) RM_SORT_STEP(1, 0) RM_SORT_END
        "v_mov_b64_e32 %[t], %[l3]\n" RM_SORT_STEP(3, 2) RM_SORT_STEP(2, 1) RM_SORT_STEP(1, 0)
        RM_SORT_END
        "v_mov_b64_e32 %[t], %[l4]\n" RM_SORT_STEP(4, 3) RM_SORT_STEP(3, 2) RM_SORT_STEP(2, 1)
        RM_SORT_STEP(1, 0) RM_SORT_END
        "v_mov_b64_e32 %[t], %[l5]\n" RM_SORT_STEP(5, 4) RM_SORT_STEP(4, 3) RM_SORT_STEP(3, 2)
        RM_SORT_STEP(2, 1) RM_SORT_STEP(1, 0) RM_SORT_END
        "v_mov_b64_e32 %[t], %[l6]\n" RM_SORT_STEP(6, 5) RM_SORT_STEP(5, 4) RM_SORT_STEP(4, 3)
        RM_SORT_STEP(3, 2) RM_SORT_STEP(2, 1) RM_SORT_STEP(1, 0) RM_SORT_END
        "v_mov_b64_e32 %[t], %[l7]\n" RM_SORT_STEP(7, 6) RM_SORT_STEP(6, 5) RM_SORT_STEP(5, 4)
        RM_SORT_STEP(4, 3) RM_SORT_STEP(3, 2) RM_SORT_STEP(2, 1) RM_SORT_STEP(1, 0) RM_SORT_END
        "v_mov_b64_e32 %[t], %[l8]\n" RM_SORT_STEP(8, 7) RM_SORT_STEP(7, 6) RM_SORT_STEP(6, 5)
        RM_SORT_STEP(5, 4) RM_SORT_STEP(4, 3) RM_SORT_STEP(3, 2) RM_SORT_STEP(2, 1)
        RM_SORT_STEP(1, 0) RM_SORT_END
        "v_mov_b64_e32 %[t], %[l9]\n" RM_SORT_STEP(9, 8) RM_SORT_STEP(8, 7) RM_SORT_STEP(7, 6)
        RM_SORT_STEP(6, 5) RM_SORT_STEP(5, 4) RM_SORT_STEP(4, 3) RM_SORT_STEP(3, 2)
        RM_SORT_STEP(2, 1) RM_SORT_STEP(1, 0) RM_SORT_END
        : [sv] "=&s"(saved), [t] "=&v"(t), [l0] "+v"(k[0]), [l1] "+v"(k[1]), [l2] "+v"(k[2]),
          [l3] "+v"(k[3]), [l4] "+v"(k[4]), [l5] "+v"(k[5]), [l6] "+v"(k[6]), [l7] "+v"(k[7]),
          [l8] "+v"(k[8]), [l9] "+v"(k[9])
        :
        : "vcc");
#undef RM_SORT_STEP
#undef RM_SORT_END
}

__device__ __forceinline__ int lane_id() { return (int)__lane_id(); }

// Workgroups are dispatched to the chip's 8 XCDs round-robin (block b runs on XCD b mod 8), and
// each XCD has its own L2. This renumbers blocks so that every XCD works on one contiguous run of
// the logical block range: neighbouring work shares an L2. A bijection on [0, gridDim.x).
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_block() {
    const int nb = (int)gridDim.x, b = (int)blockIdx.x;
    const int xcd = b % kXcds, local = b / kXcds;
    const int base = nb / kXcds, extra = nb % kXcds;
    return xcd * base + min(xcd, extra) + local;
}

// ------------------------------------------------------------------------------------------
// Stack entries (both traversals): an inner child is pushed as (node << 1), a leaf child as
// (parent << 2 | slot << 1 | 1). The per-lane walk re-tests a leaf's box against the current
// k-th distance when it is popped (the bound may have shrunk since the push); the packet walks
// scan a popped leaf directly, since the union of 64 lanes' balls nearly always still needs it.
// An inner child is expanded directly, since expanding it tests both of its children anyway.
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int child_entry(int ni, int s, int ref, int cnt) {
    return cnt > 0 ? ((ni << 2) | (s << 1) | 1) : (ref << 1);
}

// ------------------------------------------------------------------------------------------
// Packet traversal: the whole wavefront walks one node sequence. A child is entered when any
// active lane's mindist is within that lane's current k-th distance (ballot). Node records and
// leaf points are read with scalar loads (the node index is wave-uniform) and broadcast to
// every lane. The stack is wave-uniform and lives in LDS (one row of kWaveStack per wave).
// ------------------------------------------------------------------------------------------
// Node record ni of a packet walk (wave-uniform index): a 32-bit byte offset (trees have at most
// kMaxNodes records), which the scalar load takes in one register.
__device__ __forceinline__ CFloat* node_rec(CFloat* nf, int ni) {
    typedef const __attribute__((address_space(4))) char CChar;
    return (CFloat*)((CChar*)nf + ((uint32_t)ni << 6));
}

// amask: ballot of the active lanes (a wave-uniform 64-bit mask). A compare's ballot masked with
// it needs no per-lane boolean (ballot of (active && cond) would materialise one in a VGPR).
// MAYBE_EMPTY: the node may have an empty child slot (count < 0). Without culling that is only
// the root of a one-leaf tree, whose walk never expands another node, so the walk's loop skips
// the test. The one or two entries are written with one two-word store and sp moves by the
// number of children entered (near child on top).
template <int K, bool CULL = false, bool MAYBE_EMPTY = CULL>
__device__ __forceinline__ void packet_push_children(CFloat* nf, int ni, float qx, float qy,
                                                     float qz, uint64_t amask, const TopK<K>& tk,
                                                     int* wstk, int& sp, int* err,
                                                     CUint* nbits = nullptr) {
    CFloat* f = node_rec(nf, ni);
    CInt* fi = (CInt*)f;
    const int r0 = fi[12], r1 = fi[13];
    const int c0 = fi[14], c1 = fi[15];
    const float w = tk.worst_d2();
    const float m0 = mindist(f[0], f[1], f[2], f[3], f[4], f[5], qx, qy, qz);
    const float m1 = mindist(f[6], f[7], f[8], f[9], f[10], f[11], qx, qy, qz);
    uint64_t b0 = __builtin_amdgcn_ballot_w64(m0 <= w) & amask;
    uint64_t b1 = __builtin_amdgcn_ballot_w64(m1 <= w) & amask;
    if (CULL) {
        // a child slot with no point inside the view frustum is not entered (an empty slot
        // has its bit clear too)
        const uint32_t nb = nbits[ni >> 2] >> ((ni & 3) * 8);
        b0 = (nb & 1u) ? b0 : 0ull;
        b1 = (nb & 2u) ? b1 : 0ull;
    } else if (MAYBE_EMPTY) {
        b0 = c0 >= 0 ? b0 : 0ull;
        b1 = c1 >= 0 ? b1 : 0ull;
    }
    if ((b0 | b1) == 0) return;
    if (sp + 2 > kWaveStack) {
        *err = 1;
        return;
    }
    const int e0 = child_entry(ni, 0, r0, c0), e1 = child_entry(ni, 1, r1, c1);
    int lo = b0 ? e0 : e1, hi = lo, n = 1;
    if (b0 && b1) {
        // near child (popped first) = the one nearer for most active lanes
        const uint64_t pref0 = __builtin_amdgcn_ballot_w64(m0 <= m1) & amask;
        const bool near0 = 2 * __popcll(pref0) >= __popcll(amask);
        lo = near0 ? e1 : e0;
        hi = near0 ? e0 : e1;
        n = 2;
    }
    wstk[sp] = lo;
    wstk[sp + 1] = hi;
    sp += n;
}

// WIDE: a leaf of exactly kWideLeaf points (every leaf of a leaf-16 tree but the last one) is
// loaded whole before any of it is scored, so its four 64-byte scalar loads are in flight
// together instead of one after another.
constexpr int kWideLeaf = 16;

template <int K, bool NODUP, bool CULL = false, bool WIDE = false>
__device__ __forceinline__ void packet_scan_leaf(CFloat* pf, int b, int cnt, float qx, float qy,
                                                 float qz, TopK<K>& tk, CUint* mask = nullptr) {
    CFloat* p = pf + (int64_t)b * 4;
    if (WIDE && !CULL && cnt == kWideLeaf) {
        float x[kWideLeaf], y[kWideLeaf], z[kWideLeaf];
        uint32_t id[kWideLeaf];
#pragma unroll
        for (int u = 0; u < kWideLeaf; ++u) {
            x[u] = p[4 * u + 0];
            y[u] = p[4 * u + 1];
            z[u] = p[4 * u + 2];
            id[u] = __float_as_uint(p[4 * u + 3]);
        }
#pragma unroll
        for (int u = 0; u < kWideLeaf; ++u) {
            const float d = dist2(x[u], y[u], z[u], qx, qy, qz);
            if (__builtin_amdgcn_ballot_w64(d <= tk.worst_d2()) != 0)
                tk.template add<NODUP>(make_key(d, id[u]));
        }
        return;
    }
    // culling: a leaf (at most 64 points) takes its mask bits from one three-word window,
    // shifted to start at its first point (the mask array ends with a spare pair of words, so
    // the window is always inside it)
    uint64_t leaf_keep = ~0ull;
    if (CULL) {
        const int64_t w0 = (int64_t)b >> 5;
        const uint32_t sh = (uint32_t)b & 31u;
        const uint64_t w01 = (uint64_t)mask[w0] | ((uint64_t)mask[w0 + 1] << 32);
        leaf_keep = sh ? (w01 >> sh) | ((uint64_t)mask[w0 + 2] << (64u - sh)) : w01;
    }
    int j = 0;
    for (; j + 4 <= cnt; j += 4) {
        // culled points (wave-uniform bits) are skipped by the scalar unit
        const uint32_t keep = CULL ? (uint32_t)(leaf_keep >> j) & 0xfu : 0xfu;
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
        // d2 <= the k-th d2 is necessary for key < the k-th key: a 32-bit test per candidate,
        // and the key is formed and compared only when some lane passes it
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (((keep >> u) & 1u) && __builtin_amdgcn_ballot_w64(d[u] <= tk.worst_d2()) != 0)
                tk.template add<NODUP>(make_key(d[u], id[u]));
    }
    for (; j < cnt; ++j) {
        if (CULL && !((leaf_keep >> j) & 1u)) continue;
        float d = dist2(p[4 * j], p[4 * j + 1], p[4 * j + 2], qx, qy, qz);
        tk.template add<NODUP>(make_key(d, __float_as_uint(p[4 * j + 3])));
    }
}

// Exact top-K of every active lane's query. Must be called by the whole wavefront
// (convergent); inactive lanes must hold keys that reject every insert (init(0)). NODUP: the
// list was seeded with real points, so a candidate already listed is skipped.
// Leaves whose points lie in [skip_b, skip_e) of the leaf order are not scanned (the caller has
// scored those points already).
template <int K, bool NODUP = false>
__device__ void knn_packet(const DevTree& t, float qx, float qy, float qz, bool active,
                           TopK<K>& tk, int* wstk, int* err, int skip_b = 0, int skip_e = 0) {
    CFloat* nf = (CFloat*)t.nodes;
    CFloat* pf = (CFloat*)t.pts;
    const uint64_t amask = __builtin_amdgcn_ballot_w64(active);
    int sp = 0;
    packet_push_children<K, false, true>(nf, 0, qx, qy, qz, amask, tk, wstk, sp, err);
    while (sp > 0) {
        --sp;
        const int e = __builtin_amdgcn_readfirstlane(wstk[sp]);
        if (e & 1) {
            const int ni = e >> 2, s = (e >> 1) & 1;
            CInt* fi = (CInt*)node_rec(nf, ni);
            const int b = fi[12 + s];
            if (b >= skip_b && b < skip_e) continue;
            packet_scan_leaf<K, NODUP, false, true>(pf, b, fi[14 + s], qx, qy, qz, tk);
        } else {
            packet_push_children<K>(nf, e >> 1, qx, qy, qz, amask, tk, wstk, sp, err);
        }
    }
}

// Same result as knn_packet, in rounds: expand inner nodes (reading only the k-th distance)
// until the stack is empty or kLeafQueue leaves are queued, then scan the queued leaves in
// order. The list is modified in
// the scan loop only, so it keeps its registers across the node expansions. Leaves are queued
// in the near-first order of the depth-first walk; bounds tightened by a round prune the
// entries still on the stack when they are expanded.
template <int K, bool NODUP = false, bool CULL = false>
__device__ void knn_packet_rounds(const DevTree& t, float qx, float qy, float qz, bool active,
                                  TopK<K>& tk, int* wstk, int* err,
                                  CUint* mask = nullptr) {
    CFloat* nf = (CFloat*)t.nodes;
    CFloat* pf = (CFloat*)t.pts;
    CUint* nbits = CULL ? mask + cull_mask_words(t.n) : nullptr;
    const uint64_t amask = __builtin_amdgcn_ballot_w64(active);
    int sp = 0;
    packet_push_children<K, CULL, true>(nf, 0, qx, qy, qz, amask, tk, wstk, sp, err, nbits);
    static_assert(kLeafQueue == 2, "the leaf queue is two scalar registers");
    while (sp > 0) {
        int nq = 0, q0 = 0, q1 = 0;
        while (sp > 0 && nq < kLeafQueue) {
            --sp;
            const int e = __builtin_amdgcn_readfirstlane(wstk[sp]);
            if (e & 1) {
                if (nq == 0)
                    q0 = e;
                else
                    q1 = e;
                ++nq;
            } else {
                packet_push_children<K, CULL>(nf, e >> 1, qx, qy, qz, amask, tk, wstk, sp, err,
                                              nbits);
            }
        }
        for (int i = 0; i < nq; ++i) {
            const int e = i == 0 ? q0 : q1;
            const int ni = e >> 2, s = (e >> 1) & 1;
            CInt* fi = (CInt*)node_rec(nf, ni);
            packet_scan_leaf<K, NODUP, CULL>(pf, fi[12 + s], fi[14 + s], qx, qy, qz, tk, mask);
        }
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

// A lane's stack: entry i lives in LDS at lds[i * stride] while i < lds_depth, deeper entries
// in global memory at ovf[(i - lds_depth) * ostride + ocol] (column ocol of a row-major area
// with one column per thread of the grid; the area holds fewer than 2^31 ints). The column is
// obase + (lds - lds0), the lane's offset in its block's LDS stack array, so that no register
// beyond the LDS pointer is kept per lane. OVF = false: the whole stack (maxsp entries) is in
// LDS and ovf is not used.
struct LaneStack {
    int* lds;
    int stride;
    int maxsp;
    int lds_depth;
    int* ovf;
    const int* lds0;
    uint32_t obase, ostride;
};

template <bool OVF>
__device__ __forceinline__ uint32_t lane_ovf_at(const LaneStack& ls, int i) {
    return (uint32_t)(i - ls.lds_depth) * ls.ostride + ls.obase + (uint32_t)(ls.lds - ls.lds0);
}

template <bool OVF>
__device__ __forceinline__ void lane_put(const LaneStack& ls, int i, int v) {
    if (!OVF || i < ls.lds_depth)
        ls.lds[i * ls.stride] = v;
    else
        ls.ovf[lane_ovf_at<OVF>(ls, i)] = v;
}

template <bool OVF>
__device__ __forceinline__ int lane_get(const LaneStack& ls, int i) {
    if (!OVF || i < ls.lds_depth) return ls.lds[i * ls.stride];
    return ls.ovf[lane_ovf_at<OVF>(ls, i)];
}

template <bool CULL = false, bool OVF = false, class Set>
__device__ __forceinline__ void lane_push_children(const RmNode* nodes, int ni, float qx,
                                                   float qy, float qz, const Set& tk,
                                                   const LaneStack& ls, int& sp, int* err,
                                                   const uint8_t* nbits = nullptr) {
    NodeRec n = load_node(nodes, ni);
    const float w = tk.worst_d2();
    const float m0 = mindist(n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, qx, qy, qz);
    const float m1 = mindist(n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, qx, qy, qz);
    bool n0 = (n.r.z >= 0) & (m0 <= w);
    bool n1 = (n.r.w >= 0) & (m1 <= w);
    if (CULL) {
        // a child slot with no point inside the view frustum is not entered
        const uint32_t nb = nbits[ni];
        n0 = n0 & ((nb & 1u) != 0);
        n1 = n1 & ((nb & 2u) != 0);
    }
    if (sp + 2 > ls.maxsp) {
        if (n0 || n1) *err = 1;
        return;
    }
    const int e0 = child_entry(ni, 0, n.r.x, n.r.z), e1 = child_entry(ni, 1, n.r.y, n.r.w);
    if (n0 && n1) {
        const bool near0 = m0 <= m1;
        lane_put<OVF>(ls, sp++, near0 ? e1 : e0);
        lane_put<OVF>(ls, sp++, near0 ? e0 : e1);
    } else if (n0) {
        lane_put<OVF>(ls, sp++, e0);
    } else if (n1) {
        lane_put<OVF>(ls, sp++, e1);
    }
}

template <bool NODUP = false, bool CULL = false, bool OVF = false, class Set>
__device__ void knn_lane_s(const DevTree& t, float qx, float qy, float qz, Set& tk,
                           const LaneStack& ls, int* err, const uint32_t* mask = nullptr) {
    int sp = 0;
    const uint8_t* nbits = CULL ? (const uint8_t*)(mask + cull_mask_words(t.n)) : nullptr;
    lane_push_children<CULL, OVF>(t.nodes, 0, qx, qy, qz, tk, ls, sp, err, nbits);
    while (sp > 0) {
        const int e = lane_get<OVF>(ls, --sp);
        if (!(e & 1)) {
            lane_push_children<CULL, OVF>(t.nodes, e >> 1, qx, qy, qz, tk, ls, sp, err, nbits);
            continue;
        }
        const int ni = e >> 2, s = (e >> 1) & 1;
        NodeRec n = load_node(t.nodes, ni);
        float lx = s ? n.b.z : n.a.x, ly = s ? n.b.w : n.a.y, lz = s ? n.c.x : n.a.z;
        float hx = s ? n.c.y : n.a.w, hy = s ? n.c.z : n.b.x, hz = s ? n.c.w : n.b.y;
        int ref = s ? n.r.y : n.r.x, cnt = s ? n.r.w : n.r.z;
        float md = mindist(lx, ly, lz, hx, hy, hz, qx, qy, qz);
        if (md > tk.worst_d2()) continue;
        const float4* p = t.pts + ref;
        int j = 0;
        for (; j + 4 <= cnt; j += 4) {
            float4 v[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) v[u] = p[j + u];
#pragma unroll
            for (int u = 0; u < 4; ++u)
                if (!CULL || mask_bit(mask, (int64_t)ref + j + u))
                    tk.template add<NODUP>(make_key(dist2(v[u].x, v[u].y, v[u].z, qx, qy, qz),
                                                    __float_as_uint(v[u].w)));
        }
        for (; j < cnt; ++j) {
            float4 v = p[j];
            if (!CULL || mask_bit(mask, (int64_t)ref + j))
                tk.template add<NODUP>(
                    make_key(dist2(v.x, v.y, v.z, qx, qy, qz), __float_as_uint(v.w)));
        }
    }
}

// Whole stack in LDS: column `lane` of a [maxsp][stride] int array.
template <bool NODUP = false, bool CULL = false, class Set>
__device__ __forceinline__ void knn_lane(const DevTree& t, float qx, float qy, float qz, Set& tk,
                                         int* lstk, int stride, int maxsp, int* err,
                                         const uint32_t* mask = nullptr) {
    const LaneStack ls{lstk, stride, maxsp, maxsp, nullptr, lstk, 0u, 0u};
    knn_lane_s<NODUP, CULL, false>(t, qx, qy, qz, tk, ls, err, mask);
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
    knn_lane(t, qq.x, qq.y, qq.z, tk, lstk + threadIdx.x, blockDim.x, maxsp, err);
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
    knn_lane(t, qq.x, qq.y, qq.z, tk, lstk + threadIdx.x, blockDim.x, maxsp, err);
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

// Counting result set (CountingRadiusResultSet with max_neighbors -1, util/cuda/result_set.h:
// 225-267): the pruning bound stays at r2 and every candidate with d2 < r2 is counted, with no
// cap. Takes the place of TopK in knn_lane.
struct CountSet {
    float r2;
    int count;
    __device__ __forceinline__ float worst_d2() const { return r2; }
    template <bool NODUP>
    __device__ __forceinline__ void add(uint64_t kk) {
        count += key_d2(kk) < r2 ? 1 : 0;
    }
};

__global__ __launch_bounds__(256) void radius_count_kernel(DevTree t, const float* __restrict__ q,
                                                           int64_t m, float r2,
                                                           int32_t* __restrict__ counts, int maxsp,
                                                           int* err) {
    extern __shared__ int lstk[];
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    float4 qq = reinterpret_cast<const float4*>(q)[i];
    CountSet cs{r2, 0};
    knn_lane(t, qq.x, qq.y, qq.z, cs, lstk + threadIdx.x, blockDim.x, maxsp, err);
    counts[i] = cs.count;
}

// Radii: one wavefront per 64 consecutive leaf-ordered points (spatially coherent), packet
// traversal with k = 2, then the square root, written in leaf order (coalesced). The original
// order is one random permutation of the leaf order, so a scatter would pay a 32-byte sector
// per 4-byte value; radii_order_kernel instead gathers each original index's value and writes
// both outputs coalesced.
template <int WPB>
__global__ __launch_bounds__(WPB * 64) void radii_kernel(DevTree t, float* __restrict__ r_leaf,
                                                         int* err) {
    __shared__ int wstk[WPB][kWaveStack];
    int w = threadIdx.x >> 6;
    const int64_t g0 = ((int64_t)xcd_block() * WPB + w) * 64;  // first point of the group
    int64_t pos = g0 + lane_id();
    bool active = pos < t.n;
    float4 q = active ? t.pts[pos] : make_float4(0.f, 0.f, 0.f, 0.f);
    TopK<2> tk;
    tk.init(active ? kPadKey : 0ull);
    // Every lane first scores the wavefront's own 64 points, which gives each lane a tight
    // bound before the walk from the root; the walk then skips those leaves. The group is whole
    // leaves when leaf_max divides 64 (all leaves hold leaf_max points except the last one);
    // otherwise there is no pre-scan.
    const int gcnt = (64 % t.leaf_max == 0) ? (int)min((int64_t)64, t.n - g0) : 0;
    // in chunks of kWideLeaf points, each loaded whole by scalar loads (g0 is wave-uniform)
    const int gb = __builtin_amdgcn_readfirstlane((int)g0);
    const int gc = __builtin_amdgcn_readfirstlane(gcnt);
    for (int c = 0; c < gc; c += kWideLeaf)
        packet_scan_leaf<2, false, false, true>((CFloat*)t.pts, gb + c, min(kWideLeaf, gc - c),
                                                q.x, q.y, q.z, tk);
    knn_packet<2>(t, q.x, q.y, q.z, active, tk, wstk[w], err, (int)g0, (int)(g0 + gcnt));
    if (active) r_leaf[pos] = sqrtf(key_d2(tk.key[1]));
}

// Original order: r_unc[o] = r_leaf[inv[o]], r_cl[o] = the same clamped to max_radius (when
// max_radius > 0). Reads of inv and writes are coalesced; the gathers hit a table of 4 n bytes.
__global__ __launch_bounds__(256) void radii_order_kernel(const int32_t* __restrict__ inv,
                                                          const float* __restrict__ r_leaf,
                                                          int64_t n, float max_radius,
                                                          float* __restrict__ r_unc,
                                                          float* __restrict__ r_cl) {
    const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= n) return;
    const float r = r_leaf[inv[o]];
    if (r_unc) r_unc[o] = r;
    if (r_cl) r_cl[o] = (max_radius > 0.0f && r > max_radius) ? max_radius : r;
}

__global__ void tree_inverse_kernel(const float4* __restrict__ pts, int64_t n,
                                    int32_t* __restrict__ inv) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) inv[__float_as_uint(pts[i].w)] = (int32_t)i;
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

// Distance along a ray past which no sphere can be hit. The box of all centres (the root's two
// child boxes) is dilated by the largest radius, twice MIN_DIST and a float-error margin
// (10^-5 of the coordinates' scale, about 80 times the rounding of the march point o + d t and of
// the distance arithmetic); past the ray's exit from that convex box every centre is farther from
// the march point than its radius + MIN_DIST, so every SDF value there exceeds MIN_DIST and the
// march can only end as a miss, as it would at the far plane. -INF when the ray misses the box,
// +INF when the slab arithmetic is indeterminate.
__device__ __forceinline__ float ray_exit_bound(const DevCloud& c, const Ray& r) {
    CFloat* f = (CFloat*)c.tree.nodes;
    CInt* fi = (CInt*)f;
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
#pragma unroll
    for (int s = 0; s < 2; ++s)
        if (fi[14 + s] >= 0)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], f[6 * s + a]);
                hi[a] = fmaxf(hi[a], f[6 * s + 3 + a]);
            }
    float ext = 0.0f;
#pragma unroll
    for (int a = 0; a < 3; ++a) ext = fmaxf(ext, fmaxf(fabsf(lo[a]), fabsf(hi[a])));
    const float scale = fmaxf(fmaxf(fabsf(r.ox), fabsf(r.oy)), fabsf(r.oz)) + ext + 1000.0f;
    const float e = __uint_as_float(*(const __attribute__((address_space(4))) uint32_t*)c.rmax) +
                    2.0f * kMinDist + 1e-5f * scale;
    const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    float tin = -kInf, tout = kInf;
    bool nan = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float inv = 1.0f / d[a];
        const float t1 = (lo[a] - e - o[a]) * inv, t2 = (hi[a] + e - o[a]) * inv;
        nan |= (t1 != t1) | (t2 != t2);
        tin = fmaxf(tin, fminf(t1, t2));
        tout = fminf(tout, fmaxf(t1, t2));
    }
    if (nan) return kInf;
    if (tin > tout) return -kInf;
    return tout + fabsf(tout) * 1e-5f + 1e-3f;
}

// Early miss over a frontier of the tree: the union of the tight boxes of the tree's child slots
// at depth D (a leaf met above that depth stands for itself). Every centre lies in one of them,
// so past the ray's exit from the last of those boxes (each dilated as in ray_exit_bound) that
// it crosses, it is outside every dilated box for good (a ray leaves a convex box once) and no
// sphere can be hit any more; -INF when it crosses none. The frontier is an array of
// kFrontierFloats-float entries (lo.xyz, hi.xyz, valid, unused) after one header entry whose
// first float counts entries with a non-finite coordinate (then the root box is used instead).
// Rays with a direction component below 1e-20 in magnitude also use the root box, so that the
// slab arithmetic below stays finite.
__device__ __forceinline__ float frontier_exit(const DevCloud& c, const Ray& r, CFloat* fr,
                                               int nfr);

// Whether a camera at (x, y, z) is outside the box of all centres (the root's two child boxes).
// The batched march uses the frontier for such views only: from inside the cloud's box (a room
// seen from within) nearly every ray meets a surface and the frontier's cost is not repaid.
__device__ __forceinline__ bool camera_outside_cloud(const DevCloud& c, float x, float y,
                                                     float z) {
    CFloat* f = (CFloat*)c.tree.nodes;
    CInt* fi = (CInt*)f;
    float lo[3] = {kInf, kInf, kInf}, hi[3] = {-kInf, -kInf, -kInf};
#pragma unroll
    for (int s = 0; s < 2; ++s)
        if (fi[14 + s] >= 0)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], f[6 * s + a]);
                hi[a] = fmaxf(hi[a], f[6 * s + 3 + a]);
            }
    return !(x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1] && z >= lo[2] && z <= hi[2]);
}

__device__ __forceinline__ float frontier_exit(const DevCloud& c, const Ray& r, CFloat* fr,
                                               int nfr) {
    if (fr[0] != 0.0f) return ray_exit_bound(c, r);
    const float o[3] = {r.ox, r.oy, r.oz}, d[3] = {r.dx, r.dy, r.dz};
    if (!(fabsf(d[0]) >= 1e-20f && fabsf(d[1]) >= 1e-20f && fabsf(d[2]) >= 1e-20f))
        return ray_exit_bound(c, r);
    CFloat* f = (CFloat*)c.tree.nodes;
    CInt* fi = (CInt*)f;
    float ext = 0.0f;
#pragma unroll
    for (int s = 0; s < 2; ++s)
        if (fi[14 + s] >= 0)
#pragma unroll
            for (int a = 0; a < 6; ++a) ext = fmaxf(ext, fabsf(f[6 * s + a]));
    const float scale = fmaxf(fmaxf(fabsf(r.ox), fabsf(r.oy)), fabsf(r.oz)) + ext + 1000.0f;
    const float e = __uint_as_float(*(const __attribute__((address_space(4))) uint32_t*)c.rmax) +
                    2.0f * kMinDist + 1e-5f * scale;
    float inv[3], c1[3], c2[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        inv[a] = 1.0f / d[a];
        c1[a] = (-e - o[a]) * inv[a];
        c2[a] = (e - o[a]) * inv[a];
    }
    float best = -kInf;
    for (int i = 1; i <= nfr; ++i) {
        CFloat* b = fr + (int64_t)i * kFrontierFloats;
        if (b[6] == 0.0f) continue;
        float tin = -kInf, tout = kInf;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float t1 = __builtin_fmaf(b[a], inv[a], c1[a]);
            const float t2 = __builtin_fmaf(b[3 + a], inv[a], c2[a]);
            tin = fmaxf(tin, fminf(t1, t2));
            tout = fminf(tout, fmaxf(t1, t2));
        }
        if (tin <= tout) best = fmaxf(best, tout);
    }
    if (best == -kInf) return -kInf;
    return best + fabsf(best) * 1e-5f + 1e-3f;
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
            const float d2 = key_d2(tk.key[j]);
            const float r = c.xyzr[id].w;
            // The hardware square root is within 1 ulp; scaled by (1 - 2^-20) it is a lower bound
            // of sqrtf(d2) (0 for flushed tiny inputs is one too). Subtraction rounds
            // monotonically, so lb - r >= best implies sqrtf(d2) - r >= best: the entry cannot
            // become the strict minimum and its correctly rounded square root is skipped.
            const float lb = __builtin_amdgcn_sqrtf(d2) * (1.0f - 0x1p-20f);
            if (lb - r < best) {
                const float s = sqrtf(d2) - r;
                if (s < best) {
                    best = s;
                    bi = id;
                }
            }
        }
    }
    s_out = best;
    i_out = bi;
}

// Warm start: the previous step's 10 neighbours, re-measured from the new query point and
// sorted, seed the list. They are 10 valid candidates, so the list's 10th key bounds the new 10th
// key from above and pruning is tight from the first node; the traversal then only inserts
// points not already listed (insert_nodup), and its result is the exact top 10. Needs n >= 10;
// with culling fewer than 10 points may be visible, and padding entries stay padding.
template <bool CULL = false>
__device__ __forceinline__ void warm_seed(const DevCloud& c, TopK<kMarchK>& tk, float qx,
                                          float qy, float qz) {
    const float* xyzr = (const float*)c.xyzr;
    uint32_t ids[kMarchK];
#pragma unroll
    for (int j = 0; j < kMarchK; ++j) ids[j] = key_idx(tk.key[j]);
    // Two batches of five gathers. With all ten in flight at once, the 30 coordinates and the
    // list do not fit the register budget of 8 wavefronts per SIMD and the list is spilled;
    // the empty asm makes the second batch's indices depend on the first batch's keys, so the
    // second batch is issued only once the first is consumed.
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        if (h == 1)
            asm volatile("" : "+v"(ids[5]), "+v"(ids[6]), "+v"(ids[7]), "+v"(ids[8]), "+v"(ids[9])
                         : "v"(tk.key[0]), "v"(tk.key[1]), "v"(tk.key[2]), "v"(tk.key[3]),
                           "v"(tk.key[4]));
#pragma unroll
        for (int j = 5 * h; j < 5 * h + 5; ++j) {
            const uint32_t id = ids[j];
            if (CULL && id >= (uint32_t)c.tree.n) {
                tk.key[j] = kPadKey;
                continue;
            }
            const float* q = xyzr + 4 * (size_t)id;   // x, y, z only: the radius is not needed
            tk.key[j] = make_key(dist2(q[0], q[1], q[2], qx, qy, qz), id);
        }
    }
    // the sort starts once every key is formed (its first exchanges would otherwise overlap
    // the second batch of gathers and raise the register peak again)
    asm volatile("" : "+v"(tk.key[0]), "+v"(tk.key[1]), "+v"(tk.key[2]), "+v"(tk.key[3]),
                 "+v"(tk.key[4]), "+v"(tk.key[5]), "+v"(tk.key[6]), "+v"(tk.key[7]),
                 "+v"(tk.key[8]), "+v"(tk.key[9]));
    static_assert(kMarchK == 10, "the seed sort sorts ten keys");
    sort10_nearly_sorted(tk.key);
}

template <bool WARM, bool CULL>
__device__ void march_lane(const DevCloud& c, const Ray& r, const uint32_t* mask, int* lstk,
                           int stride, int maxsp, int* err, int& hit, float& depth) {
    TopK<kMarchK> tk;
    tk.init(kPadKey);
    float t = r.tnear;
    hit = -1;
    depth = 0.0f;
    for (int i = 0; i < kMaxSteps; ++i) {
        float px = r.ox + r.dx * t;
        float py = r.oy + r.dy * t;
        float pz = r.oz + r.dz * t;
        if (WARM && i > 0) {
            warm_seed<CULL>(c, tk, px, py, pz);
            knn_lane<true, CULL>(c.tree, px, py, pz, tk, lstk, stride, maxsp, err, mask);
        } else {
            tk.init(kPadKey);
            knn_lane<false, CULL>(c.tree, px, py, pz, tk, lstk, stride, maxsp, err, mask);
        }
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
// pixel each), per-lane traversal restarted from the root at every step. With CULL the kNN
// runs over the points of the view's culling mask only.
template <bool CULL>
__global__ __launch_bounds__(256) void render_plain_kernel(DevCloud c, DevView v, MarchOut o,
                                                           int maxsp, int* err,
                                                           const uint32_t* mask) {
    extern __shared__ int lstk[];
    int x = blockIdx.x * 16 + threadIdx.x;
    int y = blockIdx.y * 16 + threadIdx.y;
    int tid = threadIdx.y * 16 + threadIdx.x;
    if (x >= v.W || y >= v.H) return;
    Ray r = make_ray([&](int i) { return v.M[i]; }, v.pane, v.W, v.H, x, y);
    r.tfar = fminf(r.tfar, ray_exit_bound(c, r));
    int hit;
    float depth;
    march_lane<false, CULL>(c, r, mask, &lstk[tid], 256, maxsp, err, hit, depth);
    write_pixel(c, o, v.pix_offset + (int64_t)y * v.W + x, hit, depth);
}

// Batched path: a persistent grid over all views. A wavefront claims one screen tile
// (tile_w x tile_h pixels, a multiple of 64) from a shared 64-bit queue and keeps 64 rays of that
// tile in flight: a lane whose ray ends takes the tile's next pixel (claimed for every idle lane
// at once by ballot + prefix count, once at least `refill_min` lanes are idle), so rays in a
// wavefront stay screen neighbours and the tail of long rays is paid once per tile, not once per
// 64 pixels. One wavefront step advances every live ray by one march step (one kNN-10 query
// each), in packet (MODE 0) or per-lane (MODE 1) traversal.
//
// Occupancy: each traversal step waits on scalar loads of node records and points and on the
// LDS stack, so the vector ALUs need many resident wavefronts to stay busy. The kernel is
// compiled for 8 wavefronts per SIMD (64 VGPRs). The per-ray state that is not needed inside the
// traversal (direction, distance marched, far limit, step count, pixel) lives in LDS between the
// points where it is used, so the traversal keeps the registers and nothing goes to scratch.
template <int WPB>
struct RayStateLds {
    float dx[WPB][64], dy[WPB][64], dz[WPB][64], t[WPB][64], tfar[WPB][64];
    int steps[WPB][64], pix[WPB][64];
};

// Compiler barrier: values read from LDS after it are read again (not kept in registers from
// an earlier read or store), which keeps the per-ray state out of the registers across the
// traversal.
__device__ __forceinline__ void reread_fence() { asm volatile("" ::: "memory"); }

// v unchanged, but opaque to the compiler: what is computed from it stays where it is written
// instead of being hoisted out of the loop (and held in a register across the traversal).
__device__ __forceinline__ int opaque_s(int v) {
    asm volatile("" : "+s"(v));
    return v;
}
// kPadKey, formed where it is used rather than kept in a register pair across the march
__device__ __forceinline__ uint64_t pad_key_here() {
    uint32_t lo = 0xffffffffu, hi = 0x7f800000u;
    asm volatile("" : "+v"(lo), "+v"(hi));
    return ((uint64_t)hi << 32) | lo;
}
template <class T>
__device__ __forceinline__ T* opaque_ptr(T* p) {
    asm volatile("" : "+s"(p));
    return p;
}

template <int MODE, bool WARM, int WPB, bool CULL, bool OVF = false>
__global__ __launch_bounds__(WPB * 64) __attribute__((amdgpu_waves_per_eu(8))) void batch_kernel(
    BatchParams p, int maxsp) {
    __shared__ int wstk[WPB][kWaveStack];
    __shared__ RayStateLds<WPB> rs;
    extern __shared__ int lstk[];
    const int w = threadIdx.x >> 6;
    const int lane = lane_id();
    CFloat* vf = (CFloat*)p.views;
    constexpr int VSTRIDE = sizeof(DevView) / 4;
    const int tile_pix = p.tile_w * p.tile_h;
    while (true) {
        unsigned long long qv = 0;
        if (lane == 0) qv = atomicAdd(p.queue, 1ull);
        qv = ((unsigned long long)__builtin_amdgcn_readfirstlane((uint32_t)(qv >> 32)) << 32) |
             (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)qv);
        // the host keeps total_tiles below 2^32, so tiles are compared in 32 bits (the scalar
        // unit has no ordered 64-bit compare)
        if ((uint32_t)(qv >> 32) != 0u || (uint32_t)qv >= (uint32_t)p.total_tiles) break;
        const int64_t tile = (int64_t)(uint32_t)qv;
        // view lookup: last view whose tile_offset <= tile (scalar loads)
        int lo = 0, hi = p.nviews - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            const uint32_t off = *(const __attribute__((address_space(4))) uint32_t*)(
                vf + (int64_t)mid * VSTRIDE + offsetof(DevView, tile_offset) / 4);
            if (off <= (uint32_t)tile)
                lo = mid;
            else
                hi = mid - 1;
        }
        CFloat* V = vf + (int64_t)lo * VSTRIDE;
        CInt* VI = (CInt*)V;
        const int W = VI[17], H = VI[18], tiles_x = VI[19];
        const int64_t pix_off =
            *(const __attribute__((address_space(4))) int64_t*)(V + offsetof(DevView, pix_offset) / 4);
        const int64_t tile_off =
            *(const __attribute__((address_space(4))) int64_t*)(V + offsetof(DevView, tile_offset) / 4);
        const int64_t lt = tile - tile_off;
        const int x0 = (int)(lt % tiles_x) * p.tile_w;
        const int y0 = (int)(lt / tiles_x) * p.tile_h;
        // camera origin: the same for every lane of the tile (scalar registers)
        const float ox = V[12], oy = V[13], oz = V[14];
        const uint32_t* vmask = CULL ? p.cull_mask + (int64_t)lo * p.cull_words : nullptr;
        int next = 0;                      // wave-uniform: next unclaimed pixel of the tile
        bool has = false;
        TopK<kMarchK> tk;
        tk.init(kPadKey);
        while (true) {
            const uint64_t needm = __builtin_amdgcn_ballot_w64(!has);
            const int nneed = __popcll(needm);
            if (next < tile_pix && (nneed >= p.refill_min || nneed == 64)) {
                // idle lanes below this one (mbcnt counts the mask's bits under the lane)
                const int my = next + (int)__builtin_amdgcn_mbcnt_hi(
                                          (uint32_t)(needm >> 32),
                                          __builtin_amdgcn_mbcnt_lo((uint32_t)needm, 0u));
                if (!has && my < tile_pix) {
                    // the tile and view values are made opaque here so that what the refill
                    // derives from them is computed here, not kept live across the march
                    const int tw = opaque_s(p.tile_w), Wr = opaque_s(W), Hr = opaque_s(H);
                    const int x = x0 + my % tw, y = y0 + my / tw;
                    if (x < Wr && y < Hr) {
                        CFloat* Vr = opaque_ptr(V);
                        DevCloud cl = p.cloud;
                        cl.tree.nodes = opaque_ptr(cl.tree.nodes);
                        cl.rmax = opaque_ptr(cl.rmax);
                        const Ray r = make_ray([&](int i) { return Vr[i]; }, Vr[16], Wr, Hr, x, y);
                        rs.dx[w][lane] = r.dx;
                        rs.dy[w][lane] = r.dy;
                        rs.dz[w][lane] = r.dz;
                        rs.t[w][lane] = r.tnear;
                        const bool use_frontier =
                            p.nfrontier > 0 && camera_outside_cloud(cl, Vr[12], Vr[13], Vr[14]);
                        rs.tfar[w][lane] = fminf(
                            r.tfar, use_frontier
                                        ? frontier_exit(cl, r, (CFloat*)opaque_ptr(p.frontier),
                                                        opaque_s(p.nfrontier))
                                        : ray_exit_bound(cl, r));
                        rs.steps[w][lane] = 0;
                        rs.pix[w][lane] = y * Wr + x;
                        has = true;
                    }
                }
                next += nneed;
            }
            if (__builtin_amdgcn_ballot_w64(has) == 0) {
                if (next >= tile_pix) break;
                continue;
            }
            reread_fence();
            float px = 0.f, py = 0.f, pz = 0.f;
            bool warm_lane = false;
            if (has) {
                const float t = rs.t[w][lane];
                px = ox + rs.dx[w][lane] * t;
                py = oy + rs.dy[w][lane] * t;
                pz = oz + rs.dz[w][lane] * t;
                warm_lane = WARM && rs.steps[w][lane] > 0;
            }
            if (!has)
                tk.init(0ull);
            else if (warm_lane)
                warm_seed<CULL>(p.cloud, tk, px, py, pz);
            else
                tk.init(pad_key_here());
            if (MODE == 0) {
                // a fresh lane's pads (index 0xffffffff) never match a real candidate, so the
                // duplicate check is harmless for it and one traversal serves both kinds
                knn_packet_rounds<kMarchK, WARM, CULL>(p.cloud.tree, px, py, pz, has, tk, wstk[w],
                                                       p.err, (CUint*)vmask);
            } else if (has) {
                // the first lane_lds entries of the lane's stack in LDS, deeper ones in its
                // column of the overflow area (never reached when lane_lds == maxsp)
                const LaneStack ls{&lstk[threadIdx.x], WPB * 64, maxsp, p.lane_lds, p.lane_ovf,
                                   lstk, blockIdx.x * (WPB * 64), gridDim.x * (WPB * 64)};
                if (warm_lane)
                    knn_lane_s<true, CULL, OVF>(p.cloud.tree, px, py, pz, tk, ls, p.err, vmask);
                else
                    knn_lane_s<false, CULL, OVF>(p.cloud.tree, px, py, pz, tk, ls, p.err, vmask);
            }
            reread_fence();
            if (has) {
                float s;
                uint32_t id;
                sdf10(p.cloud, tk, s, id);
                const float t = rs.t[w][lane] + s;
                const int steps = rs.steps[w][lane] + 1;
                const bool hitnow = s < kMinDist;
                if (hitnow || t > rs.tfar[w][lane] || steps >= kMaxSteps) {
                    write_pixel(p.cloud, p.out, pix_off + rs.pix[w][lane], hitnow ? (int)id : -1,
                                hitnow ? t : 0.0f);
                    has = false;
                } else {
                    rs.t[w][lane] = t;
                    rs.steps[w][lane] = steps;
                }
            }
        }
    }
}

// Frustum predicate is_inside of src/ray_marching.cu:231-258 (commented out there, so the
// reference never culls): the point is inside the four side planes through the camera origin
// (o), beyond the near plane and before the far plane. F(i) yields the view's frustum floats: side normals 0..11, forward
// 12..14. Float, no contraction, sums left to right as written.
template <class FR>
__device__ __forceinline__ bool frustum_inside(FR F, float ox, float oy, float oz, float x,
                                               float y, float z) {
    const float px = x - ox, py = y - oy, pz = z - oz;
    bool in = true;
#pragma unroll
    for (int i = 0; i < 4; ++i)
        in &= ((F(3 * i) * px + F(3 * i + 1) * py) + F(3 * i + 2) * pz) < 0.0f;
    const float fx = F(12), fy = F(13), fz = F(14);
    const float ax = px - fx * kZNear, ay = py - fy * kZNear, az = pz - fz * kZNear;
    in &= ((ax * fx + ay * fy) + az * fz) > 0.0f;
    const float bx = ax - fx * (kZFar - kZNear), by = ay - fy * (kZFar - kZNear),
                bz = az - fz * (kZFar - kZNear);
    in &= ((bx * -fx + by * -fy) + bz * -fz) > 0.0f;
    return in;
}

// Point masks of nviews views (MANY: the device view table; otherwise the one view `one`):
// one wavefront per 64 consecutive leaf-ordered points, the predicate of every lane's point
// packed by ballot into two mask words per view. View v's data starts at mask + v * stride.
template <bool MANY>
__global__ __launch_bounds__(256) void cull_mask_kernel(DevTree t, DevView one,
                                                        const DevView* views, int nviews,
                                                        uint32_t* __restrict__ mask,
                                                        int64_t stride) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t base = i - lane_id();
    if (base >= t.n) return;
    const bool valid = i < t.n;
    const float4 q = valid ? t.pts[i] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (int v = 0; v < nviews; ++v) {
        bool in = false;
        if (MANY) {
            CFloat* V = (CFloat*)(views + v);
            CFloat* F = V + offsetof(DevView, fn) / 4;
            in = frustum_inside([&](int k) { return F[k]; }, V[12], V[13], V[14], q.x, q.y, q.z);
        } else {
            in = frustum_inside([&](int k) { return k < 12 ? one.fn[k] : one.fdir[k - 12]; },
                                one.M[12], one.M[13], one.M[14], q.x, q.y, q.z);
        }
        const uint64_t b = __builtin_amdgcn_ballot_w64(valid && in);
        if (lane_id() < 2) {
            const uint32_t w = lane_id() == 0 ? (uint32_t)b : (uint32_t)(b >> 32);
            mask[(int64_t)v * stride + (base >> 5) + lane_id()] = w;
        }
    }
}

// Node bits of one level of inner nodes for every view (blockIdx.y = view), bottom-up: a leaf
// child's bit is the OR of its points' mask bits, an inner child's the OR of its own two bits
// (written by the launch for the level below).
__global__ __launch_bounds__(256) void cull_nodes_kernel(const RmNode* __restrict__ nodes,
                                                         const int32_t* __restrict__ ids,
                                                         int count, int64_t n,
                                                         uint32_t* __restrict__ cull,
                                                         int64_t stride) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const uint32_t* pm = cull + (int64_t)blockIdx.y * stride;
    uint8_t* nb = (uint8_t*)(cull + (int64_t)blockIdx.y * stride + cull_mask_words(n));
    const int id = ids[i];
    const RmNode& nd = nodes[id];
    uint32_t bits = 0;
    for (int s = 0; s < 2; ++s) {
        const int c = nd.cnt[s];
        if (c < 0) continue;
        bool any = false;
        if (c > 0) {
            const int64_t b = nd.ref[s], e = b + c;
            for (int64_t w = b >> 5; w <= (e - 1) >> 5; ++w) {
                const int64_t lo = (b > w * 32 ? b : w * 32) - w * 32;
                const int64_t hi = (e < w * 32 + 32 ? e : w * 32 + 32) - w * 32;
                const uint32_t sel =
                    hi - lo == 32 ? ~0u : (((1u << (uint32_t)(hi - lo)) - 1u) << (uint32_t)lo);
                any |= (pm[w] & sel) != 0u;
            }
        } else {
            any = nb[nd.ref[s]] != 0;
        }
        bits |= any ? (1u << s) : 0u;
    }
    nb[id] = (uint8_t)bits;
}

__global__ void reset_queue_kernel(unsigned long long* q) { *q = 0ull; }

// Frontier of the early miss (see frontier_exit): entry 1 + f holds the box of the child slot
// reached from the root by the depth bits of f (most significant first); a leaf met above depth
// `depth` is written by the first entry of its range, the others and empty slots are invalid.
// One block of 2^depth threads (depth <= 10); thread 0 writes the header.
__global__ __launch_bounds__(1024) void frontier_kernel(const RmNode* __restrict__ nodes,
                                                        int depth, float* __restrict__ out) {
    __shared__ int nonfinite;
    const int f = threadIdx.x;
    if (f == 0) nonfinite = 0;
    __syncthreads();
    float box[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    bool valid = false;
    int ni = 0;
    for (int level = 0; level < depth; ++level) {
        const int slot = (f >> (depth - 1 - level)) & 1;
        const int rem = depth - 1 - level;
        const int cnt = nodes[ni].cnt[slot];
        if (cnt < 0 || (cnt > 0 && rem > 0 && (f & ((1 << rem) - 1)) != 0)) break;
        if (cnt > 0 || rem == 0) {
#pragma unroll
            for (int a = 0; a < 6; ++a) box[a] = nodes[ni].box[slot][a];
            valid = true;
            break;
        }
        ni = nodes[ni].ref[slot];
    }
    if (valid) {
        bool fin = true;
#pragma unroll
        for (int a = 0; a < 6; ++a) fin &= fabsf(box[a]) <= 3.0e38f;
        if (!fin) atomicAdd(&nonfinite, 1);
    }
    float* o = out + (int64_t)(1 + f) * kFrontierFloats;
#pragma unroll
    for (int a = 0; a < 6; ++a) o[a] = box[a];
    o[6] = valid ? 1.0f : 0.0f;
    o[7] = 0.0f;
    __syncthreads();
    if (f == 0) {
        out[0] = (float)nonfinite;
#pragma unroll
        for (int a = 1; a < kFrontierFloats; ++a) out[a] = 0.0f;
    }
}

__global__ void store_views_kernel(ViewChunk ch, DevView* __restrict__ table) {
    const int i = threadIdx.x;
    if (i < ch.count) table[ch.first + i] = ch.v[i];
}

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

__global__ void zero_u32_kernel(uint32_t* p) { *p = 0u; }

// Largest radius (xyzr.w) of a cloud, at least 0, merged into *out by an integer atomic
// maximum (non-negative floats order like their bit patterns; NaN radii are ignored).
__global__ __launch_bounds__(256) void max_radius_kernel(const float4* __restrict__ xyzr,
                                                         int64_t n, uint32_t* out) {
    float m = 0.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * blockDim.x)
        m = fmaxf(m, xyzr[i].w);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o));
    if (lane_id() == 0) atomicMax(out, __float_as_uint(m));
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

// Adjoint over every view of a batch at once. The pixels of all views are numbered view after
// view, j in [0, P); cum[v] is the number of the first pixel of view v (cum[nviews] = P), and
// pixel j lives at views[v].pix_offset + (j - cum[v]) in the caller's concatenated arrays.
__device__ __forceinline__ int view_of(const int64_t* __restrict__ cum, int nviews, int64_t j) {
    int lo = 0, hi = nviews - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cum[mid] <= j)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void adjoint_batch_keys_kernel(const DevView* __restrict__ views,
                                          const int64_t* __restrict__ cum, int nviews, int64_t P,
                                          const int32_t* __restrict__ hit, int32_t miss,
                                          int32_t* __restrict__ keys, int32_t* __restrict__ vals) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= P) return;
    const int v = view_of(cum, nviews, j);
    const int h = hit[views[v].pix_offset + (j - cum[v])];
    keys[j] = h >= 0 ? h : miss;
    vals[j] = (int32_t)j;
}

// One thread per run of equal keys (one hit point): it sums the gradients of the run's pixels,
// in pixel-number order, over all views, and adds them to the point. xyz (n x 3) and rad (n)
// give the centres and radii the forward marched; null takes the cloud's current ones.
__global__ void adjoint_batch_reduce_kernel(DevCloud c, const float* __restrict__ xyz,
                                            const float* __restrict__ rad,
                                            const DevView* __restrict__ views,
                                            const int64_t* __restrict__ cum, int nviews,
                                            const int32_t* __restrict__ keys,
                                            const int32_t* __restrict__ vals, int64_t P,
                                            const float* __restrict__ grad_rgb,
                                            const float* __restrict__ grad_depth, float* g_xyz,
                                            float* g_r, float* g_rgb) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= P) return;
    const int key = keys[i];
    if (key >= c.tree.n) return;
    if (i > 0 && keys[i - 1] == key) return;
    float4 cr = c.xyzr[key];
    if (xyz) cr = make_float4(xyz[3 * (int64_t)key], xyz[3 * (int64_t)key + 1],
                              xyz[3 * (int64_t)key + 2], cr.w);
    if (rad) cr.w = rad[key];
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, ar = 0.f, c0 = 0.f, c1 = 0.f, c2 = 0.f;
    for (int64_t k = i; k < P && keys[k] == key; ++k) {
        const int64_t j = vals[k];
        const int v = view_of(cum, nviews, j);
        const DevView& dv = views[v];
        const int64_t local = j - cum[v];
        const int64_t pix = dv.pix_offset + local;
        c0 += grad_rgb[3 * pix];
        c1 += grad_rgb[3 * pix + 1];
        c2 += grad_rgb[3 * pix + 2];
        const float gd = grad_depth[pix];
        if (gd == 0.0f) continue;
        const int x = (int)(local % dv.W), y = (int)(local / dv.W);
        const Ray r = make_ray([&](int u) { return dv.M[u]; }, dv.pane, dv.W, dv.H, x, y);
        const SphereHit sh = sphere_entry(r, cr);
        if (!sh.ok) continue;
        // dt/dc = d - h / sqrt(disc), dt/dr = -r / sqrt(disc)
        const float inv = 1.0f / sh.sq;
        a0 += gd * (r.dx - sh.hx * inv);
        a1 += gd * (r.dy - sh.hy * inv);
        a2 += gd * (r.dz - sh.hz * inv);
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

// ------------------------------------------------------------------------------------------
// Device tree build. The shape (segments, left counts, child refs) comes from the host; the
// device keeps three lists of point ids, each sorted by one coordinate with ties by id, and per
// level: picks each node's split axis from its tight box (the ends of the three lists), marks
// the first `nl` points of the axis list as left, and stable-partitions all three lists, which
// keeps every child segment sorted along every axis (so child boxes are again list ends).
// ------------------------------------------------------------------------------------------
__device__ __forceinline__ uint32_t ordered_key(float f) {
    uint32_t u = __float_as_uint(f);
    if (u == 0x80000000u) u = 0u;  // -0 sorts with +0
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ __forceinline__ float coord(const float4* p, int32_t id, int d) {
    const float4 v = p[id];
    return d == 0 ? v.x : (d == 1 ? v.y : v.z);
}

__global__ void build_keys_kernel(const float4* __restrict__ p, int64_t n, int d,
                                  uint32_t* __restrict__ keys, int32_t* __restrict__ ids) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    keys[i] = ordered_key(coord(p, (int32_t)i, d));
    ids[i] = (int32_t)i;
}

struct BuildLists {
    const int32_t* l[3];
};

__global__ void build_axis_kernel(const int64_t* __restrict__ seg_b,
                                  const int64_t* __restrict__ seg_e, int node0, int count,
                                  const float4* __restrict__ p, BuildLists L,
                                  int8_t* __restrict__ axis) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int node = node0 + i;
    const int64_t b = seg_b[node], e = seg_e[node];
    float ext[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) ext[d] = coord(p, L.l[d][e - 1], d) - coord(p, L.l[d][b], d);
    int dim = 0;
    if (ext[1] > ext[dim]) dim = 1;
    if (ext[2] > ext[dim]) dim = 2;
    axis[node] = (int8_t)dim;
}

__global__ void build_flag_kernel(int64_t n, const int32_t* __restrict__ owner,
                                  const int64_t* __restrict__ seg_b,
                                  const int32_t* __restrict__ nl, const int8_t* __restrict__ axis,
                                  BuildLists L, int32_t* __restrict__ flag) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int node = owner[i];
    if (node < 0) return;
    const int32_t* list = L.l[axis[node]];
    flag[list[i]] = (i - seg_b[node] < nl[node]) ? 1 : 0;
}

__global__ void build_isleft_kernel(int64_t n, const int32_t* __restrict__ owner,
                                    const int32_t* __restrict__ list,
                                    const int32_t* __restrict__ flag, int32_t* __restrict__ il) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    il[i] = owner[i] >= 0 ? flag[list[i]] : 0;
}

__global__ void build_scatter_kernel(int64_t n, const int32_t* __restrict__ owner,
                                     const int64_t* __restrict__ seg_b,
                                     const int32_t* __restrict__ nl,
                                     const int32_t* __restrict__ list,
                                     const int32_t* __restrict__ il,
                                     const int32_t* __restrict__ pl, int32_t* __restrict__ out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int node = owner[i];
    if (node < 0) {
        out[i] = list[i];
        return;
    }
    const int64_t b = seg_b[node];
    const int64_t lc = (int64_t)pl[i] - pl[b];
    const int64_t dst = il[i] ? b + lc : b + nl[node] + (i - b - lc);
    out[dst] = list[i];
}

__global__ void build_box_kernel(int node0, int count, const int64_t* __restrict__ seg_b,
                                 const int64_t* __restrict__ seg_e,
                                 const int32_t* __restrict__ nl, const float4* __restrict__ p,
                                 BuildLists L, RmNode* __restrict__ nodes) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int node = node0 + i;
    const int64_t b = seg_b[node], e = seg_e[node], m = b + nl[node];
    const int64_t cb[2] = {b, m}, ce[2] = {m, e};
    RmNode& nd = nodes[node];
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            nd.box[s][d] = coord(p, L.l[d][cb[s]], d);
            nd.box[s][3 + d] = coord(p, L.l[d][ce[s] - 1], d);
        }
}

__global__ void build_leafroot_kernel(int64_t n, const float4* __restrict__ p, BuildLists L,
                                      RmNode* __restrict__ nodes) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int d = 0; d < 3; ++d) {
        nodes[0].box[0][d] = coord(p, L.l[d][0], d);
        nodes[0].box[0][3 + d] = coord(p, L.l[d][n - 1], d);
    }
}

__global__ void build_owner_kernel(int64_t n, int32_t* __restrict__ owner,
                                   const int64_t* __restrict__ seg_b,
                                   const int32_t* __restrict__ nl,
                                   const RmNode* __restrict__ nodes) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int node = owner[i];
    if (node < 0) return;
    const int s = (i - seg_b[node] < nl[node]) ? 0 : 1;
    owner[i] = nodes[node].cnt[s] == 0 ? nodes[node].ref[s] : -1;
}

__global__ void build_fill_kernel(int64_t n, int32_t v, int32_t* __restrict__ a) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

__global__ void build_pts_kernel(int64_t n, const float4* __restrict__ p,
                                 const int32_t* __restrict__ list, float4* __restrict__ pts) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t id = list[i];
    const float4 v = p[id];
    pts[i] = make_float4(v.x, v.y, v.z, __uint_as_float((uint32_t)id));
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

static int lane_stack_depth(const DevTree& t) { return lane_stack_need(t.depth); }

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
    if (max_nn == 0) {
        hipLaunchKernelGGL(radius_count_kernel, dim3(grid_for(m, 256)), dim3(256), shmem, s, t, q,
                           m, r2, counts, maxsp, err);
        return hipGetLastError();
    }
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
                        float* r_cl, float* tmp, int* err) {
    // one wavefront per block: measured 5 % faster than 2 or 4 (finer scheduling, XCD-contiguous
    // runs of single groups), same results
    constexpr int WPB = 1;
    int64_t waves = (t.n + 63) / 64;
    hipLaunchKernelGGL(radii_kernel<WPB>, dim3(grid_for(waves, WPB)), dim3(WPB * 64), 0, s, t,
                       tmp, err);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(radii_order_kernel, dim3(grid_for(t.n, 256)), dim3(256), 0, s, t.inv, tmp,
                       t.n, max_radius, r_unc, r_cl);
    return hipGetLastError();
}

hipError_t launch_tree_inverse(hipStream_t s, const float4* pts, int64_t n, int32_t* inv) {
    hipLaunchKernelGGL(tree_inverse_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s, pts, n, inv);
    return hipGetLastError();
}

hipError_t launch_make_cloud(hipStream_t s, const float* xyzw, const float* radii,
                             const float* colors, int64_t n, float4* xyzr, float4* col) {
    hipLaunchKernelGGL(make_cloud_kernel, dim3(grid_for(n, 256)), dim3(256), 0, s,
                       reinterpret_cast<const float4*>(xyzw), radii,
                       reinterpret_cast<const float4*>(colors), n, xyzr, col);
    return hipGetLastError();
}

hipError_t launch_max_radius(hipStream_t s, const float4* xyzr, int64_t n, uint32_t* out) {
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n == 0) return e;
    const int blocks = (int)std::min<int64_t>(grid_for(n, 256), 2048);
    hipLaunchKernelGGL(max_radius_kernel, dim3(blocks), dim3(256), 0, s, xyzr, n, out);
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

static hipError_t launch_cull_nodes(hipStream_t s, const DevTree& t, uint32_t* cull,
                                    int64_t stride, int nviews, const TreeLevels& lv) {
    for (int l = lv.nlevels - 1; l >= 0; --l) {
        const int b = lv.start[l], cnt = lv.start[l + 1] - b;
        if (cnt <= 0) continue;
        hipLaunchKernelGGL(cull_nodes_kernel, dim3(grid_for(cnt, 256), nviews), dim3(256), 0, s,
                           t.nodes, lv.order + b, cnt, t.n, cull, stride);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_render_plain(hipStream_t s, const DevCloud& c, const DevView& v, MarchOut o,
                               int* err, uint32_t* cull, const TreeLevels& lv) {
    int maxsp = lane_stack_depth(c.tree);
    size_t shmem = (size_t)maxsp * 256 * sizeof(int);
    dim3 grid((v.W + 15) / 16, (v.H + 15) / 16);
    if (c.cull) {
        if (!cull) return hipErrorInvalidValue;
        hipLaunchKernelGGL(cull_mask_kernel<false>, dim3(grid_for(c.tree.n, 256)), dim3(256), 0,
                           s, c.tree, v, (const DevView*)nullptr, 1, cull, (int64_t)0);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = launch_cull_nodes(s, c.tree, cull, 0, 1, lv);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(render_plain_kernel<true>, grid, dim3(16, 16), shmem, s, c, v, o,
                           maxsp, err, (const uint32_t*)cull);
    } else {
        hipLaunchKernelGGL(render_plain_kernel<false>, grid, dim3(16, 16), shmem, s, c, v, o,
                           maxsp, err, (const uint32_t*)nullptr);
    }
    return hipGetLastError();
}

hipError_t launch_cull_masks(hipStream_t s, const DevTree& t, const DevView* views, int nviews,
                             uint32_t* cull, int64_t stride, const TreeLevels& lv) {
    if (t.n == 0 || nviews <= 0) return hipSuccess;
    hipLaunchKernelGGL(cull_mask_kernel<true>, dim3(grid_for(t.n, 256)), dim3(256), 0, s, t,
                       DevView{}, views, nviews, cull, stride);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return launch_cull_nodes(s, t, cull, stride, nviews, lv);
}

template <int MODE, bool WARM, int WPB>
static hipError_t launch_batch_t(hipStream_t s, const BatchParams& p0, int blocks) {
    BatchParams p = p0;
    const int maxsp = lane_stack_depth(p.cloud.tree);
    if (p.lane_lds <= 0 || p.lane_lds > maxsp) p.lane_lds = maxsp;
    // a stack split between LDS and the overflow area needs the area
    if (MODE == 1 && p.lane_lds < maxsp && !p.lane_ovf) return hipErrorInvalidValue;
    size_t shmem = MODE == 1 ? (size_t)p.lane_lds * WPB * 64 * sizeof(int) : 0;
    // the split stack has its own instantiations, so a whole-LDS stack pays nothing for it
    const bool ovf = MODE == 1 && p.lane_lds < maxsp;
    if (p.cloud.cull && ovf)
        hipLaunchKernelGGL((batch_kernel<MODE, WARM, WPB, true, MODE == 1>), dim3(blocks),
                           dim3(WPB * 64), shmem, s, p, maxsp);
    else if (p.cloud.cull)
        hipLaunchKernelGGL((batch_kernel<MODE, WARM, WPB, true>), dim3(blocks), dim3(WPB * 64),
                           shmem, s, p, maxsp);
    else if (ovf)
        hipLaunchKernelGGL((batch_kernel<MODE, WARM, WPB, false, MODE == 1>), dim3(blocks),
                           dim3(WPB * 64), shmem, s, p, maxsp);
    else
        hipLaunchKernelGGL((batch_kernel<MODE, WARM, WPB, false>), dim3(blocks), dim3(WPB * 64),
                           shmem, s, p, maxsp);
    return hipGetLastError();
}

hipError_t launch_batch(hipStream_t s, const BatchParams& p, int mode, int warm, int wpb,
                        int blocks) {
#define RM_B(M, W, P) \
    if (mode == M && (warm != 0) == W && wpb == P) return launch_batch_t<M, W, P>(s, p, blocks);
    RM_B(0, false, 1) RM_B(0, false, 2) RM_B(0, false, 4)
    RM_B(0, true, 1) RM_B(0, true, 2) RM_B(0, true, 4)
    RM_B(1, false, 1) RM_B(1, false, 2) RM_B(1, false, 4)
    RM_B(1, true, 1) RM_B(1, true, 2) RM_B(1, true, 4)
#undef RM_B
    return hipErrorInvalidValue;
}

hipError_t launch_frontier(hipStream_t s, const RmNode* nodes, int depth, float* out) {
    if (depth < 1 || depth > kMaxFrontierDepth) return hipErrorInvalidValue;
    hipLaunchKernelGGL(frontier_kernel, dim3(1), dim3(1 << depth), 0, s, nodes, depth, out);
    return hipGetLastError();
}

hipError_t launch_reset_queue(hipStream_t s, unsigned long long* q) {
    hipLaunchKernelGGL(reset_queue_kernel, dim3(1), dim3(1), 0, s, q);
    return hipGetLastError();
}

hipError_t launch_store_views(hipStream_t s, const ViewChunk& ch, DevView* table) {
    hipLaunchKernelGGL(store_views_kernel, dim3(1), dim3(kViewChunk), 0, s, ch, table);
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

hipError_t launch_adjoint_batch_keys(hipStream_t s, const DevView* views, const int64_t* cum,
                                     int nviews, int64_t P, const int32_t* hit, int32_t miss,
                                     int32_t* keys, int32_t* vals) {
    hipLaunchKernelGGL(adjoint_batch_keys_kernel, dim3(grid_for(P, 256)), dim3(256), 0, s, views,
                       cum, nviews, P, hit, miss, keys, vals);
    return hipGetLastError();
}

hipError_t launch_adjoint_batch_reduce(hipStream_t s, const DevCloud& c, const float* xyz,
                                       const float* rad, const DevView* views, const int64_t* cum,
                                       int nviews, const int32_t* keys_sorted,
                                       const int32_t* vals_sorted, int64_t P,
                                       const float* grad_rgb, const float* grad_depth,
                                       float* g_xyz, float* g_r, float* g_rgb) {
    hipLaunchKernelGGL(adjoint_batch_reduce_kernel, dim3(grid_for(P, 256)), dim3(256), 0, s, c,
                       xyz, rad, views, cum, nviews, keys_sorted, vals_sorted, P, grad_rgb,
                       grad_depth, g_xyz, g_r, g_rgb);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------
// Device tree build driver
// ------------------------------------------------------------------------------------------
#define RM_BUILD_CHECK(call)                        \
    do {                                            \
        hipError_t e_ = (call);                     \
        if (e_ != hipSuccess) return e_;            \
    } while (0)

hipError_t build_tree_device(hipStream_t s, const float* xyzw, const RmShape& shape,
                             const int64_t* d_seg_b, const int64_t* d_seg_e,
                             const int32_t* d_nl, RmNode* d_nodes, float4* d_pts,
                             BuildScratch& w) {
    const int64_t n = shape.n;
    const float4* p = reinterpret_cast<const float4*>(xyzw);
    const unsigned g = grid_for(n, 256);
    // three id lists, each sorted by one coordinate (stable: ties keep ascending id)
    for (int d = 0; d < 3; ++d) {
        hipLaunchKernelGGL(build_keys_kernel, dim3(g), dim3(256), 0, s, p, n, d, w.keys,
                           w.ids);
        RM_BUILD_CHECK(hipGetLastError());
        size_t bytes = w.temp_bytes;
        RM_BUILD_CHECK(hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, w.keys, w.keys_out,
                                                          w.ids, w.list[d][0], (int)n, 0, 32, s));
    }
    int cur = 0;
    auto lists = [&](int c) {
        BuildLists L;
        for (int d = 0; d < 3; ++d) L.l[d] = w.list[d][c];
        return L;
    };
    if (shape.nodes[0].cnt[1] < 0) {  // a single leaf
        hipLaunchKernelGGL(build_leafroot_kernel, dim3(1), dim3(1), 0, s, n, p, lists(cur),
                           d_nodes);
        RM_BUILD_CHECK(hipGetLastError());
    } else {
        hipLaunchKernelGGL(build_fill_kernel, dim3(g), dim3(256), 0, s, n, 0, w.owner);
        RM_BUILD_CHECK(hipGetLastError());
        const int nlev = (int)shape.level_start.size() - 1;
        for (int l = 0; l < nlev; ++l) {
            const int node0 = shape.level_start[l];
            const int count = shape.level_start[l + 1] - node0;
            const unsigned gn = grid_for(count, 256);
            hipLaunchKernelGGL(build_axis_kernel, dim3(gn), dim3(256), 0, s, d_seg_b, d_seg_e,
                               node0, count, p, lists(cur), w.axis);
            RM_BUILD_CHECK(hipGetLastError());
            hipLaunchKernelGGL(build_flag_kernel, dim3(g), dim3(256), 0, s, n, w.owner, d_seg_b,
                               d_nl, w.axis, lists(cur), w.flag);
            RM_BUILD_CHECK(hipGetLastError());
            for (int d = 0; d < 3; ++d) {
                hipLaunchKernelGGL(build_isleft_kernel, dim3(g), dim3(256), 0, s, n, w.owner,
                                   w.list[d][cur], w.flag, w.il);
                RM_BUILD_CHECK(hipGetLastError());
                size_t bytes = w.temp_bytes;
                RM_BUILD_CHECK(hipcub::DeviceScan::ExclusiveSum(w.temp, bytes, w.il, w.pl,
                                                                (int)n, s));
                hipLaunchKernelGGL(build_scatter_kernel, dim3(g), dim3(256), 0, s, n, w.owner,
                                   d_seg_b, d_nl, w.list[d][cur], w.il, w.pl,
                                   w.list[d][1 - cur]);
                RM_BUILD_CHECK(hipGetLastError());
            }
            cur = 1 - cur;
            hipLaunchKernelGGL(build_box_kernel, dim3(gn), dim3(256), 0, s, node0, count,
                               d_seg_b, d_seg_e, d_nl, p, lists(cur), d_nodes);
            RM_BUILD_CHECK(hipGetLastError());
            hipLaunchKernelGGL(build_owner_kernel, dim3(g), dim3(256), 0, s, n, w.owner,
                               d_seg_b, d_nl, d_nodes);
            RM_BUILD_CHECK(hipGetLastError());
        }
    }
    hipLaunchKernelGGL(build_pts_kernel, dim3(g), dim3(256), 0, s, n, p, w.list[0][cur], d_pts);
    return hipGetLastError();
}

size_t build_temp_bytes(int64_t n) {
    size_t a = 0, b = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, a, (uint32_t*)nullptr, (uint32_t*)nullptr,
                                             (int32_t*)nullptr, (int32_t*)nullptr, (int)n, 0, 32);
    (void)hipcub::DeviceScan::ExclusiveSum(nullptr, b, (int32_t*)nullptr, (int32_t*)nullptr,
                                           (int)n);
    return a > b ? a : b;
}

}  // namespace rm

// Internal declarations shared by the C-ABI layer (rm_api.hip) and the kernels
// (rm_kernels.hip). Not part of the public interface.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "rm_tree_host.h"

namespace rm {

// March constants of src/common.h:4-7 and src/ray_marching.h:17.
constexpr float kZFar = 100.0f;
constexpr float kZNear = 0.1f;
constexpr int kMaxSteps = 128;
constexpr float kMinDist = 0.001f;
constexpr int kMarchK = 10;

// Largest tree depth the traversals accept (stack bound); the host refuses deeper trees.
constexpr int kMaxDepth = 60;
// Most node records a tree may have: the packet walks address a record by a 32-bit byte
// offset (64 B each, under 4 GiB). 2^26 records hold over a billion points at 16 per leaf.
constexpr int64_t kMaxNodes = int64_t(1) << 26;
constexpr int kWaveStack = 64;
// Leaves queued per round of the batched march's packet traversal (two scalar registers).
constexpr int kLeafQueue = 2;
// Early-miss frontier of the batched march: entries of this many floats, depth at most this.
constexpr int kFrontierFloats = 8;
constexpr int kMaxFrontierDepth = 10;


// Device view of a tree + cloud, passed by value to kernels.
struct DevTree {
    const RmNode* nodes;
    const float4* pts;      // leaf order: x, y, z, original index bits
    const int32_t* inv;     // original order: leaf position of each point
    int64_t n;
    int depth;
    int leaf_max;           // every leaf holds leaf_max points except the last one
};

struct DevCloud {
    DevTree tree;
    const float4* xyzr;     // original order: x, y, z, radius
    const float4* colors;   // original order: r, g, b, 1
    int32_t cull;           // 1: march over the points inside the view frustum only
    const uint32_t* rmax;   // bits of the largest radius (at least 0), kept on the device
};

// One view as the kernels see it.
struct DevView {
    float M[16];            // column-major camera-to-world
    float pane;             // W / (2 tan(fov / 2)), computed on the host
    int32_t W, H;
    int32_t tiles_x;
    int64_t pix_offset;     // first pixel in the concatenated outputs
    int64_t tile_offset;    // first tile in the batch's tile numbering
    // World-space frustum: normals of the four side planes through the camera origin (x, y, z
    // each; a point p is inside plane i iff fn_i . (p - origin) < 0), then the forward direction.
    float fn[12];
    float fdir[3];
};
static_assert(offsetof(DevView, fdir) == offsetof(DevView, fn) + 12 * sizeof(float),
              "the kernels read fn and fdir as one run of 15 floats");

// Views handed to a kernel by value (kernel arguments stay within 4 KiB): the allocation-free
// batched march fills its view table with one launch per chunk.
constexpr int kViewChunk = 24;
struct ViewChunk {
    DevView v[kViewChunk];
    int32_t first;          // index of v[0] in the table
    int32_t count;          // views used, 1..kViewChunk
};
static_assert(sizeof(ViewChunk) <= 4096, "kernel arguments are limited to 4 KiB");

struct MarchOut {
    float4* rgba;
    int32_t* hit;           // may be null
};

struct BatchParams {
    DevCloud cloud;
    const DevView* views;
    int32_t nviews;
    int32_t tile_w, tile_h;   // tile_w * tile_h pixels per claimed tile (a multiple of 64)
    int32_t refill_min;       // refill idle lanes from the tile once this many are idle
    int64_t total_tiles;
    unsigned long long* queue;  // work counter (tiles, or pixels with refill)
    int* err;               // device error flag
    MarchOut out;
    const uint32_t* cull_mask;  // cloud.cull: the culling data of each view, cull_words apart
    int64_t cull_words;
    // per-lane traversal: stack entries kept in LDS per lane (0: the whole stack), and the
    // global area for the deeper ones, lane_ovf_words ints per thread of the grid (column
    // layout); null when the whole stack is in LDS
    int32_t lane_lds;
    int* lane_ovf;
    // early miss: 0 = the root box; else the frontier's entry count (2^depth) and its array
    // (a header entry, then the entries), refreshed from the tree at every launch
    int32_t nfrontier;
    const float* frontier;
};

// Per-lane stack entries a traversal of a tree of this depth needs (root children, then at most
// two per level below).
inline int lane_stack_need(int tree_depth) { return tree_depth + 3; }
// Ints of per-lane overflow area a grid of `threads` threads needs for a tree of this depth when
// `lane_lds` entries per lane are in LDS (0 when they cover the whole stack).
inline int64_t lane_ovf_ints(int tree_depth, int lane_lds, int64_t threads) {
    const int need = lane_stack_need(tree_depth);
    return (lane_lds <= 0 || lane_lds >= need) ? 0 : (int64_t)(need - lane_lds) * threads;
}

// Culling data of one view: a point mask (bit i: point i of the leaf order is inside the view
// frustum; 32 points a word, whole pairs of words plus a spare pair), then one byte per inner
// node (bit s: child slot s holds a point inside the frustum), padded to whole words.
__host__ __device__ inline int64_t cull_mask_words(int64_t n) { return 2 * ((n + 63) / 64) + 2; }
inline int64_t cull_view_words(int64_t n, int64_t nodes) {
    return cull_mask_words(n) + (nodes + 3) / 4 + 1;
}

// Levels of the inner nodes, for bottom-up passes: ids of level l are
// order[start[l] .. start[l + 1]) (order on the device, start on the host).
struct TreeLevels {
    const int32_t* order;
    const int32_t* start;
    int nlevels;
};

// Kernel launchers (rm_kernels.hip). Each returns the hipError_t of the launch.
hipError_t launch_knn(hipStream_t s, const DevTree& t, const float* q, int64_t m, int k,
                      int32_t* idx, float* d2, int* err);
hipError_t launch_radius(hipStream_t s, const DevTree& t, const float* q, int64_t m, float r2,
                         int max_nn, int32_t* idx, float* d2, int32_t* counts, int* err);
// tmp: n floats of scratch (the radii in leaf order before they are put in original order)
hipError_t launch_radii(hipStream_t s, const DevTree& t, float max_radius, float* r_unc,
                        float* r_cl, float* tmp, int* err);
// inv[original index of pts[i]] = i
hipError_t launch_tree_inverse(hipStream_t s, const float4* pts, int64_t n, int32_t* inv);
hipError_t launch_make_cloud(hipStream_t s, const float* xyzw, const float* radii,
                             const float* colors, int64_t n, float4* xyzr, float4* col);
hipError_t launch_set_radii(hipStream_t s, const float* radii, int64_t n, float4* xyzr);
hipError_t launch_max_radius(hipStream_t s, const float4* xyzr, int64_t n, uint32_t* out);
hipError_t launch_set_colors(hipStream_t s, const float* colors, int64_t n, float4* col);
hipError_t launch_render_plain(hipStream_t s, const DevCloud& c, const DevView& v, MarchOut o,
                               int* err, uint32_t* cull, const TreeLevels& lv);
hipError_t launch_batch(hipStream_t s, const BatchParams& p, int mode, int warm, int wpb,
                        int blocks);
hipError_t launch_reset_queue(hipStream_t s, unsigned long long* q);
// frontier of depth 1..kMaxFrontierDepth: (1 + 2^depth) * kFrontierFloats floats at out
hipError_t launch_frontier(hipStream_t s, const RmNode* nodes, int depth, float* out);
hipError_t launch_store_views(hipStream_t s, const ViewChunk& ch, DevView* table);
hipError_t launch_cull_masks(hipStream_t s, const DevTree& t, const DevView* views, int nviews,
                             uint32_t* cull, int64_t stride, const TreeLevels& lv);
hipError_t launch_refit(hipStream_t s, RmNode* nodes, float4* pts, float4* xyzr,
                        const float* xyz, int64_t n, const int32_t* level_order,
                        const int32_t* level_start, int nlevels);
hipError_t launch_diff_forward(hipStream_t s, const DevCloud& c, const DevView& v,
                               const float4* rgba_marched, const int32_t* hit, float* rgb,
                               float* depth);
hipError_t launch_adjoint_keys(hipStream_t s, const int32_t* hit, int64_t npix, int32_t miss,
                               int32_t* keys, int32_t* vals);
hipError_t launch_adjoint_reduce(hipStream_t s, const DevCloud& c, const DevView& v,
                                 const int32_t* keys_sorted, const int32_t* vals_sorted,
                                 int64_t npix, const float* grad_rgb, const float* grad_depth,
                                 float* g_xyz, float* g_r, float* g_rgb);

// Adjoint over all views of a batch (pixels numbered view after view; cum: nviews + 1 entries).
hipError_t launch_adjoint_batch_keys(hipStream_t s, const DevView* views, const int64_t* cum,
                                     int nviews, int64_t P, const int32_t* hit, int32_t miss,
                                     int32_t* keys, int32_t* vals);
hipError_t launch_adjoint_batch_reduce(hipStream_t s, const DevCloud& c, const float* xyz,
                                       const float* rad, const DevView* views, const int64_t* cum,
                                       int nviews, const int32_t* keys_sorted,
                                       const int32_t* vals_sorted, int64_t P,
                                       const float* grad_rgb, const float* grad_depth,
                                       float* g_xyz, float* g_r, float* g_rgb);

// Scratch of the device tree build (all n-sized unless noted).
struct BuildScratch {
    uint32_t* keys;
    uint32_t* keys_out;
    int32_t* ids;
    int32_t* list[3][2];    // per axis, double-buffered
    int32_t* flag;
    int32_t* il;
    int32_t* pl;
    int32_t* owner;
    int8_t* axis;           // per inner node
    void* temp;             // hipcub temp storage, temp_bytes
    size_t temp_bytes;
};

size_t build_temp_bytes(int64_t n);
hipError_t build_tree_device(hipStream_t s, const float* xyzw, const RmShape& shape,
                             const int64_t* d_seg_b, const int64_t* d_seg_e,
                             const int32_t* d_nl, RmNode* d_nodes, float4* d_pts,
                             BuildScratch& w);

}  // namespace rm

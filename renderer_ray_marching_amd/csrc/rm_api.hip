// C ABI of librm_hip.so (declared in include/rm.h): contexts, trees, clouds, the plain and the
// batched march (persistent kernel, HIP graph, autotuner) and the differentiable form.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "../../include/rm.h"
#include "rm_internal.h"

using rm::DevCloud;
using rm::DevTree;
using rm::DevView;

// Scratch of rm_render_batch_noalloc: the tile-queue counter (first 256 bytes), then the view
// table of at most RM_NOALLOC_MAX_VIEWS views, then the early-miss frontier of the default tune's
// depth (refreshed from the tree at every call). A file-scope device array: every GPU the library
// runs on holds its own copy, at its own address, until the process ends. In host code the name
// `scratch` is only a stand-in for the symbol: kernels, copies and fills use the device address
// that hipGetSymbolAddress returns for the current device (kept per device by scratch_lookup
// below), never the name.
constexpr size_t kScratchHeader = 256;
// the early-miss frontier of the default tune's depth follows the view table
constexpr int kDefaultExitDepth = 6;
constexpr size_t kScratchViews = (size_t)RM_NOALLOC_MAX_VIEWS * sizeof(DevView);
constexpr size_t kScratchFrontier =
    sizeof(float) * rm::kFrontierFloats * ((size_t)1 + ((size_t)1 << kDefaultExitDepth));
constexpr size_t kScratchBytes = kScratchHeader + kScratchViews + kScratchFrontier;
static_assert(kDefaultExitDepth <= rm::kMaxFrontierDepth && kScratchViews % 256 == 0,
              "the frontier follows the view table, 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

struct rm_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    std::string err;
    int* d_err = nullptr;          // device-side invariant flag
    bool timing = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool ev_pending = false;
    int num_cus = 0;
    void* ws = nullptr;            // grow-only workspace
    size_t ws_bytes = 0;
    unsigned char* scratch = nullptr;  // device address of `scratch` on this ctx's device
};

struct rm_tree {
    int device = 0;
    int64_t n = 0;
    int64_t n_nodes = 0;
    int depth = 0;
    int leaf_max = 0;
    RmNode* nodes = nullptr;
    float4* pts = nullptr;
    int32_t* inv = nullptr;                // leaf position of each original index (device)
    int32_t* level_order = nullptr;        // inner node ids grouped by level (device)
    std::vector<int32_t> level_start;      // host: offsets into level_order, nlevels + 1
    uint64_t gen = 0;                      // bumped by every refit (which moves the points)
};

struct rm_cloud {
    int device = 0;
    rm_tree* tree = nullptr;
    int64_t n = 0;
    float4* xyzr = nullptr;
    float4* colors = nullptr;
    int32_t cull = 0;
    uint32_t* mask = nullptr;      // culling mask of single-view renders (allocated with culling)
    uint32_t* rmax = nullptr;      // device: bits of the largest radius
    uint64_t gen = 0;              // the tree's refit generation this cloud's centres match
};

struct rm_batch {
    rm_ctx* ctx = nullptr;
    const rm_cloud* cloud = nullptr;
    std::vector<rm_view> views;
    DevView* d_views = nullptr;
    int64_t total_tiles = 0;
    int laid_views = 0;
    rm_tune tune{};
    unsigned long long* d_queue = nullptr;
    float* rgba = nullptr;
    int32_t* hit = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int blocks = 0;
    int32_t cull = 0;              // the cloud's culling setting at creation
    uint32_t* d_mask = nullptr;    // culling masks, one per view
    int* d_ovf = nullptr;          // per-lane stack overflow area (tune.stack_depth), grow-only
    int64_t ovf_ints = 0;
    float* d_frontier = nullptr;   // early-miss frontier (tune.exit_depth), deepest size
};

namespace {

const int kDefaultLeaf = 16;

int fail(rm_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int hip_fail(rm_ctx* ctx, hipError_t e, const char* where) {
    return fail(ctx, RM_ERR_HIP, std::string(where) + ": " + hipGetErrorString(e));
}

#define RM_HIP(ctx, call)                                   \
    do {                                                    \
        hipError_t e_ = (call);                             \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
    } while (0)

DevTree dev_tree(const rm_tree* t) {
    DevTree d;
    d.nodes = t->nodes;
    d.pts = t->pts;
    d.inv = t->inv;
    d.n = t->n;
    d.depth = t->depth;
    d.leaf_max = t->leaf_max;
    return d;
}

rm::TreeLevels tree_levels(const rm_tree* t) {
    return rm::TreeLevels{t->level_order, t->level_start.data(), (int)t->level_start.size() - 1};
}

DevCloud dev_cloud(const rm_cloud* c) {
    DevCloud d;
    d.tree = dev_tree(c->tree);
    d.xyzr = c->xyzr;
    d.colors = c->colors;
    d.cull = c->cull;
    d.rmax = c->rmax;
    return d;
}

// Pane distance of src/ray_marching.cu:74, computed once on the host (float, libm tanf).
float pane_distance(int W, float fov) { return (float)W / (2.0f * std::tan(0.5f * fov)); }

// Frustum of one view, as src/ray_marching.cu:112-124 builds its edge points (and the commented
// predicate at :243-257 would use them): corner directions E_k = M * (x W/2, y H/2, -pane, 0)
// for (x, y) = (-1,-1), (1,-1), (1,1), (-1,1), with glm's (m0 x + m1 y) + (m2 z + m3 0); side
// normals cross(E_k, E_{k+1 mod 4}); forward = -M[2]. Float, evaluated as written.
void view_frustum(const float* M, float pane, int W, int H, float* fn, float* fdir) {
    static const float cx[4] = {-1.0f, 1.0f, 1.0f, -1.0f}, cy[4] = {-1.0f, -1.0f, 1.0f, 1.0f};
    float E[4][3];
    for (int k = 0; k < 4; ++k) {
        const float vx = cx[k] * (0.5f * (float)W), vy = cy[k] * (0.5f * (float)H), vz = -pane;
        for (int i = 0; i < 3; ++i) {
            const float add0 = M[0 + i] * vx + M[4 + i] * vy;
            const float add1 = M[8 + i] * vz + M[12 + i] * 0.0f;
            E[k][i] = add0 + add1;
        }
    }
    for (int k = 0; k < 4; ++k) {
        const float* a = E[k];
        const float* b = E[(k + 1) % 4];
        fn[3 * k + 0] = a[1] * b[2] - a[2] * b[1];
        fn[3 * k + 1] = a[2] * b[0] - a[0] * b[2];
        fn[3 * k + 2] = a[0] * b[1] - a[1] * b[0];
    }
    fdir[0] = -M[8];
    fdir[1] = -M[9];
    fdir[2] = -M[10];
}

DevView dev_view(const float* pose, float fov, int W, int H, int64_t pix_offset, int tile_w,
                 int64_t tile_offset) {
    DevView v;
    std::memcpy(v.M, pose, sizeof(v.M));
    v.pane = pane_distance(W, fov);
    v.W = W;
    v.H = H;
    v.tiles_x = tile_w > 0 ? (W + tile_w - 1) / tile_w : 0;
    v.pix_offset = pix_offset;
    v.tile_offset = tile_offset;
    view_frustum(v.M, v.pane, W, H, v.fn, v.fdir);
    return v;
}

// Makes the context's device current and clears the thread's last HIP error, so that the
// launch checks of the call that follows (hipGetLastError after each launch) report its own
// launches only, not a failure some other component of the process left behind.
int set_device(rm_ctx* ctx) {
    RM_HIP(ctx, hipSetDevice(ctx->device));
    (void)hipGetLastError();
    return RM_OK;
}

// Device address of `scratch` on each device, looked up once per device with that device
// current (at the first context created on it); every context on the device shares it.
std::mutex g_scratch_mu;
std::vector<unsigned char*> g_scratch_addr;

hipError_t scratch_lookup(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device >= g_scratch_addr.size()) g_scratch_addr.resize((size_t)device + 1, nullptr);
    if (!g_scratch_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = static_cast<unsigned char*>(p);
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}

int ensure_ws(rm_ctx* ctx, size_t bytes) {
    if (bytes <= ctx->ws_bytes) return RM_OK;
    if (ctx->ws) {
        RM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        RM_HIP(ctx, hipFree(ctx->ws));
        ctx->ws = nullptr;
        ctx->ws_bytes = 0;
    }
    size_t b = std::max(bytes, (size_t)1 << 20);
    if (hipMalloc(&ctx->ws, b) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, RM_ERR_NOMEM, "workspace allocation failed");
    }
    ctx->ws_bytes = b;
    return RM_OK;
}

void timing_begin(rm_ctx* ctx) {
    if (ctx->timing) (void)hipEventRecord(ctx->ev0, ctx->stream);
}
void timing_end(rm_ctx* ctx) {
    if (ctx->timing) {
        (void)hipEventRecord(ctx->ev1, ctx->stream);
        ctx->ev_pending = true;
    }
}

int check_tree(rm_ctx* ctx, const rm_tree* t) {
    if (!t) return fail(ctx, RM_ERR_INVALID, "null tree");
    if (t->device != ctx->device) return fail(ctx, RM_ERR_INVALID, "tree is on another device");
    return RM_OK;
}

int check_cloud(rm_ctx* ctx, const rm_cloud* c) {
    if (!c) return fail(ctx, RM_ERR_INVALID, "null cloud");
    return check_tree(ctx, c->tree);
}

// A cloud marches its own centres over its tree's points; once the tree has been refit through
// another cloud the two disagree (the warm start and the signed distance would read stale
// centres), so the march refuses until this cloud is refit too.
int check_fresh(rm_ctx* ctx, const rm_cloud* c, const char* where) {
    if (int r = check_cloud(ctx, c)) return r;
    if (c->gen != c->tree->gen)
        return fail(ctx, RM_ERR_INVALID,
                    std::string(where) + ": the cloud's tree was refit through another cloud; "
                                         "refit this cloud (rm_cloud_refit) or create it again");
    return RM_OK;
}

size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------------------------
int rm_ctx_create(int device, rm_ctx** out) {
    if (!out) return RM_ERR_INVALID;
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || device < 0 || device >= count) {
        (void)hipGetLastError();
        return RM_ERR_INVALID;
    }
    rm_ctx* ctx = new (std::nothrow) rm_ctx();
    if (!ctx) return RM_ERR_NOMEM;
    ctx->device = device;
    if (hipSetDevice(device) != hipSuccess ||
        hipStreamCreateWithFlags(&ctx->own, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&ctx->d_err, sizeof(int)) != hipSuccess ||
        hipMemset(ctx->d_err, 0, sizeof(int)) != hipSuccess ||
        hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
        (void)hipGetLastError();
        rm_ctx_destroy(ctx);
        return RM_ERR_HIP;
    }
    ctx->stream = ctx->own;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) ctx->num_cus = prop.multiProcessorCount;
    if (ctx->num_cus <= 0) ctx->num_cus = 256;
    // the allocation-free entry point's scratch; on failure that entry point reports the error
    // and every other function works as before
    const hipError_t se = scratch_lookup(device, &ctx->scratch);
    if (se != hipSuccess) {
        (void)hipGetLastError();
        ctx->scratch = nullptr;
        ctx->err = std::string("scratch lookup: ") + hipGetErrorString(se);
    }
    *out = ctx;
    return RM_OK;
}

int rm_ctx_destroy(rm_ctx* ctx) {
    if (!ctx) return RM_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ws) (void)hipFree(ctx->ws);
    if (ctx->d_err) (void)hipFree(ctx->d_err);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->own) (void)hipStreamDestroy(ctx->own);
    delete ctx;
    return RM_OK;
}

int rm_ctx_set_stream(rm_ctx* ctx, void* hip_stream) {
    if (!ctx) return RM_ERR_INVALID;
    ctx->stream = (hipStream_t)hip_stream;
    return RM_OK;
}

int rm_ctx_get_stream(const rm_ctx* ctx, void** hip_stream) {
    if (!ctx || !hip_stream) return RM_ERR_INVALID;
    *hip_stream = (void*)ctx->stream;
    return RM_OK;
}

int rm_ctx_synchronize(rm_ctx* ctx) {
    if (!ctx) return RM_ERR_INVALID;
    if (int r = set_device(ctx)) return r;
    RM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int flag = 0;
    RM_HIP(ctx, hipMemcpy(&flag, ctx->d_err, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        (void)hipMemset(ctx->d_err, 0, sizeof(int));
        return fail(ctx, RM_ERR_INTERNAL, "device traversal stack overflow");
    }
    return RM_OK;
}

const char* rm_last_error(const rm_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rm_ctx_enable_timing(rm_ctx* ctx, int on) {
    if (!ctx) return RM_ERR_INVALID;
    ctx->timing = on != 0;
    ctx->ev_pending = false;
    return RM_OK;
}

float rm_ctx_last_kernel_ms(const rm_ctx* ctx) {
    if (!ctx || !ctx->timing || !ctx->ev_pending) return 0.0f;
    if (hipEventSynchronize(ctx->ev1) != hipSuccess) return 0.0f;
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess) return 0.0f;
    return ms;
}

// ---------------------------------------------------------------------------------------------
// tree
// ---------------------------------------------------------------------------------------------
}  // extern "C"

namespace {

// Host build (rm_tree_host.cpp): kept for cross-checks and as a fallback (RM_TREE_BUILD=host).
int tree_build_host(rm_ctx* ctx, const float* xyzw, int64_t n, int leaf_max, rm_tree* t) {
    std::vector<float> host((size_t)n * 4);
    RM_HIP(ctx, hipMemcpyAsync(host.data(), xyzw, (size_t)n * 16, hipMemcpyDefault, ctx->stream));
    RM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    RmHostTree ht;
    std::string msg;
    if (!rm_build_host_tree(host.data(), n, leaf_max, ht, msg)) return fail(ctx, RM_ERR_INVALID, msg);
    if (ht.depth > rm::kMaxDepth)
        return fail(ctx, RM_ERR_UNSUPPORTED, "rm_tree_build: tree deeper than supported");
    t->n_nodes = (int64_t)ht.nodes.size();
    t->depth = ht.depth;
    int nlev = 0;
    for (int32_t l : ht.level) nlev = std::max(nlev, l + 1);
    t->level_start.assign(nlev + 1, 0);
    for (int32_t l : ht.level) t->level_start[l + 1]++;
    for (int l = 0; l < nlev; ++l) t->level_start[l + 1] += t->level_start[l];
    std::vector<int32_t> order(ht.level.size());
    {
        std::vector<int32_t> fillp(t->level_start.begin(), t->level_start.end() - 1);
        for (size_t i = 0; i < ht.level.size(); ++i) order[fillp[ht.level[i]]++] = (int32_t)i;
    }
    if (hipMalloc(&t->nodes, sizeof(RmNode) * ht.nodes.size()) != hipSuccess ||
        hipMalloc(&t->pts, sizeof(float4) * (size_t)n) != hipSuccess ||
        hipMalloc(&t->inv, sizeof(int32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&t->level_order, sizeof(int32_t) * std::max<size_t>(1, order.size())) !=
            hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, RM_ERR_NOMEM, "rm_tree_build: device allocation failed");
    }
    RM_HIP(ctx, hipMemcpyAsync(t->nodes, ht.nodes.data(), sizeof(RmNode) * ht.nodes.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    RM_HIP(ctx, hipMemcpyAsync(t->pts, ht.pts.data(), sizeof(float4) * (size_t)n,
                               hipMemcpyHostToDevice, ctx->stream));
    if (!order.empty())
        RM_HIP(ctx, hipMemcpyAsync(t->level_order, order.data(), sizeof(int32_t) * order.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
    RM_HIP(ctx, rm::launch_tree_inverse(ctx->stream, t->pts, n, t->inv));
    RM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RM_OK;
}

// Device build: the host lays out the shape (counts only), the GPU sorts, partitions and fills
// in boxes level by level (rm_kernels.hip, build_tree_device).
int tree_build_device(rm_ctx* ctx, const float* xyzw, int64_t n, int leaf_max, rm_tree* t) {
    RmShape shape;
    std::string msg;
    if (!rm_tree_shape(n, leaf_max, shape, msg)) return fail(ctx, RM_ERR_INVALID, msg);
    if (shape.depth > rm::kMaxDepth)
        return fail(ctx, RM_ERR_UNSUPPORTED, "rm_tree_build: tree deeper than supported");
    const size_t nn = shape.nodes.size();
    t->n_nodes = (int64_t)nn;
    t->depth = shape.depth;
    t->level_start = shape.level_start;
    std::vector<int32_t> order(nn);
    for (size_t i = 0; i < nn; ++i) order[i] = (int32_t)i;  // breadth-first = level order
    if (hipMalloc(&t->nodes, sizeof(RmNode) * nn) != hipSuccess ||
        hipMalloc(&t->pts, sizeof(float4) * (size_t)n) != hipSuccess ||
        hipMalloc(&t->inv, sizeof(int32_t) * (size_t)n) != hipSuccess ||
        hipMalloc(&t->level_order, sizeof(int32_t) * nn) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, RM_ERR_NOMEM, "rm_tree_build: device allocation failed");
    }
    const size_t temp = rm::build_temp_bytes(n);
    const size_t n4 = align_up(sizeof(int32_t) * (size_t)n);
    const size_t total = align_up(sizeof(float4) * (size_t)n) + 15 * n4 + align_up(nn) +
                         3 * align_up(sizeof(int64_t) * nn) + align_up(temp);
    char* base = nullptr;
    if (hipMalloc(&base, total) != hipSuccess) {
        (void)hipGetLastError();
        return fail(ctx, RM_ERR_NOMEM, "rm_tree_build: scratch allocation failed");
    }
    char* c = base;
    auto take = [&](size_t bytes) {
        char* r = c;
        c += align_up(bytes);
        return r;
    };
    float* in = reinterpret_cast<float*>(take(sizeof(float4) * (size_t)n));
    rm::BuildScratch w;
    w.keys = reinterpret_cast<uint32_t*>(take(n4));
    w.keys_out = reinterpret_cast<uint32_t*>(take(n4));
    w.ids = reinterpret_cast<int32_t*>(take(n4));
    for (int d = 0; d < 3; ++d)
        for (int b = 0; b < 2; ++b) w.list[d][b] = reinterpret_cast<int32_t*>(take(n4));
    w.flag = reinterpret_cast<int32_t*>(take(n4));
    w.il = reinterpret_cast<int32_t*>(take(n4));
    w.pl = reinterpret_cast<int32_t*>(take(n4));
    w.owner = reinterpret_cast<int32_t*>(take(n4));
    w.axis = reinterpret_cast<int8_t*>(take(nn));
    int64_t* seg_b = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * nn));
    int64_t* seg_e = reinterpret_cast<int64_t*>(take(sizeof(int64_t) * nn));
    int32_t* nl = reinterpret_cast<int32_t*>(take(sizeof(int64_t) * nn));
    w.temp = take(temp);
    w.temp_bytes = temp;
    hipError_t e = hipMemcpyAsync(in, xyzw, sizeof(float4) * (size_t)n, hipMemcpyDefault,
                                  ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t->nodes, shape.nodes.data(), sizeof(RmNode) * nn,
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(t->level_order, order.data(), sizeof(int32_t) * nn,
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(seg_b, shape.seg_b.data(), sizeof(int64_t) * nn,
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(seg_e, shape.seg_e.data(), sizeof(int64_t) * nn,
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(nl, shape.nl.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess)
        e = rm::build_tree_device(ctx->stream, in, shape, seg_b, seg_e, nl, t->nodes, t->pts, w);
    if (e == hipSuccess) e = rm::launch_tree_inverse(ctx->stream, t->pts, n, t->inv);
    hipError_t e2 = hipStreamSynchronize(ctx->stream);
    (void)hipFree(base);
    if (e != hipSuccess) return hip_fail(ctx, e, "rm_tree_build (device)");
    if (e2 != hipSuccess) return hip_fail(ctx, e2, "rm_tree_build (device)");
    return RM_OK;
}

}  // namespace

extern "C" {

int rm_tree_build(rm_ctx* ctx, const float* xyzw, int64_t n, int leaf_max, rm_tree** out) {
    if (!ctx || !out) return RM_ERR_INVALID;
    *out = nullptr;
    if (!xyzw || n <= 0) return fail(ctx, RM_ERR_INVALID, "rm_tree_build: empty input");
    if (leaf_max == 0) leaf_max = kDefaultLeaf;
    // refused before any layout work: the packet walks address node records by 32-bit offsets
    if (leaf_max >= 1 && leaf_max <= 64 && rm_tree_node_count(n, leaf_max) > rm::kMaxNodes)
        return fail(ctx, RM_ERR_UNSUPPORTED, "rm_tree_build: more than 2^26 tree nodes");
    if (int r = set_device(ctx)) return r;
    rm_tree* t = new (std::nothrow) rm_tree();
    if (!t) return fail(ctx, RM_ERR_NOMEM, "host allocation failed");
    t->device = ctx->device;
    t->n = n;
    t->leaf_max = leaf_max;
    const char* env = std::getenv("RM_TREE_BUILD");
    const bool host = env && std::strcmp(env, "host") == 0;
    int r = host ? tree_build_host(ctx, xyzw, n, leaf_max, t)
                 : tree_build_device(ctx, xyzw, n, leaf_max, t);
    if (r != RM_OK) {
        rm_tree_destroy(t);
        return r;
    }
    *out = t;
    return RM_OK;
}

int rm_tree_destroy(rm_tree* t) {
    if (!t) return RM_OK;
    (void)hipSetDevice(t->device);
    if (t->nodes) (void)hipFree(t->nodes);
    if (t->pts) (void)hipFree(t->pts);
    if (t->inv) (void)hipFree(t->inv);
    if (t->level_order) (void)hipFree(t->level_order);
    delete t;
    return RM_OK;
}

int rm_tree_info(const rm_tree* t, int64_t* n, int64_t* n_nodes, int* depth, int* leaf_max) {
    if (!t) return RM_ERR_INVALID;
    if (n) *n = t->n;
    if (n_nodes) *n_nodes = t->n_nodes;
    if (depth) *depth = t->depth;
    if (leaf_max) *leaf_max = t->leaf_max;
    return RM_OK;
}

int rm_knn(rm_ctx* ctx, const rm_tree* tree, const float* q, int64_t m, int k, int32_t* idx,
           float* d2) {
    if (!ctx) return RM_ERR_INVALID;
    if (int r = check_tree(ctx, tree)) return r;
    if (k < 1) return fail(ctx, RM_ERR_INVALID, "rm_knn: k must be >= 1");
    if (k > RM_MAX_K) return fail(ctx, RM_ERR_UNSUPPORTED, "rm_knn: k above RM_MAX_K");
    if (m < 0 || (m > 0 && (!q || !idx || !d2)))
        return fail(ctx, RM_ERR_INVALID, "rm_knn: bad buffers");
    if (int r = set_device(ctx)) return r;
    timing_begin(ctx);
    RM_HIP(ctx, rm::launch_knn(ctx->stream, dev_tree(tree), q, m, k, idx, d2, ctx->d_err));
    timing_end(ctx);
    return RM_OK;
}

int rm_radius_search(rm_ctx* ctx, const rm_tree* tree, const float* q, int64_t m, float radius,
                     int max_nn, int32_t* idx, float* d2, int32_t* counts) {
    if (!ctx) return RM_ERR_INVALID;
    if (int r = check_tree(ctx, tree)) return r;
    if (max_nn < 0) return fail(ctx, RM_ERR_INVALID, "rm_radius_search: max_nn must be >= 0");
    if (max_nn > RM_MAX_K)
        return fail(ctx, RM_ERR_UNSUPPORTED, "rm_radius_search: max_nn above RM_MAX_K");
    if (!(radius >= 0.0f)) return fail(ctx, RM_ERR_INVALID, "rm_radius_search: bad radius");
    // max_nn == 0 counts only (kdtree_cuda_3d_index.cu:436-452): counts required, idx/d2 unused
    if (m < 0 || (m > 0 && (!q || (max_nn > 0 ? (!idx || !d2) : !counts))))
        return fail(ctx, RM_ERR_INVALID, "rm_radius_search: bad buffers");
    if (int r = set_device(ctx)) return r;
    float r2 = radius * radius;  // float(radius * radius), src/kdtree/kdtree_flann.cu:98
    timing_begin(ctx);
    RM_HIP(ctx, rm::launch_radius(ctx->stream, dev_tree(tree), q, m, r2, max_nn, idx, d2, counts,
                                  ctx->d_err));
    timing_end(ctx);
    return RM_OK;
}

int rm_radii(rm_ctx* ctx, const rm_tree* tree, float max_radius, float* r_unc, float* r_cl) {
    if (!ctx) return RM_ERR_INVALID;
    if (int r = check_tree(ctx, tree)) return r;
    if (int r = set_device(ctx)) return r;
    if (!r_unc && !r_cl) return RM_OK;
    if (int r = ensure_ws(ctx, sizeof(float) * (size_t)tree->n)) return r;
    timing_begin(ctx);
    RM_HIP(ctx, rm::launch_radii(ctx->stream, dev_tree(tree), max_radius, r_unc, r_cl,
                                 static_cast<float*>(ctx->ws), ctx->d_err));
    timing_end(ctx);
    return RM_OK;
}

// ---------------------------------------------------------------------------------------------
// cloud
// ---------------------------------------------------------------------------------------------
int rm_cloud_create(rm_ctx* ctx, const rm_tree* tree, const float* xyzw, const float* colors,
                    const float* radii, int64_t n, rm_cloud** out) {
    if (!ctx || !out) return RM_ERR_INVALID;
    *out = nullptr;
    if (int r = check_tree(ctx, tree)) return r;
    if (n != tree->n) return fail(ctx, RM_ERR_INVALID, "rm_cloud_create: n differs from the tree");
    if (!xyzw || !colors || !radii) return fail(ctx, RM_ERR_INVALID, "rm_cloud_create: null input");
    if (int r = set_device(ctx)) return r;
    rm_cloud* c = new (std::nothrow) rm_cloud();
    if (!c) return fail(ctx, RM_ERR_NOMEM, "host allocation failed");
    c->device = tree->device;
    c->tree = const_cast<rm_tree*>(tree);
    c->n = n;
    c->gen = tree->gen;
    if (hipMalloc(&c->xyzr, sizeof(float4) * (size_t)n) != hipSuccess ||
        hipMalloc(&c->colors, sizeof(float4) * (size_t)n) != hipSuccess ||
        hipMalloc(&c->rmax, sizeof(uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        rm_cloud_destroy(c);
        return fail(ctx, RM_ERR_NOMEM, "rm_cloud_create: device allocation failed");
    }
    hipError_t e = rm::launch_make_cloud(ctx->stream, xyzw, radii, colors, n, c->xyzr, c->colors);
    if (e == hipSuccess) e = rm::launch_max_radius(ctx->stream, c->xyzr, n, c->rmax);
    if (e != hipSuccess) {
        rm_cloud_destroy(c);
        return hip_fail(ctx, e, "rm_cloud_create");
    }
    *out = c;
    return RM_OK;
}

int rm_cloud_destroy(rm_cloud* c) {
    if (!c) return RM_OK;
    (void)hipSetDevice(c->device);
    if (c->xyzr) (void)hipFree(c->xyzr);
    if (c->colors) (void)hipFree(c->colors);
    if (c->mask) (void)hipFree(c->mask);
    if (c->rmax) (void)hipFree(c->rmax);
    delete c;
    return RM_OK;
}

int rm_cloud_set_radii(rm_ctx* ctx, rm_cloud* c, const float* radii) {
    if (!ctx || !c || !radii) return fail(ctx, RM_ERR_INVALID, "rm_cloud_set_radii: null input");
    if (int r = check_cloud(ctx, c)) return r;
    if (int r = set_device(ctx)) return r;
    RM_HIP(ctx, rm::launch_set_radii(ctx->stream, radii, c->n, c->xyzr));
    RM_HIP(ctx, rm::launch_max_radius(ctx->stream, c->xyzr, c->n, c->rmax));
    return RM_OK;
}

int rm_cloud_set_colors(rm_ctx* ctx, rm_cloud* c, const float* colors) {
    if (!ctx || !c || !colors) return fail(ctx, RM_ERR_INVALID, "rm_cloud_set_colors: null input");
    if (int r = check_cloud(ctx, c)) return r;
    if (int r = set_device(ctx)) return r;
    RM_HIP(ctx, rm::launch_set_colors(ctx->stream, colors, c->n, c->colors));
    return RM_OK;
}

int rm_cloud_set_culling(rm_ctx* ctx, rm_cloud* c, int enable) {
    if (!ctx || !c) return fail(ctx, RM_ERR_INVALID, "rm_cloud_set_culling: null input");
    if (enable != 0 && enable != 1)
        return fail(ctx, RM_ERR_INVALID, "rm_cloud_set_culling: enable must be 0 or 1");
    if (int r = check_cloud(ctx, c)) return r;
    if (enable && !c->mask) {
        if (int r = set_device(ctx)) return r;
        const size_t bytes =
            sizeof(uint32_t) * (size_t)rm::cull_view_words(c->n, c->tree->n_nodes);
        if (hipMalloc(&c->mask, bytes) != hipSuccess) {
            (void)hipGetLastError();
            c->mask = nullptr;
            return fail(ctx, RM_ERR_NOMEM, "rm_cloud_set_culling: device allocation failed");
        }
        RM_HIP(ctx, hipMemsetAsync(c->mask, 0, bytes, ctx->stream));
    }
    c->cull = enable;
    return RM_OK;
}

int rm_cloud_refit(rm_ctx* ctx, rm_cloud* c, const float* xyz) {
    if (!ctx || !c || !xyz) return fail(ctx, RM_ERR_INVALID, "rm_cloud_refit: null input");
    if (int r = check_cloud(ctx, c)) return r;
    if (int r = set_device(ctx)) return r;
    rm_tree* t = c->tree;
    RM_HIP(ctx, rm::launch_refit(ctx->stream, t->nodes, t->pts, c->xyzr, xyz, t->n,
                                 t->level_order, t->level_start.data(),
                                 (int)t->level_start.size() - 1));
    c->gen = ++t->gen;
    return RM_OK;
}

// ---------------------------------------------------------------------------------------------
// plain march
// ---------------------------------------------------------------------------------------------
int rm_render(rm_ctx* ctx, const rm_cloud* c, const float* pose, float fov, int W, int H,
              float* rgba, int32_t* hit) {
    if (!ctx) return RM_ERR_INVALID;
    if (!c || !pose || !rgba) return fail(ctx, RM_ERR_INVALID, "rm_render: null input");
    if (W <= 0 || H <= 0) return fail(ctx, RM_ERR_INVALID, "rm_render: empty image");
    if (!(fov > 0.0f && fov < 3.14159f)) return fail(ctx, RM_ERR_INVALID, "rm_render: bad fov");
    if (int r = check_fresh(ctx, c, "rm_render")) return r;
    if (int r = set_device(ctx)) return r;
    DevView v = dev_view(pose, fov, W, H, 0, 0, 0);
    timing_begin(ctx);
    RM_HIP(ctx, rm::launch_render_plain(ctx->stream, dev_cloud(c), v,
                                        rm::MarchOut{reinterpret_cast<float4*>(rgba), hit},
                                        ctx->d_err, c->mask, tree_levels(c->tree)));
    timing_end(ctx);
    return RM_OK;
}

// ---------------------------------------------------------------------------------------------
// batched march
// ---------------------------------------------------------------------------------------------
namespace {

rm_tune default_tune() {
    rm_tune t;
    t.tile_w = 8;
    t.tile_h = 8;
    t.mode = 0;
    t.warm = 1;
    t.waves_per_block = 2;
    t.blocks_per_cu = 0;
    t.refill = 1;
    t.stack_depth = 0;
    t.exit_depth = kDefaultExitDepth;
    return t;
}

static_assert(RM_MAX_EXIT_DEPTH == rm::kMaxFrontierDepth,
              "the public early-miss depth limit is the kernels' frontier limit");

bool tune_valid(const rm_tune& t) {
    const int pix = t.tile_w * t.tile_h;
    return t.tile_w > 0 && t.tile_h > 0 && pix >= 64 && pix <= 4096 && pix % 64 == 0 &&
           (t.mode == 0 || t.mode == 1) &&
           (t.warm == 0 || t.warm == 1) &&
           (t.waves_per_block == 1 || t.waves_per_block == 2 || t.waves_per_block == 4) &&
           t.blocks_per_cu >= 0 && t.blocks_per_cu <= 64 && t.refill >= 1 && t.refill <= 64 &&
           t.stack_depth >= 0 && t.stack_depth <= RM_MAX_STACK_DEPTH && t.exit_depth >= 0 &&
           t.exit_depth <= RM_MAX_EXIT_DEPTH;
}

// Lay out the views' tiles for a tune and upload the view table.
// Lays out the tiles of every `view_stride`-th view (1 = all; the autotuner times a subset of
// whole views, which keeps each view's screen locality and the mix of view sizes).
int batch_layout(rm_batch* b, int view_stride = 1) {
    rm_ctx* ctx = b->ctx;
    std::vector<DevView> dv;
    int64_t tiles = 0;
    for (size_t i = 0; i < b->views.size(); i += (size_t)view_stride) {
        const rm_view& v = b->views[i];
        dv.push_back(dev_view(v.pose, v.fov_radians, v.width, v.height, v.offset,
                              b->tune.tile_w, tiles));
        int64_t ty = (v.height + b->tune.tile_h - 1) / b->tune.tile_h;
        tiles += (int64_t)dv.back().tiles_x * ty;
    }
    b->total_tiles = tiles;
    b->laid_views = (int)dv.size();
    if (tiles >= (int64_t)UINT32_MAX) return fail(ctx, RM_ERR_UNSUPPORTED, "too many tiles");
    RM_HIP(ctx, hipMemcpyAsync(b->d_views, dv.data(), sizeof(DevView) * dv.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    RM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    int bpc = b->tune.blocks_per_cu;
    if (bpc == 0) {
        // occupancy-derived persistent grid: as many resident blocks as fit (the kernel is
        // built for 8 wavefronts per SIMD, 32 per CU)
        bpc = std::max(1, 32 / b->tune.waves_per_block);
    }
    b->blocks = ctx->num_cus * bpc;
    // per-lane traversal with part of the stack outside LDS: one column of overflow entries
    // per thread of the persistent grid
    if (b->tune.mode == 1) {
        const int64_t need = rm::lane_ovf_ints(b->cloud->tree->depth, b->tune.stack_depth,
                                               (int64_t)b->blocks * b->tune.waves_per_block * 64);
        if (need > b->ovf_ints) {
            if (b->d_ovf) (void)hipFree(b->d_ovf);
            b->d_ovf = nullptr;
            b->ovf_ints = 0;
            RM_HIP(ctx, hipMalloc(&b->d_ovf, (size_t)need * sizeof(int)));
            b->ovf_ints = need;
        }
    }
    return RM_OK;
}

rm::BatchParams batch_params(rm_batch* b) {
    rm::BatchParams p;
    p.cloud = dev_cloud(b->cloud);
    p.cloud.cull = b->cull;
    p.cull_mask = b->d_mask;
    p.cull_words = rm::cull_view_words(b->cloud->n, b->cloud->tree->n_nodes);
    p.views = b->d_views;
    p.nviews = (int32_t)b->laid_views;
    p.tile_w = b->tune.tile_w;
    p.tile_h = b->tune.tile_h;
    p.refill_min = b->tune.refill;
    p.total_tiles = b->total_tiles;
    p.queue = b->d_queue;
    p.err = b->ctx->d_err;
    p.out.rgba = reinterpret_cast<float4*>(b->rgba);
    p.out.hit = b->hit;
    p.lane_lds = b->tune.stack_depth;
    p.lane_ovf = b->d_ovf;
    p.nfrontier = b->tune.exit_depth > 0 ? (1 << b->tune.exit_depth) : 0;
    p.frontier = b->d_frontier;
    return p;
}

int batch_enqueue(rm_batch* b, hipStream_t s) {
    rm_ctx* ctx = b->ctx;
    if (b->cull)
        RM_HIP(ctx, rm::launch_cull_masks(
                        s, dev_tree(b->cloud->tree), b->d_views, b->laid_views, b->d_mask,
                        rm::cull_view_words(b->cloud->n, b->cloud->tree->n_nodes),
                        tree_levels(b->cloud->tree)));
    RM_HIP(ctx, rm::launch_reset_queue(s, b->d_queue));
    // the frontier is read from the tree at every launch, so it follows refits
    if (b->tune.exit_depth > 0)
        RM_HIP(ctx, rm::launch_frontier(s, b->cloud->tree->nodes, b->tune.exit_depth,
                                        b->d_frontier));
    RM_HIP(ctx, rm::launch_batch(s, batch_params(b), b->tune.mode, b->tune.warm,
                                 b->tune.waves_per_block, b->blocks));
    return RM_OK;
}

void batch_drop_graph(rm_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

int batch_capture(rm_batch* b) {
    rm_ctx* ctx = b->ctx;
    batch_drop_graph(b);
    // Capture on the context's own stream (the null stream cannot be captured); the graph is
    // launched on whatever stream the context works on.
    RM_HIP(ctx, hipStreamBeginCapture(ctx->own, hipStreamCaptureModeThreadLocal));
    int r = batch_enqueue(b, ctx->own);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->own, &g);
    if (r != RM_OK) {
        if (g) (void)hipGraphDestroy(g);
        return r;
    }
    if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
    b->graph = g;
    RM_HIP(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    return RM_OK;
}

}  // namespace

int rm_batch_create(rm_ctx* ctx, const rm_cloud* cloud, const rm_view* views, int nviews,
                    const rm_tune* tune, float* rgba, int32_t* hit, rm_batch** out) {
    if (!ctx || !out) return RM_ERR_INVALID;
    *out = nullptr;
    if (!cloud || !views || nviews <= 0 || !rgba)
        return fail(ctx, RM_ERR_INVALID, "rm_batch_create: bad arguments");
    for (int i = 0; i < nviews; ++i) {
        const rm_view& v = views[i];
        if (v.width <= 0 || v.height <= 0 || v.offset < 0 ||
            !(v.fov_radians > 0.0f && v.fov_radians < 3.14159f))
            return fail(ctx, RM_ERR_INVALID, "rm_batch_create: bad view " + std::to_string(i));
    }
    if (int r = check_fresh(ctx, cloud, "rm_batch_create")) return r;
    if (cloud->cull && nviews > 65535)
        return fail(ctx, RM_ERR_UNSUPPORTED,
                    "rm_batch_create: a culled batch holds at most 65535 views");
    rm_tune t = tune ? *tune : default_tune();
    if (!tune_valid(t)) return fail(ctx, RM_ERR_INVALID, "rm_batch_create: bad tune");
    if (t.warm && cloud->n < rm::kMarchK) t.warm = 0;
    if (int r = set_device(ctx)) return r;
    rm_batch* b = new (std::nothrow) rm_batch();
    if (!b) return fail(ctx, RM_ERR_NOMEM, "host allocation failed");
    b->ctx = ctx;
    b->cloud = cloud;
    b->views.assign(views, views + nviews);
    b->tune = t;
    b->rgba = rgba;
    b->hit = hit;
    b->cull = cloud->cull;
    if (hipMalloc(&b->d_views, sizeof(DevView) * nviews) != hipSuccess ||
        hipMalloc(&b->d_queue, sizeof(unsigned long long)) != hipSuccess ||
        hipMalloc(&b->d_frontier, sizeof(float) * rm::kFrontierFloats *
                                      ((size_t)1 + ((size_t)1 << rm::kMaxFrontierDepth))) !=
            hipSuccess ||
        (b->cull && hipMalloc(&b->d_mask, sizeof(uint32_t) * (size_t)nviews *
                                              (size_t)rm::cull_view_words(
                                                  cloud->n, cloud->tree->n_nodes)) !=
                        hipSuccess)) {
        (void)hipGetLastError();
        rm_batch_destroy(b);
        return fail(ctx, RM_ERR_NOMEM, "rm_batch_create: device allocation failed");
    }
    int r = batch_layout(b);
    if (r == RM_OK) r = batch_capture(b);
    if (r != RM_OK) {
        rm_batch_destroy(b);
        return r;
    }
    *out = b;
    return RM_OK;
}

int rm_batch_get_tune(const rm_batch* b, rm_tune* tune) {
    if (!b || !tune) return RM_ERR_INVALID;
    *tune = b->tune;
    return RM_OK;
}

int rm_batch_launch(rm_batch* b) {
    if (!b) return RM_ERR_INVALID;
    rm_ctx* ctx = b->ctx;
    if (int r = check_fresh(ctx, b->cloud, "rm_batch_launch")) return r;
    if (int r = set_device(ctx)) return r;
    timing_begin(ctx);
    RM_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    timing_end(ctx);
    return RM_OK;
}

int rm_batch_render_diff(rm_batch* b, float* rgb, float* depth) {
    if (!b) return RM_ERR_INVALID;
    rm_ctx* ctx = b->ctx;
    if (!rgb || !depth) return fail(ctx, RM_ERR_INVALID, "rm_batch_render_diff: null output");
    if (!b->hit)
        return fail(ctx, RM_ERR_INVALID, "rm_batch_render_diff: the batch has no hit buffer");
    if (int r = check_fresh(ctx, b->cloud, "rm_batch_render_diff")) return r;
    if (int r = set_device(ctx)) return r;
    DevCloud dc = dev_cloud(b->cloud);
    timing_begin(ctx);
    RM_HIP(ctx, hipGraphLaunch(b->exec, ctx->stream));
    for (const rm_view& v : b->views) {
        const DevView dv = dev_view(v.pose, v.fov_radians, v.width, v.height, 0, 0, 0);
        RM_HIP(ctx, rm::launch_diff_forward(ctx->stream, dc, dv,
                                            reinterpret_cast<const float4*>(b->rgba) + v.offset,
                                            b->hit + v.offset, rgb + 3 * v.offset,
                                            depth + v.offset));
    }
    timing_end(ctx);
    return RM_OK;
}

int rm_batch_render_adjoint(rm_batch* b, const float* xyz, const float* radii,
                            const int32_t* hit, const float* grad_rgb, const float* grad_depth,
                            float* g_xyz, float* g_r, float* g_rgb) {
    if (!b) return RM_ERR_INVALID;
    rm_ctx* ctx = b->ctx;
    if (!hit || !grad_rgb || !grad_depth || !g_xyz || !g_r || !g_rgb)
        return fail(ctx, RM_ERR_INVALID, "rm_batch_render_adjoint: null input");
    const rm_cloud* c = b->cloud;
    if (int r = check_cloud(ctx, c)) return r;
    if ((!xyz || !radii) && c->gen != c->tree->gen)
        return fail(ctx, RM_ERR_INVALID,
                    "rm_batch_render_adjoint: the cloud's tree was refit through another cloud; "
                    "pass the forward's centres and radii");
    if (b->laid_views != (int)b->views.size())
        return fail(ctx, RM_ERR_INTERNAL, "rm_batch_render_adjoint: view table not laid out");
    if (int r = set_device(ctx)) return r;
    const int nv = (int)b->views.size();
    std::vector<int64_t> cum((size_t)nv + 1);
    int64_t P = 0;
    for (int v = 0; v < nv; ++v) {
        cum[v] = P;
        P += (int64_t)b->views[v].width * b->views[v].height;
    }
    cum[nv] = P;
    if (P > INT32_MAX)
        return fail(ctx, RM_ERR_UNSUPPORTED, "rm_batch_render_adjoint: more than 2^31 pixels");
    int end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) <= c->n) ++end_bit;
    size_t temp_bytes = 0;
    RM_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, (int32_t*)nullptr,
                                                   (int32_t*)nullptr, (int32_t*)nullptr,
                                                   (int32_t*)nullptr, (int)P, 0, end_bit,
                                                   ctx->stream));
    const size_t a = align_up(sizeof(int32_t) * (size_t)P);
    const size_t tb = align_up(sizeof(int64_t) * cum.size());
    if (int r = ensure_ws(ctx, 4 * a + tb + align_up(temp_bytes))) return r;
    char* base = static_cast<char*>(ctx->ws);
    int32_t* keys = reinterpret_cast<int32_t*>(base);
    int32_t* vals = reinterpret_cast<int32_t*>(base + a);
    int32_t* keys_s = reinterpret_cast<int32_t*>(base + 2 * a);
    int32_t* vals_s = reinterpret_cast<int32_t*>(base + 3 * a);
    int64_t* d_cum = reinterpret_cast<int64_t*>(base + 4 * a);
    void* temp = base + 4 * a + tb;
    RM_HIP(ctx, hipMemcpyAsync(d_cum, cum.data(), sizeof(int64_t) * cum.size(),
                               hipMemcpyHostToDevice, ctx->stream));
    timing_begin(ctx);
    RM_HIP(ctx, rm::launch_adjoint_batch_keys(ctx->stream, b->d_views, d_cum, nv, P, hit,
                                              (int32_t)c->n, keys, vals));
    RM_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_s, vals, vals_s,
                                                   (int)P, 0, end_bit, ctx->stream));
    RM_HIP(ctx, rm::launch_adjoint_batch_reduce(ctx->stream, dev_cloud(c), xyz, radii, b->d_views,
                                                d_cum, nv, keys_s, vals_s, P, grad_rgb,
                                                grad_depth, g_xyz, g_r, g_rgb));
    timing_end(ctx);
    // the copy's source is this function's vector: wait for it before returning
    RM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return RM_OK;
}

int rm_batch_autotune(rm_batch* b, int sample_stride, rm_tune* chosen) {
    if (!b) return RM_ERR_INVALID;
    rm_ctx* ctx = b->ctx;
    if (sample_stride < 1) sample_stride = 1;
    if (int r = check_fresh(ctx, b->cloud, "rm_batch_autotune")) return r;
    if (int r = set_device(ctx)) return r;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        const hipError_t e = hipGetLastError();
        if (e0) (void)hipEventDestroy(e0);
        return hip_fail(ctx, e, "rm_batch_autotune: hipEventCreate");
    }
    auto measure = [&](const rm_tune& t, float& ms) -> int {
        b->tune = t;
        if (int r = batch_layout(b, sample_stride)) return r;
        float best = 1e30f;
        for (int rep = 0; rep < 5; ++rep) {
            RM_HIP(ctx, hipEventRecord(e0, ctx->stream));
            if (int r = batch_enqueue(b, ctx->stream)) return r;
            RM_HIP(ctx, hipEventRecord(e1, ctx->stream));
            RM_HIP(ctx, hipEventSynchronize(e1));
            float v = 0.0f;
            RM_HIP(ctx, hipEventElapsedTime(&v, e0, e1));
            best = std::min(best, v);
        }
        ms = best;
        return RM_OK;
    };
    const rm_tune incumbent = b->tune;
    rm_tune best = b->tune;
    float best_ms = 1e30f;
    const char* env = std::getenv("RM_TUNE_VERBOSE");
    const bool verbose = env && env[0] == '1';
    int r = RM_OK;
    auto consider = [&](rm_tune t) {
        if (r != RM_OK || !tune_valid(t)) return;
        if (t.warm && b->cloud->n < rm::kMarchK) return;
        float ms;
        r = measure(t, ms);
        if (r == RM_OK && verbose)
            std::fprintf(stderr,
                         "[rm autotune] tile %dx%d mode %d warm %d wpb %d bpc %d refill %d "
                         "stack %d exit %d: %.3f ms\n",
                         t.tile_w, t.tile_h, t.mode, t.warm, t.waves_per_block, t.blocks_per_cu,
                         t.refill, t.stack_depth, t.exit_depth, ms);
        // a candidate must beat the incumbent by 2 % (timing noise would otherwise flip
        // between equivalent settings from run to run)
        if (r == RM_OK && ms < best_ms * 0.98f) {
            best_ms = ms;
            best = t;
        }
    };
    // coordinate descent: packet size (a wavefront's 64 rays traverse together, or every ray
    // alone) x warm start x stack depth (per-lane stacks: entries kept in LDS, the rest in
    // global memory; fewer in LDS leave room for more resident wavefronts), then tile shape and
    // size (tiles above 64 pixels refill lanes locally), then block shape, then the refill
    // threshold
    rm_tune base = default_tune();
    const int need = rm::lane_stack_need(b->cloud->tree->depth);
    for (int mode = 0; mode < 2; ++mode)
        for (int warm = 0; warm < 2; ++warm)
            for (int depth : {0, 16, 8}) {
                if (depth != 0 && (mode == 0 || depth >= need)) continue;
                rm_tune t = base;
                t.mode = mode;
                t.warm = warm;
                t.stack_depth = depth;
                consider(t);
            }
    const int shapes[][2] = {{8, 8}, {16, 4}, {4, 16}, {16, 8}};
    rm_tune cur = best;
    for (auto& sh : shapes) {
        rm_tune t = cur;
        t.tile_w = sh[0];
        t.tile_h = sh[1];
        consider(t);
    }
    cur = best;
    for (int wpb : {1, 2, 4}) {
        rm_tune t = cur;
        t.waves_per_block = wpb;
        consider(t);
    }
    cur = best;
    for (int bpc : {4, 8}) {
        rm_tune t = cur;
        t.blocks_per_cu = bpc;
        consider(t);
    }
    cur = best;
    if (cur.tile_w * cur.tile_h > 64)
        for (int rf : {4, 16, 32}) {
            rm_tune t = cur;
            t.refill = rf;
            consider(t);
        }
    // early-miss frontier depth (pays where many rays leave the cloud between its parts)
    cur = best;
    for (int ed : {0, 4, 6, 8, 10}) {
        if (ed == cur.exit_depth) continue;
        rm_tune t = cur;
        t.exit_depth = ed;
        consider(t);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (r != RM_OK) {
        // the view table holds a sampled layout now: put the incumbent tune, its layout and
        // its graph back so that the batch stays launchable, and report the first error
        const std::string msg = ctx->err;
        b->tune = incumbent;
        if (batch_layout(b) == RM_OK) (void)batch_capture(b);
        ctx->err = msg;
        return r;
    }
    b->tune = best;
    if (int rr = batch_layout(b)) return rr;
    if (int rr = batch_capture(b)) return rr;
    if (chosen) *chosen = best;
    return RM_OK;
}

int rm_batch_destroy(rm_batch* b) {
    if (!b) return RM_OK;
    rm_ctx* ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    batch_drop_graph(b);
    if (b->d_views) (void)hipFree(b->d_views);
    if (b->d_queue) (void)hipFree(b->d_queue);
    if (b->d_mask) (void)hipFree(b->d_mask);
    if (b->d_ovf) (void)hipFree(b->d_ovf);
    if (b->d_frontier) (void)hipFree(b->d_frontier);
    delete b;
    return RM_OK;
}

int rm_render_batch(rm_ctx* ctx, const rm_cloud* cloud, const rm_view* views, int nviews,
                    float* rgba, int32_t* hit) {
    rm_batch* b = nullptr;
    int r = rm_batch_create(ctx, cloud, views, nviews, nullptr, rgba, hit, &b);
    if (r != RM_OK) return r;
    r = rm_batch_launch(b);
    int r2 = rm_batch_destroy(b);
    return r != RM_OK ? r : r2;
}

int rm_render_batch_noalloc(rm_ctx* ctx, const rm_cloud* cloud, const rm_view* views, int nviews,
                            float* rgba, int32_t* hit) {
    if (!ctx) return RM_ERR_INVALID;
    if (!cloud || !views || nviews <= 0 || !rgba)
        return fail(ctx, RM_ERR_INVALID, "rm_render_batch_noalloc: bad arguments");
    if (nviews > RM_NOALLOC_MAX_VIEWS)
        return fail(ctx, RM_ERR_UNSUPPORTED,
                    "rm_render_batch_noalloc: " + std::to_string(nviews) +
                        " views exceed RM_NOALLOC_MAX_VIEWS (" +
                        std::to_string(RM_NOALLOC_MAX_VIEWS) + ")");
    for (int i = 0; i < nviews; ++i) {
        const rm_view& v = views[i];
        if (v.width <= 0 || v.height <= 0 || v.offset < 0 ||
            !(v.fov_radians > 0.0f && v.fov_radians < 3.14159f))
            return fail(ctx, RM_ERR_INVALID,
                        "rm_render_batch_noalloc: bad view " + std::to_string(i));
    }
    if (int r = check_fresh(ctx, cloud, "rm_render_batch_noalloc")) return r;
    if (cloud->cull)
        return fail(ctx, RM_ERR_UNSUPPORTED,
                    "rm_render_batch_noalloc: culling data grow with the cloud and views; "
                    "use rm_render_batch for a culled cloud");
    if (!ctx->scratch)
        return fail(ctx, RM_ERR_HIP,
                    "rm_render_batch_noalloc: the scratch address lookup failed at context "
                    "creation");
    if (int r = set_device(ctx)) return r;
    rm_tune t = default_tune();
    if (cloud->n < rm::kMarchK) t.warm = 0;
    // The view table goes to the scratch array through kernel arguments, kViewChunk views per
    // launch: no staging buffer, and ordered on the stream like the march itself.
    DevView* d_views = reinterpret_cast<DevView*>(ctx->scratch + kScratchHeader);
    int64_t tiles = 0;
    rm::ViewChunk ch;
    for (int i0 = 0; i0 < nviews; i0 += rm::kViewChunk) {
        ch.first = i0;
        ch.count = std::min(rm::kViewChunk, nviews - i0);
        for (int j = 0; j < ch.count; ++j) {
            const rm_view& v = views[i0 + j];
            ch.v[j] = dev_view(v.pose, v.fov_radians, v.width, v.height, v.offset, t.tile_w, tiles);
            tiles += (int64_t)ch.v[j].tiles_x * ((v.height + t.tile_h - 1) / t.tile_h);
        }
        if (tiles >= (int64_t)UINT32_MAX)
            return fail(ctx, RM_ERR_UNSUPPORTED, "rm_render_batch_noalloc: too many tiles");
        RM_HIP(ctx, rm::launch_store_views(ctx->stream, ch, d_views));
    }
    rm::BatchParams p;
    p.cloud = dev_cloud(cloud);
    p.cull_mask = nullptr;
    p.cull_words = 0;
    p.views = d_views;
    p.nviews = nviews;
    p.tile_w = t.tile_w;
    p.tile_h = t.tile_h;
    p.refill_min = t.refill;
    p.total_tiles = tiles;
    p.queue = reinterpret_cast<unsigned long long*>(ctx->scratch);
    p.err = ctx->d_err;
    p.out.rgba = reinterpret_cast<float4*>(rgba);
    p.out.hit = hit;
    p.lane_lds = 0;          // the default tune keeps any per-lane stack whole in LDS
    p.lane_ovf = nullptr;
    float* d_frontier = reinterpret_cast<float*>(ctx->scratch + kScratchHeader + kScratchViews);
    p.nfrontier = t.exit_depth > 0 ? (1 << t.exit_depth) : 0;
    p.frontier = d_frontier;
    timing_begin(ctx);
    RM_HIP(ctx, rm::launch_reset_queue(ctx->stream, p.queue));
    if (t.exit_depth > 0)
        RM_HIP(ctx, rm::launch_frontier(ctx->stream, cloud->tree->nodes, t.exit_depth,
                                        d_frontier));
    RM_HIP(ctx, rm::launch_batch(ctx->stream, p, t.mode, t.warm, t.waves_per_block,
                                 ctx->num_cus * std::max(1, 32 / t.waves_per_block)));
    timing_end(ctx);
    return RM_OK;
}

// ---------------------------------------------------------------------------------------------
// differentiable form
// ---------------------------------------------------------------------------------------------
int rm_render_diff(rm_ctx* ctx, const rm_cloud* c, const float* pose, float fov, int W, int H,
                   float* rgb, float* depth, int32_t* hit) {
    if (!ctx) return RM_ERR_INVALID;
    if (!c || !pose || !rgb || !depth || !hit)
        return fail(ctx, RM_ERR_INVALID, "rm_render_diff: null input");
    if (W <= 0 || H <= 0) return fail(ctx, RM_ERR_INVALID, "rm_render_diff: empty image");
    if (!(fov > 0.0f && fov < 3.14159f)) return fail(ctx, RM_ERR_INVALID, "rm_render_diff: bad fov");
    if (int r = check_fresh(ctx, c, "rm_render_diff")) return r;
    if (int r = set_device(ctx)) return r;
    int64_t npix = (int64_t)W * H;
    if (int r = ensure_ws(ctx, align_up(sizeof(float4) * (size_t)npix))) return r;
    float4* marched = reinterpret_cast<float4*>(ctx->ws);
    DevView v = dev_view(pose, fov, W, H, 0, 0, 0);
    DevCloud dc = dev_cloud(c);
    timing_begin(ctx);
    RM_HIP(ctx, rm::launch_render_plain(ctx->stream, dc, v, rm::MarchOut{marched, hit}, ctx->d_err,
                                        c->mask, tree_levels(c->tree)));
    RM_HIP(ctx, rm::launch_diff_forward(ctx->stream, dc, v, marched, hit, rgb, depth));
    timing_end(ctx);
    return RM_OK;
}

int rm_render_adjoint(rm_ctx* ctx, const rm_cloud* c, const float* pose, float fov, int W, int H,
                      const int32_t* hit, const float* grad_rgb, const float* grad_depth,
                      float* g_xyz, float* g_r, float* g_rgb) {
    if (!ctx) return RM_ERR_INVALID;
    if (!c || !pose || !hit || !grad_rgb || !grad_depth || !g_xyz || !g_r || !g_rgb)
        return fail(ctx, RM_ERR_INVALID, "rm_render_adjoint: null input");
    if (W <= 0 || H <= 0) return fail(ctx, RM_ERR_INVALID, "rm_render_adjoint: empty image");
    if (!(fov > 0.0f && fov < 3.14159f))
        return fail(ctx, RM_ERR_INVALID, "rm_render_adjoint: bad fov");
    if (int r = check_fresh(ctx, c, "rm_render_adjoint")) return r;
    if (int r = set_device(ctx)) return r;
    int64_t npix = (int64_t)W * H;
    if (npix > INT32_MAX) return fail(ctx, RM_ERR_UNSUPPORTED, "rm_render_adjoint: image too large");
    int end_bit = 1;
    while (end_bit < 32 && ((int64_t)1 << end_bit) <= c->n) ++end_bit;
    size_t temp_bytes = 0;
    RM_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, (int32_t*)nullptr,
                                                   (int32_t*)nullptr, (int32_t*)nullptr,
                                                   (int32_t*)nullptr, (int)npix, 0, end_bit,
                                                   ctx->stream));
    size_t a = align_up(sizeof(int32_t) * (size_t)npix);
    if (int r = ensure_ws(ctx, 4 * a + align_up(temp_bytes))) return r;
    char* base = static_cast<char*>(ctx->ws);
    int32_t* keys = reinterpret_cast<int32_t*>(base);
    int32_t* vals = reinterpret_cast<int32_t*>(base + a);
    int32_t* keys_s = reinterpret_cast<int32_t*>(base + 2 * a);
    int32_t* vals_s = reinterpret_cast<int32_t*>(base + 3 * a);
    void* temp = base + 4 * a;
    DevView v = dev_view(pose, fov, W, H, 0, 0, 0);
    timing_begin(ctx);
    RM_HIP(ctx, rm::launch_adjoint_keys(ctx->stream, hit, npix, (int32_t)c->n, keys, vals));
    RM_HIP(ctx, hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys, keys_s, vals, vals_s,
                                                   (int)npix, 0, end_bit, ctx->stream));
    RM_HIP(ctx, rm::launch_adjoint_reduce(ctx->stream, dev_cloud(c), v, keys_s, vals_s, npix,
                                          grad_rgb, grad_depth, g_xyz, g_r, g_rgb));
    timing_end(ctx);
    return RM_OK;
}

}  // extern "C"

/*
 * librm_hip.so — MI355X (gfx950) kd-tree kNN, sphere radii and point-cloud sphere ray marching.
 *
 * C ABI. Every data pointer is a DEVICE pointer (the reference API takes
 * thrust::device_vector for all of these), except rm_view arrays, poses and tuning
 * structs, which are host memory. All work of a context is queued on the context's HIP stream;
 * functions that return results to the host (rm_tree_build, the autotuner) synchronise it.
 * Status: 0 on success, a negative RM_ERR_* code on failure; rm_last_error(ctx) then holds
 * a message. No function calls exit() or throws across the ABI.
 *
 * Threading: one rm_ctx per host thread; calls on one ctx are serialised by its stream. There
 * is no global device state: any number of trees, clouds and batches may coexist.
 * Lifetimes: a batch uses its context and cloud, a cloud its tree; destroy a batch before its
 * context and cloud, and a cloud before its tree.
 */
#ifndef RM_H
#define RM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RM_OK 0
#define RM_ERR_INVALID (-1)     /* bad argument (null, size, range) */
#define RM_ERR_HIP (-2)         /* a HIP runtime call failed */
#define RM_ERR_UNSUPPORTED (-3) /* e.g. k above RM_MAX_K */
#define RM_ERR_NOMEM (-4)       /* device allocation failed */
#define RM_ERR_INTERNAL (-5)    /* a device-side invariant failed (traversal stack overflow) */

#define RM_MAX_K 32             /* largest k served by rm_knn */
#define RM_MARCH_K 10           /* neighbours per march step (src/ray_marching.h:17) */

typedef struct rm_ctx rm_ctx;
typedef struct rm_tree rm_tree;
typedef struct rm_cloud rm_cloud;
typedef struct rm_batch rm_batch;

/* ---- context -------------------------------------------------------------------------- */

/* Create a context on HIP device `device` with its own non-blocking stream. */
int rm_ctx_create(int device, rm_ctx** out);
int rm_ctx_destroy(rm_ctx* ctx);
/* Queue work on a caller-owned stream instead (e.g. torch's current stream); NULL selects the
 * device's null stream. Batch graphs are captured on the context's own stream either way. */
int rm_ctx_set_stream(rm_ctx* ctx, void* hip_stream);
/* The stream the context queues work on now (its own stream until rm_ctx_set_stream). */
int rm_ctx_get_stream(const rm_ctx* ctx, void** hip_stream);
int rm_ctx_synchronize(rm_ctx* ctx);
const char* rm_last_error(const rm_ctx* ctx);
/* Device time (ms) of the last call of rm_knn / rm_radius_search / rm_radii / rm_render /
 * rm_render_diff / rm_batch_launch / rm_batch_render_diff, measured with HIP events on the
 * stream its kernels ran on (0 if events are off). */
int rm_ctx_enable_timing(rm_ctx* ctx, int on);
float rm_ctx_last_kernel_ms(const rm_ctx* ctx);

/* ---- kd-tree --------------------------------------------------------------------------- */

/* Build the search tree over n points given as n x 4 floats (x, y, z, w; w ignored, stride
 * 16 B). Replaces kdtree::KDTreeFlann::Build (src/kdtree/kdtree_flann.cu:101-117) and
 * KDTreeCuda3dIndex::buildIndex (third_party/flann/algorithms/kdtree_cuda_3d_index.h:124-141).
 * leaf_max: points per leaf (the reference uses 64, kdtree_cuda_3d_index.h:55-61; 0 picks
 * the library default). The tree copies the points; the input may be freed afterwards.
 * Trees of more than 2^26 nodes (over a billion points at the default 16 per leaf, 2^26 points
 * at one per leaf) are refused with RM_ERR_UNSUPPORTED. */
int rm_tree_build(rm_ctx* ctx, const float* xyzw, int64_t n, int leaf_max, rm_tree** out);
int rm_tree_destroy(rm_tree* tree);
int rm_tree_info(const rm_tree* tree, int64_t* n, int64_t* n_nodes, int* depth, int* leaf_max);

/* Exact k nearest neighbours of m queries (m x 4 floats, w ignored). Outputs m x k, row-major,
 * ascending by (squared distance, original index), padded with -1 / +INF when the tree holds
 * fewer than k points. Replaces KDTreeFlann::Search / SearchKNN
 * (src/kdtree/kdtree_flann.cu:5-28, 65-80) and KDTreeCuda3dIndex::knnSearchGpu
 * (kdtree_cuda_3d_index.cu:261-399) including the index remap map_indices (:160-173).
 * Squared distance is fma(dz, dz, fma(dy, dy, dx*dx)) in float (each step rounded once). */
int rm_knn(rm_ctx* ctx, const rm_tree* tree, const float* q_xyzw, int64_t m, int k,
           int32_t* idx, float* d2);

/* Radius search: up to max_nn neighbours with d2 < radius^2 (strict, as the reference result
 * set KnnRadiusResultSet, util/cuda/result_set.h:372-476) per query, ascending by
 * (d2, index), padded with -1 / +INF; counts[i] = number found (<= max_nn, may be NULL).
 * max_nn == 0 is the counting mode (CountingRadiusResultSet with no cap, result_set.h:225-267,
 * kdtree_cuda_3d_index.cu:436-452): counts[i] = number of points with d2 < radius^2, uncapped;
 * counts is required and idx, d2 are not touched (may be NULL).
 * Replaces KDTreeFlann::SearchRadius (src/kdtree/kdtree_flann.cu:82-99) and radiusSearchGpu
 * (kdtree_cuda_3d_index.cu:410-557). 0 <= max_nn <= RM_MAX_K. */
int rm_radius_search(rm_ctx* ctx, const rm_tree* tree, const float* q_xyzw, int64_t m,
                     float radius, int max_nn, int32_t* idx, float* d2, int32_t* counts);

/* Per-point sphere radius: r_i = sqrtf(second smallest squared distance from point i to the
 * cloud, self included), i.e. the distance to the nearest other point (0 for duplicates).
 * r_clamped_i = (max_radius > 0 && r_i > max_radius) ? max_radius : r_i. Both outputs are in
 * the original point order; either may be NULL. Replaces main.cu:99-149 (k = 2 search, the
 * host sqrt loop :110-112, and the clamp :144-149) with two kernels: the search writes the
 * square roots in the tree's leaf order, then a reorder writes both outputs in point order
 * (coalesced). */
int rm_radii(rm_ctx* ctx, const rm_tree* tree, float max_radius, float* r_unclamped,
             float* r_clamped);

/* ---- cloud + march --------------------------------------------------------------------- */

/* Bind colours and radii to a tree for marching. xyzw: n x 4 (the same points the tree was
 * built from, original order), colors: n x 4 floats (r, g, b in [0,1], a ignored),
 * radii: n floats (already clamped). The cloud copies what it needs.
 * Replaces PointcloudRayMarcher::get_instance / its constructor
 * (src/ray_marching.cu:159-200); unlike the reference it is not a singleton. */
int rm_cloud_create(rm_ctx* ctx, const rm_tree* tree, const float* xyzw, const float* colors,
                    const float* radii, int64_t n, rm_cloud** out);
int rm_cloud_destroy(rm_cloud* cloud);
/* Replace the radii (n floats, original order) of a cloud, e.g. after optimisation. */
int rm_cloud_set_radii(rm_ctx* ctx, rm_cloud* cloud, const float* radii);
/* Replace the colours (n x 4 floats) of a cloud. */
int rm_cloud_set_colors(rm_ctx* ctx, rm_cloud* cloud, const float* colors);
/* Frustum culling (enable = 1) or none (0, the default, as the reference). With culling, each
 * view marches over the points whose centres are inside its frustum: inside the four side
 * planes through the camera origin and the image corner directions, beyond the near plane (0.1)
 * and before the far plane (100) along the forward axis, i.e. the predicate is_inside of
 * src/ray_marching.cu:231-258 with the edge points of :112-124, which the reference computes but
 * has commented out. The kNN-10 of every march step is then exact over the inside points.
 * Applies to rm_render, rm_render_diff and to batches created after the call (a batch keeps the
 * setting it was created with). Enabling allocates one view's culling data on the cloud, which
 * single-view renders share: culled rm_render / rm_render_diff calls on one cloud must be
 * ordered (same stream), while each batch has its own (a culled batch holds at most 65535
 * views: its per-view node pass uses the grid's second dimension). */
int rm_cloud_set_culling(rm_ctx* ctx, rm_cloud* cloud, int enable);
/* Move the points of a cloud (xyz: n x 3 floats, original order) and refit the boxes of the
 * tree it was created from, keeping the topology. Search stays exact because every bound is
 * recomputed from the moved points. The tree is modified in place: other clouds created on the
 * same tree no longer match its points, and their march calls (rm_render, rm_render_diff,
 * rm_render_adjoint, the batch calls) fail with RM_ERR_INVALID until they are refit too. */
int rm_cloud_refit(rm_ctx* ctx, rm_cloud* cloud, const float* xyz);

/* Render one view. pose: 16 floats, column-major camera-to-world (OpenGL, -z forward), as the
 * glm::mat4 of render_to_texture. fov_radians: horizontal field of view. Output rgba: H x W x 4
 * floats, row 0 = bottom (GL texture order); a hit writes (r, g, b, depth = ray length), a
 * miss (0, 0, 0, 0). hit: H x W int32 original point index, -1 on a miss (may be NULL).
 * Replaces PointcloudRayMarcher::render_to_texture (src/ray_marching.cu:93-154) and the
 * render kernel (:63-91, 202-228, 260-292). */
int rm_render(rm_ctx* ctx, const rm_cloud* cloud, const float* pose, float fov_radians, int W,
              int H, float* rgba, int32_t* hit);

/* ---- batched march (no counterpart in the reference) ------------------------------------ */

typedef struct rm_view {
    float pose[16];    /* column-major camera-to-world */
    float fov_radians;
    int32_t width, height;
    int32_t reserved;
    int64_t offset;    /* first pixel of this view in the concatenated outputs */
} rm_view;

/* Launch parameters picked by the autotuner. */
typedef struct rm_tune {
    int32_t tile_w, tile_h;   /* pixels per claimed tile: tile_w * tile_h, a multiple of 64
                                 (<= 4096); a wavefront keeps 64 rays of its tile in flight */
    int32_t mode;             /* packet size: 0 = 64-ray packets (a wavefront's rays traverse
                                 the tree together, wave-uniform stack), 1 = 1-ray packets
                                 (every lane traverses alone with its own stack) */
    int32_t warm;             /* 1 = warm-start bound from the previous step's neighbours */
    int32_t waves_per_block;  /* 1, 2 or 4 */
    int32_t blocks_per_cu;    /* persistent grid = CUs * blocks_per_cu (0 = 32 / waves_per_block,
                                 8 wavefronts per SIMD) */
    int32_t refill;           /* lanes whose ray ended take the tile's next pixels once at
                                 least this many (1..64) are idle */
    int32_t stack_depth;      /* mode 1: entries of each lane's traversal stack kept in LDS
                                 (1..RM_MAX_STACK_DEPTH), the deeper ones going to a global
                                 area the batch allocates; 0, or at least the tree's need
                                 (depth + 3), keeps the whole stack in LDS. Fewer LDS entries
                                 leave room for more resident wavefronts. Mode 0 ignores it.
                                 Results are identical for every value. */
    int32_t exit_depth;       /* early miss of views whose camera is outside the box of all
                                 centres: a ray ends as a miss once it has left, for good, every
                                 box of the tree's 2^exit_depth slots at that depth, each dilated
                                 by the largest radius and a margin (0..RM_MAX_EXIT_DEPTH; 0: the
                                 box of the whole cloud, which views from inside always use).
                                 Pays where many rays pass between the cloud's parts (an object
                                 seen from around it). Results are identical for every value. */
} rm_tune;
#define RM_MAX_STACK_DEPTH 64
#define RM_MAX_EXIT_DEPTH 10

/* A batch of views marched by one persistent kernel over a shared tile queue, launched as a
 * captured HIP graph. rgba / hit (hit may be NULL) are the concatenated outputs, sized by the
 * views' offsets and sizes; they are bound at creation (the graph captures them). tune may be
 * NULL: the default tune. */
int rm_batch_create(rm_ctx* ctx, const rm_cloud* cloud, const rm_view* views, int nviews,
                    const rm_tune* tune, float* rgba, int32_t* hit, rm_batch** out);
/* Time candidate rm_tune settings on every sample_stride-th view (whole views, so the sample
 * keeps each view's screen locality and the mix of view sizes) and keep the
 * fastest (re-captures the graph). chosen may be NULL. On an error the batch keeps the tune,
 * layout and graph it had before the call. */
int rm_batch_autotune(rm_batch* batch, int sample_stride, rm_tune* chosen);
int rm_batch_get_tune(const rm_batch* batch, rm_tune* tune);
int rm_batch_launch(rm_batch* batch);
int rm_batch_destroy(rm_batch* batch);
/* One-call form: create with the default tune, launch once, destroy. */
int rm_render_batch(rm_ctx* ctx, const rm_cloud* cloud, const rm_view* views, int nviews,
                    float* rgba, int32_t* hit);

/* Allocation-free form of rm_render_batch. It has no counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). Same outputs as rm_render_batch (default tune, one
 * launch of the persistent march, no graph), but the tile-queue counter, the view table and the
 * default tune's early-miss frontier live in a static device array of the library, one per GPU,
 * sized for RM_NOALLOC_MAX_VIEWS views (256 B + 160 B per view + 2,080 B, 162.3 KiB); it calls
 * no hipMalloc, hipFree or allocating library
 * routine. More views than that, or a cloud with culling enabled (its culling data grow with
 * the cloud), are rejected with RM_ERR_UNSUPPORTED. Every call on a GPU shares that GPU's
 * array and overwrites it: calls ordered on one stream are safe, calls that overlap on one GPU
 * (other streams or contexts) race. */
#define RM_NOALLOC_MAX_VIEWS 1024
int rm_render_batch_noalloc(rm_ctx* ctx, const rm_cloud* cloud, const rm_view* views, int nviews,
                            float* rgba, int32_t* hit);

/* ---- differentiable march (no counterpart in the reference) ----------------------------- */

/* Forward of the differentiable form for one view: the plain march picks the hit point h of
 * each pixel (hit, -1 on a miss); then rgb = colour of h (H x W x 3) and depth = analytic
 * entry distance of the ray into the sphere (centre, radius) of h (H x W), or the marched
 * distance where the ray only grazes that sphere (discriminant <= 0). Misses give 0. */
int rm_render_diff(rm_ctx* ctx, const rm_cloud* cloud, const float* pose, float fov_radians,
                   int W, int H, float* rgb, float* depth, int32_t* hit);

/* Forward of the differentiable form for every view of a batch (no counterpart in the
 * reference): one launch of the batch's march (its rgba / hit buffers receive the marched
 * output; hit must have been given at creation), then rgb (total pixels x 3) and depth (total
 * pixels) as rm_render_diff defines them, at each view's offset. Equal to rm_render_diff view
 * by view. */
int rm_batch_render_diff(rm_batch* batch, float* rgb, float* depth);

/* Adjoint of rm_render_diff with the hit map held fixed, at the cloud's current centres and
 * radii. grad_rgb: H x W x 3, grad_depth:
 * H x W upstream gradients. Accumulates (adds) into g_xyz (n x 3), g_r (n), g_rgb (n x 3) in
 * the original point order. Per-point sums are formed by sorting the hit pixels by point
 * index and reducing each run, not by atomics on the point arrays. */
int rm_render_adjoint(rm_ctx* ctx, const rm_cloud* cloud, const float* pose, float fov_radians,
                      int W, int H, const int32_t* hit, const float* grad_rgb,
                      const float* grad_depth, float* g_xyz, float* g_r, float* g_rgb);

/* Adjoint of rm_batch_render_diff over every view of the batch at once (no counterpart in the
 * reference), with the hit map held fixed. hit (total pixels), grad_rgb (total pixels x 3) and
 * grad_depth (total pixels) are laid out as the batch's outputs (each view at its offset; views
 * may differ in size). xyz (n x 3) and radii (n) are the centres and radii the forward marched,
 * so an earlier forward can be differentiated after the cloud has moved; NULL takes the cloud's
 * current ones. Accumulates into g_xyz (n x 3), g_r (n), g_rgb (n x 3), original point order:
 * one sort of all views' hit pixels by point, then one reduction per point in pixel order
 * (deterministic, no atomics). At most 2^31 pixels. */
int rm_batch_render_adjoint(rm_batch* batch, const float* xyz, const float* radii,
                            const int32_t* hit, const float* grad_rgb, const float* grad_depth,
                            float* g_xyz, float* g_r, float* g_rgb);

#ifdef __cplusplus
}
#endif
#endif /* RM_H */

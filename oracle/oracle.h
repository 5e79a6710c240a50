/*
 * CPU oracle for the point-cloud ray-marching hot path.
 *
 * TEST INFRASTRUCTURE ONLY. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or the timed CPU baseline). The product path (librm_hip.so) never links,
 * loads or calls it.
 *
 * It restates, in plain C++17 + OpenMP, the reference algorithm:
 *   - kd-tree build rule     third_party/flann/algorithms/kdtree_cuda_builder.h:215-223, 297-349
 *   - stackless traversal    third_party/flann/algorithms/kdtree_cuda_3d_index.cu:52-129
 *   - distance expressions   kdtree_cuda_3d_index.cu:176-193, util/cutil_math.h:1127-1130
 *   - top-k result set       util/cuda/result_set.h:121-223 (with a (d2, index) total order)
 *   - radius rule            main.cu:99-149
 *   - ray generation/march   src/ray_marching.cu:63-91, 202-228, 260-292; src/common.h:4-11
 *
 * Pinning (what checks each function):
 *   orc_tree_*      structural invariants (tests/test_oracle.py), traversal exactness vs brute force
 *   orc_knn         brute force (orc_knn_brute) on seeded clouds, closed-form lattices,
 *                   FLANN CPU kd-tree (oracle/_ref) fixtures
 *   orc_knn_brute   textbook definition, numpy argsort cross-check on tiny inputs
 *   orc_radii       lattice closed forms (r = spacing), duplicates (r = 0), brute force
 *   orc_render      single-sphere analytic depth, tiny scenes vs brute-force SDF (kNN-10 = all
 *                   points), miss/background, start-inside case
 *   orc_render_grad central finite differences of orc_render_loss_fixed (fp64)
 *   orc_subsample_indices  is a permutation; the generator's standard-mandated 10000th value;
 *                   the shuffle itself is libstdc++'s (parity with the reference's GCC 9 build
 *                   unpinned)
 *
 * All float arithmetic is IEEE single precision, built with -ffp-contract=off so no
 * multiply-add is fused implicitly; the squared distance uses explicit fused multiply-adds,
 * fma(dz, dz, fma(dy, dy, dx * dx)). The HIP kernels use the same expressions.
 */
#ifndef RM_ORACLE_H
#define RM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Counters of the reference traversal (kdtree_cuda_3d_index.cu:52-129):
 *   c = candidate distance evaluations (points of visited leaves),
 *   v = loop iterations, each one a node-record load (SplitInfo 8 B + child1 4 B + parent 4 B),
 *   a = far-child AABB tests (backtrack iterations),
 *   q = queries, s = march steps (kNN queries issued by the march). */
typedef struct orc_counters {
    uint64_t c, v, a, q, s;
} orc_counters;

/* Build a kd-tree by the reference rule. xyzw: n x 4 floats (stride 16 B, w ignored). */
void* orc_tree_build(const float* xyzw, int64_t n, int leaf_max);
void orc_tree_free(void* tree);
int64_t orc_tree_num_nodes(const void* tree);
int orc_tree_depth(const void* tree);
/* Copy out node arrays (each of length num_nodes) and the leaf-order permutation (n). */
void orc_tree_export(const void* tree, int32_t* child1, int32_t* parent, int32_t* begin,
                     int32_t* end, float* lo /* nn x 3 */, float* hi /* nn x 3 */,
                     int32_t* order);

/* Exact kNN by the reference traversal with a (d2, index) total order.
 * q: m x 4 floats. idx/d2: m x k, ascending, padded with -1 / +INF. */
void orc_knn(const void* tree, const float* q_xyzw, int64_t m, int k, int32_t* idx, float* d2,
             orc_counters* cnt, int nthreads);

/* Brute force kNN, same total order. */
void orc_knn_brute(const float* xyzw, int64_t n, const float* q_xyzw, int64_t m, int k,
                   int32_t* idx, float* d2, int nthreads);

/* Radius rule of main.cu:99-149: r = sqrtf(second-smallest d2 including self);
 * clamped copy: r > max_radius ? max_radius : r, only when max_radius > 0. */
void orc_radii(const void* tree, float max_radius, float* r_unclamped, float* r_clamped,
               orc_counters* cnt, int nthreads);

/* Subsample order of main.cu:75-87: 0..n-1 shuffled by std::shuffle(std::mt19937(42));
 * the -s subsample is the first mp entries. */
void orc_subsample_indices(int64_t n, int32_t* out);
/* 10000th output of a default-seeded std::mt19937 (4123659995 by the C++ standard). */
uint32_t orc_mt19937_10000th(void);

/* Host-side pane distance W / (2 tan(fov/2)) (src/ray_marching.cu:74), float, libm tanf. */
float orc_pane_distance(int W, float fov_radians);

/* Render one view by the reference march.
 * pose: 16 floats column-major camera->world (OpenGL, -z forward).
 * radii: n, colors: n x 4 floats (rgb in [0,1], a ignored).
 * pix: optional list of npix linear pixel ids y*W+x (NULL = all W*H pixels, npix ignored).
 * rgba: npix x 4 (a = depth, 0 on a miss), hit: npix (-1 on a miss). */
void orc_render(const void* tree, const float* radii, const float* colors, const float* pose,
                float pane, int W, int H, const int32_t* pix, int64_t npix, float* rgba,
                int32_t* hit, orc_counters* cnt, int nthreads);
/* Same, also reporting the number of march steps of every pixel (steps: npix). */
void orc_render_steps(const void* tree, const float* radii, const float* colors,
                      const float* pose, float pane, int W, int H, const int32_t* pix,
                      int64_t npix, float* rgba, int32_t* hit, int32_t* steps,
                      orc_counters* cnt, int nthreads);

/* Frustum predicate as src/ray_marching.cu:112-124 (corner directions) and :231-258
 * (is_inside, commented out in the reference) define it: inside[i] = 1 iff point i (xyzw,
 * n x 4) is inside the view frustum. A culled render is orc_render over the inside points. */
void orc_frustum_inside(const float* pose, float pane, int W, int H, const float* xyzw,
                        int64_t n, uint8_t* inside);

/* Differentiable form, fp64 reference. With the hit map fixed, per hit pixel:
 *   depth = analytic ray/sphere entry distance of the hit sphere (centre c, radius r),
 *   colour = rgb of the hit point.
 * Loss = sum_pix w_rgb * |colour - target_rgb|^2 + w_depth * |depth - target_depth|
 * over hit pixels whose discriminant is positive. Rays are regenerated from pose/pane
 * (same float expressions as the march, then promoted to double).
 * Returns the loss; grads (double) are accumulated into g_xyz (n x 3), g_r (n), g_rgb (n x 3). */
double orc_loss_grad_fixed(const double* xyz, const double* r, const double* rgb, int64_t n,
                           const float* pose, float pane, int W, int H, const int32_t* hit,
                           const float* target_rgb, const float* target_depth, double w_rgb,
                           double w_depth, double* g_xyz, double* g_r, double* g_rgb);

#ifdef __cplusplus
}
#endif
#endif

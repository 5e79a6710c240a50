// Driver for the FLANN 1.9.1 CPU kd-tree vendored in the reference
// (third_party/flann/flann.hpp, algorithms/kdtree_index.h). TEST INFRASTRUCTURE ONLY.
// Compiled by oracle/Makefile from the headers where they lie in the reference checkout,
// into oracle/_ref/libflannref.so; used only to cross-check the oracle's kNN.
//
// Exact mode: one tree (KDTreeIndexParams(1)) and unlimited checks (kdtree_index.h:228-245,
// 528-540). FLANN_FIRST_MATCH breaks ties among inserted candidates by index
// (util/result_set.h:158-162); a candidate tying the current worst of a full set is still
// rejected (:153), so ties at the k-th place depend on the visit order.
#include <cstdint>
#include <flann/flann.hpp>
#include <flann/algorithms/kdtree_index.h>

extern "C" int flannref_knn(const float* xyzw, int64_t n, const float* q_xyzw, int64_t m,
                            int k, int32_t* idx, float* d2) {
    flann::Matrix<float> data(const_cast<float*>(xyzw), (size_t)n, 3, 4 * sizeof(float));
    flann::Matrix<float> queries(const_cast<float*>(q_xyzw), (size_t)m, 3, 4 * sizeof(float));
    // The vendored index factory (algorithms/all_indices.h) only knows the CUDA index, so the
    // CPU kd-tree index is constructed directly.
    flann::KDTreeIndex<flann::L2<float>> index(data, flann::KDTreeIndexParams(1));
    index.buildIndex();
    flann::Matrix<int> indices(reinterpret_cast<int*>(idx), (size_t)m, (size_t)k);
    flann::Matrix<float> dists(d2, (size_t)m, (size_t)k);
    flann::SearchParams params(flann::FLANN_CHECKS_UNLIMITED, 0.0f, true);
    return index.knnSearch(queries, indices, dists, (size_t)k, params);
}

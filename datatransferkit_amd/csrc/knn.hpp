// k-nearest-neighbour search over a linear BVH (one GPU).
//
// Replaces the reference's ArborX::DistributedTree construction and
// Nearest<Point> query (DTK_MovingLeastSquaresOperator_def.hpp:47-60,
// DTK_DetailsMovingLeastSquaresOperatorImpl.hpp:28-47) on a one-rank
// communicator. Results are the min(k, n_src) nearest sources of each target
// in ascending (distance, source index) order; under float geometry the
// distance is the float distance of the float-rounded points, formed as
// tmp = b - a; s += tmp*tmp over x, y, z, then sqrt, with no FMA.
#ifndef DTKX_KNN_HPP
#define DTKX_KNN_HPP

#include "common.hpp"

namespace dtkx
{

struct KnnInfo
{
    int tree_height = 0; // internal levels on the longest root-leaf path
    int stack_used = 0;  // traversal stack capacity chosen
};

// Results of a search. indices: [n_tgt][kk], kk = min(k, n_src); dist
// (nullable): matching distances. query_order (nullable): the targets in
// the Morton order the queries ran in (row r of a Morton layout holds
// target query_order[r]). src_perm / src_pos (nullable): the sources in
// Morton order and its inverse. With morton_layout, row r of indices/dist
// is the r-th query and every index is the source's Morton position
// (src_pos[source]); otherwise row t is target t and indices are source
// indices. Either way a row is ordered by (distance, source index).
struct KnnOutput
{
    int32_t *indices = nullptr;
    double *dist = nullptr;
    int32_t *query_order = nullptr;
    int32_t *src_perm = nullptr;
    int32_t *src_pos = nullptr;
    bool morton_layout = false;
};

// src/tgt: float4 (geometry FLOAT) or double4 (geometry DOUBLE) device
// points.
void knn_search( int geometry, void const *src, int64_t n_src,
                 void const *tgt, int64_t n_tgt, int k, KnnOutput const &out,
                 KnnInfo *info, hipStream_t s );

// Largest k the query kernels are instantiated for.
constexpr int knn_max_k = 64;

} // namespace dtkx

#endif

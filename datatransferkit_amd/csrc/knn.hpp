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

// src/tgt: float4 (geometry FLOAT) or double4 (geometry DOUBLE) device
// points. indices: [n_tgt][kk], kk = min(k, n_src). dist (nullable):
// matching distances. query_order (nullable): receives the targets in the
// Morton order the queries were processed in (a permutation of 0..n_tgt-1).
void knn_search( int geometry, void const *src, int64_t n_src,
                 void const *tgt, int64_t n_tgt, int k, int32_t *indices,
                 double *dist, int32_t *query_order, KnnInfo *info,
                 hipStream_t s );

// Largest k the query kernels are instantiated for.
constexpr int knn_max_k = 64;

} // namespace dtkx

#endif

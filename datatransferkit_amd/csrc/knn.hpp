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

#include <vector>

namespace dtkx
{

struct KnnInfo
{
    int n_overflow = 0; // queries re-run with the global traversal stack
    // (optional) device counter the re-run count is added to instead,
    // leaving n_overflow alone: the search then never waits on the stream
    int32_t *n_overflow_dev = nullptr;
    // (optional) recorded on the stream around the query kernels
    hipEvent_t query_begin = nullptr, query_end = nullptr;
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
    // nullable, Morton layout: each row's largest neighbour distance (the
    // k-th, as the search computes it)
    float *row_max = nullptr;
    // nullable: the inverse of query_order (query_pos[target] = row)
    int32_t *query_pos = nullptr;
    // nullable: sorted_src[position] = src4[source], a double4 copy of the
    // sources in Morton order (src4 may be the search's own points)
    double4 const *src4 = nullptr;
    double4 *sorted_src = nullptr;
};

// src/tgt: float4 (geometry FLOAT) or double4 (geometry DOUBLE) device
// points.
void knn_search( int geometry, void const *src, int64_t n_src,
                 void const *tgt, int64_t n_tgt, int k, KnnOutput const &out,
                 KnnInfo *info, hipStream_t s );

// Many independent problems in one build and one query pass per launch
// group. Problem p owns sources [src_off[p], src_off[p + 1]) and targets
// [tgt_off[p], tgt_off[p + 1]) of the concatenated arrays (every problem has
// at least one source). Each problem's hierarchy is built over its own
// sorted range inside one set of kernels, and each query starts at its
// problem's root. Output rows are ordered by (row_key[problem], Morton code
// of the target in its problem's box); row r holds kk_prob[p] neighbours at
// CSR offset row_off[r] = row_base[p] + (r - row_start[p]) kk_prob[p], as
// global source positions (the sorted order is problem-major, so problem p's
// positions are [src_off[p], src_off[p + 1]) as well). Ties are broken by
// the source index, so each row equals the one-problem search.
struct KnnBatch
{
    int n_problems = 0;
    int32_t const *src_off = nullptr;   // device [P + 1]
    int32_t const *src_prob = nullptr;  // device [n_src]
    int32_t const *tgt_prob = nullptr;  // device [n_tgt]
    int32_t const *row_key = nullptr;   // device [P]
    int key_bits = 1;                   // row_key values < 2^key_bits
    int32_t const *kk_prob = nullptr;   // device [P], min(k_p, n_src_p)
    int64_t const *row_base = nullptr;  // device [P]
    int32_t const *row_start = nullptr; // device [P]
    int64_t nnz = 0;                    // row_off[n_tgt]
    // query launches over row ranges: {row_begin, row_end, max k in range}
    struct Launch
    {
        int64_t begin, end;
        int kmax;
    };
    std::vector<Launch> launches;
};

struct KnnBatchOutput
{
    int32_t *order = nullptr;    // [n_tgt] row -> target
    int32_t *src_perm = nullptr; // [n_src] position -> source
    int32_t *src_pos = nullptr;  // [n_src] source -> position
    int64_t *row_off = nullptr;  // [n_tgt + 1]
    int32_t *indices = nullptr;  // [nnz]
    float *row_max = nullptr;    // nullable [n_tgt]: largest distance per row
};

void knn_search_batched( int geometry, void const *src, int64_t n_src,
                         void const *tgt, int64_t n_tgt, KnnBatch const &b,
                         KnnBatchOutput const &out, KnnInfo *info,
                         hipStream_t s );

// Largest k the query kernels are instantiated for.
constexpr int knn_max_k = 64;

} // namespace dtkx

#endif

// CRS coefficient apply of the MLS operator and its transpose.
//
// apply:     y[f][i] = sum_j c(i,j) x[f][idx(i,j)], the row sum starting at 0
//            and running j ascending (computeTargetValues,
//            DTK_DetailsMovingLeastSquaresOperatorImpl.hpp:58-66), after an
//            in-kernel gather that replaces the reference's fetch()
//            (DTK_DetailsNearestNeighborOperatorImpl.hpp:152-192) on one rank.
// transpose: xbar[f][s] = sum over entries e with idx(e) = s, in ascending
//            entry order, of c(e) ybar[f][row(e)]; gathered through a
//            column index (CSC) built once, so no atomics and a fixed order.
#ifndef DTKX_APPLY_HPP
#define DTKX_APPLY_HPP

#include "common.hpp"

namespace dtkx
{

// order (nullable): the rows are processed in this order (the targets'
// Morton order from the kNN), which only changes which rows run together.
void crs_apply_f64( int32_t const *idx, double const *coeffs, int64_t n_tgt,
                    int kk, double const *x, int64_t ld_src, double *y,
                    int64_t ld_tgt, int n_fields, int32_t const *order,
                    hipStream_t s );

void crs_apply_f32( int32_t const *idx, float const *coeffs, int64_t n_tgt,
                    int kk, float const *x, int64_t ld_src, float *y,
                    int64_t ld_tgt, int n_fields, int32_t const *order,
                    hipStream_t s );

// FP32 coefficients applied to double values at the boundary: values are
// rounded to float on load, accumulated in float, widened on store.
void crs_apply_f32_dvals( int32_t const *idx, float const *coeffs,
                          int64_t n_tgt, int kk, double const *x,
                          int64_t ld_src, double *y, int64_t ld_tgt,
                          int n_fields, int32_t const *order, hipStream_t s );

struct CscIndex
{
    DeviceBuffer<int32_t> col_offsets; // [n_src + 1]
    DeviceBuffer<int32_t> entries;     // [nnz], entry ids sorted by column
};

void build_csc( int32_t const *idx, int64_t nnz, int64_t n_src,
                CscIndex &csc, hipStream_t s );

void crs_apply_transpose_f64( CscIndex const &csc, double const *coeffs,
                              int kk, int64_t n_src, double const *ybar,
                              int64_t ld_tgt, double *xbar, int64_t ld_src,
                              int n_fields, hipStream_t s );

} // namespace dtkx

#endif

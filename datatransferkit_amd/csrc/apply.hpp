// CRS coefficient apply of the MLS operator and its transpose.
//
// apply:     y[f][i] = sum_j c(i,j) x[f][idx(i,j)], the row sum starting at 0
//            and running j ascending (computeTargetValues,
//            DTK_DetailsMovingLeastSquaresOperatorImpl.hpp:58-66), after an
//            in-kernel gather that replaces the reference's fetch()
//            (DTK_DetailsNearestNeighborOperatorImpl.hpp:152-192) on one rank.
// transpose: xbar[f][s] = sum over entries e with idx(e) = s, in ascending
//            canonical entry order (target, then neighbour rank), of
//            c(e) ybar[f][row(e)]; gathered through a column index (CSC)
//            built once, so no atomics and a fixed order.
//
// Row layout. Rows are stored in the spatial (Morton) order of the targets
// and column indices name source *positions* in the Morton order of the
// sources: row r belongs to target order[r], position p is source perm[p].
// Neighbouring rows then gather from neighbouring positions, so one pass
// over the rows reads the values nearly sequentially. An apply first
// gathers the caller's fields into a position-ordered, field-interleaved
// copy (one random read per source and field instead of one per entry),
// runs the row sums over it and scatters each row's results to its target.
// Null order / perm mean the identity (rows and columns in caller order).
#ifndef DTKX_APPLY_HPP
#define DTKX_APPLY_HPP

#include "common.hpp"

namespace dtkx
{

struct RowLayout
{
    int32_t const *idx = nullptr;   // [n_tgt * kk] source positions
    int32_t const *order = nullptr; // row -> target (nullable)
    int32_t const *perm = nullptr;  // position -> source (nullable)
    int32_t const *trow = nullptr;  // target -> row (nullable, inverse of order)
    int32_t const *spos = nullptr;  // source -> position (nullable, inverse of perm)
    int64_t n_src = 0, n_tgt = 0;
    int kk = 0;
};

// Nearest-neighbour transfer: y[f][t] = x[f][idx[t]] for every field, a
// copy of each value (signed zeros and NaN payloads included).
void nn_gather_f64( int32_t const *idx, int64_t n_tgt, double const *x,
                    int64_t ld_src, double *y, int64_t ld_tgt, int n_fields,
                    hipStream_t s );

void crs_apply_f64( RowLayout const &g, double const *coeffs, double const *x,
                    int64_t ld_src, double *y, int64_t ld_tgt, int n_fields,
                    hipStream_t s );

void crs_apply_f32( RowLayout const &g, float const *coeffs, float const *x,
                    int64_t ld_src, float *y, int64_t ld_tgt, int n_fields,
                    hipStream_t s );

// FP32 coefficients applied to double values at the boundary: values are
// rounded to float on load, accumulated in float, widened on store.
void crs_apply_f32_dvals( RowLayout const &g, float const *coeffs,
                          double const *x, int64_t ld_src, double *y,
                          int64_t ld_tgt, int n_fields, hipStream_t s );

// Device scratch bytes the two applies above need for n_fields fields
// (FP64 coefficients, or FP32 coefficients with double values).
std::size_t crs_apply_scratch_bytes( RowLayout const &g, int n_fields,
                                     bool fp32_coeffs );

// The same applies with their scratch in ws (at least
// crs_apply_scratch_bytes, 256-byte aligned): they allocate nothing.
void crs_apply_f64_ws( RowLayout const &g, double const *coeffs,
                       double const *x, int64_t ld_src, double *y,
                       int64_t ld_tgt, int n_fields, void *ws, hipStream_t s );
void crs_apply_f32_dvals_ws( RowLayout const &g, float const *coeffs,
                             double const *x, int64_t ld_src, double *y,
                             int64_t ld_tgt, int n_fields, void *ws,
                             hipStream_t s );

struct CscIndex
{
    DeviceBuffer<int32_t> col_offsets; // [n_src + 1], by source position
    DeviceBuffer<int32_t> entries;     // [nnz] storage slots, canonical order
    // per CSC entry: its row (storage slot / kk) and its coefficient
    DeviceBuffer<int32_t> rows;
    DeviceBuffer<double> coefs;
};

// coeffs: the handle's FP64 coefficients, copied into column order
void build_csc( RowLayout const &g, double const *coeffs, CscIndex &csc,
                hipStream_t s );

void crs_apply_transpose_f64( RowLayout const &g, CscIndex const &csc,
                              double const *coeffs, double const *ybar,
                              int64_t ld_tgt, double *xbar, int64_t ld_src,
                              int n_fields, hipStream_t s );

// out[p * pitch + f] = x[f][perm[p]] for f < nf (pitch >= nf), doubles.
// With spos (perm's inverse) the lanes run over sources instead: coalesced
// reads of x and one contiguous record written per source.
void gather_fields_f64( double const *x, int64_t ld, int nf,
                        int32_t const *perm, int32_t const *spos, int64_t n,
                        double *out, int pitch, hipStream_t s );

// Batched operators, one field each: rows in CSR form (row_off), columns
// as global positions of the concatenated problem-major sources. Problem q
// reads its values from x[q] (local source index = perm[p] - src_off[q])
// and writes y[q] (local target index = order[r] - tgt_off[q]); x and y are
// device arrays of device pointers, read when the kernels run.
struct BatchLayout
{
    int32_t const *idx = nullptr;
    int64_t const *row_off = nullptr;
    int32_t const *order = nullptr, *perm = nullptr;
    int32_t const *src_off = nullptr, *tgt_off = nullptr;
    int32_t const *src_prob = nullptr, *tgt_prob = nullptr;
    int64_t n_src = 0, n_tgt = 0;
};

// xs: scratch of n_src doubles (or floats for float coefficients)
void batch_apply_f64( BatchLayout const &g, double const *coeffs,
                      double const *const *x, double *const *y, double *xs,
                      hipStream_t s );
void batch_apply_f32( BatchLayout const &g, float const *coeffs,
                      double const *const *x, double *const *y, float *xs,
                      hipStream_t s );

// The graph in caller order: row t = target t, indices = source indices.
template <typename C>
void canonical_graph( RowLayout const &g, C const *coeffs, int32_t *idx_out,
                      C *coeffs_out, hipStream_t s );

} // namespace dtkx

#endif

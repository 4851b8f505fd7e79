// CRS apply kernels (HBM-bound gather-dot) and the transpose.
#include "apply.hpp"

#include <hipcub/hipcub.hpp>

namespace dtkx
{
namespace
{

constexpr int kBlock = 256;

// Accumulators: plain (FP64 path, the reference's sequential row sum) or
// compensated (FP32 path: the error of each float addition is carried in a
// second float, so a long row sums to float accuracy, not k times worse).
template <typename T>
struct PlainSum
{
    T s = 0;
    __device__ void add( T v ) { s += v; }
    __device__ T value() const { return s; }
};

struct KahanSum
{
    float s = 0.f, c = 0.f;
    __device__ void add( float v )
    {
#pragma clang fp contract( off )
        float const y = v - c;
        float const t = s + y;
        c = ( t - s ) - y;
        s = t;
    }
    __device__ float value() const { return s; }
};

template <typename Acc>
struct SumOf
{
    using type = PlainSum<Acc>;
};
template <>
struct SumOf<float>
{
    using type = KahanSum;
};

// One lane per target row; every field of the row reuses the row's
// coefficients and indices from registers. Sequential sum per row.
template <typename C, typename X, typename Acc, int NF>
__global__ void __launch_bounds__( kBlock )
    crs_apply_kernel( int32_t const *__restrict__ idx,
                      C const *__restrict__ coeffs, int64_t n_tgt, int kk,
                      X const *__restrict__ x, int64_t ld_src,
                      X *__restrict__ y, int64_t ld_tgt, int n_fields,
                      int32_t const *__restrict__ order )
{
#pragma clang fp contract( off )
    using Sum = typename SumOf<Acc>::type;
    int64_t const slot = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( slot >= n_tgt )
        return;
    // rows in the spatial (Morton) order of the targets when known: a
    // wave's gathers then touch overlapping source sets, which stay in L2
    int64_t const i = order ? order[slot] : slot;
    int64_t const row = i * kk;
    if ( NF > 0 )
    {
        Sum acc[NF > 0 ? NF : 1];
        for ( int j = 0; j < kk; ++j )
        {
            int32_t const s = idx[row + j];
            Acc const c = (Acc)coeffs[row + j];
#pragma unroll
            for ( int f = 0; f < NF; ++f )
                acc[f].add( c * (Acc)x[f * ld_src + s] );
        }
#pragma unroll
        for ( int f = 0; f < NF; ++f )
            y[f * ld_tgt + i] = (X)acc[f].value();
    }
    else
    {
        for ( int f = 0; f < n_fields; ++f )
        {
            Sum acc;
            for ( int j = 0; j < kk; ++j )
                acc.add( (Acc)coeffs[row + j] * (Acc)x[f * ld_src + idx[row + j]] );
            y[f * ld_tgt + i] = (X)acc.value();
        }
    }
}

template <typename C, typename X, typename Acc>
void launch_apply( int32_t const *idx, C const *coeffs, int64_t n_tgt, int kk,
                   X const *x, int64_t ld_src, X *y, int64_t ld_tgt,
                   int n_fields, int32_t const *order, hipStream_t s )
{
    if ( n_tgt == 0 || n_fields == 0 )
        return;
    dim3 grid( ceil_div( n_tgt, kBlock ) ), block( kBlock );
    switch ( n_fields )
    {
    case 1:
        crs_apply_kernel<C, X, Acc, 1><<<grid, block, 0, s>>>(
            idx, coeffs, n_tgt, kk, x, ld_src, y, ld_tgt, n_fields, order );
        break;
    case 2:
        crs_apply_kernel<C, X, Acc, 2><<<grid, block, 0, s>>>(
            idx, coeffs, n_tgt, kk, x, ld_src, y, ld_tgt, n_fields, order );
        break;
    case 4:
        crs_apply_kernel<C, X, Acc, 4><<<grid, block, 0, s>>>(
            idx, coeffs, n_tgt, kk, x, ld_src, y, ld_tgt, n_fields, order );
        break;
    case 8:
        crs_apply_kernel<C, X, Acc, 8><<<grid, block, 0, s>>>(
            idx, coeffs, n_tgt, kk, x, ld_src, y, ld_tgt, n_fields, order );
        break;
    default:
        crs_apply_kernel<C, X, Acc, 0><<<grid, block, 0, s>>>(
            idx, coeffs, n_tgt, kk, x, ld_src, y, ld_tgt, n_fields, order );
    }
    DTKX_HIP( hipGetLastError() );
}

// ------------------------------------------------------------ transpose
__global__ void __launch_bounds__( kBlock )
    iota_kernel( int32_t *v, int64_t n )
{
    int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i < n )
        v[i] = (int32_t)i;
}

// col_offsets[s] = first position p in the sorted keys with key[p] >= s
__global__ void __launch_bounds__( kBlock )
    col_offsets_kernel( int32_t const *keys, int64_t nnz, int64_t n_src,
                        int32_t *off )
{
    int64_t p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( p > nnz )
        return;
    int64_t const prev = p == 0 ? -1 : keys[p - 1];
    int64_t const cur = p == nnz ? n_src : keys[p];
    for ( int64_t s = prev + 1; s <= cur; ++s )
        off[s] = (int32_t)p;
}

__global__ void __launch_bounds__( kBlock )
    transpose_kernel( int32_t const *__restrict__ off,
                      int32_t const *__restrict__ ent,
                      double const *__restrict__ coeffs, int kk, int64_t n_src,
                      double const *__restrict__ ybar, int64_t ld_tgt,
                      double *__restrict__ xbar, int64_t ld_src, int n_fields )
{
#pragma clang fp contract( off )
    int64_t const s = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( s >= n_src )
        return;
    int32_t const b = off[s], e = off[s + 1];
    for ( int f = 0; f < n_fields; ++f )
    {
        double acc = 0.;
        for ( int32_t p = b; p < e; ++p )
        {
            int32_t const en = ent[p];
            acc += coeffs[en] * ybar[f * ld_tgt + en / kk];
        }
        xbar[f * ld_src + s] = acc;
    }
}

} // namespace

void crs_apply_f64( int32_t const *idx, double const *coeffs, int64_t n_tgt,
                    int kk, double const *x, int64_t ld_src, double *y,
                    int64_t ld_tgt, int n_fields, int32_t const *order,
                    hipStream_t s )
{
    launch_apply<double, double, double>( idx, coeffs, n_tgt, kk, x, ld_src,
                                          y, ld_tgt, n_fields, order, s );
}

void crs_apply_f32( int32_t const *idx, float const *coeffs, int64_t n_tgt,
                    int kk, float const *x, int64_t ld_src, float *y,
                    int64_t ld_tgt, int n_fields, int32_t const *order,
                    hipStream_t s )
{
    launch_apply<float, float, float>( idx, coeffs, n_tgt, kk, x, ld_src, y,
                                       ld_tgt, n_fields, order, s );
}

void crs_apply_f32_dvals( int32_t const *idx, float const *coeffs,
                          int64_t n_tgt, int kk, double const *x,
                          int64_t ld_src, double *y, int64_t ld_tgt,
                          int n_fields, int32_t const *order, hipStream_t s )
{
    // values are rounded to float on load, products and sums in float
    launch_apply<float, double, float>( idx, coeffs, n_tgt, kk, x, ld_src, y,
                                        ld_tgt, n_fields, order, s );
}

void build_csc( int32_t const *idx, int64_t nnz, int64_t n_src,
                CscIndex &csc, hipStream_t s )
{
    DTKX_REQUIRE( nnz < INT32_MAX, DTKX_ERR_UNSUPPORTED );
    csc.col_offsets.resize( n_src + 1, s );
    csc.entries.resize( std::max<int64_t>( nnz, 1 ), s );
    if ( nnz == 0 )
    {
        DTKX_HIP( hipMemsetAsync( csc.col_offsets.get(), 0,
                                  sizeof( int32_t ) * ( n_src + 1 ), s ) );
        return;
    }
    DeviceBuffer<int32_t> ids( nnz, s ), keys_sorted( nnz, s );
    iota_kernel<<<ceil_div( nnz, kBlock ), kBlock, 0, s>>>( ids.get(), nnz );
    int end_bit = 1;
    while ( ( int64_t( 1 ) << end_bit ) < n_src )
        ++end_bit;
    std::size_t temp_bytes = 0;
    DTKX_HIP( hipcub::DeviceRadixSort::SortPairs(
        nullptr, temp_bytes, idx, keys_sorted.get(), ids.get(),
        csc.entries.get(), static_cast<int>( nnz ), 0, end_bit, s ) );
    DeviceBuffer<char> temp( temp_bytes, s );
    DTKX_HIP( hipcub::DeviceRadixSort::SortPairs(
        temp.get(), temp_bytes, idx, keys_sorted.get(), ids.get(),
        csc.entries.get(), static_cast<int>( nnz ), 0, end_bit, s ) );
    col_offsets_kernel<<<ceil_div( nnz + 1, kBlock ), kBlock, 0, s>>>(
        keys_sorted.get(), nnz, n_src, csc.col_offsets.get() );
    DTKX_HIP( hipGetLastError() );
}

void crs_apply_transpose_f64( CscIndex const &csc, double const *coeffs,
                              int kk, int64_t n_src, double const *ybar,
                              int64_t ld_tgt, double *xbar, int64_t ld_src,
                              int n_fields, hipStream_t s )
{
    if ( n_src == 0 || n_fields == 0 )
        return;
    transpose_kernel<<<ceil_div( n_src, kBlock ), kBlock, 0, s>>>(
        csc.col_offsets.get(), csc.entries.get(), coeffs, kk, n_src, ybar,
        ld_tgt, xbar, ld_src, n_fields );
    DTKX_HIP( hipGetLastError() );
}

} // namespace dtkx

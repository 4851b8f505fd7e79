// Fused MLS setup with a certified Cholesky solve (DTKX_SOLVER_CHOLESKY).
//
// The reference inverts the moment matrix A = P^T Phi P through a Jacobi
// SVD and drops singular values below DBL_EPSILON (DTK_DetailsSVDImpl.hpp
// :281-297). When every eigenvalue of A is provably above that threshold,
// the pseudo-inverse is the inverse, and a Cholesky factorisation gives
// a0 = A^-1 e_0 with m^3/3 flops instead of several Jacobi sweeps. The
// proof is a bound: with A = U^T U (U upper triangular) and V = U^-1,
//   lambda_min(A) = sigma_min(U)^2 = 1 / |V|_2^2 >= 1 / |V|_F^2,
// compared against the truncation threshold (absolute DBL_EPSILON, or
// rcond * trace(A) >= rcond * lambda_max) times 4 plus 64 m eps trace(A),
// a margin for the rounding of the eigenvalues an eigen-solver would
// compute. A system that fails the test, or whose factorisation meets a
// non-positive pivot, is appended to a list that the cyclic eigen-solver
// then sets up; it is counted there, so the counters mean what they mean
// for the cyclic solver.
//
// One lane per system, everything in registers: the packed triangle of A
// is overwritten by U, then by V (all indices compile-time constants).
#include "cyclic_device.hpp"
#include "setup.hpp"

#include <algorithm>
#include <type_traits>

namespace dtkx
{
namespace
{

using namespace cyc;
constexpr int kBlock = 64;
// neighbours gathered per lane and round (0.69 ms for one at a time, 0.48 ms
// for four, about the same for eight: 1M rows of 32, the FP32-bulk form)
#ifndef DTKX_SETUP_GATHER
#define DTKX_SETUP_GATHER 4
#endif
constexpr int kGather = DTKX_SETUP_GATHER;

// The coefficients of a wave's rows are written with the lanes walking the
// wave's entries in memory order (consecutive rows hold consecutive
// entries), the rows' radius, target and a0 read from LDS: contiguous index
// loads and coefficient stores instead of one strided row per lane.
template <int M>
struct RowShare
{
    double x[kBlock][M];
    double geo[kBlock][4]; // radius, target x, y, z
    int64_t beg[kBlock];
    int kk[kBlock];
};

// two waves per SIMD: the compiler otherwise budgets for one (194 VGPRs;
// at three waves 23 of them spill and the 1M x 1M, k = 32 FP32-bulk setup
// takes 0.61 ms against 0.48)
template <int M, bool F32, bool STORE>
__global__ void __launch_bounds__( kBlock, 2 )
    setup_cholesky_kernel( double4 const *__restrict__ src,
                           double4 const *__restrict__ tgt, int64_t n_tgt,
                           int32_t const *__restrict__ indices,
                           int32_t const *__restrict__ order, int64_t row0,
                           int64_t const *__restrict__ row_off, int kk_uniform,
                           int rbf, int geometry, int truncation, double rcond,
                           double *__restrict__ coeffs,
                           float *__restrict__ coeffs32,
                           double *__restrict__ pinv_out,
                           int32_t *__restrict__ fallback,
                           int32_t *__restrict__ n_fallback,
                           unsigned long long *__restrict__ counters,
                           float const *__restrict__ row_max )
{
    using S = SymIdx<M>;
    __shared__ RowShare<M> sh;
    int64_t const slot0 = row0 + xcd_tile() * kBlock;
    int64_t const slot = slot0 + threadIdx.x;
    bool const valid = slot < n_tgt;
    int64_t const t = valid ? ( order ? order[slot] : slot ) : 0;
    int64_t const rbeg =
        valid ? ( row_off ? row_off[slot] : slot * kk_uniform ) : 0;
    int const kk =
        valid ? ( row_off ? (int)( row_off[slot + 1] - rbeg ) : kk_uniform )
              : 0;
    int32_t const *nbr = indices + rbeg;
    double4 const tp = valid ? tgt[t] : double4{0, 0, 0, 0};

    // the neighbours are gathered kGather at a time, so that many loads of
    // a lane are in flight together (one at a time, the two passes wait on
    // 2 kk dependent index-then-point loads)
    auto gather = [&]( int l0, double4 ( &p )[kGather] ) {
#pragma unroll
        for ( int j = 0; j < kGather; ++j )
            p[j] = src[nbr[l0 + j < kk ? l0 + j : kk - 1]];
    };

    // radius (Impl.hpp:118-134); the FP32-bulk form may take the largest
    // distance from the kNN
    double dmax = 10. * DBL_EPSILON;
    if ( F32 && row_max )
    {
        double const d = valid ? (double)row_max[slot] : 0.0;
        dmax = d > dmax ? d : dmax;
    }
    else
    {
        for ( int l0 = 0; l0 < kk; l0 += kGather )
        {
            double4 p[kGather];
            gather( l0, p );
#pragma unroll
            for ( int j = 0; j < kGather; ++j )
            {
                double const d = nbr_distance<M, F32>( geometry, p[j], tp );
                // a repeated last entry changes nothing
                dmax = d > dmax ? d : dmax;
            }
        }
    }
    double const radius = F32 ? (double)( 1.1f * (float)dmax ) : 1.1 * dmax;

    // moments, upper triangle
    double A[S::size];
#pragma unroll
    for ( int e = 0; e < S::size; ++e )
        A[e] = 0.0;
    for ( int l0 = 0; l0 < kk; l0 += kGather )
    {
        double4 p[kGather];
        gather( l0, p );
#pragma unroll
        for ( int j = 0; j < kGather; ++j )
        {
            if ( l0 + j >= kk )
                break;
            double P[M];
            double const phi =
                nbr_row<M, F32>( geometry, rbf, radius, p[j], tp, P );
#pragma unroll
            for ( int r = 0; r < M; ++r )
            {
                double const pr = P[r] * phi;
#pragma unroll
                for ( int c = r; c < M; ++c )
                    A[S::at( r, c )] = fma( pr, P[c], A[S::at( r, c )] );
            }
        }
    }
    double trace = 0.0;
#pragma unroll
    for ( int i = 0; i < M; ++i )
        trace += A[S::at( i, i )];
    if ( !valid )
    {
        // an identity keeps the factorisation below finite
#pragma unroll
        for ( int i = 0; i < M; ++i )
            A[S::at( i, i )] = 1.0;
    }

    bool const ok = cholesky_inverse<M>( A, trace, truncation, rcond );
    if ( valid && !ok )
    {
        // set up by the eigen-solver afterwards (it also rewrites this
        // row's coefficients, which the shared loop below fills regardless)
        fallback[atomicAdd( n_fallback, 1 )] = (int32_t)slot;
        if ( counters )
            atomicAdd( counters + 3, 1ull );
    }

    // a0 = V V^T e_0: a0_i = sum_{k >= i} V(i, k) V(0, k)
#pragma unroll
    for ( int i = 0; i < M; ++i )
    {
        double s = 0.0;
#pragma unroll
        for ( int k = i; k < M; ++k )
            s = fma( A[S::at( i, k )], A[S::at( 0, k )], s );
        sh.x[threadIdx.x][i] = s;
    }
    if constexpr ( STORE )
    {
        // A^-1 = V V^T, packed upper triangle
        if ( valid && ok )
#pragma unroll
            for ( int i = 0; i < M; ++i )
#pragma unroll
                for ( int j = i; j < M; ++j )
                {
                    double s = 0.0;
#pragma unroll
                    for ( int k = j; k < M; ++k )
                        s = fma( A[S::at( i, k )], A[S::at( j, k )], s );
                    pinv_out[slot * S::size + S::at( i, j )] = s;
                }
    }
    sh.geo[threadIdx.x][0] = radius;
    sh.geo[threadIdx.x][1] = tp.x;
    sh.geo[threadIdx.x][2] = tp.y;
    sh.geo[threadIdx.x][3] = tp.z;
    sh.beg[threadIdx.x] = rbeg;
    sh.kk[threadIdx.x] = kk;
    __syncthreads();

    // coefficients c_l = phi_l (P_l . x), entries of the wave in order
    int const nrows = n_tgt - slot0 < kBlock ? (int)( n_tgt - slot0 ) : kBlock;
    int64_t const wbeg = sh.beg[0];
    int64_t const wend = sh.beg[nrows - 1] + sh.kk[nrows - 1];
    // (kGather entries per lane and round, their loads issued together)
    int row = 0;
    for ( int64_t g0 = wbeg + threadIdx.x; g0 < wend;
          g0 += kGather * kBlock )
    {
        double4 p[kGather];
#pragma unroll
        for ( int j = 0; j < kGather; ++j )
        {
            int64_t const g = g0 + j * kBlock;
            p[j] = src[indices[g < wend ? g : wend - 1]];
        }
#pragma unroll
        for ( int j = 0; j < kGather; ++j )
        {
            int64_t const g = g0 + j * kBlock;
            if ( g >= wend )
                break;
            // the row holding entry g: the last row beginning at or before g
            while ( row + 1 < nrows && sh.beg[row + 1] <= g )
                ++row;
            double const *xr = sh.x[row];
            double const rr = sh.geo[row][0];
            double4 const tr = {sh.geo[row][1], sh.geo[row][2], sh.geo[row][3],
                                0.0};
            double P[M];
            double const phi = nbr_row<M, F32>( geometry, rbf, rr, p[j], tr, P );
            double acc = 0.0;
#pragma unroll
            for ( int q = 0; q < M; ++q )
                acc = fma( xr[q], P[q], acc );
            if constexpr ( F32 )
                coeffs32[g] = (float)( acc * phi );
            else
                coeffs[g] = acc * phi;
        }
    }
}

// The eigen-solver's untruncated rows again, one lane per listed row: the
// moments, their Cholesky inverse factor and the row's coefficients (a
// non-positive pivot leaves the eigen-solver's coefficients in place). A
// bounded grid strides over the list, whose length only the device knows.
constexpr int kRefineBlocks = 1024;

template <int M, bool F32>
__global__ void __launch_bounds__( kBlock )
    refine_kernel( double4 const *__restrict__ src,
                   double4 const *__restrict__ tgt,
                   int32_t const *__restrict__ indices,
                   int32_t const *__restrict__ order,
                   int64_t const *__restrict__ row_off, int kk_uniform,
                   int rbf, int geometry, int32_t const *__restrict__ rows,
                   int32_t const *__restrict__ n_rows,
                   double *__restrict__ coeffs, float *__restrict__ coeffs32,
                   float const *__restrict__ row_max,
                   int32_t *__restrict__ full_rows,
                   int32_t *__restrict__ n_full )
{
    using S = SymIdx<M>;
    int64_t const n = *n_rows;
    for ( int64_t li = blockIdx.x * (int64_t)kBlock + threadIdx.x; li < n;
          li += (int64_t)gridDim.x * kBlock )
    {
        int64_t const slot = rows[li];
        int64_t const t = order ? order[slot] : slot;
        int64_t const rbeg = row_off ? row_off[slot] : slot * kk_uniform;
        int const kk =
            row_off ? (int)( row_off[slot + 1] - rbeg ) : kk_uniform;
        int32_t const *nbr = indices + rbeg;
        double4 const tp = tgt[t];
        double dmax = 10. * DBL_EPSILON;
        if ( F32 && row_max )
        {
            double const d = (double)row_max[slot];
            dmax = d > dmax ? d : dmax;
        }
        else
            for ( int l = 0; l < kk; ++l )
            {
                double const d =
                    nbr_distance<M, F32>( geometry, src[nbr[l]], tp );
                dmax = d > dmax ? d : dmax;
            }
        double const radius =
            F32 ? (double)( 1.1f * (float)dmax ) : 1.1 * dmax;
        double A[S::size];
#pragma unroll
        for ( int e = 0; e < S::size; ++e )
            A[e] = 0.0;
        for ( int l = 0; l < kk; ++l )
        {
            double P[M];
            double const phi =
                nbr_row<M, F32>( geometry, rbf, radius, src[nbr[l]], tp, P );
#pragma unroll
            for ( int r = 0; r < M; ++r )
            {
                double const pr = P[r] * phi;
#pragma unroll
                for ( int c = r; c < M; ++c )
                    A[S::at( r, c )] = fma( pr, P[c], A[S::at( r, c )] );
            }
        }
        if ( !cholesky_factor_inverse<M>( A ) )
        {
            // to the eigen-solver with eigenvectors, which runs next
            full_rows[atomicAdd( n_full, 1 )] = (int32_t)slot;
            continue;
        }
        double x[M];
#pragma unroll
        for ( int i = 0; i < M; ++i )
        {
            double acc = 0.0;
#pragma unroll
            for ( int k = i; k < M; ++k )
                acc = fma( A[S::at( i, k )], A[S::at( 0, k )], acc );
            x[i] = acc;
        }
        for ( int l = 0; l < kk; ++l )
        {
            double P[M];
            double const phi =
                nbr_row<M, F32>( geometry, rbf, radius, src[nbr[l]], tp, P );
            double acc = 0.0;
#pragma unroll
            for ( int q = 0; q < M; ++q )
                acc = fma( x[q], P[q], acc );
            if constexpr ( F32 )
                coeffs32[rbeg + l] = (float)( acc * phi );
            else
                coeffs[rbeg + l] = acc * phi;
        }
    }
}

template <int M>
void launch_fast( double4 const *src, double4 const *tgt, int64_t n_tgt,
                  int32_t const *indices, int32_t const *order,
                  SetupParams const &p, double *coeffs, float *coeffs32,
                  double *pinv, int32_t *list, int32_t *count,
                  unsigned long long *counters, hipStream_t s )
{
    dim3 const grid( ceil_div( n_tgt - p.row0, kBlock ) ), block( kBlock );
    auto run = [&]( auto f32, auto store ) {
        setup_cholesky_kernel<M, decltype( f32 )::value,
                              decltype( store )::value>
            <<<grid, block, 0, s>>>( src, tgt, n_tgt, indices, order, p.row0,
                                     p.row_off, p.kk, p.rbf, p.geometry,
                                     p.truncation, p.rcond, coeffs, coeffs32,
                                     pinv, list, count, counters, p.row_max );
    };
    using T = std::true_type;
    using F = std::false_type;
    if ( coeffs32 )
        pinv ? run( T{}, T{} ) : run( T{}, F{} );
    else
        pinv ? run( F{}, T{} ) : run( F{}, F{} );
    DTKX_HIP( hipGetLastError() );
}

} // namespace

void setup_fast( double4 const *src, double4 const *tgt, int64_t n_tgt,
                 int32_t const *indices, int32_t const *order,
                 SetupParams const &p, double *coeffs, float *coeffs32,
                 double *pinv, unsigned long long *counters, hipStream_t s )
{
    if ( n_tgt <= p.row0 )
        return;
    DTKX_REQUIRE( p.m == 1 || p.m == 4 || p.m == 10,
                  DTKX_ERR_INVALID_ARGUMENT );
    DeviceBuffer<int32_t> list( n_tgt - p.row0, s ), count( 1, s );
    zero_async( count.get(), sizeof( int32_t ), s );
    if ( p.m == 1 )
        launch_fast<1>( src, tgt, n_tgt, indices, order, p, coeffs, coeffs32,
                        pinv, list.get(), count.get(), counters, s );
    else if ( p.m == 4 )
        launch_fast<4>( src, tgt, n_tgt, indices, order, p, coeffs, coeffs32,
                        pinv, list.get(), count.get(), counters, s );
    else
        launch_fast<10>( src, tgt, n_tgt, indices, order, p, coeffs, coeffs32,
                         pinv, list.get(), count.get(), counters, s );
    // the listed systems: with a stored pseudo-inverse to keep consistent,
    // all through the eigen-solver (a grid for every row; blocks past the
    // device-side count return at once)
    SetupParams q = p;
    q.rows = list.get();
    q.n_rows = count.get();
    if ( pinv )
    {
        setup_cyclic( src, tgt, n_tgt, indices, order, q, coeffs, coeffs32,
                      pinv, counters, s );
        return;
    }
    // otherwise their eigenvalues first: the rows with none below the
    // threshold have A^+ = A^-1 and are solved again by Cholesky, the
    // others by the eigen-solver with eigenvectors, on a forked stream at
    // the same time (disjoint rows); rows whose Cholesky meets a
    // non-positive pivot go to the eigen-solver afterwards. Each row ends
    // as it did when the eigen-solver ran over the whole list before the
    // Cholesky of its untruncated rows.
    int64_t const nl = n_tgt - p.row0;
    DeviceBuffer<int32_t> rlist( nl, s ), tlist( nl, s ), flist( nl, s ),
        counts( 3, s ); // refine, truncated, refine failures
    zero_async( counts.get(), 3 * sizeof( int32_t ), s );
    q.refine_rows = rlist.get();
    q.n_refine = counts.get();
    setup_screen( src, tgt, n_tgt, indices, order, q, coeffs32 != nullptr,
                  tlist.get(), counts.get() + 1, counters, s );
    SetupParams f = p;
    f.screened = true;
    {
        Fork fork( s );
        f.rows = tlist.get();
        f.n_rows = counts.get() + 1;
        setup_cyclic( src, tgt, n_tgt, indices, order, f, coeffs, coeffs32,
                      nullptr, counters, fork.side() );
        int64_t const blocks = std::min<int64_t>(
            ceil_div( n_tgt - p.row0, kBlock ), kRefineBlocks );
        auto refine = [&]( auto m, auto f32 ) {
            refine_kernel<decltype( m )::value, decltype( f32 )::value>
                <<<blocks, kBlock, 0, s>>>(
                    src, tgt, indices, order, p.row_off, p.kk, p.rbf,
                    p.geometry, rlist.get(), counts.get(), coeffs, coeffs32,
                    p.row_max, flist.get(), counts.get() + 2 );
        };
        using T = std::true_type;
        using F = std::false_type;
        using M1 = std::integral_constant<int, 1>;
        using M4 = std::integral_constant<int, 4>;
        using M10 = std::integral_constant<int, 10>;
        if ( p.m == 1 )
            coeffs32 ? refine( M1{}, T{} ) : refine( M1{}, F{} );
        else if ( p.m == 4 )
            coeffs32 ? refine( M4{}, T{} ) : refine( M4{}, F{} );
        else
            coeffs32 ? refine( M10{}, T{} ) : refine( M10{}, F{} );
        DTKX_HIP( hipGetLastError() );
        fork.join();
    }
    f.rows = flist.get();
    f.n_rows = counts.get() + 2;
    setup_cyclic( src, tgt, n_tgt, indices, order, f, coeffs, coeffs32,
                  nullptr, counters, s );
}

} // namespace dtkx

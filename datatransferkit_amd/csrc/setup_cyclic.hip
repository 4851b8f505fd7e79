// Fused MLS setup with a cyclic symmetric Jacobi eigen-solver: the fast
// FP64 path (DTKX_SOLVER_CYCLIC_JACOBI).
//
// One target system per lane (per lane pair for the quadratic basis, see
// setup_cyclic_kernel). The moment matrix A = P^T Phi P is symmetric
// positive semi-definite, so A^+ e_0 = V Lambda^+ V^T e_0 with A = V Lambda
// V^T; the pseudo-inverse is the same as the reference's two-sided SVD
// (DTK_DetailsSVDImpl.hpp) up to rounding. Every (p, q) index is a
// compile-time constant (fully unrolled sweeps), so A (upper triangle) and
// V stay in registers: no LDS, no barriers, no divergence except the
// per-wave sweep count. Rotations use the Rutishauser form (t = sgn(theta)
// / (|theta| + sqrt(theta^2 + 1)), c = 1/sqrt(1 + t^2), tau = s/(1 + c)).
// Each system sweeps until its off-diagonal mass is below (DBL_EPSILON/8)^2
// of its diagonal mass, or to the sweep cap (a converged system idles while
// the rest of its wave finishes, so its bits do not depend on its wave).
// Truncation keeps the reference's semantics: eigenvalues with |l| >=
// DBL_EPSILON (or rcond * max|l|) contribute. Rows listed by the Cholesky
// setup (systems it could not certify) whose eigenvalues all pass are
// solved again by Cholesky (setup_fast.hip).
// Neighbour data is re-gathered from L2 in three passes (radius, moments,
// coefficients) instead of being staged.
#include "cyclic_device.hpp"
#include "setup.hpp"

#include <type_traits>

namespace dtkx
{
namespace
{

using namespace cyc;
constexpr int kBlock = 64;

// One listed or ranged row's neighbour list, target and radius
// (Impl.hpp:118-134: the largest neighbour distance, for the FP32-bulk form
// possibly taken from the kNN), with its moments (upper triangle) written
// to A. Shared by the eigen-solver and the eigenvalue screen, so both form
// the same bits.
struct SystemRow
{
    int64_t rbeg;
    int kk;
    int32_t const *nbr;
    double4 tp;
    double radius;
};

template <int M, bool F32>
__device__ __forceinline__ SystemRow
    system_moments( bool valid, int64_t slot, double4 const *__restrict__ src,
                    double4 const *__restrict__ tgt,
                    int32_t const *__restrict__ indices,
                    int32_t const *__restrict__ order,
                    int64_t const *__restrict__ row_off, int kk_uniform,
                    int rbf, int geometry, float const *__restrict__ row_max,
                    double *A )
{
    using S = SymIdx<M>;
    int64_t const t = valid ? ( order ? order[slot] : slot ) : 0;
    int64_t const rbeg =
        valid ? ( row_off ? row_off[slot] : slot * kk_uniform ) : 0;
    int const kk =
        valid ? ( row_off ? (int)( row_off[slot + 1] - rbeg ) : kk_uniform )
              : 0;
    int32_t const *nbr = indices + rbeg;
    double4 const tp = valid ? tgt[t] : double4{0, 0, 0, 0};

    double dmax = 10. * DBL_EPSILON;
    if ( F32 && row_max )
    {
        double const d = valid ? (double)row_max[slot] : 0.0;
        dmax = d > dmax ? d : dmax;
    }
    else if ( valid )
        for ( int l = 0; l < kk; ++l )
        {
            double const d = nbr_distance<M, F32>( geometry, src[nbr[l]], tp );
            dmax = d > dmax ? d : dmax;
        }
    double const radius = F32 ? (double)( 1.1f * (float)dmax ) : 1.1 * dmax;

#pragma unroll
    for ( int e = 0; e < S::size; ++e )
        A[e] = 0.0;
    if ( valid )
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
    return SystemRow{rbeg, kk, nbr, tp, radius};
}

// Truncation threshold of a system's eigenvalues (DBL_EPSILON, or rcond
// times the largest |eigenvalue|) and whether any falls below it.
template <int M>
__device__ __forceinline__ bool truncated( double const *A, int truncation,
                                           double rcond, double &thr )
{
    using S = SymIdx<M>;
    thr = DBL_EPSILON;
    if ( truncation == DTKX_TRUNCATION_RELATIVE )
    {
        double lmax = 0.;
#pragma unroll
        for ( int k = 0; k < M; ++k )
            lmax = fmax( lmax, fabs( A[S::at( k, k )] ) );
        thr = rcond * lmax;
    }
    bool under = false;
#pragma unroll
    for ( int k = 0; k < M; ++k )
        under = under || !( fabs( A[S::at( k, k )] ) >= thr );
    return under;
}

// H lanes per system (H = 1, or 2 for M = 10 so A and half of V fit the
// 256 VGPRs): every lane of the pair keeps the whole moment matrix and runs
// the same rotations on it (identical values, so both agree), and owns
// R = M / H rows of V.
// refine_rows (the Cholesky setup's fallback rows, nullable): systems whose
// eigenvalues all pass the truncation test are listed there, to be solved
// again by Cholesky (setup_fast.hip, refine_kernel).
template <int M, int H, bool F32, bool STORE>
__device__ __forceinline__ void
    cyclic_system( int64_t const li, double4 const *__restrict__ src,
                         double4 const *__restrict__ tgt, int64_t n_tgt,
                         int32_t const *__restrict__ indices,
                         int32_t const *__restrict__ order, int64_t row0,
                         int64_t const *__restrict__ row_off,
                         int32_t const *__restrict__ rows,
                         int32_t const *__restrict__ n_rows, int kk_uniform,
                         int rbf,
                         int geometry, int truncation, double rcond,
                         int max_sweeps, double *__restrict__ coeffs,
                         float *__restrict__ coeffs32,
                         double *__restrict__ pinv_out,
                         unsigned long long *__restrict__ counters,
                         float const *__restrict__ row_max,
                         int32_t *__restrict__ refine_rows,
                         int32_t *__restrict__ n_refine, bool screened )
{
    using S = SymIdx<M>;
    constexpr int R = M / H;
    int const h = threadIdx.x % H;
    bool valid;
    int64_t slot;
    if ( rows )
    {
        valid = li < *n_rows;
        slot = valid ? rows[li] : 0;
    }
    else
    {
        slot = row0 + li;
        valid = slot < n_tgt;
    }
    double A[S::size];
    SystemRow const sr =
        system_moments<M, F32>( valid, slot, src, tgt, indices, order, row_off,
                                kk_uniform, rbf, geometry, row_max, A );
    int const kk = sr.kk;
    int64_t const rbeg = sr.rbeg;
    int32_t const *nbr = sr.nbr;
    double4 const tp = sr.tp;
    double const radius = sr.radius;

    // cyclic Jacobi; this lane owns rows h*R .. h*R + R - 1 of V
    double V[R * M];
#pragma unroll
    for ( int e = 0; e < R * M; ++e )
        V[e] = ( h * R + e / M == e % M ) ? 1.0 : 0.0;
    // each system stops at its own convergence, so its result does not
    // depend on which systems share its wave
    int sweeps = 0;
    bool done = !valid || converged<M>( A );
    for ( int pass = 0; !__all( done ) && pass < max_sweeps; ++pass )
    {
        sweep_all<M, R>( A, V, done );
        sweeps += done ? 0 : 1;
        done = done || converged<M>( A );
    }

    // x = A^+ e_0 = sum_k [|l_k| >= thr] V(0,k) V(:,k) / l_k
    double thr = DBL_EPSILON;
    if ( truncation == DTKX_TRUNCATION_RELATIVE )
    {
        double lmax = 0.;
#pragma unroll
        for ( int k = 0; k < M; ++k )
            lmax = fmax( lmax, fabs( A[S::at( k, k )] ) );
        thr = rcond * lmax;
    }
    double w[M];
    bool under = false;
#pragma unroll
    for ( int k = 0; k < M; ++k )
    {
        double const lk = A[S::at( k, k )];
        bool const keep = fabs( lk ) >= thr;
        under = under || !keep;
        // V(0, k) lives in the lane owning row 0 (h == 0)
        double v0k = V[k];
        if ( H > 1 )
            v0k = __shfl( v0k, ( threadIdx.x / H ) * H, kBlock );
        w[k] = keep ? v0k / lk : 0.0;
    }
    double x[M];
#pragma unroll
    for ( int r = 0; r < R; ++r )
    {
        double acc = 0.0;
#pragma unroll
        for ( int k = 0; k < M; ++k )
            acc = fma( w[k], V[r * M + k], acc );
        x[r] = acc;
    }
    if ( H > 1 )
    {
        // gather the partner's rows: x[R..M) from lane h^1 (H == 2)
        double xo[R];
#pragma unroll
        for ( int r = 0; r < R; ++r )
            xo[r] = __shfl_xor( x[r], 1, kBlock );
#pragma unroll
        for ( int r = 0; r < R; ++r )
        {
            double const mine = x[r];
            x[h * R + r] = mine;
            x[( 1 - h ) * R + r] = xo[r];
        }
    }

    if constexpr ( STORE )
    {
        // the whole pseudo-inverse V diag(w) V^T, packed upper triangle;
        // each lane writes the rows it owns, with the partner's V rows
        // fetched by shuffles (H == 2)
        double wk[M];
#pragma unroll
        for ( int k = 0; k < M; ++k )
        {
            double const lk = A[S::at( k, k )];
            wk[k] = fabs( lk ) >= thr ? 1.0 / lk : 0.0;
        }
#pragma unroll
        for ( int jr = 0; jr < M; ++jr )
        {
            // row jr of V, from the lane that owns it
            double vj[M];
#pragma unroll
            for ( int k = 0; k < M; ++k )
            {
                double const mine = V[( jr % R ) * M + k];
                double const other =
                    H > 1 ? __shfl_xor( mine, 1, kBlock ) : mine;
                vj[k] = ( jr / R == h ) ? mine : other;
            }
#pragma unroll
            for ( int r = 0; r < R; ++r )
            {
                int const i = h * R + r;
                if ( valid && i <= jr )
                {
                    double acc = 0.0;
#pragma unroll
                    for ( int k = 0; k < M; ++k )
                        acc = fma( V[r * M + k] * wk[k], vj[k], acc );
                    pinv_out[slot * S::size + S::at( i, jr )] = acc;
                }
            }
        }
    }
    if ( !valid )
        return;
    // coefficients c_l = phi_l (P_l . x), neighbours split over the H lanes
    for ( int l = h; l < kk; l += H )
    {
        double P[M];
        double const phi =
            nbr_row<M, F32>( geometry, rbf, radius, src[nbr[l]], tp, P );
        if constexpr ( F32 )
        {
            // the float basis row and weight are exact in double: the short
            // dot product with x (large cancelling terms, the monomials span
            // several scales) is formed in double and rounded once
            double acc = 0.0;
#pragma unroll
            for ( int j = 0; j < M; ++j )
                acc = fma( x[j], P[j], acc );
            coeffs32[rbeg + l] = (float)( acc * phi );
        }
        else
        {
            double acc = 0.0;
#pragma unroll
            for ( int j = 0; j < M; ++j )
                acc = fma( x[j], P[j], acc );
            coeffs[rbeg + l] = acc * phi;
        }
    }
    if ( refine_rows && h == 0 && !under )
        refine_rows[atomicAdd( n_refine, 1 )] = (int32_t)slot;
    // (a screened list's untruncated rows had their sweeps counted by the
    // screen)
    if ( counters && h == 0 )
    {
        if ( under )
            atomicAdd( counters + 0, 1ull );
        if ( !screened || under )
        {
            if ( !done )
                atomicAdd( counters + 1, 1ull );
            atomicAdd( counters + 2, (unsigned long long)sweeps );
        }
    }
}

// Row-list mode (a device-side list of rows, its length known only on the
// device) runs a bounded grid that strides over the list, so an empty list
// costs a few thousand empty blocks, not one per row of the operator.
constexpr int kListBlocks = 2048;
constexpr int kScreenBlocks = 16384;

template <int M, int H, bool F32, bool STORE, bool LIST>
__global__ void __launch_bounds__( kBlock )
    setup_cyclic_kernel( double4 const *__restrict__ src,
                         double4 const *__restrict__ tgt, int64_t n_tgt,
                         int32_t const *__restrict__ indices,
                         int32_t const *__restrict__ order, int64_t row0,
                         int64_t const *__restrict__ row_off,
                         int32_t const *__restrict__ rows,
                         int32_t const *__restrict__ n_rows, int kk_uniform,
                         int rbf, int geometry, int truncation, double rcond,
                         int max_sweeps, double *__restrict__ coeffs,
                         float *__restrict__ coeffs32,
                         double *__restrict__ pinv_out,
                         unsigned long long *__restrict__ counters,
                         float const *__restrict__ row_max,
                         int32_t *__restrict__ refine_rows,
                         int32_t *__restrict__ n_refine, bool screened )
{
    constexpr int SPB = kBlock / H;
    auto one = [&]( int64_t tile ) {
        cyclic_system<M, H, F32, STORE>(
            tile * SPB + threadIdx.x / H, src, tgt, n_tgt, indices, order, row0,
            row_off, rows, n_rows, kk_uniform, rbf, geometry, truncation, rcond,
            max_sweeps, coeffs, coeffs32, pinv_out, counters, row_max,
            refine_rows, n_refine, screened );
    };
    if constexpr ( LIST )
    {
        // the trip count is uniform over the block (the pair shuffles and
        // the wave-wide convergence vote need every lane); the loop costs
        // registers, so the whole-operator form below has none
        // (a list's rows sit in its first tiles and are scattered over the
        // operator: tiles in dispatch order spread them over the XCDs)
        int64_t const n = *n_rows;
        for ( int64_t tile = blockIdx.x; tile * SPB < n; tile += gridDim.x )
            one( tile );
    }
    else if ( rows )
    {
        // a row list over a grid for its whole capacity: blocks past the
        // device-side count return at once
        int64_t const tile = blockIdx.x;
        if ( tile * SPB >= *n_rows )
            return;
        one( tile );
    }
    else
        one( xcd_tile() );
}

template <int M, int H>
void launch_cyclic( double4 const *src, double4 const *tgt, int64_t n_tgt,
                    int32_t const *indices, int32_t const *order,
                    SetupParams const &p, double *coeffs, float *coeffs32,
                    double *pinv, unsigned long long *counters, hipStream_t s )
{
    int const sweeps = p.max_rot > 0 ? static_cast<int>( p.max_rot ) : 30;
    // A screened list (the rows the eigenvalue screen found truncated, a
    // few percent of the listed rows) runs without the strided loop, whose
    // registers spilled (364 bytes a lane for m = 10), over a grid for the
    // list's whole capacity; other lists stride over a bounded grid.
    bool const strided = p.rows && !p.screened;
    int64_t blocks = ceil_div( n_tgt - p.row0, kBlock / H );
    if ( strided )
        blocks = blocks < kListBlocks ? blocks : kListBlocks;
    dim3 const grid( blocks ), block( kBlock );
    auto run = [&]( auto f32, auto store ) {
        auto kernel = strided ? setup_cyclic_kernel<M, H, decltype( f32 )::value,
                                                    decltype( store )::value, true>
                              : setup_cyclic_kernel<M, H, decltype( f32 )::value,
                                                    decltype( store )::value, false>;
        kernel<<<grid, block, 0, s>>>( src, tgt, n_tgt, indices, order, p.row0,
                                     p.row_off, p.rows, p.n_rows, p.kk,
                                     p.rbf, p.geometry, p.truncation, p.rcond,
                                     sweeps, coeffs, coeffs32, pinv, counters,
                                     p.row_max, p.refine_rows, p.n_refine,
                                     p.screened );
    };
    using T = std::true_type;
    using F = std::false_type;
    if ( coeffs32 )
        pinv ? run( T{}, T{} ) : run( T{}, F{} );
    else
        pinv ? run( F{}, T{} ) : run( F{}, F{} );
    DTKX_HIP( hipGetLastError() );
}

// Eigenvalues only, one lane per listed row: the cyclic Jacobi on A
// without eigenvectors. A rotation of A never reads V, and a converged
// system's rotations are exact no-ops whatever its wave, so the
// eigenvalues, the sweep count and the convergence flag are those
// cyclic_system computes for the row. Rows whose eigenvalues all pass the
// truncation test go to refine_rows (solved again by Cholesky, with their
// sweeps and convergence counted here), the others to full_rows (the
// eigen-solver with eigenvectors). Most rows the Cholesky setup lists are
// of the first kind: on 10-neighbour quadratic systems (1M targets), 466k
// of the 493k listed rows.
template <int M, bool F32>
__global__ void __launch_bounds__( kBlock )
    eigen_screen_kernel( double4 const *__restrict__ src,
                         double4 const *__restrict__ tgt,
                         int32_t const *__restrict__ indices,
                         int32_t const *__restrict__ order,
                         int64_t const *__restrict__ row_off,
                         int32_t const *__restrict__ rows,
                         int32_t const *__restrict__ n_rows, int kk_uniform,
                         int rbf, int geometry, int truncation, double rcond,
                         int max_sweeps,
                         unsigned long long *__restrict__ counters,
                         float const *__restrict__ row_max,
                         int32_t *__restrict__ refine_rows,
                         int32_t *__restrict__ n_refine,
                         int32_t *__restrict__ full_rows,
                         int32_t *__restrict__ n_full )
{
    using S = SymIdx<M>;
    // the trip count is uniform over the block (the wave-wide vote); tiles
    // in dispatch order spread the list's first tiles over the XCDs
    int64_t const n = *n_rows;
    for ( int64_t tile = blockIdx.x; tile * kBlock < n; tile += gridDim.x )
    {
        int64_t const li = tile * kBlock + threadIdx.x;
        bool const valid = li < n;
        int64_t const slot = valid ? rows[li] : 0;
        double A[S::size];
        (void)system_moments<M, F32>( valid, slot, src, tgt, indices, order,
                                      row_off, kk_uniform, rbf, geometry,
                                      row_max, A );
        double no_v[1];
        int sweeps = 0;
        bool done = !valid || converged<M>( A );
        for ( int pass = 0; !__all( done ) && pass < max_sweeps; ++pass )
        {
            sweep_all<M, 0>( A, no_v, done );
            sweeps += done ? 0 : 1;
            done = done || converged<M>( A );
        }
        if ( !valid )
            continue;
        double thr;
        if ( truncated<M>( A, truncation, rcond, thr ) )
            full_rows[atomicAdd( n_full, 1 )] = (int32_t)slot;
        else
        {
            refine_rows[atomicAdd( n_refine, 1 )] = (int32_t)slot;
            if ( counters )
            {
                if ( !done )
                    atomicAdd( counters + 1, 1ull );
                atomicAdd( counters + 2, (unsigned long long)sweeps );
            }
        }
    }
}

} // namespace

void setup_screen( double4 const *src, double4 const *tgt, int64_t n_tgt,
                   int32_t const *indices, int32_t const *order,
                   SetupParams const &p, bool f32, int32_t *full_rows,
                   int32_t *n_full, unsigned long long *counters,
                   hipStream_t s )
{
    DTKX_REQUIRE( p.rows && p.n_rows && p.refine_rows && p.n_refine,
                  DTKX_ERR_INVALID_ARGUMENT );
    if ( n_tgt <= p.row0 )
        return;
    int const sweeps = p.max_rot > 0 ? static_cast<int>( p.max_rot ) : 30;
    // one wave a block and up to three waves a SIMD: a grid of 2,048
    // blocks would leave the SIMDs two waves each
    int64_t blocks = ceil_div( n_tgt - p.row0, kBlock );
    blocks = blocks < kScreenBlocks ? blocks : kScreenBlocks;
    auto run = [&]( auto m, auto f32 ) {
        eigen_screen_kernel<decltype( m )::value, decltype( f32 )::value>
            <<<blocks, kBlock, 0, s>>>(
                src, tgt, indices, order, p.row_off, p.rows, p.n_rows, p.kk,
                p.rbf, p.geometry, p.truncation, p.rcond, sweeps, counters,
                p.row_max, p.refine_rows, p.n_refine, full_rows, n_full );
    };
    using T = std::true_type;
    using F = std::false_type;
    using M1 = std::integral_constant<int, 1>;
    using M4 = std::integral_constant<int, 4>;
    using M10 = std::integral_constant<int, 10>;
    if ( p.m == 1 )
        f32 ? run( M1{}, T{} ) : run( M1{}, F{} );
    else if ( p.m == 4 )
        f32 ? run( M4{}, T{} ) : run( M4{}, F{} );
    else if ( p.m == 10 )
        f32 ? run( M10{}, T{} ) : run( M10{}, F{} );
    else
        DTKX_REQUIRE( p.m == 1 || p.m == 4 || p.m == 10,
                      DTKX_ERR_INVALID_ARGUMENT );
    DTKX_HIP( hipGetLastError() );
}

void setup_cyclic( double4 const *src, double4 const *tgt, int64_t n_tgt,
                   int32_t const *indices, int32_t const *order,
                   SetupParams const &p, double *coeffs, float *coeffs32,
                   double *pinv, unsigned long long *counters, hipStream_t s )
{
    if ( n_tgt <= p.row0 )
        return;
    switch ( p.m )
    {
    case 1:
        launch_cyclic<1, 1>( src, tgt, n_tgt, indices, order, p, coeffs,
                                 coeffs32, pinv, counters, s );
        return;
    case 4:
        launch_cyclic<4, 1>( src, tgt, n_tgt, indices, order, p, coeffs,
                                 coeffs32, pinv, counters, s );
        return;
    case 10:
        launch_cyclic<10, 2>( src, tgt, n_tgt, indices, order, p, coeffs,
                                 coeffs32, pinv, counters, s );
        return;
    }
    DTKX_REQUIRE( p.m == 1 || p.m == 4 || p.m == 10, DTKX_ERR_INVALID_ARGUMENT );
}

} // namespace dtkx

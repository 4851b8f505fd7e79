// Fused MLS setup with the reference's classical Jacobi pseudo-inverse.
//
// Mapping: a workgroup is one wavefront; each target system is owned by a
// group of G lanes (64/G systems per wave). Per system, LDS holds the
// moment matrix E, the left rotations U, column 0 of V, row 0 of A^+, the
// Vandermonde rows P (kk x M) and the weights. Every step follows the
// reference's operation order with no FMA contraction, and the rotation
// angles use correctly rounded atan2/cos/sin, so the result rounds as the
// serial reference does with a correctly rounded libm:
//   transform + Vandermonde + distances  Impl.hpp:92-95, 184-187, 121-125
//   radius                               Impl.hpp:118-134
//   weights                              Impl.hpp:161-165
//   moments (l ascending per entry)      Impl.hpp:242-253
//   Jacobi: argmax (first strict max, row-major), swap p<q, svd_2x2,
//     givens_left(E), givens_right(E), givens_right(U), givens_left(V),
//     off-diagonal norm recomputed row-major, stop at norm <= DBL_EPSILON
//                                        SVD.hpp:224-274
//   pinv row 0 (k ascending)             SVD.hpp:281-297
//   coefficients (j ascending)           Impl.hpp:326-333
// Lanes of a group split the independent parts (neighbour rows, moment
// entries, rotation rows/columns); the argmax, the norm and the 2x2 SVD are
// evaluated redundantly by every lane of the group in the reference's
// sequential order, so their results do not depend on the lane split.
// The rotation loop is capped (the reference's has no cap); systems that
// hit the cap are counted as unconverged.
#include "mls_device.hpp"
#include "setup.hpp"

namespace dtkx
{
namespace
{

constexpr int kWave = 64;

template <int M>
__device__ inline double norm_off_diag( double const *E )
{
#pragma clang fp contract( off )
    double norm = 0.0;
#pragma unroll
    for ( int i = 0; i < M; i++ )
#pragma unroll
        for ( int j = 0; j < M; j++ )
            if ( i != j )
                norm += E[i * M + j] * E[i * M + j];
    return __builtin_sqrt( norm );
}

template <int M>
__device__ inline void argmax_off_diag( double const *E, int &p, int &q )
{
    p = -1;
    q = -1;
    double mx = -1;
#pragma unroll
    for ( int i = 0; i < M; i++ )
#pragma unroll
        for ( int j = 0; j < M; j++ )
            if ( i != j )
            {
                double v = fabs( E[i * M + j] );
                if ( v > mx )
                {
                    p = i;
                    q = j;
                    mx = v;
                }
            }
}

template <int M, int G>
__global__ void __launch_bounds__( kWave )
    setup_classical_kernel( double4 const *__restrict__ src,
                            double4 const *__restrict__ tgt, int64_t n_tgt,
                            int32_t const *__restrict__ indices,
                            int32_t const *__restrict__ order, int kk, int rbf,
                            int geometry, int truncation, double rcond,
                            long max_rot, double *__restrict__ coeffs,
                            double *__restrict__ a0_out,
                            unsigned long long *__restrict__ counters )
{
#pragma clang fp contract( off )
    constexpr int S = kWave / G;
    extern __shared__ double lds[];
    int const sys = threadIdx.x / G, lane = threadIdx.x % G;
    int64_t const slot = blockIdx.x * (int64_t)S + sys;
    bool const valid = slot < n_tgt;
    int64_t const t = valid ? ( order ? order[slot] : slot ) : 0;

    int const per_sys = 2 * M * M + 2 * M + kk * ( M + 1 );
    double *E = lds + sys * per_sys;
    double *U = E + M * M;
    double *V0 = U + M * M;
    double *A0 = V0 + M;
    double *P = A0 + M;
    double *PHI = P + kk * M;

    // neighbours: relative coordinates -> basis rows and distances
    double dmax = 10. * DBL_EPSILON;
    if ( valid )
    {
        double4 const tp = tgt[t];
        for ( int l = lane; l < kk; l += G )
        {
            int32_t const s = indices[t * kk + l];
            double4 const sp = src[s];
            double const xr[3] = {sp.x - tp.x, sp.y - tp.y, sp.z - tp.z};
            dev::basis_row<M>( geometry, xr, P + l * M );
            double const d = dev::rel_distance( geometry, xr );
            PHI[l] = d;
            if ( d > dmax )
                dmax = d;
        }
    }
#pragma unroll
    for ( int off = G / 2; off > 0; off >>= 1 )
        dmax = fmax( dmax, __shfl_xor( dmax, off, G ) );
    double const radius = 1.1 * dmax;
    __syncthreads();
    if ( valid )
        for ( int l = lane; l < kk; l += G )
            PHI[l] = dev::rbf_eval( rbf, PHI[l] / radius );
    __syncthreads();

    // moments, identity U, V e_0
    if ( valid )
    {
        for ( int e = lane; e < M * M; e += G )
        {
            int const r = e / M, c = e % M;
            double tmp = 0.;
            for ( int l = 0; l < kk; ++l )
                tmp += P[l * M + r] * PHI[l] * P[l * M + c];
            E[e] = tmp;
            U[e] = ( r == c ) ? 1.0 : 0.0;
        }
        for ( int j = lane; j < M; j += G )
            V0[j] = ( j == 0 ) ? 1.0 : 0.0;
    }
    __syncthreads();

    // classical Jacobi
    double norm = valid ? norm_off_diag<M>( E ) : 0.0;
    bool active = valid && norm > DBL_EPSILON;
    long rot = 0;
    while ( __syncthreads_or( active ) )
    {
        int p = 0, q = 0;
        double cl = 1, sl = 0, cr = 1, sr = 0;
        if ( active )
        {
            argmax_off_diag<M>( E, p, q );
            if ( p < 0 || q < 0 )
            {
                // unreachable (the loop runs only for a finite norm above
                // DBL_EPSILON); keeps every LDS index in range regardless
                p = 0;
                q = M - 1;
            }
            if ( p > q )
            {
                int const tt = p;
                p = q;
                q = tt;
            }
            dev::svd_2x2_rot( E[p * M + p], E[p * M + q], E[q * M + p],
                              E[q * M + q], cl, sl, cr, sr );
        }
        __syncthreads();
        if ( active )
            for ( int j = lane; j < M; j += G )
            {
                double const a = E[p * M + j], b = E[q * M + j];
                E[p * M + j] = cl * a - sl * b;
                E[q * M + j] = sl * a + cl * b;
            }
        __syncthreads();
        if ( active )
        {
            for ( int j = lane; j < M; j += G )
            {
                double const a = E[j * M + p], b = E[j * M + q];
                E[j * M + p] = cr * a - sr * b;
                E[j * M + q] = sr * a + cr * b;
                double const u = U[j * M + p], v = U[j * M + q];
                U[j * M + p] = cl * u - sl * v;
                U[j * M + q] = sl * u + cl * v;
            }
            if ( lane == 0 )
            {
                double const a = V0[p], b = V0[q];
                V0[p] = cr * a - sr * b;
                V0[q] = sr * a + cr * b;
            }
        }
        __syncthreads();
        if ( active )
        {
            norm = norm_off_diag<M>( E );
            ++rot;
            active = norm > DBL_EPSILON && rot < max_rot;
        }
    }

    // pseudo-inverse row 0 and coefficients
    bool under = false;
    if ( valid )
    {
        double thr = DBL_EPSILON;
        if ( truncation == DTKX_TRUNCATION_RELATIVE )
        {
            double smax = 0.;
            for ( int k = 0; k < M; ++k )
                smax = fmax( smax, fabs( E[k * M + k] ) );
            thr = rcond * smax;
        }
        for ( int k = 0; k < M; ++k )
            if ( !( fabs( E[k * M + k] ) >= thr ) )
                under = true;
        for ( int j = lane; j < M; j += G )
        {
            double value = 0;
            for ( int k = 0; k < M; k++ )
            {
                double const ekk = E[k * M + k];
                if ( fabs( ekk ) >= thr )
                    value += V0[k] * U[j * M + k] / ekk;
            }
            A0[j] = value;
            if ( a0_out )
                a0_out[t * M + j] = value;
        }
    }
    __syncthreads();
    if ( valid )
    {
        for ( int l = lane; l < kk; l += G )
        {
            double c = 0.;
            for ( int j = 0; j < M; j++ )
                c += A0[j] * P[l * M + j] * PHI[l];
            coeffs[t * kk + l] = c;
        }
        if ( lane == 0 && counters )
        {
            if ( under )
                atomicAdd( counters + 0, 1ull );
            if ( norm > DBL_EPSILON )
                atomicAdd( counters + 1, 1ull );
            atomicAdd( counters + 2, (unsigned long long)rot );
        }
    }
}

template <int M, int G>
void launch( double4 const *src, double4 const *tgt, int64_t n_tgt,
             int32_t const *indices, int32_t const *order,
             SetupParams const &p, double *coeffs, double *a0,
             unsigned long long *counters, hipStream_t s )
{
    constexpr int S = kWave / G;
    std::size_t const lds =
        sizeof( double ) * S * ( 2 * M * M + 2 * M + p.kk * ( M + 1 ) );
    DTKX_REQUIRE( lds <= 64 * 1024, DTKX_ERR_UNSUPPORTED );
    setup_classical_kernel<M, G><<<ceil_div( n_tgt, S ), kWave, lds, s>>>(
        src, tgt, n_tgt, indices, order, p.kk, p.rbf, p.geometry, p.truncation,
        p.rcond, p.max_rot, coeffs, a0, counters );
    DTKX_HIP( hipGetLastError() );
}

std::size_t lds_bytes( int m, int kk, int g )
{
    return sizeof( double ) * ( kWave / g ) *
           ( 2 * m * m + 2 * m + kk * ( m + 1 ) );
}

} // namespace

int setup_classical_group( int m, int kk )
{
    if ( m == 1 )
        return 1;
    // smallest group (most systems per wave) whose LDS fits 32 KB, so that
    // several waves share a CU
    int const opts[] = {m == 4 ? 4 : 8, 16};
    for ( int g : opts )
        if ( lds_bytes( m, kk, g ) <= 32 * 1024 )
            return g;
    return 16;
}

int setup_classical( double4 const *src, double4 const *tgt, int64_t n_tgt,
                     int32_t const *indices, int32_t const *order,
                     SetupParams const &p, double *coeffs, double *a0,
                     unsigned long long *counters, hipStream_t s )
{
    if ( n_tgt == 0 )
        return 0;
    int g = p.group_size > 0 ? p.group_size : setup_classical_group( p.m, p.kk );
    switch ( p.m )
    {
    case 1:
        launch<1, 1>( src, tgt, n_tgt, indices, order, p, coeffs, a0, counters,
                      s );
        return 1;
    case 4:
        if ( g <= 1 )
        {
            launch<4, 1>( src, tgt, n_tgt, indices, order, p, coeffs, a0,
                          counters, s );
            return 1;
        }
        if ( g == 2 )
        {
            launch<4, 2>( src, tgt, n_tgt, indices, order, p, coeffs, a0,
                          counters, s );
            return 2;
        }
        launch<4, 4>( src, tgt, n_tgt, indices, order, p, coeffs, a0, counters,
                      s );
        return 4;
    case 10:
        if ( g <= 4 )
        {
            launch<10, 4>( src, tgt, n_tgt, indices, order, p, coeffs, a0,
                           counters, s );
            return 4;
        }
        if ( g == 8 )
        {
            launch<10, 8>( src, tgt, n_tgt, indices, order, p, coeffs, a0,
                           counters, s );
            return 8;
        }
        launch<10, 16>( src, tgt, n_tgt, indices, order, p, coeffs, a0,
                        counters, s );
        return 16;
    }
    DTKX_REQUIRE( p.m == 1 || p.m == 4 || p.m == 10, DTKX_ERR_INVALID_ARGUMENT );
    return 0;
}

// ------------------------------------------------- general batched SVD
// SVDFunctor on n x n general matrices (n <= 64): one matrix per wave,
// E, U, V in LDS, full pseudo-inverse. Same operation order as above.
namespace
{

__global__ void __launch_bounds__( kWave )
    svd_general_kernel( int n, int64_t n_mat, double const *__restrict__ a,
                        double *__restrict__ pinv, long max_rot,
                        unsigned long long *__restrict__ counters )
{
#pragma clang fp contract( off )
    extern __shared__ double lds[];
    int64_t const mi = blockIdx.x;
    int const lane = threadIdx.x;
    int const nn = n * n;
    double *E = lds, *U = E + nn, *V = U + nn;
    double const *A = a + mi * nn;
    for ( int e = lane; e < nn; e += kWave )
    {
        E[e] = A[e];
        U[e] = ( e / n == e % n ) ? 1.0 : 0.0;
        V[e] = U[e];
    }
    __syncthreads();
    auto norm_off = [&]() {
        double norm = 0.0;
        for ( int i = 0; i < n; i++ )
            for ( int j = 0; j < n; j++ )
                if ( i != j )
                    norm += E[i * n + j] * E[i * n + j];
        return __builtin_sqrt( norm );
    };
    // the norm and argmax are computed by lane 0 and broadcast through LDS
    __shared__ double sh_norm;
    __shared__ int sh_pq[2];
    if ( lane == 0 )
        sh_norm = norm_off();
    __syncthreads();
    double norm = sh_norm;
    long rot = 0;
    while ( norm > DBL_EPSILON && rot < max_rot )
    {
        if ( lane == 0 )
        {
            int p = -1, q = -1;
            double mx = -1;
            for ( int i = 0; i < n; i++ )
                for ( int j = 0; j < n; j++ )
                    if ( i != j && fabs( E[i * n + j] ) > mx )
                    {
                        p = i;
                        q = j;
                        mx = fabs( E[i * n + j] );
                    }
            if ( p < 0 || q < 0 )
            {
                p = 0; // unreachable, see setup_classical_kernel
                q = n - 1;
            }
            if ( p > q )
            {
                int const tt = p;
                p = q;
                q = tt;
            }
            sh_pq[0] = p;
            sh_pq[1] = q;
        }
        __syncthreads();
        int const p = sh_pq[0], q = sh_pq[1];
        double cl, sl, cr, sr;
        dev::svd_2x2_rot( E[p * n + p], E[p * n + q], E[q * n + p],
                          E[q * n + q], cl, sl, cr, sr );
        __syncthreads();
        for ( int j = lane; j < n; j += kWave )
        {
            double const x = E[p * n + j], y = E[q * n + j];
            E[p * n + j] = cl * x - sl * y;
            E[q * n + j] = sl * x + cl * y;
        }
        __syncthreads();
        for ( int j = lane; j < n; j += kWave )
        {
            double const x = E[j * n + p], y = E[j * n + q];
            E[j * n + p] = cr * x - sr * y;
            E[j * n + q] = sr * x + cr * y;
            double const u = U[j * n + p], v = U[j * n + q];
            U[j * n + p] = cl * u - sl * v;
            U[j * n + q] = sl * u + cl * v;
            double const g = V[p * n + j], h = V[q * n + j];
            V[p * n + j] = cr * g - sr * h;
            V[q * n + j] = sr * g + cr * h;
        }
        __syncthreads();
        if ( lane == 0 )
            sh_norm = norm_off();
        __syncthreads();
        norm = sh_norm;
        ++rot;
    }
    bool under = false;
    for ( int k = 0; k < n; ++k )
        if ( !( fabs( E[k * n + k] ) >= DBL_EPSILON ) )
            under = true;
    for ( int e = lane; e < nn; e += kWave )
    {
        int const i = e / n, j = e % n;
        double value = 0;
        for ( int k = 0; k < n; k++ )
        {
            double const ekk = E[k * n + k];
            if ( fabs( ekk ) >= DBL_EPSILON )
                value += V[k * n + i] * U[j * n + k] / ekk;
        }
        pinv[mi * nn + e] = value;
    }
    if ( lane == 0 && counters )
    {
        if ( under )
            atomicAdd( counters + 0, 1ull );
        if ( norm > DBL_EPSILON )
            atomicAdd( counters + 1, 1ull );
        atomicAdd( counters + 2, (unsigned long long)rot );
    }
}

} // namespace

void svd_pinv_batched( int n, int64_t n_mat, double const *a, double *pinv,
                       long max_rot, unsigned long long *counters,
                       hipStream_t s )
{
    DTKX_REQUIRE( n >= 1 && n <= 64, DTKX_ERR_UNSUPPORTED );
    if ( n_mat == 0 )
        return;
    std::size_t const lds = sizeof( double ) * 3 * n * n;
    svd_general_kernel<<<static_cast<unsigned>( n_mat ), kWave, lds, s>>>(
        n, n_mat, a, pinv, max_rot, counters );
    DTKX_HIP( hipGetLastError() );
}

} // namespace dtkx

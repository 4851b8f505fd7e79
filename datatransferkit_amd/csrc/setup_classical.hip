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

// neighbours staged in LDS at a time: M of them fill exactly the space of
// U and V(:,0) they share
constexpr int chunk_of( int m ) { return m == 1 ? 16 : m; }

// Per-system LDS (doubles): E [M*M], a region W shared by the staged
// Vandermonde chunk (P [chunk*M] + weights [chunk]) during the moments and
// by U [M*M] + V(:,0) [M] during the Jacobi, and A^+ row 0 [M].
constexpr int region_w( int m )
{
    return m * m + m > chunk_of( m ) * ( m + 1 ) ? m * m + m
                                                 : chunk_of( m ) * ( m + 1 );
}
constexpr int per_system( int m ) { return m * m + region_w( m ) + m; }

// Exact off-diagonal norm of the reference (row-major sequential sum, then
// a correctly rounded sqrt), compared with DBL_EPSILON.
template <int M>
__device__ inline bool norm_above_tol_exact( double const *E )
{
#pragma clang fp contract( off )
    double norm = 0.0;
    for ( int i = 0; i < M; i++ )
        for ( int j = 0; j < M; j++ )
            if ( i != j )
                norm += E[i * M + j] * E[i * M + j];
    return __builtin_sqrt( norm ) > DBL_EPSILON;
}

// The neighbour-independent geometry of one neighbour: basis row and the
// distance used by the radius and the weights (Impl.hpp:121-125, 162-165,
// 184-187), recomputed bit-identically wherever it is needed.
template <int M>
__device__ inline double neighbour_row( int geometry, double4 const &sp,
                                        double4 const &tp, double *P )
{
#pragma clang fp contract( off )
    double const xr[3] = {sp.x - tp.x, sp.y - tp.y, sp.z - tp.z};
    dev::basis_row<M>( geometry, xr, P );
    return dev::rel_distance( geometry, xr );
}

template <int M, int G>
__global__ void __launch_bounds__( kWave )
    setup_classical_kernel( double4 const *__restrict__ src,
                            double4 const *__restrict__ tgt, int64_t n_tgt,
                            int32_t const *__restrict__ indices,
                            int32_t const *__restrict__ order,
                            int64_t row0, int64_t const *__restrict__ row_off,
                            int kk_uniform, int rbf,
                            int geometry, int truncation, double rcond,
                            long max_rot, double *__restrict__ coeffs,
                            unsigned long long *__restrict__ counters )
{
#pragma clang fp contract( off )
    constexpr int S = kWave / G;
    constexpr int MM = M * M;
    // contiguous row-major chunk of E scanned by each lane
    constexpr int CS = ( MM + G - 1 ) / G;
    extern __shared__ double lds[];
    int const sys = threadIdx.x / G, lane = threadIdx.x % G;
    int const gbase = threadIdx.x - lane; // first lane of the group
    int64_t const slot = row0 + xcd_tile() * S + sys;
    bool const valid = slot < n_tgt;
    int64_t const t = valid ? ( order ? order[slot] : slot ) : 0;
    int64_t const rbeg =
        valid ? ( row_off ? row_off[slot] : slot * kk_uniform ) : 0;
    int const kk =
        valid ? ( row_off ? (int)( row_off[slot + 1] - rbeg ) : kk_uniform )
              : 0;
    int32_t const *nbr = indices + rbeg;
    double4 const tp = valid ? tgt[t] : double4{0, 0, 0, 0};
    // the staged moment loop has barriers: its trip count is the block's
    // largest row length (rows of one launch may differ in a batch)
    __shared__ int kk_block;
    if ( threadIdx.x == 0 )
        kk_block = 0;
    __syncthreads();
    if ( lane == 0 && valid )
        atomicMax( &kk_block, kk );
    __syncthreads();
    int const kk_max = kk_block;

    double *E = lds + sys * per_system( M );
    double *W = E + MM;
    double *U = W, *V0 = W + MM; // Jacobi view of W
    constexpr int kChunk = chunk_of( M );
    double *P = W, *PHI = W + kChunk * M; // moments view of W
    double *A0 = W + region_w( M );

    // radius (Impl.hpp:118-134): max is exact, so the lane split is free
    double dmax = 10. * DBL_EPSILON;
    if ( valid )
        for ( int l = lane; l < kk; l += G )
        {
            double Pl[M];
            double const d = neighbour_row<M>( geometry, src[nbr[l]], tp, Pl );
            if ( d > dmax )
                dmax = d;
        }
#pragma unroll
    for ( int off = G / 2; off > 0; off >>= 1 )
        dmax = fmax( dmax, __shfl_xor( dmax, off, G ) );
    double const radius = 1.1 * dmax;

    // moments A = P^T Phi P, each entry summed over l ascending
    // (Impl.hpp:242-253), neighbours staged kChunk at a time; E holds the
    // running sums between chunks, so the order of additions is unchanged
    if ( valid )
        for ( int e = lane; e < MM; e += G )
            E[e] = 0.0;
    for ( int l0 = 0; l0 < kk_max; l0 += kChunk )
    {
        int const nl =
            kk - l0 < kChunk ? ( kk - l0 > 0 ? kk - l0 : 0 ) : kChunk;
        if ( valid )
            for ( int l = lane; l < nl; l += G )
            {
                double const d = neighbour_row<M>( geometry, src[nbr[l0 + l]],
                                                   tp, P + l * M );
                PHI[l] = dev::rbf_eval( rbf, d / radius );
            }
        __syncthreads();
        if ( valid )
            for ( int e = lane; e < MM; e += G )
            {
                int const r = e / M, c = e % M;
                double tmp = E[e];
                for ( int l = 0; l < nl; ++l )
                    tmp += P[l * M + r] * PHI[l] * P[l * M + c];
                E[e] = tmp;
            }
        __syncthreads();
    }
    if ( valid )
    {
        for ( int e = lane; e < MM; e += G )
            U[e] = ( e / M == e % M ) ? 1.0 : 0.0;
        for ( int j = lane; j < M; j += G )
            V0[j] = ( j == 0 ) ? 1.0 : 0.0;
    }
    __syncthreads();

    // classical Jacobi (SVD.hpp:236-274)
    bool active = valid;
    long rot = 0;
    bool unconverged = false;
    // the off-diagonal entries of this lane's row-major chunk (e < M*M and
    // e != r (M + 1)), fixed for the whole iteration
    int const e0 = lane * CS;
    uint64_t offdiag = 0;
#pragma unroll
    for ( int i = 0; i < CS; ++i )
        if ( e0 + i < MM && ( e0 + i ) % ( M + 1 ) != 0 )
            offdiag |= uint64_t( 1 ) << i;
    while ( true )
    {
        // fused pass over this lane's row-major chunk: partial off-diagonal
        // norm and the first strict maximum of |E(i,j)|, i != j (ascending
        // e with a strict compare keeps the first). Branch-free: diagonal
        // and out-of-range entries add zero to the norm and enter the
        // maximum as -1, its starting value. The partial norm only decides
        // against a margin (below), so it may use fma.
        double part = 0.0, mx = -1.0;
        int mi = INT32_MAX;
        if ( active )
        {
            // compile-time trip count: all LDS loads of the chunk issue
            // back to back instead of one load latency per entry
            double v[CS];
#pragma unroll
            for ( int i = 0; i < CS; ++i )
                v[i] = ( e0 + i < MM ) ? E[e0 + i] : 0.0;
#pragma unroll
            for ( int i = 0; i < CS; ++i )
            {
                bool const od = ( offdiag >> i ) & 1;
                double const vo = od ? v[i] : 0.0; // a NaN still propagates
                part = fma( vo, vo, part );
                double const av = od ? fabs( v[i] ) : -1.0;
                bool const gt = av > mx;
                mx = gt ? av : mx;
                mi = gt ? e0 + i : mi;
            }
        }
        // group reductions: the sum in a fixed tree order; the maximum with
        // ties to the lower row-major index, which is the reference's first
        // strict maximum
#pragma unroll
        for ( int off = 1; off < G; off <<= 1 )
        {
            double const op = __shfl_xor( part, off, G );
            double const om = __shfl_xor( mx, off, G );
            int const oi = __shfl_xor( mi, off, G );
            part = ( lane & off ) ? op + part : part + op;
            if ( om > mx || ( om == mx && oi < mi ) )
            {
                mx = om;
                mi = oi;
            }
        }
        if ( active )
        {
            // norm > DBL_EPSILON decided on the tree sum unless it lies within
            // a relative 1e-12 of DBL_EPSILON^2 (the two sums of the same
            // non-negative terms differ by less than 2e-14 relative); then the
            // reference's sequential sum decides
            double const tol2 = DBL_EPSILON * DBL_EPSILON;
            bool above;
            if ( part > tol2 * ( 1.0 + 1e-12 ) )
                above = true;
            else if ( part < tol2 * ( 1.0 - 1e-12 ) )
                above = false;
            else
                above = norm_above_tol_exact<M>( E );
            if ( !above )
                active = false;
            else if ( rot >= max_rot )
            {
                active = false;
                unconverged = true;
            }
        }
        if ( !__syncthreads_or( active ) )
            break;
        int p = 0, q = 1;
        double cl = 1, sl = 0, cr = 1, sr = 0;
        if ( active )
        {
            p = mi / M;
            q = mi % M;
            if ( mi >= MM || mi < 0 )
            {
                p = 0; // unreachable: an active system has a finite maximum
                q = M - 1;
            }
            if ( p > q )
            {
                int const tt = p;
                p = q;
                q = tt;
            }
            double const a00 = E[p * M + p], a01 = E[p * M + q];
            double const a10 = E[q * M + p], a11 = E[q * M + q];
            // the two angle chains of svd_2x2 run on lanes 0 and 1 of the
            // group (identical code on different inputs); with four lanes
            // or more, lanes 2 and 3 repeat the angles and each of the four
            // takes one of sin/cos (one polynomial, see crm::sin_or_cos)
            double y, x;
            dev::svd_2x2_angle_args( a00, a01, a10, a11, ( lane & 1 ) != 0, y,
                                     x );
            double const ang = 0.5 * crm::atan2( y, x );
            double cphi, sphi, ctheta, stheta;
            if ( G >= 4 )
            {
                double const sc = crm::sin_or_cos( ang, ( lane >> 1 ) & 1 );
                sphi = __shfl( sc, gbase, kWave );
                stheta = __shfl( sc, gbase + 1, kWave );
                cphi = __shfl( sc, gbase + 2, kWave );
                ctheta = __shfl( sc, gbase + 3, kWave );
            }
            else if ( G > 1 )
            {
                double sa, ca;
                crm::sincos( ang, &sa, &ca );
                cphi = __shfl( ca, gbase, kWave );
                sphi = __shfl( sa, gbase, kWave );
                ctheta = __shfl( ca, gbase + 1, kWave );
                stheta = __shfl( sa, gbase + 1, kWave );
            }
            else
            {
                double sa, ca;
                crm::sincos( ang, &sa, &ca );
                cphi = ca;
                sphi = sa;
                double y2, x2;
                dev::svd_2x2_angle_args( a00, a01, a10, a11, true, y2, x2 );
                crm::sincos( 0.5 * crm::atan2( y2, x2 ), &stheta, &ctheta );
            }
            dev::svd_2x2_from_angles( a00, a01, a10, a11, cphi, sphi, ctheta,
                                      stheta, cl, sl, cr, sr );
        }
        constexpr int JR = ( M + G - 1 ) / G; // rows/columns per lane
        if ( active )
#pragma unroll
            for ( int jj = 0; jj < JR; ++jj )
            {
                int const j = lane + jj * G;
                if ( j < M )
                {
                    double const a = E[p * M + j], b = E[q * M + j];
                    E[p * M + j] = cl * a - sl * b;
                    E[q * M + j] = sl * a + cl * b;
                }
            }
        __syncthreads();
        if ( active )
        {
#pragma unroll
            for ( int jj = 0; jj < JR; ++jj )
            {
                int const j = lane + jj * G;
                if ( j < M )
                {
                    double const a = E[j * M + p], b = E[j * M + q];
                    E[j * M + p] = cr * a - sr * b;
                    E[j * M + q] = sr * a + cr * b;
                    double const u = U[j * M + p], v = U[j * M + q];
                    U[j * M + p] = cl * u - sl * v;
                    U[j * M + q] = sl * u + cl * v;
                }
            }
            if ( lane == 0 )
            {
                double const a = V0[p], b = V0[q];
                V0[p] = cr * a - sr * b;
                V0[q] = sr * a + cr * b;
            }
            ++rot;
        }
        __syncthreads();
    }

    // pseudo-inverse row 0 (SVD.hpp:281-297)
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
        }
    }
    __syncthreads();
    if ( valid )
    {
        // coefficients (Impl.hpp:326-333), the basis row recomputed
        double a0[M];
#pragma unroll
        for ( int j = 0; j < M; ++j )
            a0[j] = A0[j];
        for ( int l = lane; l < kk; l += G )
        {
            double Pl[M];
            double const d = neighbour_row<M>( geometry, src[nbr[l]], tp, Pl );
            double const phi = dev::rbf_eval( rbf, d / radius );
            double c = 0.;
#pragma unroll
            for ( int j = 0; j < M; j++ )
                c += a0[j] * Pl[j] * phi;
            coeffs[rbeg + l] = c;
        }
        if ( lane == 0 && counters )
        {
            if ( under )
                atomicAdd( counters + 0, 1ull );
            if ( unconverged )
                atomicAdd( counters + 1, 1ull );
            atomicAdd( counters + 2, (unsigned long long)rot );
        }
    }
}

std::size_t lds_bytes( int m, int g )
{
    return sizeof( double ) * ( kWave / g ) * per_system( m );
}

template <int M, int G>
void launch( double4 const *src, double4 const *tgt, int64_t n_tgt,
             int32_t const *indices, int32_t const *order,
             SetupParams const &p, double *coeffs,
             unsigned long long *counters, hipStream_t s )
{
    constexpr int S = kWave / G;
    std::size_t const lds = lds_bytes( M, G );
    DTKX_REQUIRE( lds <= 64 * 1024, DTKX_ERR_UNSUPPORTED );
    setup_classical_kernel<M, G>
        <<<ceil_div( n_tgt - p.row0, S ), kWave, lds, s>>>(
            src, tgt, n_tgt, indices, order, p.row0, p.row_off, p.kk, p.rbf,
            p.geometry, p.truncation, p.rcond, p.max_rot, coeffs, counters );
    DTKX_HIP( hipGetLastError() );
}

} // namespace

int setup_classical_group( int m, int kk )
{
    (void)kk; // the staged chunk bounds the LDS independently of kk
    if ( m == 1 )
        return 1;
    return 4;
}

int setup_classical( double4 const *src, double4 const *tgt, int64_t n_tgt,
                     int32_t const *indices, int32_t const *order,
                     SetupParams const &p, double *coeffs,
                     unsigned long long *counters, hipStream_t s )
{
    if ( n_tgt <= p.row0 )
        return 0;
    int g = p.group_size > 0 ? p.group_size : setup_classical_group( p.m, p.kk );
    switch ( p.m )
    {
    case 1:
        launch<1, 1>( src, tgt, n_tgt, indices, order, p, coeffs, counters,
                      s );
        return 1;
    case 4:
        if ( g <= 1 )
        {
            launch<4, 1>( src, tgt, n_tgt, indices, order, p, coeffs,
                          counters, s );
            return 1;
        }
        if ( g == 2 )
        {
            launch<4, 2>( src, tgt, n_tgt, indices, order, p, coeffs,
                          counters, s );
            return 2;
        }
        launch<4, 4>( src, tgt, n_tgt, indices, order, p, coeffs, counters,
                      s );
        return 4;
    case 10:
        if ( g <= 2 )
        {
            launch<10, 2>( src, tgt, n_tgt, indices, order, p, coeffs,
                           counters, s );
            return 2;
        }
        if ( g <= 4 )
        {
            launch<10, 4>( src, tgt, n_tgt, indices, order, p, coeffs,
                           counters, s );
            return 4;
        }
        if ( g == 8 )
        {
            launch<10, 8>( src, tgt, n_tgt, indices, order, p, coeffs,
                           counters, s );
            return 8;
        }
        launch<10, 16>( src, tgt, n_tgt, indices, order, p, coeffs,
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

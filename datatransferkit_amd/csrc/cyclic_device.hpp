// Register-resident cyclic symmetric Jacobi and the per-neighbour geometry
// shared by the cyclic setup (setup_cyclic.hip) and the adjoint kernels
// (gradient.hip). Every (p, q) index is a compile-time constant (fully
// unrolled sweeps), so the packed upper triangle of A and the rows of V a
// lane owns stay in registers.
#ifndef DTKX_CYCLIC_DEVICE_HPP
#define DTKX_CYCLIC_DEVICE_HPP

#include "mls_device.hpp"

namespace dtkx
{
namespace cyc
{

template <int M>
struct SymIdx
{
    // index of (r, c), r <= c, in the packed upper triangle
    static constexpr int at( int r, int c )
    {
        return r <= c ? r * M - r * ( r - 1 ) / 2 + ( c - r )
                      : c * M - c * ( c - 1 ) / 2 + ( r - c );
    }
    static constexpr int size = M * ( M + 1 ) / 2;
};

template <int M, typename T>
__device__ inline void basis_fast( T x, T y, T z, double *P )
{
    P[0] = 1.;
    if ( M >= 4 )
    {
        P[1] = x;
        P[2] = y;
        P[3] = z;
    }
    if ( M >= 10 )
    {
        P[4] = x * x;
        P[5] = x * y;
        P[6] = x * z;
        P[7] = y * y;
        P[8] = y * z;
        P[9] = z * z;
    }
}

template <int M>
__device__ inline void basis_any( int geometry, double const *xr, double *P )
{
    if ( geometry == DTKX_GEOMETRY_FLOAT )
        basis_fast<M, float>( (float)xr[0], (float)xr[1], (float)xr[2], P );
    else
        basis_fast<M, double>( xr[0], xr[1], xr[2], P );
}

// Geometry of one neighbour: the distance (radius pass) and, given the
// radius, the basis row and weight. FP64 form: the reference's operations
// (float or double geometry). FP32 form: float relative coordinates,
// distance, monomials and weight; the basis row is returned widened so the
// moments are formed in FP64.
template <int M, bool F32>
__device__ inline double nbr_distance( int geometry, double4 const &sp,
                                       double4 const &tp )
{
    if constexpr ( F32 )
    {
        float const x = (float)( sp.x - tp.x ), y = (float)( sp.y - tp.y ),
                    z = (float)( sp.z - tp.z );
        return (double)__builtin_sqrtf( x * x + y * y + z * z );
    }
    else
    {
        double const xr[3] = {sp.x - tp.x, sp.y - tp.y, sp.z - tp.z};
        return dev::rel_distance( geometry, xr );
    }
}

template <int M, bool F32>
__device__ inline double nbr_row( int geometry, int rbf, double radius,
                                  double4 const &sp, double4 const &tp,
                                  double *P )
{
    if constexpr ( F32 )
    {
        float const x = (float)( sp.x - tp.x ), y = (float)( sp.y - tp.y ),
                    z = (float)( sp.z - tp.z );
        basis_fast<M, float>( x, y, z, P );
        float const d = __builtin_sqrtf( x * x + y * y + z * z );
        return (double)dev::rbf_eval_f( rbf, d / (float)radius );
    }
    else
    {
        double const xr[3] = {sp.x - tp.x, sp.y - tp.y, sp.z - tp.z};
        basis_any<M>( geometry, xr, P );
        return dev::rbf_eval( rbf, dev::rel_distance( geometry, xr ) / radius );
    }
}

// V holds R rows of the eigenvector matrix (all of it when R == M).
// skip: a system that has converged while its wave goes on. t = 0 makes
// every update of V and of A's diagonal and other entries add an exact
// zero, so the eigenvalues and vectors keep their bits without a branch
// around the straight-line sweep; only A(p, q), already negligible and read
// by nothing after convergence, is cleared as usual.
template <int M, int R, int P_, int Q_>
__device__ inline void rotate( double *A, double *V, bool skip )
{
    using S = SymIdx<M>;
    double const apq = A[S::at( P_, Q_ )];
    double const app = A[S::at( P_, P_ )];
    double const aqq = A[S::at( Q_, Q_ )];
    double const theta = ( aqq - app ) / ( 2.0 * apq );
    double const at = fabs( theta );
    // huge |theta|: theta^2 overflows, t -> 1/(2 theta), formed as
    // 0.5 / |theta| with the sign of theta (the same bits as 0.5 / theta).
    // One division for both cases: a select between two divisions became a
    // branch per rotation, and the unrolled sweep's registers tripled.
    bool const huge = at > 1e150;
    double const den =
        huge ? at : at + __builtin_sqrt( fma( theta, theta, 1.0 ) );
    double t = ( huge ? 0.5 : 1.0 ) / den;
    t = theta < 0 ? -t : t;
    // apq == 0: no-op
    t = ( apq == 0.0 || skip ) ? 0.0 : t;
    double const c = 1.0 / __builtin_sqrt( fma( t, t, 1.0 ) );
    double const s = t * c;
    double const tau = s / ( 1.0 + c );
    double const h = t * apq;
    A[S::at( P_, P_ )] = app - h;
    A[S::at( Q_, Q_ )] = aqq + h;
    A[S::at( P_, Q_ )] = 0.0;
#pragma unroll
    for ( int r = 0; r < M; ++r )
    {
        if ( r == P_ || r == Q_ )
            continue;
        double const arp = A[S::at( r, P_ )];
        double const arq = A[S::at( r, Q_ )];
        A[S::at( r, P_ )] = arp - s * ( arq + tau * arp );
        A[S::at( r, Q_ )] = arq + s * ( arp - tau * arq );
    }
#pragma unroll
    for ( int r = 0; r < R; ++r )
    {
        double const vrp = V[r * M + P_];
        double const vrq = V[r * M + Q_];
        V[r * M + P_] = vrp - s * ( vrq + tau * vrp );
        V[r * M + Q_] = vrq + s * ( vrp - tau * vrq );
    }
}

// one cyclic sweep in row order (0,1), (0,2), ..., (M-2, M-1)
template <int M, int R, int P_, int Q_>
struct Sweep
{
    __device__ static inline void run( double *A, double *V, bool skip )
    {
        rotate<M, R, P_, Q_>( A, V, skip );
        if constexpr ( Q_ + 1 < M )
            Sweep<M, R, P_, Q_ + 1>::run( A, V, skip );
        else if constexpr ( P_ + 2 < M )
            Sweep<M, R, P_ + 1, P_ + 2>::run( A, V, skip );
    }
};

template <int M, int R>
__device__ inline void sweep_all( double *A, double *V, bool skip )
{
    if constexpr ( M >= 2 )
        Sweep<M, R, 0, 1>::run( A, V, skip );
}

template <int M>
__device__ inline bool converged( double const *A )
{
    using S = SymIdx<M>;
    double off = 0.0, diag = 0.0;
#pragma unroll
    for ( int r = 0; r < M; ++r )
    {
        diag = fma( A[S::at( r, r )], A[S::at( r, r )], diag );
#pragma unroll
        for ( int c = r + 1; c < M; ++c )
            off = fma( A[S::at( r, c )], A[S::at( r, c )], off );
    }
    constexpr double tol = ( DBL_EPSILON / 8 ) * ( DBL_EPSILON / 8 );
    return !( off > tol * diag );
}

// Cholesky inverse factor. A (packed upper triangle) becomes V = U^-1 in
// place, where A = U^T U (left-looking, each pivot's reciprocal kept so a
// column divides once; non-positive pivots are replaced by one and
// reported: the return value is whether every pivot was positive). Then
// A^-1 = V V^T. Cholesky's rounding error is bounded relative to
// sqrt(a_ii a_jj), so its accuracy is that of the equilibrated matrix
// D A D (D = diag(a_ii^-1/2)), whatever the monomials' scales.
template <int M>
__device__ inline bool cholesky_factor_inverse( double *A )
{
    using S = SymIdx<M>;
    bool ok = true;
    double rinv[M];
#pragma unroll
    for ( int j = 0; j < M; ++j )
    {
#pragma unroll
        for ( int i = 0; i <= j; ++i )
        {
            double s = A[S::at( i, j )];
#pragma unroll
            for ( int k = 0; k < i; ++k )
                s = fma( -A[S::at( k, i )], A[S::at( k, j )], s );
            if ( i < j )
                A[S::at( i, j )] = s * rinv[i];
            else
            {
                ok = ok && s > 0.0;
                double const u = __builtin_sqrt( s > 0.0 ? s : 1.0 );
                A[S::at( j, j )] = u;
                rinv[j] = 1.0 / u;
            }
        }
    }
    // V = U^-1 in place: column j from the columns before it
#pragma unroll
    for ( int j = 0; j < M; ++j )
    {
#pragma unroll
        for ( int i = 0; i < j; ++i )
        {
            double s = 0.0;
#pragma unroll
            for ( int k = i; k < j; ++k )
                s = fma( A[S::at( i, k )], A[S::at( k, j )], s );
            A[S::at( i, j )] = -s * rinv[j];
        }
        A[S::at( j, j )] = rinv[j];
    }
    return ok;
}

// Certified Cholesky inverse: cholesky_factor_inverse, then whether every
// eigenvalue of A (trace tr) is proven above the truncation threshold:
// lambda_min(A) = 1 / |V|_2^2 >= 1 / |V|_F^2 > 4 thr + 64 m eps trace,
// thr = DBL_EPSILON or rcond * trace (>= rcond * lambda_max), the margin
// covering the rounding of the eigenvalues an eigen-solver would compute.
// Then A^+ = A^-1 = V V^T.
template <int M>
__device__ inline bool cholesky_inverse( double *A, double tr, int truncation,
                                         double rcond )
{
    using S = SymIdx<M>;
    bool const ok = cholesky_factor_inverse<M>( A );
    double fro = 0.0;
#pragma unroll
    for ( int e = 0; e < S::size; ++e )
        fro = fma( A[e], A[e], fro );
    double const lambda_lb = 1.0 / fro;
    double const thr =
        truncation == DTKX_TRUNCATION_RELATIVE ? rcond * tr : DBL_EPSILON;
    return ok && lambda_lb > 4.0 * thr + 64.0 * M * DBL_EPSILON * tr &&
           lambda_lb == lambda_lb;
}

// y = V V^T x for the inverse factor V (upper triangle, packed)
template <int M>
__device__ inline void apply_vvt( double const *V, double const *x, double *y )
{
    using S = SymIdx<M>;
    double w[M];
#pragma unroll
    for ( int k = 0; k < M; ++k )
    {
        double s = 0.0;
#pragma unroll
        for ( int i = 0; i <= k; ++i )
            s = fma( V[S::at( i, k )], x[i], s );
        w[k] = s;
    }
#pragma unroll
    for ( int i = 0; i < M; ++i )
    {
        double s = 0.0;
#pragma unroll
        for ( int k = i; k < M; ++k )
            s = fma( V[S::at( i, k )], w[k], s );
        y[i] = s;
    }
}

} // namespace cyc
} // namespace dtkx

#endif

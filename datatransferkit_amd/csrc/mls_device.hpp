// Device building blocks of the MLS setup, written to round exactly as the
// reference's serial code does: no FMA contraction, the reference's
// evaluation order, correctly rounded elementary functions (crmath.hpp).
#ifndef DTKX_MLS_DEVICE_HPP
#define DTKX_MLS_DEVICE_HPP

#include "common.hpp"
#include "crmath.hpp"

#include <cfloat>

namespace dtkx
{
namespace dev
{

// Compactly supported RBFs (DTK_CompactlySupportedRadialBasisFunctions.hpp
// :45-147), evaluated left to right as written there.
__device__ inline double rbf_eval( int rbf, double x )
{
#pragma clang fp contract( off )
    switch ( rbf )
    {
    case DTKX_RBF_WENDLAND0:
        return ( 1.0 - x ) * ( 1.0 - x );
    case DTKX_RBF_WENDLAND2:
        return ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) *
               ( 4.0 * x + 1.0 );
    case DTKX_RBF_WENDLAND4:
        return ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) *
               ( 1.0 - x ) * ( 1.0 - x ) * ( 35.0 * x * x + 18.0 * x + 3.0 );
    case DTKX_RBF_WENDLAND6:
        return ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) *
               ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) *
               ( 32.0 * x * x * x + 25.0 * x * x + 8.0 * x + 1.0 );
    case DTKX_RBF_WU2:
        return ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) *
               ( 3.0 * x * x * x + 12.0 * x * x + 16.0 * x + 4.0 );
    case DTKX_RBF_WU4:
        return ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) * ( 1.0 - x ) *
               ( 1.0 - x ) * ( 1.0 - x ) *
               ( 5.0 * x * x * x * x * x + 30.0 * x * x * x * x +
                 72.0 * x * x * x + 82.0 * x * x + 36.0 * x + 6.0 );
    case DTKX_RBF_BUHMANN2: // x = 0 gives 0 * log(0) = NaN, as the reference
        return 2.0 * x * x * x * x * crm::log( x ) - 7.0 / 2.0 * x * x * x * x +
               16 / 3.0 * x * x * x - 2 * x * x + 1.0 / 6.0;
    case DTKX_RBF_BUHMANN3:
        return x * x * x * x * x * x * x * x -
               84.0 / 5.0 * x * x * x * x * x * x +
               1024.0 / 5.0 * x * x * x * x * sqrt( x ) -
               378.0 * x * x * x * x + 1024.0 / 5.0 * x * x * x * sqrt( x ) -
               84.0 / 5.0 * x * x + 1.0;
    default: // DTKX_RBF_BUHMANN4 (rejected at create; kept for the table)
        return 99.0 / 35.0 * x * x * x * x * x * x * x * x -
               132.0 * x * x * x * x * x * x +
               9216.0 / 35.0 * x * x * x * x * x * sqrt( x ) -
               11264.0 / 35.0 * x * x * x * x * sqrt( x ) +
               198.0 * x * x * x * x - 396.0 / 5.0 * x * x + 1.0;
    }
}

// The same functions in float arithmetic for the FP32-bulk form.
__device__ inline float rbf_eval_f( int rbf, float x )
{
    float const a = 1.0f - x;
    float const a2 = a * a, a4 = a2 * a2;
    switch ( rbf )
    {
    case DTKX_RBF_WENDLAND0:
        return a2;
    case DTKX_RBF_WENDLAND2:
        return a4 * ( 4.0f * x + 1.0f );
    case DTKX_RBF_WENDLAND4:
        return a4 * a2 * ( ( 35.0f * x + 18.0f ) * x + 3.0f );
    case DTKX_RBF_WENDLAND6:
        return a4 * a4 * ( ( ( 32.0f * x + 25.0f ) * x + 8.0f ) * x + 1.0f );
    case DTKX_RBF_WU2:
        return a4 * ( ( ( 3.0f * x + 12.0f ) * x + 16.0f ) * x + 4.0f );
    case DTKX_RBF_WU4:
        return a4 * a2 *
               ( ( ( ( ( 5.0f * x + 30.0f ) * x + 72.0f ) * x + 82.0f ) * x +
                   36.0f ) *
                     x +
                 6.0f );
    default:
        return (float)rbf_eval( rbf, (double)x );
    }
}

// Basis row P(l, 0..M-1) of a relative coordinate (DTK_MultivariatePolynomial
// Basis.hpp:249-276). Float geometry: the coordinate is rounded to float and
// products are float products, stored as double.
template <int M, typename T>
__device__ inline void basis_row_t( T x, T y, T z, double *P )
{
#pragma clang fp contract( off )
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
__device__ inline void basis_row( int geometry, double const *xr, double *P )
{
    if ( geometry == DTKX_GEOMETRY_FLOAT )
        basis_row_t<M, float>( (float)xr[0], (float)xr[1], (float)xr[2], P );
    else
        basis_row_t<M, double>( xr[0], xr[1], xr[2], P );
}

// |relative coordinate| as computeRadius/computeWeights form it
// (Impl.hpp:121-125, 162-165): distance(Point{x}, Point{0,0,0}).
__device__ inline double rel_distance( int geometry, double const *xr )
{
#pragma clang fp contract( off )
    if ( geometry == DTKX_GEOMETRY_FLOAT )
    {
        float s = 0.f;
        for ( int d = 0; d < 3; ++d )
        {
            float tmp = 0.f - (float)xr[d];
            s += tmp * tmp;
        }
        return (double)__builtin_sqrtf( s ); // correctly rounded (not __fsqrt_rn)
    }
    double s = 0.;
    for ( int d = 0; d < 3; ++d )
    {
        double tmp = 0. - xr[d];
        s += tmp * tmp;
    }
    return __builtin_sqrt( s );
}

// The 2x2 SVD of DTK_DetailsSVDImpl.hpp:119-162 reduced to what the Jacobi
// consumes (:258-263): cl = L00, sl = L01, cr = R00 and the sign-fixed sr.
// Split in two so the two independent angle chains can run on two lanes:
// svd_2x2_angle_args gives the atan2 arguments of phi (AA^T, second = false)
// or theta (A^T A, second = true); svd_2x2_from_angles finishes.
__device__ inline void svd_2x2_angle_args( double a00, double a01, double a10,
                                           double a11, bool second, double &y,
                                           double &x )
{
#pragma clang fp contract( off )
    // mult_2x2 order: two products, one add
    double const aat00 = a00 * a00 + a01 * a01;
    double const aat01 = a00 * a10 + a01 * a11;
    double const aat10 = a10 * a00 + a11 * a01;
    double const aat11 = a10 * a10 + a11 * a11;
    double const ata00 = a00 * a00 + a10 * a10;
    double const ata01 = a00 * a01 + a10 * a11;
    double const ata10 = a01 * a00 + a11 * a10;
    double const ata11 = a01 * a01 + a11 * a11;
    y = second ? ata01 + ata10 : aat01 + aat10;
    x = second ? ata00 - ata11 : aat00 - aat11;
}

__device__ inline void svd_2x2_from_angles( double a00, double a01, double a10,
                                            double a11, double cphi,
                                            double sphi, double ctheta,
                                            double stheta, double &cl,
                                            double &sl, double &cr, double &sr )
{
#pragma clang fp contract( off )
    // W = [[ct, -st], [st, ct]]; AW = A W; S = U^T AW with U^T = [[cp, sp],
    // [-sp, cp]]
    double const w00 = ctheta, w01 = -stheta, w10 = stheta, w11 = ctheta;
    double const aw00 = a00 * w00 + a01 * w10;
    double const aw01 = a00 * w01 + a01 * w11;
    double const aw10 = a10 * w00 + a11 * w10;
    double const aw11 = a10 * w01 + a11 * w11;
    double const ut00 = cphi, ut01 = sphi, ut10 = -sphi, ut11 = cphi;
    double const s00 = ut00 * aw00 + ut01 * aw10;
    double const s11 = ut10 * aw01 + ut11 * aw11;
    double const c00 = copysign( 1., s00 );
    double const c11 = copysign( 1., s11 );
    // V = W C (C diagonal: the off-diagonal zeros kept for signed zeros)
    double const v00 = w00 * c00 + w01 * 0.0;
    double const v01 = w00 * 0.0 + w01 * c11;
    double const v10 = w10 * c00 + w11 * 0.0;
    cl = cphi;
    sl = -sphi;
    cr = v00;
    int const sg01 = ( v01 > 0 ) - ( v01 < 0 );
    int const sg10 = ( v10 > 0 ) - ( v10 < 0 );
    sr = ( sg01 == sg10 ) ? -v01 : v01;
}

__device__ inline void svd_2x2_rot( double a00, double a01, double a10,
                                    double a11, double &cl, double &sl,
                                    double &cr, double &sr )
{
    double y, x, cphi, sphi, ctheta, stheta;
    svd_2x2_angle_args( a00, a01, a10, a11, false, y, x );
    crm::sincos( 0.5 * crm::atan2( y, x ), &sphi, &cphi );
    svd_2x2_angle_args( a00, a01, a10, a11, true, y, x );
    crm::sincos( 0.5 * crm::atan2( y, x ), &stheta, &ctheta );
    svd_2x2_from_angles( a00, a01, a10, a11, cphi, sphi, ctheta, stheta, cl,
                         sl, cr, sr );
}

} // namespace dev
} // namespace dtkx

#endif

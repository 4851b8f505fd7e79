// Correctly rounded double-precision atan2, sin, cos and log for host and
// device.
//
// The classical Jacobi of the reference forms its rotations from
// 0.5 * atan2(...), cos(...) and sin(...) (DTK_DetailsSVDImpl.hpp:129-140).
// When the moment matrices are ill conditioned (k equal to the basis size),
// a one-ulp difference in any of these calls moves interpolated values far
// beyond 1e-12. The CPU reference calls glibc, whose results are correctly
// rounded in all but rare hard cases, so the GPU kernels use these
// correctly rounded versions to round identically.
//
// Method: each function is evaluated in double-double arithmetic (about 104
// significant bits, error-free transformations two_sum/two_prod built on
// fma) and the double-double result is rounded once to double. The result
// is the correctly rounded value unless the exact value lies within about
// 2^-100 (relative) of a rounding boundary.
//   sin/cos: Cody-Waite reduction by a four-part pi/2 (|k| < 2^20 exact),
//            Taylor series of sin and cos on |r| <= pi/4, high-order terms
//            in double, low-order terms in double-double.
//   atan2:   ratio t = min/max of |y|, |x| as a double-double quotient,
//            table reduction atan(t) = atan(j/16) + atan((t - j/16) /
//            (1 + t j/16)), Taylor series of atan on |u| <= 1/32, quadrant
//            fix-up in double-double. C99 special cases (signed zeros,
//            infinities, NaN) are handled explicitly.
//   log:     x = 2^e m, m in [sqrt(1/2), sqrt(2)), the atanh series of
//            s = (m - 1) / (m + 1) in double-double plus e ln 2 (Buhmann<2>
//            RBF, DTK_CompactlySupportedRadialBasisFunctions.hpp:118).
// Constants were produced with mpmath at 400-bit precision.
//
// Domain note: sin/cos reduce exactly for |x| < 2^20 * pi/2; larger
// arguments fall back to the platform sin/cos (never reached by the
// Jacobi, whose angles lie in [-pi/2, pi/2]).
#ifndef DTKX_CRMATH_HPP
#define DTKX_CRMATH_HPP

#include <cmath>
#include <limits>

#if defined( __HIPCC__ )
#define DTKX_HD __host__ __device__ inline
#else
#define DTKX_HD inline
#endif

namespace dtkx
{
namespace crm
{

struct dd
{
    double hi, lo;
};

// Error-free transformations require that the compiler forms no FMAs:
// every function body opens with DTKX_NO_CONTRACT (host builds with g++ use
// -ffp-contract=off instead).
#if defined( __clang__ )
#define DTKX_NO_CONTRACT _Pragma( "clang fp contract(off)" )
#else
#define DTKX_NO_CONTRACT
#endif

DTKX_HD dd fast_two_sum( double a, double b )
{
    DTKX_NO_CONTRACT
    double s = a + b;
    double e = b - ( s - a );
    return {s, e};
}

DTKX_HD dd two_sum( double a, double b )
{
    DTKX_NO_CONTRACT
    double s = a + b;
    double bb = s - a;
    double e = ( a - ( s - bb ) ) + ( b - bb );
    return {s, e};
}

DTKX_HD dd two_prod( double a, double b )
{
    DTKX_NO_CONTRACT
    double p = a * b;
    double e = std::fma( a, b, -p );
    return {p, e};
}

DTKX_HD dd dd_add( dd a, dd b )
{
    DTKX_NO_CONTRACT
    dd s = two_sum( a.hi, b.hi );
    dd t = two_sum( a.lo, b.lo );
    s.lo += t.hi;
    s = fast_two_sum( s.hi, s.lo );
    s.lo += t.lo;
    return fast_two_sum( s.hi, s.lo );
}

DTKX_HD dd dd_add_d( dd a, double b )
{
    DTKX_NO_CONTRACT
    dd s = two_sum( a.hi, b );
    s.lo += a.lo;
    return fast_two_sum( s.hi, s.lo );
}

// a + b for |a| >= |b| (or a = 0): one fast_two_sum of the leading parts
DTKX_HD dd dd_add_ordered( dd a, dd b )
{
    DTKX_NO_CONTRACT
    dd s = fast_two_sum( a.hi, b.hi );
    s.lo += a.lo + b.lo;
    return fast_two_sum( s.hi, s.lo );
}

DTKX_HD dd dd_neg( dd a ) { return {-a.hi, -a.lo}; }

// Ziv's rounding test: v approximates the exact value to a relative error
// below rel; if both ends of the interval round to the same double, that
// double is the correctly rounded result.
DTKX_HD bool dd_round_certain( dd v, double rel, double *out )
{
    DTKX_NO_CONTRACT
    double const e = rel * std::fabs( v.hi );
    double const up = v.hi + ( v.lo + e );
    double const dn = v.hi + ( v.lo - e );
    *out = v.hi + v.lo;
    return up == dn;
}

DTKX_HD dd dd_mul( dd a, dd b )
{
    DTKX_NO_CONTRACT
    dd p = two_prod( a.hi, b.hi );
    p.lo += a.hi * b.lo + a.lo * b.hi;
    return fast_two_sum( p.hi, p.lo );
}

DTKX_HD dd dd_mul_d( dd a, double b )
{
    DTKX_NO_CONTRACT
    dd p = two_prod( a.hi, b );
    p.lo += a.lo * b;
    return fast_two_sum( p.hi, p.lo );
}

DTKX_HD dd dd_div( dd a, dd b )
{
    DTKX_NO_CONTRACT
    double q1 = a.hi / b.hi;
    dd r = dd_add( a, dd_neg( dd_mul_d( b, q1 ) ) );
    double q2 = r.hi / b.hi;
    r = dd_add( r, dd_neg( dd_mul_d( b, q2 ) ) );
    double q3 = r.hi / b.hi;
    dd q = fast_two_sum( q1, q2 );
    return dd_add_d( q, q3 );
}

// a / b with one double division (a reciprocal): each correction term is
// the remainder times 1/b.hi, so the terms' errors shrink geometrically as
// in dd_div and the double-double is as accurate, with a third of the
// (long, dependent) division sequences.
DTKX_HD dd dd_div_rcp( dd a, dd b )
{
    DTKX_NO_CONTRACT
    double const inv = 1.0 / b.hi;
    double q1 = a.hi * inv;
    dd r = dd_add( a, dd_neg( dd_mul_d( b, q1 ) ) );
    double q2 = r.hi * inv;
    r = dd_add( r, dd_neg( dd_mul_d( b, q2 ) ) );
    double q3 = r.hi * inv;
    dd q = fast_two_sum( q1, q2 );
    return dd_add_d( q, q3 );
}

// ------------------------------------------------------------ constants
// pi/2 = PIO2_1 + PIO2_2 + PIO2_3 + PIO2_4 (+ < 2^-176); the first two
// carry 33 significant bits so k * PIO2_{1,2} is exact for |k| < 2^20.
#define DTKX_PIO2_1 0x1.921fb54400000p+0
#define DTKX_PIO2_2 0x1.0b4611a600000p-34
#define DTKX_PIO2_3 0x1.3198a2e037073p-69
#define DTKX_PIO2_4 0x1.129024e088a68p-123
#define DTKX_PI_HI 0x1.921fb54442d18p+1
#define DTKX_PI_LO 0x1.1a62633145c07p-53
#define DTKX_PIO2_HI 0x1.921fb54442d18p+0
#define DTKX_PIO2_LO 0x1.1a62633145c07p-54
#define DTKX_PIO4_HI 0x1.921fb54442d18p-1
#define DTKX_3PIO4_HI 0x1.2d97c7f3321d2p+1
#define DTKX_TWO_OVER_PI 0x1.45f306dc9c883p-1

// 1/n! as double-double, n = 0..29
DTKX_HD dd inv_fact( int n )
{
    DTKX_NO_CONTRACT
    switch ( n )
    {
    case 0:
    case 1: return {1.0, 0.0};
    case 2: return {0x1.0000000000000p-1, 0.0};
    case 3: return {0x1.5555555555555p-3, 0x1.5555555555555p-57};
    case 4: return {0x1.5555555555555p-5, 0x1.5555555555555p-59};
    case 5: return {0x1.1111111111111p-7, 0x1.1111111111111p-63};
    case 6: return {0x1.6c16c16c16c17p-10, -0x1.f49f49f49f49fp-65};
    case 7: return {0x1.a01a01a01a01ap-13, 0x1.a01a01a01a01ap-73};
    case 8: return {0x1.a01a01a01a01ap-16, 0x1.a01a01a01a01ap-76};
    case 9: return {0x1.71de3a556c734p-19, -0x1.c154f8ddc6c00p-73};
    case 10: return {0x1.27e4fb7789f5cp-22, 0x1.cbbc05b4fa99ap-76};
    case 11: return {0x1.ae64567f544e4p-26, -0x1.c062e06d1f209p-80};
    case 12: return {0x1.1eed8eff8d898p-29, -0x1.2aec959e14c06p-83};
    case 13: return {0x1.6124613a86d09p-33, 0x1.f28e0cc748ebep-87};
    case 14: return {0x1.93974a8c07c9dp-37, 0x1.05d6f8a2efd1fp-92};
    case 15: return {0x1.ae7f3e733b81fp-41, 0x1.1d8656b0ee8cbp-97};
    case 16: return {0x1.ae7f3e733b81fp-45, 0x1.1d8656b0ee8cbp-101};
    case 17: return {0x1.952c77030ad4ap-49, 0x1.ac981465ddc6cp-103};
    case 18: return {0x1.6827863b97d97p-53, 0x1.eec01221a8b0bp-107};
    case 19: return {0x1.2f49b46814157p-57, 0x1.2650f61dbdcb4p-112};
    case 20: return {0x1.e542ba4020225p-62, 0x1.ea72b4afe3c2fp-120};
    case 21: return {0x1.71b8ef6dcf572p-66, -0x1.d043ae40c4647p-120};
    case 22: return {0x1.0ce396db7f853p-70, -0x1.aebcdbd20331cp-124};
    case 23: return {0x1.761b41316381ap-75, -0x1.3423c7d91404fp-130};
    case 24: return {0x1.f2cf01972f578p-80, -0x1.9ada5fcc1ab14p-135};
    case 25: return {0x1.3f3ccdd165fa9p-84, -0x1.58ddadf344487p-139};
    case 26: return {0x1.88e85fc6a4e5ap-89, -0x1.71c37ebd16540p-143};
    case 27: return {0x1.d1ab1c2dccea3p-94, 0x1.054d0c78aea14p-149};
    case 28: return {0x1.0a18a2635085dp-98, 0x1.b9e2e28e1aa54p-153};
    default: return {0x1.259f98b4358adp-103, 0x1.eaf8c39dd9bc5p-157};
    }
}

// atan(j/16) as double-double, j = 0..16 (a table in constant memory: an
// index that varies across lanes loads one entry instead of selecting among
// all of them in registers)
#if defined( __HIPCC__ )
__device__ __constant__
#endif
static const double kAtanJ16[17][2] = {
    {0.0, 0.0},
    {0x1.ff55bb72cfdeap-5, -0x1.c934d86d23f1dp-60},
    {0x1.fd5ba9aac2f6ep-4, -0x1.cd37686760c17p-59},
    {0x1.7b97b4bce5b02p-3, 0x1.347b0b4f881cap-58},
    {0x1.f5b75f92c80ddp-3, 0x1.8ab6e3cf7afbdp-57},
    {0x1.362773707ebccp-2, -0x1.963a544b672d8p-57},
    {0x1.6f61941e4def1p-2, -0x1.c63aae6f6e918p-56},
    {0x1.a64eec3cc23fdp-2, -0x1.24dec1b50b7ffp-56},
    {0x1.dac670561bb4fp-2, 0x1.a2b7f222f65e2p-56},
    {0x1.0657e94db30d0p-1, -0x1.d5b495f6349e6p-56},
    {0x1.1e00babdefeb4p-1, -0x1.928df287a668fp-58},
    {0x1.345f01cce37bbp-1, 0x1.1021137c71102p-55},
    {0x1.4978fa3269ee1p-1, 0x1.2419a87f2a458p-56},
    {0x1.5d58987169b18p-1, 0x1.0028e4bc5e7cap-57},
    {0x1.700a7c5784634p-1, -0x1.8c34d25aadef6p-56},
    {0x1.819d0b7158a4dp-1, -0x1.bf76229d3b917p-56},
    {0x1.921fb54442d18p-1, 0x1.1a62633145c07p-55}
};

DTKX_HD dd atan_j16( int j ) { return {kAtanJ16[j][0], kAtanJ16[j][1]}; }

// 1/(2n+1) as double-double, n = 1..4 (higher terms use double)
DTKX_HD dd inv_odd( int n )
{
    DTKX_NO_CONTRACT
    switch ( n )
    {
    case 1: return {0x1.5555555555555p-2, 0x1.5555555555555p-56};
    case 2: return {0x1.999999999999ap-3, -0x1.999999999999ap-57};
    case 3: return {0x1.2492492492492p-3, 0x1.2492492492492p-57};
    default: return {0x1.c71c71c71c71cp-4, 0x1.c71c71c71c71cp-58};
    }
}

// -------------------------------------------------------------- sin/cos
// sin(r) and cos(r) for a double-double |r| <= pi/4 (slightly more).
// w = -r^2. Terms with n >= 8 are below 2^-54 relative and use double.
DTKX_HD dd sin_poly( dd r )
{
    DTKX_NO_CONTRACT
    dd w = dd_neg( dd_mul( r, r ) );
    double wd = w.hi;
    // sum_{n>=8} w^n / (2n+1)! / w^8, n = 8..14, in double
    double t = inv_fact( 29 ).hi;
    for ( int n = 13; n >= 8; --n )
        t = t * wd + inv_fact( 2 * n + 1 ).hi;
    dd acc = {t, 0.0};
    for ( int n = 7; n >= 0; --n )
        acc = dd_add( dd_mul( acc, w ), inv_fact( 2 * n + 1 ) );
    return dd_mul( acc, r );
}

DTKX_HD dd cos_poly( dd r )
{
    DTKX_NO_CONTRACT
    dd w = dd_neg( dd_mul( r, r ) );
    double wd = w.hi;
    double t = inv_fact( 28 ).hi;
    for ( int n = 13; n >= 8; --n )
        t = t * wd + inv_fact( 2 * n ).hi;
    dd acc = {t, 0.0};
    for ( int n = 7; n >= 0; --n )
        acc = dd_add( dd_mul( acc, w ), inv_fact( 2 * n ) );
    return acc;
}

// x = k pi/2 + r, r as double-double; requires |k| < 2^20.
DTKX_HD dd reduce_pio2( double x, int &q )
{
    DTKX_NO_CONTRACT
    double kd = std::nearbyint( x * DTKX_TWO_OVER_PI );
    q = static_cast<int>( kd );
    double t1 = x - kd * DTKX_PIO2_1; // exact (Sterbenz, exact product)
    dd r = two_sum( t1, -( kd * DTKX_PIO2_2 ) );
    r = dd_add( r, dd_neg( two_prod( kd, DTKX_PIO2_3 ) ) );
    r = dd_add( r, dd_neg( two_prod( kd, DTKX_PIO2_4 ) ) );
    return r;
}

// sin (shift = 0) or cos (shift = 1) of x: cos(x) = sin(x + pi/2) is a
// quadrant shift of the same reduced argument, so both share one path.
DTKX_HD double sin_cos_q( double x, int shift )
{
    DTKX_NO_CONTRACT
    if ( !( std::fabs( x ) < 1.6e6 ) ) // inf, nan, or outside the domain
        return shift ? std::cos( x ) : std::sin( x );
    if ( shift == 0 && std::fabs( x ) < 0x1p-27 )
        return x; // sin(x) = x (1 - x^2/6) rounds to x; keeps -0
    int q;
    dd r = reduce_pio2( x, q );
    q = ( q + shift ) & 3;
    dd v = ( q & 1 ) ? cos_poly( r ) : sin_poly( r );
    if ( q & 2 )
        v = dd_neg( v );
    return v.hi + v.lo;
}


// sin (shift = 0) or cos (shift = 1) of x from one polynomial, whatever the
// quadrant: with w = -r^2, sin(r) = r sum_n w^n / (2n+1)! and cos(r) =
// sum_n w^n / (2n)! differ only in their coefficients and the final factor
// r, so lanes of one wave that need different functions (or quadrants) run
// the same instructions. Terms n <= 4 in double-double (quick ordered
// additions: each term is below the next lower one), n >= 5 in double:
// the double tail's rounding is below 2^-77 of the result on |r| <= pi/4
// (r^10 / 10! < 2.5e-8), so with the rounding test at 2^-70 the result is
// the correctly rounded one, or the full evaluation (sin_cos_q) decides.
DTKX_HD double sin_or_cos( double x, int shift )
{
    DTKX_NO_CONTRACT
    if ( !( std::fabs( x ) < 1.6e6 ) || ( shift == 0 && std::fabs( x ) < 0x1p-27 ) )
        return sin_cos_q( x, shift );
    int q;
    dd const r = reduce_pio2( x, q );
    q = ( q + shift ) & 3;
    bool const c = ( q & 1 ) != 0; // cos polynomial in this quadrant
    dd const w = dd_neg( dd_mul( r, r ) );
    double const wd = w.hi;
    // coefficients picked by a select between two compile-time constants
    // (a table index that varies across lanes would branch)
    auto coef = [c]( int n ) -> dd {
        dd const kc = inv_fact( 2 * n ), ks = inv_fact( 2 * n + 1 );
        return {c ? kc.hi : ks.hi, c ? kc.lo : ks.lo};
    };
    // the double tail sum_{n=5..13} c_n w^(n-5) in Estrin's order (depth 4
    // instead of 9)
    double const w2 = wd * wd, w4 = w2 * w2;
    double const p56 = coef( 5 ).hi + coef( 6 ).hi * wd;
    double const p78 = coef( 7 ).hi + coef( 8 ).hi * wd;
    double const p910 = coef( 9 ).hi + coef( 10 ).hi * wd;
    double const p1112 = coef( 11 ).hi + coef( 12 ).hi * wd;
    double const t = ( p56 + p78 * w2 ) + ( p910 + p1112 * w2 ) * w4 +
                     coef( 13 ).hi * ( w4 * w4 );
    dd acc = {t, 0.0};
#pragma unroll
    for ( int n = 4; n >= 0; --n )
        acc = dd_add_ordered( coef( n ), dd_mul( acc, w ) );
    dd v = c ? acc : dd_mul( acc, r );
    if ( q & 2 )
        v = dd_neg( v );
    double res;
    if ( dd_round_certain( v, 0x1p-70, &res ) )
        return res;
    return sin_cos_q( x, shift );
}

DTKX_HD double sin( double x ) { return sin_or_cos( x, 0 ); }
DTKX_HD double cos( double x ) { return sin_or_cos( x, 1 ); }

// sin and cos of one argument, each the correctly rounded value sin() and
// cos() return
DTKX_HD void sincos( double x, double *s, double *c )
{
    *s = sin_or_cos( x, 0 );
    *c = sin_or_cos( x, 1 );
}

// ---------------------------------------------------------------- atan2
// atan(t) for a double-double 0 <= t <= 1.
DTKX_HD dd atan_01( dd t )
{
    DTKX_NO_CONTRACT
    int j = static_cast<int>( t.hi * 16.0 + 0.5 );
    j = j < 0 ? 0 : ( j > 16 ? 16 : j );
    double c = j * 0.0625;
    dd u;
    if ( j == 0 )
        u = t;
    else
    {
        dd num = dd_add_d( t, -c );
        dd den = dd_add_d( dd_mul_d( t, c ), 1.0 );
        u = dd_div( num, den );
    }
    // atan(u) = u * sum_n w^n / (2n+1), w = -u^2, |u| <= 1/32
    dd w = dd_neg( dd_mul( u, u ) );
    double wd = w.hi;
    double s = 1.0 / 23.0;
    for ( int n = 10; n >= 5; --n )
        s = s * wd + 1.0 / ( 2 * n + 1 );
    dd acc = {s, 0.0};
    for ( int n = 4; n >= 1; --n )
        acc = dd_add( dd_mul( acc, w ), inv_odd( n ) );
    acc = dd_add_d( dd_mul( acc, w ), 1.0 );
    dd at = dd_mul( acc, u );
    return dd_add( atan_j16( j ), at );
}

// atan(t) for a double-double 0 <= t <= 1 with fewer double-double terms:
// atan(u) = u - u^3/3 + u^5 S(u^2), the last in double. |u| <= 1/32, so
// u^5 S is below 2^-25 u and its rounding below 2^-76 of the result; the
// caller's rounding test (2^-70) sends the rare undecided case to atan_01.
DTKX_HD dd atan_01_quick( dd t )
{
    DTKX_NO_CONTRACT
    int j = static_cast<int>( t.hi * 16.0 + 0.5 );
    j = j < 0 ? 0 : ( j > 16 ? 16 : j );
    double c = j * 0.0625;
    dd u;
    if ( j == 0 )
        u = t;
    else
    {
        dd num = dd_add_d( t, -c );
        dd den = dd_add_d( dd_mul_d( t, c ), 1.0 );
        u = dd_div_rcp( num, den );
    }
    double const uh = u.hi, w = uh * uh, w2 = w * w, w4 = w2 * w2;
    // s = sum_{n=2..11} (-1)^n w^(n-2) / (2n+1) = 1/5 - w/7 + w^2/9 - ...
    // in Estrin's order (depth 4 instead of 10: the chain is the cost)
    auto k = []( int n ) { return ( ( n & 1 ) ? -1.0 : 1.0 ) / ( 2 * n + 1 ); };
    double const p23 = k( 2 ) + k( 3 ) * w, p45 = k( 4 ) + k( 5 ) * w;
    double const p67 = k( 6 ) + k( 7 ) * w, p89 = k( 8 ) + k( 9 ) * w;
    double const p1011 = k( 10 ) + k( 11 ) * w;
    double const s = ( p23 + p45 * w2 ) + ( p67 + p89 * w2 ) * w4 + p1011 * ( w4 * w4 );
    double const tail = uh * w2 * s;
    dd const u3 = dd_mul( dd_mul( u, u ), u );
    dd const c3 = dd_mul( u3, inv_odd( 1 ) );
    dd a = dd_add_ordered( dd_neg( c3 ), dd{tail, 0.0} );
    a = dd_add_ordered( u, a );
    return dd_add( atan_j16( j ), a );
}

DTKX_HD double atan2_core( double y, double x, bool quick, bool *certain )
{
    DTKX_NO_CONTRACT
    *certain = true;
    if ( std::isnan( x ) || std::isnan( y ) )
        return x + y;
    bool const xneg = std::signbit( x );
    if ( y == 0.0 )
        return xneg ? std::copysign( DTKX_PI_HI, y ) : y;
    if ( x == 0.0 )
        return std::copysign( DTKX_PIO2_HI, y );
    if ( std::isinf( x ) || std::isinf( y ) )
    {
        double v;
        if ( std::isinf( x ) && std::isinf( y ) )
            v = xneg ? DTKX_3PIO4_HI : DTKX_PIO4_HI;
        else if ( std::isinf( y ) )
            v = DTKX_PIO2_HI;
        else
            v = xneg ? DTKX_PI_HI : 0.0;
        return std::copysign( v, y );
    }
    double ax = std::fabs( x ), ay = std::fabs( y );
    bool const swap = ay > ax;
    double num = swap ? ax : ay, den = swap ? ay : ax;
    // scale both by the same power of two (exact) so the double-double
    // division neither overflows nor loses bits to underflow
    int e;
    std::frexp( den, &e );
    num = std::ldexp( num, -e );
    den = std::ldexp( den, -e );
    dd v;
    if ( num < 0x1p-900 )
    {
        // the ratio is below 2^-900: atan(t) = t to far below an ulp, and
        // the exact quotient of the unscaled values rounds correctly
        double q = ( swap ? ax : ay ) / ( swap ? ay : ax );
        v = {q, 0.0};
    }
    else if ( quick )
    {
        // one reciprocal: q1 within two ulps, the fma remainder exact
        double const inv = 1.0 / den;
        double q1 = num * inv;
        double r = std::fma( -q1, den, num );
        double q2 = r * inv;
        v = atan_01_quick( fast_two_sum( q1, q2 ) );
    }
    else
    {
        double q1 = num / den;
        double r = std::fma( -q1, den, num ); // exact remainder
        double q2 = r / den;
        dd t = fast_two_sum( q1, q2 );
        v = atan_01( t );
    }
    if ( swap )
        v = dd_add( dd{DTKX_PIO2_HI, DTKX_PIO2_LO}, dd_neg( v ) );
    if ( xneg )
        v = dd_add( dd{DTKX_PI_HI, DTKX_PI_LO}, dd_neg( v ) );
    double res = v.hi + v.lo;
    if ( quick )
        *certain = dd_round_certain( v, 0x1p-70, &res );
    return std::copysign( res, y );
}

// Correctly rounded: the quick evaluation when its rounding test decides,
// else the full one (both give the same double whenever the quick one
// accepts).
DTKX_HD double atan2( double y, double x )
{
    bool certain;
    double const v = atan2_core( y, x, true, &certain );
    if ( certain )
        return v;
    return atan2_core( y, x, false, &certain );
}

// ------------------------------------------------------------------ log
// ln 2 = LN2_HI + LN2_MID + LN2_LO (+ < 2^-165)
#define DTKX_LN2_HI 0x1.62e42fefa39efp-1
#define DTKX_LN2_MID 0x1.abc9e3b39803fp-56
#define DTKX_LN2_LO 0x1.7b57a079a1934p-111

// 1/d as double-double for a small odd integer d (q1 + exact remainder / d)
DTKX_HD dd dd_recip_int( double d )
{
    DTKX_NO_CONTRACT
    double q1 = 1.0 / d;
    double r = std::fma( -q1, d, 1.0 ); // exact
    return fast_two_sum( q1, r / d );
}

// Natural logarithm. x = 2^e m with m in [sqrt(1/2), sqrt(2)); then
// log(m) = 2 atanh(s) = 2 s sum_n s^2n / (2n+1), s = (m - 1) / (m + 1),
// |s| < 0.1716 (s^2 < 0.0295). Terms n >= 11 are below 2^-54 relative to
// the sum and use double; n <= 10 use double-double. e ln 2 is a
// three-part product. The Buhmann<2> RBF is the only caller.
DTKX_HD double log( double x )
{
    DTKX_NO_CONTRACT
    if ( std::isnan( x ) )
        return x + x;
    if ( x < 0.0 )
        return std::numeric_limits<double>::quiet_NaN();
    if ( x == 0.0 )
        return -std::numeric_limits<double>::infinity();
    if ( std::isinf( x ) )
        return x;
    if ( x == 1.0 )
        return 0.0;
    int e;
    double m = std::frexp( x, &e ); // m in [0.5, 1), exact for subnormals
    if ( m < 0.70710678118654752 )
    {
        m *= 2.0;
        --e;
    }
    // m - 1 is exact (Sterbenz); m + 1 is carried as a double-double
    dd const s = dd_div( dd{m - 1.0, 0.0}, two_sum( m, 1.0 ) );
    dd const w = dd_mul( s, s );
    double const wd = w.hi;
    double t = 1.0 / 45.0;
    for ( int n = 21; n >= 11; --n )
        t = t * wd + 1.0 / ( 2 * n + 1 );
    dd acc = {t, 0.0};
    for ( int n = 10; n >= 1; --n )
        acc = dd_add( dd_mul( acc, w ), dd_recip_int( 2 * n + 1 ) );
    acc = dd_add_d( dd_mul( acc, w ), 1.0 );
    dd v = dd_mul( acc, dd{2.0 * s.hi, 2.0 * s.lo} );
    if ( e != 0 )
    {
        double const ed = static_cast<double>( e );
        dd l2 = two_prod( ed, DTKX_LN2_HI );
        l2.lo += ed * DTKX_LN2_MID + ed * DTKX_LN2_LO;
        v = dd_add( fast_two_sum( l2.hi, l2.lo ), v );
    }
    return v.hi + v.lo;
}

} // namespace crm
} // namespace dtkx

#endif

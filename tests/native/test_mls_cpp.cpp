// The reference's MLS operator tests, restated against the C++ drop-in
// (include/dtkx/MovingLeastSquaresOperator.hpp) with the same inputs and
// tolerances:
//   packages/Meshfree/test/tstMeshfreeOperators.cpp:130-448
//       same_npoints_and_basis (1e-7), grid (1e-14), line (1e-14),
//       single_point_in_radius (finite), for Wendland<0> x {Constant3,
//       Linear3, Quadratic3} (:464-503)
//   packages/Meshfree/test/tstMeshfreeOperatorsSimpleProblem.cpp:113-147
//       corner_cases (1e-12)
//   the same four tests for SplineOperator<Wendland<0>, Linear3>
//       (tstMeshfreeOperators.cpp:504-516) with the Spline tolerances:
//       same_npoints_and_basis 2e-7, grid 1e-9, line 2e-9, finite
// The random point sets are the fixtures regenerated bit-for-bit from the
// reference's std::default_random_engine (tests/golden/). Host pointers are
// used, as the reference's Serial node would.
// Build: g++ -std=c++17 -I include test_mls_cpp.cpp -L<dir> -ldtkx
#include "dtkx/MovingLeastSquaresOperator.hpp"
#include "dtkx/SplineOperator.hpp"

#include <array>
#include <cmath>
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

using Point = std::array<double, 3>;

static int g_failures = 0;

static void check_close( std::vector<double> const &got,
                         std::vector<double> const &ref, double tol,
                         char const *what )
{
    bool ok = got.size() == ref.size();
    for ( std::size_t i = 0; ok && i < got.size(); ++i )
    {
        double const err = std::abs( got[i] - ref[i] );
        // TEST_COMPARE_FLOATING_ARRAYS: relative to the larger magnitude
        double const scale = std::max( std::abs( got[i] ), std::abs( ref[i] ) );
        if ( err > tol * scale )
            ok = false;
    }
    std::printf( "%s %s\n", ok ? "PASS" : "FAIL", what );
    if ( !ok )
        ++g_failures;
}

static std::vector<Point> read_points( std::string const &path )
{
    std::vector<Point> pts;
    FILE *f = std::fopen( path.c_str(), "rb" );
    if ( !f )
    {
        std::printf( "FAIL cannot open %s\n", path.c_str() );
        ++g_failures;
        return pts;
    }
    double v[3];
    while ( std::fread( v, sizeof( double ), 3, f ) == 3 )
        pts.push_back( {v[0], v[1], v[2]} );
    std::fclose( f );
    return pts;
}

static std::vector<Point> grid( std::array<int, 3> n, Point off )
{
    std::vector<Point> g;
    for ( int i = 0; i < n[0]; ++i )
        for ( int j = 0; j < n[1]; ++j )
            for ( int k = 0; k < n[2]; ++k )
                g.push_back( {off[0] + i, off[1] + j, off[2] + k} );
    return g;
}

static std::function<double( Point const & )> poly( int size )
{
    switch ( size )
    {
    case 1:
        return []( Point const & ) { return 3.0; };
    case 4:
        return []( Point const &p ) {
            return 4 + 2 * p[0] + 3 * p[1] - 2 * p[2];
        };
    default:
        return []( Point const &p ) {
            return 2 + 3 * p[0] - 5 * p[1] + 2 * p[2] + 3 * p[0] * p[0] +
                   4 * p[0] * p[1] - 2 * p[0] * p[2] + p[1] * p[1] -
                   3 * p[1] * p[2] + 4 * p[2] * p[2];
        };
    }
}

template <typename Operator>
static std::vector<double> transfer( std::vector<Point> const &src,
                                     std::vector<Point> const &tgt,
                                     std::vector<double> const &vals )
{
    Operator op( src.empty() ? nullptr : src[0].data(), (int64_t)src.size(),
                 tgt.empty() ? nullptr : tgt[0].data(), (int64_t)tgt.size() );
    std::vector<double> out( tgt.size(), 0.0 );
    op.apply( vals.data(), out.data() );
    return out;
}

// tolerances of tstMeshfreeOperators.cpp:214-217, 290-293, 366-369
struct Tolerances
{
    double same_npoints, grid, line;
};

template <typename Op>
static void run_operator( std::string const &golden, Tolerances tol,
                          std::string const &name )
{
    int const m = Op::polynomial_basis::size;
    auto const f = poly( m );
    auto eval = [&]( std::vector<Point> const &p ) {
        std::vector<double> v;
        for ( auto const &x : p )
            v.push_back( f( x ) );
        return v;
    };
    std::string const tag = " " + name + " m=" + std::to_string( m );
    {
        auto src = read_points( golden + "/mls_points_m" + std::to_string( m ) + "_src.f64" );
        auto tgt = read_points( golden + "/mls_points_m" + std::to_string( m ) + "_tgt.f64" );
        check_close( transfer<Op>( src, tgt, eval( src ) ), eval( tgt ),
                     tol.same_npoints,
                     ( "same_npoints_and_basis" + tag ).c_str() );
    }
    {
        auto src = grid( {40, 40, 1}, {0., 0., 0.} );
        auto tgt = grid( {1, 1, 1}, {19., 19., 0.} );
        check_close( transfer<Op>( src, tgt, eval( src ) ), eval( tgt ),
                     tol.grid, ( "grid" + tag ).c_str() );
    }
    {
        auto src = grid( {40, 1, 1}, {0., 0., 0.} );
        auto tgt = grid( {39, 1, 1}, {0.5, 0., 0.} );
        check_close( transfer<Op>( src, tgt, eval( src ) ), eval( tgt ),
                     tol.line, ( "line" + tag ).c_str() );
    }
    {
        auto src = read_points( golden + "/mls_points_m1_src.f64" );
        auto tgt = read_points( golden + "/mls_points_m1_tgt.f64" );
        auto out = transfer<Op>( src, tgt, eval( src ) );
        bool ok = out.size() == tgt.size();
        for ( double v : out )
            ok = ok && std::isfinite( v );
        std::printf( "%s single_point_in_radius%s\n", ok ? "PASS" : "FAIL",
                     tag.c_str() );
        if ( !ok )
            ++g_failures;
    }
}

int main( int argc, char **argv )
{
    std::string const golden = argc > 1 ? argv[1] : "tests/golden";
    Tolerances const mls{1e-7, 1e-14, 1e-14}, spline{2e-7, 1e-9, 2e-9};
    using W0 = dtkx::Wendland<0>;
    using dtkx::MultivariatePolynomialBasis;
    run_operator<dtkx::MovingLeastSquaresOperator<
        W0, MultivariatePolynomialBasis<dtkx::Constant, 3>>>( golden, mls, "MLS" );
    run_operator<dtkx::MovingLeastSquaresOperator<
        W0, MultivariatePolynomialBasis<dtkx::Linear, 3>>>( golden, mls, "MLS" );
    run_operator<dtkx::MovingLeastSquaresOperator<
        W0, MultivariatePolynomialBasis<dtkx::Quadratic, 3>>>( golden, mls, "MLS" );
    run_operator<dtkx::SplineOperator<>>( golden, spline, "Spline" );

    using Op = dtkx::MovingLeastSquaresOperator<>; // Wendland<0>, Linear3
    check_close( transfer<Op>( {{1., 1., 1.}, {2., 2., 2.}}, {{1.5, 1.5, 1.5}},
                               {255., 255.} ),
                 {255.}, 1e-12, "corner_cases single point" );
    check_close( transfer<Op>( {{1., 1., 1.}}, {}, {255.} ), {}, 1e-12,
                 "corner_cases no target" );

    // preconditions raise DataTransferKitException (a std::logic_error)
    bool threw = false;
    try
    {
        Op op( nullptr, 0, nullptr, 0 );
    }
    catch ( std::logic_error const &e )
    {
        threw = std::string( e.what() ).find( "failed in" ) != std::string::npos;
    }
    std::printf( "%s empty source throws\n", threw ? "PASS" : "FAIL" );
    if ( !threw )
        ++g_failures;
    threw = false;
    try
    {
        dtkx::SplineOperator<> sp( nullptr, 0, nullptr, 0 );
    }
    catch ( std::logic_error const &e )
    {
        threw = std::string( e.what() ).find( "failed in" ) != std::string::npos;
    }
    std::printf( "%s spline empty source throws\n", threw ? "PASS" : "FAIL" );
    if ( !threw )
        ++g_failures;

    std::printf( "%d failure(s)\n", g_failures );
    return g_failures == 0 ? 0 : 1;
}

// The reference's NearestNeighborOperator tests, restated against the C++
// drop-in (include/dtkx/NearestNeighborOperator.hpp) for one rank:
//   packages/Meshfree/test/tstNearestNeighborOperator.cpp
//       unique_source_point (:85-138): one source (value 255) -> 255
//       structured_clouds (:140-200): target cloud == source cloud, values
//           = x coordinate, 1e-14
//       mixed_clouds (:202-266): structured 29x31x37 source grid over
//           17x19x23, 41 random targets; value = x of the nearest grid
//           point, round(x / Lx * nx) * Lx / nx (the last column clamped)
// The random targets are drawn with the reference's generator
// (std::default_random_engine seeded with the rank, 0, and
// uniform_real_distribution over [0, L) per axis).
// Build: g++ -std=c++17 -I include test_nn_cpp.cpp -L<dir> -ldtkx
#include "dtkx/NearestNeighborOperator.hpp"

#include <array>
#include <cmath>
#include <cstdio>
#include <functional>
#include <random>
#include <string>
#include <vector>

using Point = std::array<double, 3>;

static int g_failures = 0;

static void report( bool ok, char const *what )
{
    std::printf( "%s %s\n", ok ? "PASS" : "FAIL", what );
    if ( !ok )
        ++g_failures;
}

// tstNearestNeighborOperator.cpp:23-47 (the point order of the reference)
static std::vector<Point> structured( double lx, double ly, double lz, int nx,
                                      int ny, int nz )
{
    std::vector<Point> cloud( nx * ny * nz );
    for ( int i = 0; i < nx; ++i )
        for ( int j = 0; j < ny; ++j )
            for ( int k = 0; k < nz; ++k )
                cloud[i + j * nx + k * nx * ny] = {
                    {i * lx / nx, j * ly / ny, k * lz / nz}};
    return cloud;
}

// tstNearestNeighborOperator.cpp:49-70
static std::vector<Point> random_cloud( double lx, double ly, double lz, int n,
                                        double seed )
{
    std::vector<Point> cloud( n );
    std::default_random_engine generator( seed );
    std::uniform_real_distribution<double> dx( 0.0, lx ), dy( 0.0, ly ),
        dz( 0.0, lz );
    for ( int i = 0; i < n; ++i )
    {
        double const x = dx( generator );
        double const y = dy( generator );
        double const z = dz( generator );
        cloud[i] = {{x, y, z}};
    }
    return cloud;
}

static std::vector<double> transfer( std::vector<Point> const &src,
                                     std::vector<Point> const &tgt,
                                     std::vector<double> const &values )
{
    dtkx::NearestNeighborOperator op( src.empty() ? nullptr : src[0].data(),
                                      (std::int64_t)src.size(),
                                      tgt.empty() ? nullptr : tgt[0].data(),
                                      (std::int64_t)tgt.size() );
    std::vector<double> out( tgt.size(), -1.0 );
    op.apply( values.data(), out.empty() ? nullptr : out.data() );
    return out;
}

int main()
{
    // unique_source_point: one rank, source at (1, 1, 1), target at origin
    {
        auto out = transfer( {{{1., 1., 1.}}}, {{{0., 0., 0.}}}, {255.} );
        report( out.size() == 1 && out[0] == 255., "unique_source_point" );
    }
    // structured_clouds: the target cloud is the source cloud
    {
        auto cloud = structured( 2., 3., 5., 7, 11, 13 );
        std::vector<double> x( cloud.size() );
        for ( std::size_t i = 0; i < cloud.size(); ++i )
            x[i] = cloud[i][0];
        auto out = transfer( cloud, cloud, x );
        bool ok = true;
        for ( std::size_t i = 0; i < cloud.size(); ++i )
            ok = ok && std::fabs( out[i] - cloud[i][0] ) <=
                           1e-14 * std::fabs( cloud[i][0] );
        report( ok, "structured_clouds" );
    }
    // mixed_clouds
    {
        double const lx = 17., ly = 19., lz = 23.;
        int const nx = 29, ny = 31, nz = 37;
        auto src = structured( lx, ly, lz, nx, ny, nz );
        auto tgt = random_cloud( lx, ly, lz, 41, 0. );
        std::vector<double> x( src.size() );
        for ( std::size_t i = 0; i < src.size(); ++i )
            x[i] = src[i][0];
        auto out = transfer( src, tgt, x );
        bool ok = true;
        for ( std::size_t i = 0; i < tgt.size(); ++i )
        {
            double ref = std::round( tgt[i][0] / lx * nx ) * lx / nx;
            if ( ref == lx )
                ref -= lx / nx;
            ok = ok && std::fabs( out[i] - ref ) <= 1e-14 * std::fabs( ref );
        }
        report( ok, "mixed_clouds" );
    }
    // several fields at once equal the fields one by one
    {
        auto src = structured( 1., 1., 1., 9, 9, 9 );
        auto tgt = random_cloud( 1., 1., 1., 500, 3. );
        std::vector<double> x( 3 * src.size() );
        for ( std::size_t i = 0; i < x.size(); ++i )
            x[i] = std::sin( 0.37 * i );
        dtkx::NearestNeighborOperator op( src[0].data(), src.size(),
                                          tgt[0].data(), tgt.size() );
        std::vector<double> all( 3 * tgt.size() ), one( tgt.size() );
        op.apply( x.data(), all.data(), 3 );
        bool ok = true;
        for ( int f = 0; f < 3; ++f )
        {
            op.apply( x.data() + f * src.size(), one.data() );
            for ( std::size_t i = 0; i < tgt.size(); ++i )
                ok = ok && one[i] == all[f * tgt.size() + i];
        }
        report( ok, "fields together equal fields alone" );
    }
    // an empty source set raises (DTK_CHECK(!search_tree.empty()))
    {
        bool threw = false;
        try
        {
            dtkx::NearestNeighborOperator op( nullptr, 0, nullptr, 0 );
        }
        catch ( std::logic_error const &e )
        {
            threw = std::string( e.what() ).find( "failed in" ) !=
                    std::string::npos;
        }
        report( threw, "empty source throws" );
    }
    std::printf( "%d failure(s)\n", g_failures );
    return g_failures == 0 ? 0 : 1;
}

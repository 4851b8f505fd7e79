// Regenerates the random inputs of DataTransferKit's own Meshfree tests with
// the same engine and distribution (libstdc++ std::default_random_engine =
// minstd_rand0, default seed; std::uniform_real_distribution<double>), so the
// fixtures in this directory hold the very numbers those tests use:
//
//   svd_full_rank:      packages/Meshfree/test/tstSVD.cpp:79-84
//                       (10 matrices 32x32, uniform(-1000, 1000))
//   svd_rank_deficient: tstSVD.cpp:114-134 (10 matrices 40x40, column 3 = 0,
//                       the engine is not advanced for column 3)
//   mls_points_m<M>:    tstMeshfreeOperators.cpp:41-77 makeSourceTargetPoints
//                       with 10 targets, M sources per target, radius 0.5,
//                       rank 0 (same_npoints_and_basis, :160-176), and
//                       M = 1 for single_point_in_radius (:390-405)
//
// Output: raw little-endian float64 files named <case>.f64 in the directory
// given as argv[1]. Build: g++ -O0 -std=c++17 gen_reference_inputs.cpp
#include <array>
#include <cmath>
#include <cstdio>
#include <random>
#include <string>
#include <vector>

static void write( std::string const &dir, std::string const &name,
                   std::vector<double> const &v )
{
    std::string path = dir + "/" + name + ".f64";
    FILE *f = std::fopen( path.c_str(), "wb" );
    if ( !f )
    {
        std::perror( path.c_str() );
        return;
    }
    std::fwrite( v.data(), sizeof( double ), v.size(), f );
    std::fclose( f );
}

int main( int argc, char **argv )
{
    std::string dir = argc > 1 ? argv[1] : ".";
    {
        int const n = 10, s = 32;
        std::vector<double> m( n * s * s );
        std::default_random_engine g;
        std::uniform_real_distribution<double> dist( -1000, 1000 );
        for ( auto &x : m )
            x = dist( g );
        write( dir, "svd_full_rank", m );
    }
    {
        int const n = 10, s = 40;
        std::vector<double> m( n * s * s );
        std::default_random_engine g;
        std::uniform_real_distribution<double> dist( -1000, 1000 );
        std::size_t pos = 0;
        for ( int i = 0; i < n; ++i )
            for ( int j = 0; j < s; ++j )
                for ( int k = 0; k < s; ++k )
                    m[pos++] = ( k != 3 ) ? dist( g ) : 0.;
        write( dir, "svd_rank_deficient", m );
    }
    for ( int M : {1, 4, 10} )
    {
        int const n_target = 10;
        double const radius = 0.5;
        int const rank = 0;
        std::vector<double> src( 3 * n_target * M ), tgt( 3 * n_target );
        std::default_random_engine g;
        std::uniform_real_distribution<double> rx( -radius, radius );
        double const offset = n_target * radius * 10. * rank;
        for ( int i = 0; i < n_target; i++ )
        {
            std::array<double, 3> c = {{i * std::ceil( 4 * radius ) + offset,
                                        i * std::ceil( 4 * radius ) + offset,
                                        i * std::ceil( 4 * radius ) + offset}};
            // braced lists evaluate left to right, as in the reference test
            std::array<double, 3> t = {{c[0] + 0.5 * rx( g ),
                                        c[1] + 0.5 * rx( g ),
                                        c[2] + 0.5 * rx( g )}};
            for ( int d = 0; d < 3; ++d )
                tgt[3 * i + d] = t[d];
            for ( int k = 0; k < M; k++ )
            {
                std::array<double, 3> p = {
                    {c[0] + rx( g ), c[1] + rx( g ), c[2] + rx( g )}};
                for ( int d = 0; d < 3; ++d )
                    src[3 * ( i * M + k ) + d] = p[d];
            }
        }
        write( dir, "mls_points_m" + std::to_string( M ) + "_src", src );
        write( dir, "mls_points_m" + std::to_string( M ) + "_tgt", tgt );
    }
    return 0;
}

// Driver for the sanitizer build of the CPU oracle (oracle/dtkx_oracle.cpp)
// and of the host build of crmath.hpp: exercises every exported oracle
// function on small, ragged and edge-case inputs, so that AddressSanitizer
// and UndefinedBehaviorSanitizer see each code path. The test compiles it
// with g++ -fsanitize=address,undefined and fails on the patterns the
// reference's tests fail on ("runtime error", "leak", "data race";
// packages/Meshfree/test/CMakeLists.txt:10) or on a non-zero exit.
#include "../../datatransferkit_amd/csrc/crmath.hpp"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
struct dtkx_oracle_counters
{
    int64_t n_underdetermined, n_unconverged, total_rotations;
};
void dtkx_oracle_set_libm_mode( int );
void dtkx_oracle_elementary( int, double const *, double const *, double *,
                             int64_t );
int dtkx_oracle_rbf( int, double, double const *, int64_t, double * );
int dtkx_oracle_basis( int, int, int, double const *, int64_t, double * );
void dtkx_oracle_svd_2x2( double const *, double *, double * );
int dtkx_oracle_svd_pseudo_inverse( int, int64_t, double const *, double *,
                                    int, double, long, dtkx_oracle_counters * );
int dtkx_oracle_knn( double const *, int64_t, double const *, int64_t, int,
                     int, int, int32_t *, double * );
int dtkx_oracle_knn_brute( double const *, int64_t, double const *, int64_t,
                           int, int, int32_t *, double * );
int dtkx_oracle_mls_create( double const *, int64_t, double const *, int64_t,
                            int, int, int, int, int, int, double, long,
                            int32_t *, double *, dtkx_oracle_counters * );
int dtkx_oracle_basis_weights( double const *, int64_t, double const *,
                               int64_t, int, int32_t const *, int, int, int,
                               double * );
int dtkx_oracle_apply( int64_t, int, int32_t const *, double const *,
                       double const *, int64_t, double *, int64_t, int );
int dtkx_oracle_apply_transpose( int64_t, int, int32_t const *,
                                 double const *, double const *, int64_t,
                                 double *, int64_t, int64_t, int );
}

static int g_fail = 0;
#define CHECK( c )                                                             \
    do                                                                         \
    {                                                                          \
        if ( !( c ) )                                                          \
        {                                                                      \
            std::printf( "check failed: %s (line %d)\n", #c, __LINE__ );       \
            g_fail = 1;                                                        \
        }                                                                      \
    } while ( 0 )

int main()
{
    std::mt19937_64 rng( 7 );
    std::uniform_real_distribution<double> u( 0., 1. );
    auto cloud = [&]( int n ) {
        std::vector<double> p( 3 * n );
        for ( auto &v : p )
            v = u( rng );
        return p;
    };
    // elementary functions, both modes, and the host crmath
    std::vector<double> x( 257 ), y( 257 ), out( 257 );
    for ( int i = 0; i < 257; ++i )
    {
        x[i] = -3.0 + 6.0 * i / 256.0;
        y[i] = std::sin( 1.0 + i );
    }
    for ( int mode = 0; mode < 2; ++mode )
    {
        dtkx_oracle_set_libm_mode( mode );
        for ( int op = 0; op < 4; ++op )
            dtkx_oracle_elementary( op, y.data(), x.data(), out.data(), 257 );
    }
    for ( int i = 0; i < 257; ++i )
    {
        volatile double s = dtkx::crm::sin( x[i] ) + dtkx::crm::cos( x[i] ) +
                            dtkx::crm::atan2( y[i], x[i] ) +
                            dtkx::crm::log( std::fabs( x[i] ) + 1e-300 );
        (void)s;
    }
    // RBFs (Buhmann<2> at x = 0 is NaN by the reference's formula) and bases
    std::vector<double> r( 11 ), phi( 11 );
    for ( int i = 0; i < 11; ++i )
        r[i] = i / 10.0;
    for ( int rbf = 0; rbf <= 8; ++rbf )
        CHECK( dtkx_oracle_rbf( rbf, 1.0, r.data(), 11, phi.data() ) == 0 );
    auto pts = cloud( 20 );
    std::vector<double> basis( 20 * 10 );
    for ( int m : {1, 4, 10} )
        for ( int fp = 0; fp < 2; ++fp )
            CHECK( dtkx_oracle_basis( m, 3, fp, pts.data(), 20, basis.data() ) == 0 );
    CHECK( dtkx_oracle_basis( 6, 2, 0, pts.data(), 20, basis.data() ) == 0 );
    // SVD: 2x2, square systems with a zero column, relative truncation
    double a2[4] = {1., 2., 3., 4.}, L[4], R[4];
    dtkx_oracle_svd_2x2( a2, L, R );
    for ( int n : {1, 4, 10, 17} )
    {
        std::vector<double> a( 3 * n * n ), p( 3 * n * n );
        for ( auto &v : a )
            v = u( rng ) - 0.5;
        for ( int i = 0; i < n; ++i )
            a[i * n + ( n > 1 ? 1 : 0 )] = 0.0; // rank-deficient first matrix
        dtkx_oracle_counters c{};
        CHECK( dtkx_oracle_svd_pseudo_inverse( n, 3, a.data(), p.data(), 0, 0.,
                                               0, &c ) == 0 );
        CHECK( dtkx_oracle_svd_pseudo_inverse( n, 3, a.data(), p.data(), 1,
                                               1e-12, 0, &c ) == 0 );
    }
    // kNN: tree and brute force, both geometries and layouts, k above the
    // source count, duplicates
    for ( int ns : {1, 3, 50, 700} )
    {
        auto src = cloud( ns );
        if ( ns > 10 )
            for ( int d = 0; d < 3; ++d )
                src[3 * 5 + d] = src[3 * 4 + d];
        auto tgt = cloud( 33 );
        for ( int k : {1, 4, 16, 64} )
        {
            int const kk = k < ns ? k : ns;
            std::vector<int32_t> idx( 33 * kk ), idx2( 33 * kk );
            std::vector<double> dist( 33 * kk ), dist2( 33 * kk );
            for ( int geo = 0; geo < 2; ++geo )
                for ( int layout = 0; layout < 2; ++layout )
                {
                    CHECK( dtkx_oracle_knn( src.data(), ns, tgt.data(), 33, k,
                                            geo, layout, idx.data(),
                                            dist.data() ) == 0 );
                }
            CHECK( dtkx_oracle_knn_brute( src.data(), ns, tgt.data(), 33, k, 0,
                                          idx2.data(), dist2.data() ) == 0 );
            CHECK( dtkx_oracle_knn( src.data(), ns, tgt.data(), 33, k, 0, 0,
                                    idx.data(), dist.data() ) == 0 );
            CHECK( idx == idx2 );
        }
    }
    // MLS create + apply + transpose: every basis, solver truncation, both
    // geometries, k = m and k > m, no targets, fewer sources than k
    for ( int m : {1, 4, 10} )
        for ( int k : {0, 16} )
            for ( int geo = 0; geo < 2; ++geo )
                for ( int trunc = 0; trunc < 2; ++trunc )
                    for ( int nt : {0, 1, 40} )
                    {
                        int const ns = 300;
                        auto src = cloud( ns );
                        auto tgt = cloud( nt > 0 ? nt : 1 );
                        int const kk = k > 0 ? k : m;
                        std::vector<int32_t> idx( nt * kk + 1 );
                        std::vector<double> cf( nt * kk + 1 );
                        dtkx_oracle_counters c{};
                        CHECK( dtkx_oracle_mls_create(
                                   src.data(), ns, tgt.data(), nt, 0, k, m,
                                   geo == 0 ? 1 : 7, geo, trunc,
                                   trunc ? 1e-12 : 0., 0, idx.data(),
                                   cf.data(), &c ) == 0 );
                        std::vector<double> fx( 2 * ns ), fy( 2 * nt + 1 ),
                            fb( 2 * ns );
                        for ( auto &v : fx )
                            v = u( rng );
                        CHECK( dtkx_oracle_apply( nt, kk, idx.data(), cf.data(),
                                                  fx.data(), ns, fy.data(), nt,
                                                  2 ) == 0 );
                        CHECK( dtkx_oracle_apply_transpose(
                                   nt, kk, idx.data(), cf.data(), fy.data(),
                                   nt, fb.data(), ns, ns, 2 ) == 0 );
                        std::vector<double> w( nt * kk + 1 );
                        CHECK( dtkx_oracle_basis_weights(
                                   src.data(), ns, tgt.data(), nt, 0,
                                   idx.data(), kk, 0, geo, w.data() ) == 0 );
                    }
    {
        auto src = cloud( 3 );
        auto tgt = cloud( 5 );
        std::vector<int32_t> idx( 5 * 3 );
        std::vector<double> cf( 5 * 3 );
        dtkx_oracle_counters c{};
        CHECK( dtkx_oracle_mls_create( src.data(), 3, tgt.data(), 5, 0, 10, 4, 0,
                                       0, 0, 0., 0, idx.data(), cf.data(),
                                       &c ) == 0 );
    }
    std::printf( g_fail ? "sanitize driver: FAILED\n" : "sanitize driver: ok\n" );
    return g_fail;
}

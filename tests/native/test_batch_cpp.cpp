// The C++ batch wrapper (include/dtkx/MovingLeastSquaresBatch.hpp) against
// single operators built through the C ABI (dtkx_mls_create) with the same
// options: every problem's target values must have the same bits, for the
// per-problem and the concatenated forms, and again after the points move
// in place and the batch is rebuilt. Uneven problems: 1 to 2000 sources,
// one without targets, constant, linear and quadratic bases, k from 0 (the
// basis size) to 40. Points and values are device arrays (HIP runtime).
// Build: g++ -std=c++17 -D__HIP_PLATFORM_AMD__ -I include -I <rocm>/include
//        test_batch_cpp.cpp -L<dir> -ldtkx -L<rocm>/lib -lamdhip64
#include "dtkx/MovingLeastSquaresBatch.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

static int g_failures = 0;

static void report( bool ok, char const *what )
{
    std::printf( "%s %s\n", ok ? "PASS" : "FAIL", what );
    if ( !ok )
        ++g_failures;
}

#define HIP_OK( call )                                                         \
    do                                                                         \
    {                                                                          \
        if ( ( call ) != hipSuccess )                                          \
        {                                                                      \
            std::printf( "FAIL %s\n", #call );                                 \
            std::exit( 2 );                                                    \
        }                                                                      \
    } while ( 0 )

struct Problem
{
    std::vector<double> src, tgt, vals; // host copies, [n][3] and [n]
    int basis, k;
};

static std::vector<Problem> make_problems( unsigned seed )
{
    std::mt19937_64 rng( seed );
    std::uniform_real_distribution<double> u( 0.0, 1.0 );
    int const n_src[] = {300, 1, 2000, 64, 900, 17};
    int const n_tgt[] = {250, 5, 1800, 0, 1200, 30};
    int const basis[] = {DTKX_BASIS_LINEAR, DTKX_BASIS_CONSTANT,
                         DTKX_BASIS_QUADRATIC, DTKX_BASIS_LINEAR,
                         DTKX_BASIS_QUADRATIC, DTKX_BASIS_LINEAR};
    int const k[] = {0, 4, 40, 16, 16, 8};
    std::vector<Problem> out;
    for ( int p = 0; p < 6; ++p )
    {
        Problem q;
        double const shift = 3.0 * p;
        for ( int i = 0; i < 3 * n_src[p]; ++i )
            q.src.push_back( shift + u( rng ) );
        for ( int i = 0; i < 3 * n_tgt[p]; ++i )
            q.tgt.push_back( shift + 0.05 + 0.9 * u( rng ) );
        for ( int i = 0; i < n_src[p]; ++i )
            q.vals.push_back( std::sin( 2 * q.src[3 * i] ) +
                              q.src[3 * i + 1] * q.src[3 * i + 2] );
        q.basis = basis[p];
        q.k = k[p];
        out.push_back( q );
    }
    return out;
}

// One problem through dtkx_mls_create (host arrays, staged by the library).
static std::vector<double> single( Problem const &q, dtkx_mls_options o )
{
    std::int64_t const ns = q.src.size() / 3, nt = q.tgt.size() / 3;
    std::vector<double> y( nt );
    if ( nt == 0 )
        return y;
    o.basis = q.basis;
    o.k = q.k;
    dtkx_mls *h = nullptr;
    dtkx::check( dtkx_mls_create( q.src.data(), ns, q.tgt.data(), nt, &o, &h ) );
    dtkx::check( dtkx_mls_apply( h, q.vals.data(), y.data(), 1, ns, nt ) );
    dtkx_mls_destroy( h );
    return y;
}

template <typename T>
static T *upload( std::vector<T> const &v )
{
    T *d = nullptr;
    HIP_OK( hipMalloc( &d, sizeof( T ) * ( v.empty() ? 1 : v.size() ) ) );
    if ( !v.empty() )
        HIP_OK( hipMemcpy( d, v.data(), sizeof( T ) * v.size(),
                           hipMemcpyHostToDevice ) );
    return d;
}

static std::vector<double> download( double const *d, std::size_t n )
{
    std::vector<double> h( n );
    if ( n )
        HIP_OK( hipMemcpy( h.data(), d, sizeof( double ) * n,
                           hipMemcpyDeviceToHost ) );
    return h;
}

// y = 0 on the null stream, finished before the batch's own stream runs
static void clear( double *d, std::size_t n )
{
    HIP_OK( hipMemset( d, 0, sizeof( double ) * n ) );
    HIP_OK( hipDeviceSynchronize() );
}

static bool same_bits( std::vector<double> const &a, std::vector<double> const &b )
{
    return a.size() == b.size() &&
           ( a.empty() ||
             std::memcmp( a.data(), b.data(), sizeof( double ) * a.size() ) == 0 );
}

static void run_form( dtkx_mls_options o, char const *name )
{
    char what[160];
    auto probs = make_problems( 5 );
    std::size_t const n = probs.size();
    // concatenated device arrays, and per-problem views into them
    std::vector<double> src, tgt, vals;
    std::vector<std::int64_t> ns, nt;
    std::vector<int> bases, ks;
    for ( auto const &q : probs )
    {
        src.insert( src.end(), q.src.begin(), q.src.end() );
        tgt.insert( tgt.end(), q.tgt.begin(), q.tgt.end() );
        vals.insert( vals.end(), q.vals.begin(), q.vals.end() );
        ns.push_back( q.src.size() / 3 );
        nt.push_back( q.tgt.size() / 3 );
        bases.push_back( q.basis );
        ks.push_back( q.k );
    }
    double *d_src = upload( src ), *d_tgt = upload( tgt ), *d_x = upload( vals );
    std::size_t const n_tgt_total = tgt.size() / 3;
    double *d_y = upload( std::vector<double>( n_tgt_total + 1, 7.0 ) );

    std::vector<dtkx_mls_problem> recs;
    std::vector<double const *> xs;
    std::vector<double *> ys;
    std::int64_t s = 0, t = 0;
    for ( std::size_t p = 0; p < n; ++p )
    {
        recs.push_back( {d_src + 3 * s, ns[p], d_tgt + 3 * t, nt[p], bases[p],
                         ks[p]} );
        xs.push_back( d_x + s );
        ys.push_back( d_y + t );
        s += ns[p];
        t += nt[p];
    }

    // reference: one single operator per problem
    auto expect = [&]() {
        std::vector<double> all;
        for ( auto const &q : probs )
        {
            auto y = single( q, o );
            all.insert( all.end(), y.begin(), y.end() );
        }
        return all;
    };
    std::vector<double> ref = expect();

    dtkx::MovingLeastSquaresBatch<dtkx::Wendland<2>> per( recs, o );
    per.apply( xs, ys );
    std::snprintf( what, sizeof what, "%s: per-problem batch == single operators", name );
    report( same_bits( download( d_y, n_tgt_total ), ref ), what );

    clear( d_y, n_tgt_total );
    dtkx::MovingLeastSquaresBatch<dtkx::Wendland<2>> flat( d_src, ns, d_tgt, nt,
                                                          bases, ks, o );
    flat.apply( d_x, d_y );
    auto y_flat = download( d_y, n_tgt_total + 1 );
    std::snprintf( what, sizeof what,
                   "%s: concatenated batch == single operators, nothing past the end",
                   name );
    report( same_bits( std::vector<double>( y_flat.begin(), y_flat.end() - 1 ), ref ) &&
                y_flat.back() == 7.0,
            what );
    dtkx_mls_stats const st = flat.stats();
    std::snprintf( what, sizeof what, "%s: batch stats count every target", name );
    report( st.n_target == (std::int64_t)n_tgt_total && flat.size() == n, what );

    // move the points in place, rebuild both batches, compare again
    for ( auto &q : probs )
    {
        for ( auto &v : q.src )
            v = 1.5 * v + 0.25;
        for ( auto &v : q.tgt )
            v = 1.5 * v + 0.25;
    }
    for ( auto &v : src )
        v = 1.5 * v + 0.25;
    for ( auto &v : tgt )
        v = 1.5 * v + 0.25;
    HIP_OK( hipMemcpy( d_src, src.data(), sizeof( double ) * src.size(),
                       hipMemcpyHostToDevice ) );
    HIP_OK( hipMemcpy( d_tgt, tgt.data(), sizeof( double ) * tgt.size(),
                       hipMemcpyHostToDevice ) );
    ref = expect();
    for ( int rep = 0; rep < 2; ++rep )
    {
        clear( d_y, n_tgt_total );
        flat.rebuild();
        flat.apply( d_x, d_y );
        std::snprintf( what, sizeof what,
                       "%s: rebuild %d after moving the points == single operators",
                       name, rep + 1 );
        report( same_bits( download( d_y, n_tgt_total ), ref ), what );
    }
    clear( d_y, n_tgt_total );
    per.rebuild();
    per.apply( xs, ys );
    std::snprintf( what, sizeof what, "%s: per-problem rebuild == single operators",
                   name );
    report( same_bits( download( d_y, n_tgt_total ), ref ), what );

    HIP_OK( hipFree( d_src ) );
    HIP_OK( hipFree( d_tgt ) );
    HIP_OK( hipFree( d_x ) );
    HIP_OK( hipFree( d_y ) );
}

int main()
{
    dtkx_mls_options o = dtkx::default_options();
    o.rbf = DTKX_RBF_WENDLAND2; // what the batch's template policy sets
    run_form( o, "classical fp64" );
    o.solver = DTKX_SOLVER_CHOLESKY;
    o.precision = DTKX_PRECISION_FP32_BULK;
    run_form( o, "cholesky fp32_bulk" );

    // errors: counts of different lengths, and a problem without sources
    double *d = upload( std::vector<double>( 30, 0.5 ) );
    int status = 0;
    try
    {
        dtkx::MovingLeastSquaresBatch<> b( d, {5, 5}, d, {5}, {4, 4}, {0, 0} );
    }
    catch ( dtkx::DataTransferKitException const &e )
    {
        status = e.status();
    }
    report( status == DTKX_ERR_SIZE_MISMATCH, "mismatched counts throw" );
    status = 0;
    try
    {
        dtkx::MovingLeastSquaresBatch<> b( d, {5, 0}, d, {5, 5}, {4, 4}, {0, 0} );
    }
    catch ( dtkx::DataTransferKitException const &e )
    {
        status = e.status();
    }
    report( status == DTKX_ERR_EMPTY_SOURCE, "a problem without sources throws" );
    HIP_OK( hipFree( d ) );
    std::printf( "%d failure(s)\n", g_failures );
    return g_failures ? 1 : 0;
}

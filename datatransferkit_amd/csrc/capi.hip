// C ABI of the dtkx MLS transfer path (include/dtkx/mls.h).
//
// dtkx_mls_create mirrors the MovingLeastSquaresOperator constructor
// (DTK_MovingLeastSquaresOperator_def.hpp:24-119) and dtkx_mls_apply its
// apply (:121-141). Preconditions the reference checks with DTK_REQUIRE /
// DTK_CHECK (dimension 3 :39-42, non-empty source tree :49, value sizes
// :129-130) are always checked here and reported as status codes.
#include "apply.hpp"
#include "common.hpp"
#include "gradient.hpp"
#include "knn.hpp"
#include "setup.hpp"
#include "spline.hpp"

#include <algorithm>
#include <vector>
#include <type_traits>
#include <cmath>
#include <cstring>
#include <cstring>
#include <cstdio>
#include <memory>
#include <mutex>
#include <string>

// Events around the create stages (0..2 around the stages, 3..4 around
// the query kernels), read when the statistics are first asked for, so
// timing adds no host-device round trip to the create.
struct StageEvents
{
    hipEvent_t e[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    bool marked[5] = {false, false, false, false, false};
    hipEvent_t done = nullptr; // after the create's last launch
    hipStream_t s;
    explicit StageEvents( hipStream_t st )
        : s( st )
    {
        for ( auto &x : e )
            DTKX_HIP( hipEventCreate( &x ) );
        DTKX_HIP( hipEventCreateWithFlags( &done, hipEventDisableTiming ) );
    }
    void mark( int i )
    {
        DTKX_HIP( hipEventRecord( e[i], s ) );
        marked[i] = true;
    }
    // events 3 and 4 bracket the kNN query kernels (recorded by the search)
    void pass_query_events( dtkx::KnnInfo &info )
    {
        info.query_begin = e[3];
        info.query_end = e[4];
        marked[3] = marked[4] = true;
    }
    double ms( int i, int j )
    {
        if ( !marked[i] || !marked[j] )
            return 0.0;
        DTKX_HIP( hipEventSynchronize( e[j] ) );
        float t = 0;
        DTKX_HIP( hipEventElapsedTime( &t, e[i], e[j] ) );
        return t;
    }
    ~StageEvents()
    {
        for ( auto x : e )
            if ( x )
                (void)hipEventDestroy( x );
        if ( done )
            (void)hipEventDestroy( done );
    }
};

struct dtkx_mls
{
    dtkx_mls_options opts{};
    int64_t n_src = 0, n_tgt = 0;
    int m = 0, kk = 0;
    hipStream_t stream = nullptr;
    // CRS rows in the Morton order of the targets (row r: target order[r],
    // trow the inverse), columns as Morton positions of the sources
    // (position p: source perm[p]); empty buffers mean caller order
    dtkx::DeviceBuffer<int32_t> indices, order, trow, perm, spos;
    // create only: each row's largest neighbour distance (FP32 bulk)
    dtkx::DeviceBuffer<float> row_max;
    dtkx::DeviceBuffer<double> coeffs;
    // per-row pseudo-inverse, packed upper triangle (store_pinv)
    dtkx::DeviceBuffer<double> pinv;
    dtkx::DeviceBuffer<float> coeffs32;
    // src_d in position order, tgt_d in caller order
    dtkx::DeviceBuffer<double4> src_d, tgt_d;
    // create only, float geometry: the points rounded to float (caller order)
    dtkx::DeviceBuffer<float4> src_f, tgt_f;
    mutable dtkx::CscIndex csc;
    mutable bool csc_built = false;
    // the graph in caller order, built on the first dtkx_mls_get_graph
    mutable dtkx::DeviceBuffer<int32_t> graph_idx;
    mutable dtkx::DeviceBuffer<double> graph_c64;
    mutable dtkx::DeviceBuffer<float> graph_c32;
    mutable bool graph_built = false;
    // the CSR offsets of the rows (kk entries each), made on first request
    mutable dtkx::DeviceBuffer<int32_t> graph_off;
    // The create returns without waiting for the GPU: the counters its
    // setup and search accumulate and its stage events are read on the
    // first dtkx_mls_get_stats (resolve_stats), which waits for them.
    mutable dtkx_mls_stats stats{};
    mutable dtkx::DeviceBuffer<unsigned long long> counters;
    mutable std::unique_ptr<StageEvents> events;
    mutable bool stats_pending = false;
    void resolve_stats() const;

    dtkx::RowLayout layout() const
    {
        dtkx::RowLayout g;
        g.idx = indices.get();
        g.order = order.get();
        g.perm = perm.get();
        g.trow = trow.get();
        g.spos = spos.get();
        g.n_src = n_src;
        g.n_tgt = n_tgt;
        g.kk = kk;
        return g;
    }
};

// 1: the FP32-bulk setup takes each row's radius from the search's k-th
// distance instead of a pass over the neighbours. That distance is formed
// from float-rounded absolute coordinates, so far from the origin it is
// off by the coordinates' rounding: on a shell of radius 1 with k = 32
// (neighbours ~4e-3 apart) the FP32-bulk error against FP64 rose from
// 2.3e-7 to 1.04e-6; at the 1M x 1M cube it saved 2.7 % of a step.
#ifndef DTKX_FP32_KNN_RADIUS
#define DTKX_FP32_KNN_RADIUS 0
#endif

namespace
{

thread_local std::string g_last_error;

int fail( int status, std::string const &msg )
{
    g_last_error = msg;
    return status;
}

template <typename F>
int guarded( F &&f )
{
    try
    {
        g_last_error.clear();
        return f();
    }
    catch ( dtkx::Error const &e )
    {
        return fail( e.status, e.what() );
    }
    catch ( std::bad_alloc const & )
    {
        return fail( DTKX_ERR_OUT_OF_MEMORY, "dtkx: host allocation failed" );
    }
    catch ( std::exception const &e )
    {
        return fail( DTKX_ERR_INVALID_ARGUMENT, e.what() );
    }
}

#define REQ( c, st ) DTKX_REQUIRE( c, st )

void validate_options( dtkx_mls_options const &o )
{
    REQ( o.dim == 3, DTKX_ERR_DIMENSION );
    REQ( o.basis == DTKX_BASIS_CONSTANT || o.basis == DTKX_BASIS_LINEAR ||
             o.basis == DTKX_BASIS_QUADRATIC,
         DTKX_ERR_INVALID_ARGUMENT );
    REQ( o.rbf >= DTKX_RBF_WENDLAND0 && o.rbf <= DTKX_RBF_BUHMANN4,
         DTKX_ERR_INVALID_ARGUMENT );
    // Buhmann<4> as the reference writes it
    // (DTK_CompactlySupportedRadialBasisFunctions.hpp:139-146) is -67.886 at
    // x = 1, so it is not compactly supported and nothing pins its values
    REQ( o.rbf != DTKX_RBF_BUHMANN4, DTKX_ERR_UNSUPPORTED );
    REQ( o.k >= 0, DTKX_ERR_INVALID_ARGUMENT );
    REQ( o.k <= dtkx::knn_max_k, DTKX_ERR_UNSUPPORTED );
    REQ( o.precision == DTKX_PRECISION_FP64 ||
             o.precision == DTKX_PRECISION_FP32_BULK,
         DTKX_ERR_INVALID_ARGUMENT );
    REQ( o.geometry == DTKX_GEOMETRY_FLOAT ||
             o.geometry == DTKX_GEOMETRY_DOUBLE,
         DTKX_ERR_INVALID_ARGUMENT );
    REQ( o.solver == DTKX_SOLVER_CLASSICAL_JACOBI ||
             o.solver == DTKX_SOLVER_CYCLIC_JACOBI ||
             o.solver == DTKX_SOLVER_CHOLESKY,
         DTKX_ERR_INVALID_ARGUMENT );
    REQ( o.truncation == DTKX_TRUNCATION_ABSOLUTE ||
             o.truncation == DTKX_TRUNCATION_RELATIVE,
         DTKX_ERR_INVALID_ARGUMENT );
    REQ( o.layout == DTKX_LAYOUT_AOS || o.layout == DTKX_LAYOUT_SOA,
         DTKX_ERR_INVALID_ARGUMENT );
    REQ( o.group_size >= 0 && o.group_size <= 16, DTKX_ERR_INVALID_ARGUMENT );
    REQ( o.max_rotations >= 0, DTKX_ERR_INVALID_ARGUMENT );
    // the stored pseudo-inverse comes from the cyclic or Cholesky solvers
    REQ( !o.store_pinv || o.solver != DTKX_SOLVER_CLASSICAL_JACOBI ||
             o.precision == DTKX_PRECISION_FP32_BULK,
         DTKX_ERR_UNSUPPORTED );
}

// Device view of a caller array: used in place if it is device memory,
// otherwise copied into a staging buffer.
template <typename T>
struct Staged
{
    dtkx::DeviceBuffer<T> buf;
    T const *ptr = nullptr;
    Staged( T const *p, std::size_t n, hipStream_t s )
    {
        if ( n == 0 || !p )
        {
            ptr = p;
            return;
        }
        if ( dtkx::is_device_pointer( p ) )
        {
            ptr = p;
            return;
        }
        buf.resize( n, s );
        DTKX_HIP( hipMemcpyAsync( buf.get(), p, n * sizeof( T ),
                                  hipMemcpyHostToDevice, s ) );
        ptr = buf.get();
    }
};

// Output view: written in place if device memory, else through a staging
// buffer copied back by finish(). With gaps (fields at a leading dimension
// above their length) the staging buffer starts as a copy of the caller's
// array, so the entries between fields come back unchanged.
template <typename T>
struct StagedOut
{
    dtkx::DeviceBuffer<T> buf;
    T *user = nullptr;
    T *ptr = nullptr;
    std::size_t n = 0;
    StagedOut( T *p, std::size_t count, hipStream_t s, bool gaps = false )
        : user( p )
        , n( count )
    {
        if ( n == 0 || !p || dtkx::is_device_pointer( p ) )
        {
            ptr = p;
            return;
        }
        buf.resize( n, s );
        ptr = buf.get();
        if ( gaps )
            DTKX_HIP( hipMemcpyAsync( ptr, p, n * sizeof( T ),
                                      hipMemcpyHostToDevice, s ) );
    }
    void finish( hipStream_t s )
    {
        if ( buf.get() )
            DTKX_HIP( hipMemcpyAsync( user, buf.get(), n * sizeof( T ),
                                      hipMemcpyDeviceToHost, s ) );
    }
};

// End of an apply. With device arrays the work stays queued on the
// handle's stream, in order with the caller's later work there (the
// reference's apply launches Kokkos kernels, which do not wait either);
// staged host arrays are complete when the call returns.
void finish_apply( hipStream_t s, bool staged )
{
    if ( staged )
        DTKX_HIP( hipStreamSynchronize( s ) );
}

long default_rotation_cap( int m ) { return 200L * m * m + 1000; }

__global__ void offsets_kernel( int32_t *off, int64_t n, int kk )
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( i <= n )
        off[i] = (int32_t)( i * kk );
}

__global__ void to_f32_kernel( double const *a, float *b, int64_t n )
{
    int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if ( i < n )
        b[i] = (float)a[i];
}

// Create counters: the setup's four (underdetermined, unconverged,
// rotations, eigen fallbacks) and the search's re-run count, read back
// together at the end of the create.
constexpr int kCounters = 5;

// Shared tail of the two create paths: neighbours are in h->indices.
void run_setup( dtkx_mls *h, unsigned long long *counters_d )
{
    dtkx::SetupParams p;
    p.m = h->m;
    p.kk = h->kk;
    p.rbf = h->opts.rbf;
    p.geometry = h->opts.geometry;
    p.truncation = h->opts.truncation;
    p.rcond = h->opts.rcond;
    p.group_size = h->opts.group_size;
    if ( h->opts.store_pinv )
        h->pinv.resize( h->n_tgt * ( h->m * ( h->m + 1 ) / 2 ), h->stream );
    if ( h->opts.precision == DTKX_PRECISION_FP32_BULK ||
         h->opts.solver != DTKX_SOLVER_CLASSICAL_JACOBI )
    {
        // the FP32-bulk form uses the cyclic solver unless Cholesky is asked
        p.max_rot = h->opts.max_rotations; // sweeps; 0 = default
        bool const f32 = h->opts.precision == DTKX_PRECISION_FP32_BULK;
        p.row_max = DTKX_FP32_KNN_RADIUS && f32 ? h->row_max.get() : nullptr;
        auto *setup = h->opts.solver == DTKX_SOLVER_CHOLESKY
                          ? &dtkx::setup_fast
                          : &dtkx::setup_cyclic;
        setup( h->src_d.get(), h->tgt_d.get(), h->n_tgt, h->indices.get(),
               h->order.get(), p, f32 ? nullptr : h->coeffs.get(),
               f32 ? h->coeffs32.get() : nullptr, h->pinv.get(), counters_d,
               h->stream );
        h->stats.group_size_used = h->m == 10 ? 2 : 1;
        return;
    }
    p.max_rot = h->opts.max_rotations > 0 ? h->opts.max_rotations
                                          : default_rotation_cap( h->m );
    int g = dtkx::setup_classical(
        h->src_d.get(), h->tgt_d.get(), h->n_tgt, h->indices.get(),
        h->order.get(), p, h->coeffs.get(), counters_d,
        h->stream );
    h->stats.group_size_used = g;
}

// The create's last launch; its counters and events stay with the handle
// until the statistics are asked for (no wait here).
void finish_create( dtkx_mls *h,
                    dtkx::DeviceBuffer<unsigned long long> &&counters,
                    std::unique_ptr<StageEvents> &&events )
{
    h->stats.n_source = h->n_src;
    h->stats.n_target = h->n_tgt;
    h->stats.nnz = h->n_tgt * h->kk;
    h->stats.basis_size = h->m;
    h->stats.k = h->kk;
    DTKX_HIP( hipEventRecord( events->done, h->stream ) );
    h->counters = std::move( counters );
    h->events = std::move( events );
    h->stats_pending = true;
}

} // namespace

void dtkx_mls::resolve_stats() const
{
    if ( !stats_pending )
        return;
    // read on the library's non-blocking side stream after the create's
    // last event: the wait covers this create only, not work the caller
    // has queued on the handle's stream since
    unsigned long long c[kCounters] = {0, 0, 0, 0, 0};
    hipStream_t const side = dtkx::side_stream();
    DTKX_HIP( hipStreamWaitEvent( side, events->done, 0 ) );
    DTKX_HIP( hipMemcpyAsync( c, counters.get(), sizeof( c ),
                              hipMemcpyDeviceToHost, side ) );
    DTKX_HIP( hipStreamSynchronize( side ) );
    stats.n_underdetermined = (int64_t)c[0];
    stats.n_unconverged = (int64_t)c[1];
    stats.total_rotations = (int64_t)c[2];
    stats.n_eigen_fallback = (int64_t)c[3];
    // an int32 count in the low half of a zeroed slot
    stats.n_stack_overflow = (int64_t)( c[4] & 0xffffffffull );
    stats.ms_knn = events->ms( 0, 1 );
    stats.ms_setup = events->ms( 1, 2 );
    stats.ms_knn_query = events->ms( 3, 4 );
    counters = {};
    events.reset();
    stats_pending = false;
}

namespace
{

// coefficient storage of the handle's precision only
void alloc_coeffs( dtkx_mls *h, int64_t nnz )
{
    if ( h->opts.precision == DTKX_PRECISION_FP32_BULK )
    {
        h->coeffs.release();
        h->coeffs32.resize( std::max<int64_t>( nnz, 1 ), h->stream );
    }
    else
    {
        h->coeffs32.release();
        h->coeffs.resize( std::max<int64_t>( nnz, 1 ), h->stream );
    }
}

// float_copies: also the float points the search of float geometry needs
std::unique_ptr<dtkx_mls> make_handle( double const *src, int64_t n_src,
                                       double const *tgt, int64_t n_tgt,
                                       dtkx_mls_options const *opts,
                                       bool float_copies = false )
{
    REQ( opts != nullptr, DTKX_ERR_INVALID_ARGUMENT );
    validate_options( *opts );
    REQ( n_src >= 0 && n_tgt >= 0, DTKX_ERR_INVALID_ARGUMENT );
    REQ( n_src > 0, DTKX_ERR_EMPTY_SOURCE );
    REQ( src != nullptr, DTKX_ERR_INVALID_ARGUMENT );
    REQ( n_tgt == 0 || tgt != nullptr, DTKX_ERR_INVALID_ARGUMENT );
    REQ( n_src < INT32_MAX && n_tgt < INT32_MAX, DTKX_ERR_UNSUPPORTED );
    auto h = std::make_unique<dtkx_mls>();
    h->opts = *opts;
    h->n_src = n_src;
    h->n_tgt = n_tgt;
    h->m = opts->basis;
    h->stream = static_cast<hipStream_t>( opts->stream );
    int const k = opts->k > 0 ? opts->k : h->m;
    h->kk = static_cast<int>( k < n_src ? k : n_src );
    REQ( (int64_t)h->kk * n_tgt < INT32_MAX, DTKX_ERR_UNSUPPORTED );
    Staged<double> s_src( src, 3 * n_src, h->stream );
    Staged<double> s_tgt( tgt, 3 * n_tgt, h->stream );
    h->src_d.resize( n_src, h->stream );
    h->tgt_d.resize( std::max<int64_t>( n_tgt, 1 ), h->stream );
    if ( float_copies )
    {
        h->src_f.resize( n_src, h->stream );
        h->tgt_f.resize( std::max<int64_t>( n_tgt, 1 ), h->stream );
    }
    dtkx::points_to_double4( s_src.ptr, n_src, opts->layout, h->src_d.get(),
                             h->stream, h->src_f.get() );
    dtkx::points_to_double4( s_tgt.ptr, n_tgt, opts->layout, h->tgt_d.get(),
                             h->stream, h->tgt_f.get() );
    h->indices.resize( std::max<int64_t>( n_tgt * h->kk, 1 ), h->stream );
    alloc_coeffs( h.get(), n_tgt * h->kk );
    h->order.resize( std::max<int64_t>( n_tgt, 1 ), h->stream );
    // host point arrays are read by the time the create returns; the
    // staging buffers are freed in stream order at scope exit
    if ( s_src.buf.get() || s_tgt.buf.get() )
        DTKX_HIP( hipStreamSynchronize( h->stream ) );
    return h;
}

} // namespace

extern "C" {

const char *dtkx_last_error( void ) { return g_last_error.c_str(); }

const char *dtkx_version( void ) { return "dtkx 0.1 (gfx950)"; }

int dtkx_release_cached_memory( void )
{
    return guarded( [&]() {
        dtkx::cache_release();
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_options_init( dtkx_mls_options *o )
{
    if ( !o )
        return fail( DTKX_ERR_INVALID_ARGUMENT, "dtkx: opts != nullptr" );
    std::memset( o, 0, sizeof( *o ) );
    o->dim = 3;
    o->basis = DTKX_BASIS_LINEAR;
    o->rbf = DTKX_RBF_WENDLAND0;
    o->k = 0;
    o->precision = DTKX_PRECISION_FP64;
    o->geometry = DTKX_GEOMETRY_FLOAT;
    o->solver = DTKX_SOLVER_CLASSICAL_JACOBI;
    o->truncation = DTKX_TRUNCATION_ABSOLUTE;
    o->rcond = 1e-12;
    o->max_rotations = 0;
    o->layout = DTKX_LAYOUT_AOS;
    o->group_size = 0;
    o->store_pinv = 0;
    o->stream = nullptr;
    return DTKX_SUCCESS;
}

int dtkx_mls_create( const double *src_xyz, int64_t n_src,
                     const double *tgt_xyz, int64_t n_tgt,
                     const dtkx_mls_options *opts, dtkx_mls **out )
{
    return guarded( [&]() {
        REQ( out != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        *out = nullptr;
        REQ( opts != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        auto h = make_handle( src_xyz, n_src, tgt_xyz, n_tgt, opts,
                              opts->geometry == DTKX_GEOMETRY_FLOAT &&
                                  n_tgt > 0 );
        auto events = std::make_unique<StageEvents>( h->stream );
        StageEvents &ev = *events;
        dtkx::DeviceBuffer<unsigned long long> counters( kCounters,
                                                         h->stream );
        DTKX_HIP( hipMemsetAsync( counters.get(), 0,
                                  kCounters * sizeof( uint64_t ), h->stream ) );
        if ( n_tgt > 0 )
        {
            ev.mark( 0 );
            dtkx::KnnInfo info;
            ev.pass_query_events( info );
            info.n_overflow_dev =
                reinterpret_cast<int32_t *>( counters.get() + 4 );
            int const k = opts->k > 0 ? opts->k : h->m;
            dtkx::KnnOutput kout;
            h->perm.resize( n_src, h->stream );
            h->spos.resize( n_src, h->stream );
            // the FP32-bulk setup may take its radius from the search
            if ( DTKX_FP32_KNN_RADIUS &&
                 opts->precision == DTKX_PRECISION_FP32_BULK )
            {
                h->row_max.resize( n_tgt, h->stream );
                kout.row_max = h->row_max.get();
            }
            kout.indices = h->indices.get();
            kout.query_order = h->order.get();
            kout.src_perm = h->perm.get();
            kout.src_pos = h->spos.get();
            kout.morton_layout = true;
            // the search's last sort pass also writes the target rows and
            // the sources in position order (for the setup's gathers)
            h->trow.resize( n_tgt, h->stream );
            kout.query_pos = h->trow.get();
            dtkx::DeviceBuffer<double4> sorted( n_src, h->stream );
            kout.src4 = h->src_d.get();
            kout.sorted_src = sorted.get();
            if ( opts->geometry == DTKX_GEOMETRY_FLOAT )
            {
                // float points rounded from the caller's doubles (made
                // with the double4 copies), freed after the search
                dtkx::knn_search( opts->geometry, h->src_f.get(), n_src,
                                  h->tgt_f.get(), n_tgt, k, kout, &info,
                                  h->stream );
                h->src_f = {};
                h->tgt_f = {};
            }
            else
                dtkx::knn_search( opts->geometry, h->src_d.get(), n_src,
                                  h->tgt_d.get(), n_tgt, k, kout, &info,
                                  h->stream );
            h->src_d = std::move( sorted );
            ev.mark( 1 );
            run_setup( h.get(), counters.get() );
            h->row_max = {};
            ev.mark( 2 );
        }
        finish_create( h.get(), std::move( counters ), std::move( events ) );
        *out = h.release();
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_create_with_neighbors( const double *src_xyz, int64_t n_src,
                                    const double *tgt_xyz, int64_t n_tgt,
                                    const int32_t *indices, int32_t k_row,
                                    const dtkx_mls_options *opts,
                                    dtkx_mls **out )
{
    return guarded( [&]() {
        REQ( out != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        *out = nullptr;
        REQ( k_row >= 1 && k_row <= 1024, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_tgt == 0 || indices != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        dtkx_mls_options o = *opts;
        o.k = 0;
        auto h = make_handle( src_xyz, n_src, tgt_xyz, n_tgt, &o );
        h->opts.k = k_row;
        h->kk = k_row;
        h->indices.resize( std::max<int64_t>( n_tgt * k_row, 1 ), h->stream );
        alloc_coeffs( h.get(), n_tgt * k_row );
        h->order.release(); // natural target order
        if ( n_tgt > 0 )
        {
            std::size_t const nn = (std::size_t)n_tgt * k_row;
            // indices must lie in [0, n_src): checked on the host copy
            std::unique_ptr<int32_t[]> host( new int32_t[nn] );
            DTKX_HIP( hipMemcpy( host.get(), indices, nn * sizeof( int32_t ),
                                 hipMemcpyDefault ) );
            for ( std::size_t i = 0; i < nn; ++i )
                REQ( host[i] >= 0 && host[i] < n_src,
                     DTKX_ERR_INVALID_ARGUMENT );
            DTKX_HIP( hipMemcpyAsync( h->indices.get(), host.get(),
                                      nn * sizeof( int32_t ),
                                      hipMemcpyHostToDevice, h->stream ) );
            DTKX_HIP( hipStreamSynchronize( h->stream ) );
        }
        auto events = std::make_unique<StageEvents>( h->stream );
        dtkx::DeviceBuffer<unsigned long long> counters( kCounters,
                                                         h->stream );
        DTKX_HIP( hipMemsetAsync( counters.get(), 0,
                                  kCounters * sizeof( uint64_t ), h->stream ) );
        if ( n_tgt > 0 )
        {
            events->mark( 1 );
            run_setup( h.get(), counters.get() );
            events->mark( 2 );
        }
        finish_create( h.get(), std::move( counters ), std::move( events ) );
        *out = h.release();
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_apply( const dtkx_mls *h, const double *src_vals,
                    double *tgt_vals, int32_t n_fields, int64_t ld_src,
                    int64_t ld_tgt )
{
    return guarded( [&]() {
        REQ( h != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_fields >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( ld_src >= h->n_src, DTKX_ERR_SIZE_MISMATCH );
        REQ( ld_tgt >= h->n_tgt, DTKX_ERR_SIZE_MISMATCH );
        REQ( src_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->n_tgt == 0 || tgt_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        if ( h->n_tgt == 0 )
            return DTKX_SUCCESS;
        std::size_t const nin = (std::size_t)( n_fields - 1 ) * ld_src + h->n_src;
        std::size_t const nout = (std::size_t)( n_fields - 1 ) * ld_tgt + h->n_tgt;
        Staged<double> x( src_vals, nin, h->stream );
        StagedOut<double> y( tgt_vals, nout, h->stream, ld_tgt > h->n_tgt );
        if ( h->opts.precision == DTKX_PRECISION_FP32_BULK )
            dtkx::crs_apply_f32_dvals( h->layout(), h->coeffs32.get(), x.ptr,
                                       ld_src, y.ptr, ld_tgt, n_fields,
                                       h->stream );
        else
            dtkx::crs_apply_f64( h->layout(), h->coeffs.get(), x.ptr, ld_src,
                                 y.ptr, ld_tgt, n_fields, h->stream );
        y.finish( h->stream );
        finish_apply( h->stream, x.buf.get() || y.buf.get() );
        return DTKX_SUCCESS;
    } );
}

// ------------------------------------------------ apply without allocation
// The scratch of dtkx_mls_apply_noalloc: one array per GPU, allocated with
// the library's code object and held until the process ends. Host code
// never passes the array's name (a host-side stand-in) to a kernel: it
// takes the device address from hipGetSymbolAddress once per device and
// uses that.
namespace
{
constexpr std::size_t kNoAllocScratchBytes = DTKX_NOALLOC_SCRATCH_BYTES;
static __device__ __align__( 256 ) unsigned char
    g_noalloc_scratch[kNoAllocScratchBytes];

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mutex;
void *g_noalloc_ptr[kNoAllocMaxDevices] = {};

// the current device's copy of g_noalloc_scratch (looked up on its first use)
void *noalloc_scratch_ptr()
{
    int dev = 0;
    DTKX_HIP( hipGetDevice( &dev ) );
    DTKX_REQUIRE( dev >= 0 && dev < kNoAllocMaxDevices, DTKX_ERR_UNSUPPORTED );
    std::lock_guard<std::mutex> lock( g_noalloc_mutex );
    if ( !g_noalloc_ptr[dev] )
    {
        void *p = nullptr;
        DTKX_HIP( hipGetSymbolAddress( &p, HIP_SYMBOL( g_noalloc_scratch ) ) );
        DTKX_REQUIRE( p != nullptr, DTKX_ERR_HIP );
        g_noalloc_ptr[dev] = p;
    }
    return g_noalloc_ptr[dev];
}
} // namespace

int dtkx_mls_apply_noalloc( const dtkx_mls *h, const double *src_vals,
                            double *tgt_vals, int32_t n_fields,
                            int64_t ld_src, int64_t ld_tgt )
{
    return guarded( [&]() {
        REQ( h != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_fields >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( ld_src >= h->n_src, DTKX_ERR_SIZE_MISMATCH );
        REQ( ld_tgt >= h->n_tgt, DTKX_ERR_SIZE_MISMATCH );
        REQ( src_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->n_tgt == 0 || tgt_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        if ( h->n_tgt == 0 )
            return DTKX_SUCCESS;
        // host arrays would need staging buffers
        REQ( dtkx::is_device_pointer( src_vals ), DTKX_ERR_INVALID_ARGUMENT );
        REQ( dtkx::is_device_pointer( tgt_vals ), DTKX_ERR_INVALID_ARGUMENT );
        bool const fp32 = h->opts.precision == DTKX_PRECISION_FP32_BULK;
        std::size_t const need =
            dtkx::crs_apply_scratch_bytes( h->layout(), n_fields, fp32 );
        REQ( need <= kNoAllocScratchBytes, DTKX_ERR_UNSUPPORTED );
        void *const ws = noalloc_scratch_ptr();
        if ( fp32 )
            dtkx::crs_apply_f32_dvals_ws( h->layout(), h->coeffs32.get(),
                                          src_vals, ld_src, tgt_vals, ld_tgt,
                                          n_fields, ws, h->stream );
        else
            dtkx::crs_apply_f64_ws( h->layout(), h->coeffs.get(), src_vals,
                                    ld_src, tgt_vals, ld_tgt, n_fields, ws,
                                    h->stream );
        DTKX_HIP( hipStreamSynchronize( h->stream ) );
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_apply_fp32( const dtkx_mls *h, const float *src_vals,
                         float *tgt_vals, int32_t n_fields, int64_t ld_src,
                         int64_t ld_tgt )
{
    return guarded( [&]() {
        REQ( h != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->opts.precision == DTKX_PRECISION_FP32_BULK,
             DTKX_ERR_UNSUPPORTED );
        REQ( n_fields >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( ld_src >= h->n_src, DTKX_ERR_SIZE_MISMATCH );
        REQ( ld_tgt >= h->n_tgt, DTKX_ERR_SIZE_MISMATCH );
        REQ( src_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->n_tgt == 0 || tgt_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        if ( h->n_tgt == 0 )
            return DTKX_SUCCESS;
        std::size_t const nin = (std::size_t)( n_fields - 1 ) * ld_src + h->n_src;
        std::size_t const nout = (std::size_t)( n_fields - 1 ) * ld_tgt + h->n_tgt;
        Staged<float> x( src_vals, nin, h->stream );
        StagedOut<float> y( tgt_vals, nout, h->stream, ld_tgt > h->n_tgt );
        dtkx::crs_apply_f32( h->layout(), h->coeffs32.get(), x.ptr, ld_src,
                             y.ptr, ld_tgt, n_fields, h->stream );
        y.finish( h->stream );
        finish_apply( h->stream, x.buf.get() || y.buf.get() );
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_apply_transpose( const dtkx_mls *h, const double *tgt_bar,
                              double *src_bar, int32_t n_fields,
                              int64_t ld_tgt, int64_t ld_src )
{
    return guarded( [&]() {
        REQ( h != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_fields >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( ld_src >= h->n_src, DTKX_ERR_SIZE_MISMATCH );
        REQ( ld_tgt >= h->n_tgt, DTKX_ERR_SIZE_MISMATCH );
        REQ( src_bar != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->n_tgt == 0 || tgt_bar != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->opts.precision == DTKX_PRECISION_FP64, DTKX_ERR_UNSUPPORTED );
        if ( !h->csc_built )
        {
            dtkx::build_csc( h->layout(), h->coeffs.get(), h->csc,
                             h->stream );
            h->csc_built = true;
        }
        std::size_t const nin = (std::size_t)( n_fields - 1 ) * ld_tgt + h->n_tgt;
        std::size_t const nout = (std::size_t)( n_fields - 1 ) * ld_src + h->n_src;
        Staged<double> yb( tgt_bar, nin, h->stream );
        StagedOut<double> xb( src_bar, nout, h->stream, ld_src > h->n_src );
        dtkx::crs_apply_transpose_f64( h->layout(), h->csc, h->coeffs.get(),
                                       yb.ptr, ld_tgt, xb.ptr, ld_src,
                                       n_fields, h->stream );
        xb.finish( h->stream );
        DTKX_HIP( hipStreamSynchronize( h->stream ) );
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_target_gradient( const dtkx_mls *h, const double *src_vals,
                              const double *tgt_bar, int32_t n_fields,
                              int64_t ld_src, int64_t ld_tgt, double *d_tgt )
{
    return guarded( [&]() {
        REQ( h != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_fields >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( ld_src >= h->n_src, DTKX_ERR_SIZE_MISMATCH );
        REQ( ld_tgt >= h->n_tgt, DTKX_ERR_SIZE_MISMATCH );
        REQ( src_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->n_tgt == 0 || ( tgt_bar != nullptr && d_tgt != nullptr ),
             DTKX_ERR_INVALID_ARGUMENT );
        // float-rounded coordinates make the operator piecewise constant
        REQ( h->opts.geometry == DTKX_GEOMETRY_DOUBLE, DTKX_ERR_UNSUPPORTED );
        REQ( h->opts.precision == DTKX_PRECISION_FP64, DTKX_ERR_UNSUPPORTED );
        if ( h->n_tgt == 0 )
            return DTKX_SUCCESS;
        std::size_t const nin = (std::size_t)( n_fields - 1 ) * ld_src + h->n_src;
        std::size_t const nbar = (std::size_t)( n_fields - 1 ) * ld_tgt + h->n_tgt;
        Staged<double> x( src_vals, nin, h->stream );
        Staged<double> yb( tgt_bar, nbar, h->stream );
        StagedOut<double> g( d_tgt, (std::size_t)h->n_tgt * 3, h->stream );
        dtkx::GradientArgs a;
        a.src = h->src_d.get();
        a.tgt = h->tgt_d.get();
        a.indices = h->indices.get();
        a.order = h->order.get();
        a.perm = h->perm.get();
        a.spos = h->spos.get();
        a.pinv = h->pinv.get();
        a.n_src = h->n_src;
        a.n_tgt = h->n_tgt;
        a.m = h->m;
        a.kk = h->kk;
        a.rbf = h->opts.rbf;
        a.truncation = h->opts.truncation;
        a.rcond = h->opts.rcond;
        a.max_sweeps = h->opts.solver == DTKX_SOLVER_CYCLIC_JACOBI
                           ? h->opts.max_rotations
                           : 0;
        a.solver = h->opts.solver;
        a.ld_tgt = ld_tgt;
        a.d_tgt = g.ptr;
        dtkx::target_gradient( a, x.ptr, ld_src, n_fields, yb.ptr, h->stream );
        g.finish( h->stream );
        DTKX_HIP( hipStreamSynchronize( h->stream ) );
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_get_graph( const dtkx_mls *h, const int32_t **offsets,
                        const int32_t **indices, const void **coeffs,
                        int64_t *nnz )
{
    return guarded( [&]() {
        REQ( h != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        bool const f32 = h->opts.precision == DTKX_PRECISION_FP32_BULK;
        bool const permuted = h->order.get() || h->perm.get();
        int64_t const n_entries = h->n_tgt * h->kk;
        if ( permuted && !h->graph_built )
        {
            std::size_t const len = std::max<int64_t>( n_entries, 1 );
            h->graph_idx.resize( len, h->stream );
            if ( f32 )
            {
                h->graph_c32.resize( len, h->stream );
                dtkx::canonical_graph( h->layout(), h->coeffs32.get(),
                                       h->graph_idx.get(), h->graph_c32.get(),
                                       h->stream );
            }
            else
            {
                h->graph_c64.resize( len, h->stream );
                dtkx::canonical_graph( h->layout(), h->coeffs.get(),
                                       h->graph_idx.get(), h->graph_c64.get(),
                                       h->stream );
            }
            DTKX_HIP( hipStreamSynchronize( h->stream ) );
            h->graph_built = true;
        }
        if ( offsets )
        {
            if ( !h->graph_off.get() )
            {
                h->graph_off.resize( h->n_tgt + 1, h->stream );
                offsets_kernel<<<dtkx::ceil_div( h->n_tgt + 1, 256 ), 256, 0,
                                 h->stream>>>( h->graph_off.get(), h->n_tgt,
                                               h->kk );
                DTKX_HIP( hipGetLastError() );
                DTKX_HIP( hipStreamSynchronize( h->stream ) );
            }
            *offsets = h->graph_off.get();
        }
        if ( indices )
            *indices = permuted ? h->graph_idx.get() : h->indices.get();
        if ( coeffs )
        {
            if ( f32 )
                *coeffs = permuted ? h->graph_c32.get() : h->coeffs32.get();
            else
                *coeffs = permuted ? h->graph_c64.get() : h->coeffs.get();
        }
        if ( nnz )
            *nnz = n_entries;
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_write_graph( const dtkx_mls *h, const char *path,
                          int64_t col_base )
{
    return guarded( [&]() {
        REQ( h != nullptr && path != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        int32_t const *offsets = nullptr, *indices = nullptr;
        void const *coeffs = nullptr;
        int64_t nnz = 0;
        int const st = dtkx_mls_get_graph( h, &offsets, &indices, &coeffs, &nnz );
        if ( st != DTKX_SUCCESS )
            return static_cast<dtkx_status>( st );
        bool const f32 = h->opts.precision == DTKX_PRECISION_FP32_BULK;
        int32_t const cbytes = f32 ? 4 : 8;
        std::vector<int32_t> off( h->n_tgt + 1 ), idx( nnz );
        std::vector<char> co( (std::size_t)nnz * cbytes );
        DTKX_HIP( hipMemcpyAsync( off.data(), offsets,
                                  off.size() * sizeof( int32_t ),
                                  hipMemcpyDeviceToHost, h->stream ) );
        if ( nnz > 0 )
        {
            DTKX_HIP( hipMemcpyAsync( idx.data(), indices,
                                      idx.size() * sizeof( int32_t ),
                                      hipMemcpyDeviceToHost, h->stream ) );
            DTKX_HIP( hipMemcpyAsync( co.data(), coeffs, co.size(),
                                      hipMemcpyDeviceToHost, h->stream ) );
        }
        DTKX_HIP( hipStreamSynchronize( h->stream ) );
        std::vector<int64_t> off64( off.begin(), off.end() ), col( nnz );
        for ( int64_t e = 0; e < nnz; ++e )
            col[e] = col_base + idx[e];
        std::FILE *f = std::fopen( path, "wb" );
        REQ( f != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        int64_t const head[4] = {h->n_tgt, h->n_src, nnz, col_base};
        int32_t const head2[2] = {cbytes, 0};
        bool ok = std::fwrite( "DTKXCSR1", 1, 8, f ) == 8 &&
                  std::fwrite( head, sizeof( head ), 1, f ) == 1 &&
                  std::fwrite( head2, sizeof( head2 ), 1, f ) == 1 &&
                  std::fwrite( off64.data(), sizeof( int64_t ), off64.size(),
                               f ) == off64.size() &&
                  ( nnz == 0 ||
                    ( std::fwrite( col.data(), sizeof( int64_t ), col.size(),
                                   f ) == col.size() &&
                      std::fwrite( co.data(), 1, co.size(), f ) ==
                          co.size() ) );
        ok = ( std::fclose( f ) == 0 ) && ok;
        REQ( ok, DTKX_ERR_INVALID_ARGUMENT );
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_get_stats( const dtkx_mls *h, dtkx_mls_stats *stats )
{
    return guarded( [&]() {
        REQ( h != nullptr && stats != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        h->resolve_stats();
        *stats = h->stats;
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_destroy( dtkx_mls *h )
{
    return guarded( [&]() {
        delete h;
        return DTKX_SUCCESS;
    } );
}

int dtkx_knn( const double *src_xyz, int64_t n_src, const double *tgt_xyz,
              int64_t n_tgt, int32_t k, int32_t geometry, int32_t layout,
              int32_t *indices, double *dist, void *stream )
{
    return guarded( [&]() {
        hipStream_t s = static_cast<hipStream_t>( stream );
        REQ( n_src > 0, DTKX_ERR_EMPTY_SOURCE );
        REQ( n_tgt >= 0 && src_xyz != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_tgt == 0 || ( tgt_xyz != nullptr && indices != nullptr ),
             DTKX_ERR_INVALID_ARGUMENT );
        REQ( k >= 1 && k <= dtkx::knn_max_k, DTKX_ERR_UNSUPPORTED );
        REQ( geometry == DTKX_GEOMETRY_FLOAT ||
                 geometry == DTKX_GEOMETRY_DOUBLE,
             DTKX_ERR_INVALID_ARGUMENT );
        REQ( layout == DTKX_LAYOUT_AOS || layout == DTKX_LAYOUT_SOA,
             DTKX_ERR_INVALID_ARGUMENT );
        if ( n_tgt == 0 )
            return DTKX_SUCCESS;
        int const kk = k < n_src ? k : static_cast<int>( n_src );
        Staged<double> sx( src_xyz, 3 * n_src, s );
        Staged<double> tx( tgt_xyz, 3 * n_tgt, s );
        StagedOut<int32_t> oi( indices, (std::size_t)n_tgt * kk, s );
        StagedOut<double> od( dist, dist ? (std::size_t)n_tgt * kk : 0, s );
        dtkx::KnnOutput kout;
        kout.indices = oi.ptr;
        kout.dist = od.ptr;
        if ( geometry == DTKX_GEOMETRY_FLOAT )
        {
            dtkx::DeviceBuffer<float4> sf( n_src, s ), tf( n_tgt, s );
            dtkx::points_to_float4( sx.ptr, n_src, layout, sf.get(), s );
            dtkx::points_to_float4( tx.ptr, n_tgt, layout, tf.get(), s );
            dtkx::knn_search( geometry, sf.get(), n_src, tf.get(), n_tgt, k,
                              kout, nullptr, s );
        }
        else
        {
            dtkx::DeviceBuffer<double4> sd( n_src, s ), td( n_tgt, s );
            dtkx::points_to_double4( sx.ptr, n_src, layout, sd.get(), s );
            dtkx::points_to_double4( tx.ptr, n_tgt, layout, td.get(), s );
            dtkx::knn_search( geometry, sd.get(), n_src, td.get(), n_tgt, k,
                              kout, nullptr, s );
        }
        oi.finish( s );
        od.finish( s );
        DTKX_HIP( hipStreamSynchronize( s ) );
        return DTKX_SUCCESS;
    } );
}

// ------------------------------------------ nearest-neighbour operator
struct dtkx_nn
{
    int64_t n_src = 0, n_tgt = 0;
    hipStream_t stream = nullptr;
    dtkx::DeviceBuffer<int32_t> idx; // nearest source of each target
};

int dtkx_nn_create( const double *src_xyz, int64_t n_src,
                    const double *tgt_xyz, int64_t n_tgt, int32_t geometry,
                    int32_t layout, void *stream, dtkx_nn **out )
{
    return guarded( [&]() {
        REQ( out != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        *out = nullptr;
        hipStream_t s = static_cast<hipStream_t>( stream );
        REQ( n_src > 0, DTKX_ERR_EMPTY_SOURCE );
        REQ( n_tgt >= 0 && src_xyz != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_tgt == 0 || tgt_xyz != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_src <= INT32_MAX && n_tgt <= INT32_MAX, DTKX_ERR_UNSUPPORTED );
        REQ( geometry == DTKX_GEOMETRY_FLOAT ||
                 geometry == DTKX_GEOMETRY_DOUBLE,
             DTKX_ERR_INVALID_ARGUMENT );
        REQ( layout == DTKX_LAYOUT_AOS || layout == DTKX_LAYOUT_SOA,
             DTKX_ERR_INVALID_ARGUMENT );
        auto h = std::make_unique<dtkx_nn>();
        h->n_src = n_src;
        h->n_tgt = n_tgt;
        h->stream = s;
        if ( n_tgt > 0 )
        {
            h->idx.resize( n_tgt, s );
            Staged<double> sx( src_xyz, 3 * n_src, s );
            Staged<double> tx( tgt_xyz, 3 * n_tgt, s );
            dtkx::KnnOutput kout;
            kout.indices = h->idx.get();
            if ( geometry == DTKX_GEOMETRY_FLOAT )
            {
                dtkx::DeviceBuffer<float4> sf( n_src, s ), tf( n_tgt, s );
                dtkx::points_to_float4( sx.ptr, n_src, layout, sf.get(), s );
                dtkx::points_to_float4( tx.ptr, n_tgt, layout, tf.get(), s );
                dtkx::knn_search( geometry, sf.get(), n_src, tf.get(), n_tgt,
                                  1, kout, nullptr, s );
            }
            else
            {
                dtkx::DeviceBuffer<double4> sd( n_src, s ), td( n_tgt, s );
                dtkx::points_to_double4( sx.ptr, n_src, layout, sd.get(), s );
                dtkx::points_to_double4( tx.ptr, n_tgt, layout, td.get(), s );
                dtkx::knn_search( geometry, sd.get(), n_src, td.get(), n_tgt,
                                  1, kout, nullptr, s );
            }
        }
        DTKX_HIP( hipStreamSynchronize( s ) );
        *out = h.release();
        return DTKX_SUCCESS;
    } );
}

int dtkx_nn_apply( const dtkx_nn *h, const double *src_vals, double *tgt_vals,
                   int32_t n_fields, int64_t ld_src, int64_t ld_tgt )
{
    return guarded( [&]() {
        REQ( h != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_fields >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( ld_src >= h->n_src, DTKX_ERR_SIZE_MISMATCH );
        REQ( ld_tgt >= h->n_tgt, DTKX_ERR_SIZE_MISMATCH );
        REQ( src_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->n_tgt == 0 || tgt_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        if ( h->n_tgt == 0 )
            return DTKX_SUCCESS;
        std::size_t const nin = (std::size_t)( n_fields - 1 ) * ld_src + h->n_src;
        std::size_t const nout = (std::size_t)( n_fields - 1 ) * ld_tgt + h->n_tgt;
        Staged<double> x( src_vals, nin, h->stream );
        StagedOut<double> y( tgt_vals, nout, h->stream, ld_tgt > h->n_tgt );
        dtkx::nn_gather_f64( h->idx.get(), h->n_tgt, x.ptr, ld_src, y.ptr,
                             ld_tgt, n_fields, h->stream );
        y.finish( h->stream );
        DTKX_HIP( hipStreamSynchronize( h->stream ) );
        return DTKX_SUCCESS;
    } );
}

int dtkx_nn_get_indices( const dtkx_nn *h, const int32_t **indices,
                         int64_t *n_tgt )
{
    return guarded( [&]() {
        REQ( h != nullptr && indices != nullptr && n_tgt != nullptr,
             DTKX_ERR_INVALID_ARGUMENT );
        *indices = h->idx.get();
        *n_tgt = h->n_tgt;
        return DTKX_SUCCESS;
    } );
}

int dtkx_nn_destroy( dtkx_nn *h )
{
    return guarded( [&]() {
        delete h;
        return DTKX_SUCCESS;
    } );
}

// ------------------------------------------------------- RBF spline
struct dtkx_spline
{
    dtkx_spline_options opts{};
    int64_t n_src = 0, n_tgt = 0;
    int kk = 0;
    hipStream_t stream = nullptr;
    // points [sources | targets]: double copy (geometry DOUBLE) or float
    // copy (FLOAT); the targets' part holds the Vandermonde rows of Q
    dtkx::DeviceBuffer<double4> pts_d;
    dtkx::DeviceBuffer<float4> pts_f;
    // basis rows of M (the sources' rows) then N (the targets'), in source
    // order, as dtkx_spline_get_basis returns them
    dtkx::DeviceBuffer<int32_t> nbr;
    dtkx::DeviceBuffer<double> phi;
    // the system the solver runs on: sources renumbered by Morton position
    dtkx::DeviceBuffer<int32_t> perm; // position -> source
    dtkx::DeviceBuffer<int32_t> m_idx_p, n_idx_p;
    dtkx::DeviceBuffer<double> m_phi_p;
    dtkx::DeviceBuffer<double4> src_d_p;
    dtkx::DeviceBuffer<float4> src_f_p;
    mutable dtkx::SplineSolver solver;
    mutable dtkx_spline_info info{};
    bool float_geometry() const { return opts.geometry == DTKX_GEOMETRY_FLOAT; }
    int32_t const *m_idx() const { return nbr.get(); }
    int32_t const *n_idx() const { return nbr.get() + (size_t)n_src * kk; }
    double const *m_phi() const { return phi.get(); }
    double const *n_phi() const { return phi.get() + (size_t)n_src * kk; }
    dtkx::SplineSystem system() const
    {
        dtkx::SplineSystem y;
        y.n_src = n_src;
        y.n_tgt = n_tgt;
        y.kk = kk;
        y.geometry = opts.geometry;
        if ( float_geometry() )
        {
            y.src_pts = src_f_p.get();
            y.tgt_pts = pts_f.get() + n_src;
        }
        else
        {
            y.src_pts = src_d_p.get();
            y.tgt_pts = pts_d.get() + n_src;
        }
        y.m_idx = m_idx_p.get();
        y.m_phi = m_phi_p.get();
        y.n_idx = n_idx_p.get();
        y.n_phi = n_phi();
        return y;
    }
    std::size_t n_ext() const { return n_src + dtkx::spline_k; }
    dtkx::GmresOptions gmres() const
    {
        dtkx::GmresOptions g;
        g.tolerance = opts.tolerance;
        g.restart = opts.restart;
        g.max_iterations = opts.max_iterations;
        return g;
    }
};

int dtkx_spline_options_init( dtkx_spline_options *opts )
{
    return guarded( [&]() {
        REQ( opts != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        *opts = dtkx_spline_options{};
        opts->rbf = DTKX_RBF_WENDLAND0;
        opts->geometry = DTKX_GEOMETRY_FLOAT;
        opts->layout = DTKX_LAYOUT_AOS;
        opts->restart = 300;
        opts->max_iterations = 1000;
        opts->tolerance = 1e-10;
        opts->stream = nullptr;
        return DTKX_SUCCESS;
    } );
}

int dtkx_spline_create( const double *src_xyz, int64_t n_src,
                        const double *tgt_xyz, int64_t n_tgt,
                        const dtkx_spline_options *opts, dtkx_spline **out )
{
    return guarded( [&]() {
        REQ( out != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        *out = nullptr;
        dtkx_spline_options o;
        dtkx_spline_options_init( &o );
        if ( opts )
            o = *opts;
        REQ( o.rbf >= DTKX_RBF_WENDLAND0 && o.rbf <= DTKX_RBF_BUHMANN4,
             DTKX_ERR_INVALID_ARGUMENT );
        REQ( o.rbf != DTKX_RBF_BUHMANN4, DTKX_ERR_UNSUPPORTED );
        REQ( o.geometry == DTKX_GEOMETRY_FLOAT ||
                 o.geometry == DTKX_GEOMETRY_DOUBLE,
             DTKX_ERR_INVALID_ARGUMENT );
        REQ( o.layout == DTKX_LAYOUT_AOS || o.layout == DTKX_LAYOUT_SOA,
             DTKX_ERR_INVALID_ARGUMENT );
        REQ( o.restart >= 1 && o.restart <= 512, DTKX_ERR_INVALID_ARGUMENT );
        REQ( o.max_iterations >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( o.tolerance >= 0. && std::isfinite( o.tolerance ),
             DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_src > 0, DTKX_ERR_EMPTY_SOURCE );
        REQ( n_tgt >= 0 && src_xyz != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_tgt == 0 || tgt_xyz != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_src <= INT32_MAX - dtkx::spline_k && n_tgt <= INT32_MAX,
             DTKX_ERR_UNSUPPORTED );
        hipStream_t s = static_cast<hipStream_t>( o.stream );
        auto h = std::make_unique<dtkx_spline>();
        h->opts = o;
        h->n_src = n_src;
        h->n_tgt = n_tgt;
        h->stream = s;
        h->kk = (int)std::min<int64_t>( dtkx::spline_k, n_src );
        int const kk = h->kk;
        Staged<double> sx( src_xyz, 3 * n_src, s );
        Staged<double> tx( tgt_xyz, 3 * n_tgt, s );
        int64_t const n_all = n_src + n_tgt;
        h->pts_d.resize( n_all, s );
        dtkx::points_to_double4( sx.ptr, n_src, o.layout, h->pts_d.get(), s );
        if ( n_tgt > 0 )
            dtkx::points_to_double4( tx.ptr, n_tgt, o.layout,
                                     h->pts_d.get() + n_src, s );
        bool const fl = h->float_geometry();
        if ( fl )
        {
            h->pts_f.resize( n_all, s );
            dtkx::double4_to_float4( h->pts_d.get(), n_all, h->pts_f.get(), s );
        }
        void const *pts = fl ? (void const *)h->pts_f.get()
                             : (void const *)h->pts_d.get();
        // One search over the queries [sources | targets]: rows below n_src
        // are M's, the sources' neighbours among the sources
        // (def.hpp:211-212), the rest N's, the targets' (def.hpp:215-216).
        h->nbr.resize( (size_t)n_all * kk, s );
        h->phi.resize( (size_t)n_all * kk, s );
        h->perm.resize( n_src, s );
        dtkx::DeviceBuffer<int32_t> pos( n_src, s );
        dtkx::KnnOutput ko;
        ko.indices = h->nbr.get();
        ko.src_perm = h->perm.get();
        ko.src_pos = pos.get();
        dtkx::knn_search( o.geometry, pts, n_src, pts, n_all, dtkx::spline_k,
                          ko, nullptr, s );
        dtkx::spline_basis_weights( h->pts_d.get(), h->pts_d.get(), n_all, kk,
                                    h->nbr.get(), o.rbf, o.geometry,
                                    h->phi.get(), s );
        // the solver's system: sources in Morton position order
        h->m_idx_p.resize( (size_t)n_src * kk, s );
        h->m_phi_p.resize( (size_t)n_src * kk, s );
        h->n_idx_p.resize( (size_t)n_tgt * kk, s );
        void *src_p;
        if ( fl )
        {
            h->src_f_p.resize( n_src, s );
            src_p = h->src_f_p.get();
        }
        else
        {
            h->src_d_p.resize( n_src, s );
            src_p = h->src_d_p.get();
        }
        dtkx::spline_renumber( o.geometry, n_src, kk, h->perm.get(), pos.get(),
                               h->m_idx(), h->m_phi(), pts, h->m_idx_p.get(),
                               h->m_phi_p.get(), src_p, n_tgt, h->n_idx(),
                               h->n_idx_p.get(), s );
        DTKX_HIP( hipStreamSynchronize( s ) );
        if ( fl )
            h->pts_d.release();
        *out = h.release();
        return DTKX_SUCCESS;
    } );
}

int dtkx_spline_apply( const dtkx_spline *h, const double *src_vals,
                       double *tgt_vals, int32_t n_fields, int64_t ld_src,
                       int64_t ld_tgt )
{
    return guarded( [&]() {
        REQ( h != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_fields >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( ld_src >= h->n_src, DTKX_ERR_SIZE_MISMATCH );
        REQ( ld_tgt >= h->n_tgt, DTKX_ERR_SIZE_MISMATCH );
        REQ( src_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->n_tgt == 0 || tgt_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        h->info = dtkx_spline_info{};
        if ( h->n_tgt == 0 )
            return DTKX_SUCCESS;
        hipStream_t s = h->stream;
        std::size_t const nin = (std::size_t)( n_fields - 1 ) * ld_src + h->n_src;
        std::size_t const nout = (std::size_t)( n_fields - 1 ) * ld_tgt + h->n_tgt;
        Staged<double> x( src_vals, nin, s );
        StagedOut<double> y( tgt_vals, nout, s, ld_tgt > h->n_tgt );
        dtkx::DeviceBuffer<double> z( h->n_ext(), s ), xp( h->n_src, s );
        auto const sys = h->system();
        dtkx_spline_info agg{};
        agg.converged = 1;
        for ( int f = 0; f < n_fields; ++f )
        {
            dtkx::spline_to_positions( x.ptr + f * ld_src, h->perm.get(),
                                       h->n_src, h->n_src, xp.get(), s );
            auto const r = h->solver.solve( sys, xp.get(), z.get(),
                                            h->gmres(), s );
            agg.iterations = std::max( agg.iterations, r.iterations );
            agg.reorthogonalized =
                std::max( agg.reorthogonalized, r.reorthogonalized );
            agg.restarts = std::max( agg.restarts, r.restarts );
            agg.residual = std::max( agg.residual, r.residual );
            agg.converged = agg.converged && r.converged;
            dtkx::spline_evaluate( sys, z.get(), y.ptr + f * ld_tgt, s );
        }
        y.finish( s );
        DTKX_HIP( hipStreamSynchronize( s ) );
        h->info = agg;
        return DTKX_SUCCESS;
    } );
}

int dtkx_spline_solve( const dtkx_spline *h, const double *src_vals,
                       double *coeffs )
{
    return guarded( [&]() {
        REQ( h != nullptr && src_vals != nullptr && coeffs != nullptr,
             DTKX_ERR_INVALID_ARGUMENT );
        hipStream_t s = h->stream;
        std::size_t const nz = h->n_ext();
        Staged<double> x( src_vals, h->n_src, s );
        StagedOut<double> z( coeffs, nz, s );
        dtkx::DeviceBuffer<double> xp( h->n_src, s ), zp( nz, s );
        dtkx::spline_to_positions( x.ptr, h->perm.get(), h->n_src, h->n_src,
                                   xp.get(), s );
        auto const r = h->solver.solve( h->system(), xp.get(), zp.get(),
                                        h->gmres(), s );
        dtkx::spline_to_sources( zp.get(), h->perm.get(), h->n_src, nz, z.ptr,
                                 s );
        z.finish( s );
        DTKX_HIP( hipStreamSynchronize( s ) );
        h->info.iterations = r.iterations;
        h->info.reorthogonalized = r.reorthogonalized;
        h->info.restarts = r.restarts;
        h->info.residual = r.residual;
        h->info.converged = r.converged;
        return DTKX_SUCCESS;
    } );
}

int dtkx_spline_evaluate( const dtkx_spline *h, const double *coeffs,
                          double *tgt_vals )
{
    return guarded( [&]() {
        REQ( h != nullptr && coeffs != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        REQ( h->n_tgt == 0 || tgt_vals != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        if ( h->n_tgt == 0 )
            return DTKX_SUCCESS;
        hipStream_t s = h->stream;
        Staged<double> z( coeffs, h->n_ext(), s );
        StagedOut<double> y( tgt_vals, h->n_tgt, s );
        dtkx::DeviceBuffer<double> zp( h->n_ext(), s );
        dtkx::spline_to_positions( z.ptr, h->perm.get(), h->n_src, h->n_ext(),
                                   zp.get(), s );
        dtkx::spline_evaluate( h->system(), zp.get(), y.ptr, s );
        y.finish( s );
        DTKX_HIP( hipStreamSynchronize( s ) );
        return DTKX_SUCCESS;
    } );
}

int dtkx_spline_system_apply( const dtkx_spline *h, const double *z,
                              double *y )
{
    return guarded( [&]() {
        REQ( h != nullptr && z != nullptr && y != nullptr,
             DTKX_ERR_INVALID_ARGUMENT );
        hipStream_t s = h->stream;
        std::size_t const n = h->n_ext();
        Staged<double> zi( z, n, s );
        StagedOut<double> yo( y, n, s );
        dtkx::DeviceBuffer<double> zp( n, s ), yp( n, s );
        dtkx::spline_to_positions( zi.ptr, h->perm.get(), h->n_src, n,
                                   zp.get(), s );
        dtkx::spline_system_apply( h->system(), zp.get(), yp.get(), s );
        dtkx::spline_to_sources( yp.get(), h->perm.get(), h->n_src, n, yo.ptr,
                                 s );
        yo.finish( s );
        DTKX_HIP( hipStreamSynchronize( s ) );
        return DTKX_SUCCESS;
    } );
}

int dtkx_spline_get_basis( const dtkx_spline *h, int32_t which,
                           const int32_t **indices, const double **values,
                           int64_t *n_rows, int32_t *kk )
{
    return guarded( [&]() {
        REQ( h != nullptr && ( which == 0 || which == 1 ),
             DTKX_ERR_INVALID_ARGUMENT );
        if ( indices )
            *indices = which == 0 ? h->m_idx() : h->n_idx();
        if ( values )
            *values = which == 0 ? h->m_phi() : h->n_phi();
        if ( n_rows )
            *n_rows = which == 0 ? h->n_src : h->n_tgt;
        if ( kk )
            *kk = h->kk;
        return DTKX_SUCCESS;
    } );
}

int dtkx_spline_get_info( const dtkx_spline *h, dtkx_spline_info *info )
{
    return guarded( [&]() {
        REQ( h != nullptr && info != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        *info = h->info;
        return DTKX_SUCCESS;
    } );
}

int dtkx_spline_destroy( dtkx_spline *h )
{
    return guarded( [&]() {
        delete h;
        return DTKX_SUCCESS;
    } );
}

int dtkx_svd_pseudo_inverse( int32_t n, int64_t n_matrices, const double *a,
                             double *pinv, int64_t *n_underdetermined,
                             int64_t *n_unconverged, int32_t max_rotations,
                             void *stream )
{
    return guarded( [&]() {
        hipStream_t s = static_cast<hipStream_t>( stream );
        REQ( n >= 1, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n <= 52, DTKX_ERR_UNSUPPORTED );
        REQ( n_matrices >= 0, DTKX_ERR_INVALID_ARGUMENT );
        REQ( n_matrices == 0 || ( a != nullptr && pinv != nullptr ),
             DTKX_ERR_INVALID_ARGUMENT );
        REQ( max_rotations >= 0, DTKX_ERR_INVALID_ARGUMENT );
        std::size_t const cnt = (std::size_t)n_matrices * n * n;
        Staged<double> sa( a, cnt, s );
        StagedOut<double> sp( pinv, cnt, s );
        dtkx::DeviceBuffer<unsigned long long> counters( 3, s );
        DTKX_HIP( hipMemsetAsync( counters.get(), 0, 3 * sizeof( uint64_t ), s ) );
        long const cap =
            max_rotations > 0 ? max_rotations : default_rotation_cap( n );
        dtkx::svd_pinv_batched( n, n_matrices, sa.ptr, sp.ptr, cap,
                                counters.get(), s );
        sp.finish( s );
        unsigned long long c[3];
        DTKX_HIP( hipMemcpyAsync( c, counters.get(), sizeof( c ),
                                  hipMemcpyDeviceToHost, s ) );
        DTKX_HIP( hipStreamSynchronize( s ) );
        if ( n_underdetermined )
            *n_underdetermined = (int64_t)c[0];
        if ( n_unconverged )
            *n_unconverged = (int64_t)c[1];
        return DTKX_SUCCESS;
    } );
}

} // extern "C"

// ------------------------------------------------------------- batched
struct dtkx_mls_batch
{
    dtkx_mls_options opts{};
    int n_prob = 0;
    hipStream_t stream = nullptr;
    int64_t n_src = 0, n_tgt = 0, nnz = 0;
    std::vector<int64_t> tgt_count; // per problem
    // concatenated problems: offsets, problem ids, Morton layout, CSR rows
    dtkx::DeviceBuffer<int32_t> src_off, tgt_off, src_prob, tgt_prob;
    dtkx::DeviceBuffer<int32_t> order, perm, pos, indices;
    dtkx::DeviceBuffer<int64_t> row_off;
    dtkx::DeviceBuffer<double> coeffs;
    dtkx::DeviceBuffer<float> coeffs32;
    // the build's fixed inputs, kept for rebuilds: per-problem neighbours,
    // row keys and CSR bases, the callers' point arrays, the query launches
    // and the setup's row groups (with the lanes per system chosen once)
    dtkx::DeviceBuffer<int32_t> d_kk, d_key, d_rstart;
    dtkx::DeviceBuffer<int64_t> d_rbase;
    dtkx::DeviceBuffer<double const *> d_sp, d_tp;
    dtkx::KnnBatch kb;
    struct Group
    {
        int m;
        int64_t r0, r1;
        int group_size; // classical solver; 0 until chosen
    };
    std::vector<Group> groups;
    dtkx::DeviceBuffer<float> row_max;
    // [0..3] the setup's counters, [4] the search's overflow count (int32)
    dtkx::DeviceBuffer<unsigned long long> counters;
    // apply: gathered values and the per-problem pointer arrays the
    // captured graph reads
    mutable dtkx::DeviceBuffer<double> xs64;
    mutable dtkx::DeviceBuffer<float> xs32;
    mutable dtkx::DeviceBuffer<double const *> xptr;
    mutable dtkx::DeviceBuffer<double *> yptr;
    mutable hipGraph_t graph = nullptr;
    mutable hipGraphExec_t exec = nullptr;
    // rebuild: search and setup captured on the first rebuild, with the
    // blocks of its temporaries
    hipGraph_t build_graph = nullptr;
    hipGraphExec_t build_exec = nullptr;
    dtkx::GraphBlocks build_blocks;
    dtkx_mls_stats stats{};

    dtkx::BatchLayout layout() const
    {
        dtkx::BatchLayout g;
        g.idx = indices.get();
        g.row_off = row_off.get();
        g.order = order.get();
        g.perm = perm.get();
        g.src_off = src_off.get();
        g.tgt_off = tgt_off.get();
        g.src_prob = src_prob.get();
        g.tgt_prob = tgt_prob.get();
        g.n_src = n_src;
        g.n_tgt = n_tgt;
        return g;
    }
    void launch_apply() const
    {
        if ( opts.precision == DTKX_PRECISION_FP32_BULK )
            dtkx::batch_apply_f32( layout(), coeffs32.get(), xptr.get(),
                                   yptr.get(), xs32.get(), stream );
        else
            dtkx::batch_apply_f64( layout(), coeffs.get(), xptr.get(),
                                   yptr.get(), xs64.get(), stream );
    }
    ~dtkx_mls_batch()
    {
        // the kept blocks return to the cache of this stream: its later
        // users are ordered after the graphs' launches
        if ( exec )
            (void)hipGraphExecDestroy( exec );
        if ( graph )
            (void)hipGraphDestroy( graph );
        if ( build_exec )
            (void)hipGraphExecDestroy( build_exec );
        if ( build_graph )
            (void)hipGraphDestroy( build_graph );
    }
};

namespace
{

// the K of the query kernel that holds up to k results
int query_bucket( int k )
{
    int const ks[] = {1, 4, 8, 10, 16, 24, 32, 48, 64};
    for ( int v : ks )
        if ( k <= v )
            return v;
    return 64;
}

// Lanes per system of the classical solver for rows [r0, r1): the fastest
// of the candidates on a sample of the rows (every candidate gives the same
// bits, so the sample's outputs are simply overwritten afterwards).
int autotune_group( dtkx_mls_batch *b, dtkx::SetupParams p, double4 const *src,
                    double4 const *tgt, int64_t r0, int64_t r1 )
{
    std::vector<int> cand;
    if ( p.m == 1 )
        cand = {1};
    else if ( p.m == 4 )
        cand = {1, 2, 4};
    else
        cand = {2, 4, 8, 16};
    if ( cand.size() == 1 )
        return cand[0];
    int64_t const end = std::min<int64_t>( r1, r0 + 65536 );
    dtkx::DeviceBuffer<unsigned long long> scratch( 3, b->stream );
    p.row0 = r0;
    int best = cand[0];
    double best_ms = 1e300;
    for ( int g : cand )
    {
        p.group_size = g;
        StageEvents ev( b->stream );
        dtkx::setup_classical( src, tgt, end, b->indices.get(),
                               b->order.get(), p, b->coeffs.get(),
                               scratch.get(), b->stream ); // warm
        ev.mark( 0 );
        dtkx::setup_classical( src, tgt, end, b->indices.get(),
                               b->order.get(), p, b->coeffs.get(),
                               scratch.get(), b->stream );
        ev.mark( 1 );
        double const ms = ev.ms( 0, 1 );
        if ( ms < best_ms )
        {
            best_ms = ms;
            best = g;
        }
    }
    return best;
}

// Search and setup of every problem from the callers' point arrays, queued
// on the batch's stream: the create's build (with stage events, and the
// classical solver's lanes per system chosen by timing where not given)
// and the body of the rebuild graph (ev null: no events, no timing, no host
// synchronisation, the lanes chosen at create). Writes the batch's CSR
// rows, order, permutations and coefficients in place, so the captured
// apply reads the new operator.
void enqueue_build( dtkx_mls_batch *b, StageEvents *ev )
{
    hipStream_t const s = b->stream;
    int const P = b->n_prob;
    int64_t const ns = b->n_src, nt = b->n_tgt;
    dtkx_mls_options const &o = b->opts;
    dtkx::zero_async( b->counters.get(), 5 * sizeof( uint64_t ), s );
    dtkx::DeviceBuffer<double4> src_d( ns, s ),
        tgt_d( std::max<int64_t>( nt, 1 ), s );
    dtkx::concat_points( b->d_sp.get(), b->src_off.get(), b->src_prob.get(),
                         ns, src_d.get(), s );
    dtkx::concat_points( b->d_tp.get(), b->tgt_off.get(), b->tgt_prob.get(),
                         nt, tgt_d.get(), s );
    if ( ev )
        ev->mark( 0 );
    dtkx::KnnBatchOutput ko;
    ko.order = b->order.get();
    ko.src_perm = b->perm.get();
    ko.src_pos = b->pos.get();
    ko.row_off = b->row_off.get();
    ko.indices = b->indices.get();
    ko.row_max = b->row_max.get(); // null unless the FP32-bulk radius
    dtkx::KnnInfo info;
    info.n_overflow_dev = reinterpret_cast<int32_t *>( b->counters.get() + 4 );
    if ( ev && nt > 0 )
        ev->pass_query_events( info );
    if ( o.geometry == DTKX_GEOMETRY_FLOAT )
    {
        dtkx::DeviceBuffer<float4> sf( ns, s ), tf( std::max<int64_t>( nt, 1 ), s );
        dtkx::double4_to_float4( src_d.get(), ns, sf.get(), s );
        dtkx::double4_to_float4( tgt_d.get(), nt, tf.get(), s );
        dtkx::knn_search_batched( o.geometry, sf.get(), ns, tf.get(), nt, b->kb,
                                  ko, &info, s );
    }
    else
        dtkx::knn_search_batched( o.geometry, src_d.get(), ns, tgt_d.get(),
                                  nt, b->kb, ko, &info, s );
    {
        dtkx::DeviceBuffer<double4> sorted( ns, s );
        dtkx::gather_points( src_d.get(), b->perm.get(), ns, sorted.get(), s );
        src_d = std::move( sorted );
    }
    if ( ev )
        ev->mark( 1 );

    // setup: one launch per basis over its rows
    bool const f32 = o.precision == DTKX_PRECISION_FP32_BULK;
    bool const cyclic = f32 || o.solver != DTKX_SOLVER_CLASSICAL_JACOBI;
    int group_used = 0;
    for ( auto &g : b->groups )
    {
        dtkx::SetupParams p;
        p.m = g.m;
        p.kk = 0;
        p.rbf = o.rbf;
        p.geometry = o.geometry;
        p.truncation = o.truncation;
        p.rcond = o.rcond;
        p.row0 = g.r0;
        p.row_off = b->row_off.get();
        if ( cyclic )
        {
            p.max_rot = o.max_rotations;
            p.row_max = b->row_max.get();
            auto *setup = o.solver == DTKX_SOLVER_CHOLESKY ? &dtkx::setup_fast
                                                            : &dtkx::setup_cyclic;
            setup( src_d.get(), tgt_d.get(), g.r1, b->indices.get(),
                   b->order.get(), p, f32 ? nullptr : b->coeffs.get(),
                   f32 ? b->coeffs32.get() : nullptr, nullptr,
                   b->counters.get(), s );
            group_used = g.m == 10 ? 2 : 1;
        }
        else
        {
            p.max_rot = o.max_rotations > 0 ? o.max_rotations
                                            : default_rotation_cap( g.m );
            if ( g.group_size == 0 )
            {
                DTKX_REQUIRE( ev != nullptr, DTKX_ERR_INVALID_ARGUMENT );
                g.group_size = o.group_size > 0
                                   ? o.group_size
                                   : autotune_group( b, p, src_d.get(),
                                                     tgt_d.get(), g.r0, g.r1 );
            }
            p.group_size = g.group_size;
            group_used = dtkx::setup_classical(
                src_d.get(), tgt_d.get(), g.r1, b->indices.get(),
                b->order.get(), p, b->coeffs.get(), b->counters.get(), s );
        }
    }
    if ( ev )
        ev->mark( 2 );
    b->stats.group_size_used = group_used;
}

// The counters of the last build into the statistics (synchronises).
void read_build_counters( dtkx_mls_batch *b )
{
    unsigned long long c[5] = {0, 0, 0, 0, 0};
    DTKX_HIP( hipMemcpyAsync( c, b->counters.get(), sizeof( c ),
                              hipMemcpyDeviceToHost, b->stream ) );
    DTKX_HIP( hipStreamSynchronize( b->stream ) );
    b->stats.n_underdetermined = (int64_t)c[0];
    b->stats.n_unconverged = (int64_t)c[1];
    b->stats.total_rotations = (int64_t)c[2];
    b->stats.n_eigen_fallback = (int64_t)c[3];
    int32_t over = 0;
    std::memcpy( &over, &c[4], sizeof( over ) );
    b->stats.n_stack_overflow = over;
}

} // namespace

extern "C" {

int dtkx_mls_batch_create( int32_t n_problems, const dtkx_mls_problem *problems,
                           const dtkx_mls_options *opts, dtkx_mls_batch **out )
{
    return guarded( [&]() {
        REQ( out != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        *out = nullptr;
        REQ( n_problems >= 1 && problems != nullptr && opts != nullptr,
             DTKX_ERR_INVALID_ARGUMENT );
        validate_options( *opts );
        REQ( !opts->store_pinv, DTKX_ERR_UNSUPPORTED );
        int const P = n_problems;
        auto b = std::make_unique<dtkx_mls_batch>();
        b->opts = *opts;
        b->n_prob = P;
        b->stream = static_cast<hipStream_t>( opts->stream );
        if ( !b->stream )
        {
            // graph capture needs a stream other than the null stream; the
            // null stream's earlier work is waited for first
            DTKX_HIP( hipDeviceSynchronize() );
            b->stream = dtkx::library_stream();
        }
        hipStream_t const s = b->stream;

        // host bookkeeping: sizes, basis, neighbours, row order
        std::vector<int32_t> so( P + 1, 0 ), to( P + 1, 0 ), m( P ), kk( P ),
            key( P ), rstart( P );
        std::vector<int64_t> rbase( P );
        std::vector<double const *> sp( P ), tp( P );
        int64_t ns = 0, nt = 0;
        for ( int p = 0; p < P; ++p )
        {
            dtkx_mls_problem const &q = problems[p];
            REQ( q.n_src >= 1, DTKX_ERR_EMPTY_SOURCE );
            REQ( q.n_tgt >= 0, DTKX_ERR_INVALID_ARGUMENT );
            REQ( q.basis == DTKX_BASIS_CONSTANT || q.basis == DTKX_BASIS_LINEAR ||
                     q.basis == DTKX_BASIS_QUADRATIC,
                 DTKX_ERR_INVALID_ARGUMENT );
            REQ( q.k >= 0 && q.k <= dtkx::knn_max_k, DTKX_ERR_UNSUPPORTED );
            REQ( q.src_xyz != nullptr && dtkx::is_device_pointer( q.src_xyz ),
                 DTKX_ERR_INVALID_ARGUMENT );
            REQ( q.n_tgt == 0 || ( q.tgt_xyz != nullptr &&
                                   dtkx::is_device_pointer( q.tgt_xyz ) ),
                 DTKX_ERR_INVALID_ARGUMENT );
            m[p] = q.basis;
            int const k = q.k > 0 ? q.k : q.basis;
            kk[p] = static_cast<int32_t>( std::min<int64_t>( k, q.n_src ) );
            ns += q.n_src;
            nt += q.n_tgt;
            REQ( ns < INT32_MAX && nt < INT32_MAX, DTKX_ERR_UNSUPPORTED );
            so[p + 1] = static_cast<int32_t>( ns );
            to[p + 1] = static_cast<int32_t>( nt );
            sp[p] = q.src_xyz;
            tp[p] = q.tgt_xyz;
        }
        b->n_src = ns;
        b->n_tgt = nt;
        b->tgt_count.resize( P );
        for ( int p = 0; p < P; ++p )
            b->tgt_count[p] = problems[p].n_tgt;
        // rows grouped by (basis, query kernel, problem): one setup launch
        // per basis, one query launch per kernel
        std::vector<int> rank( P );
        for ( int p = 0; p < P; ++p )
            rank[p] = p;
        std::stable_sort( rank.begin(), rank.end(), [&]( int a, int c ) {
            if ( m[a] != m[c] )
                return m[a] < m[c];
            return query_bucket( kk[a] ) < query_bucket( kk[c] );
        } );
        dtkx::KnnBatch &kb = b->kb;
        auto &groups = b->groups;
        int64_t row = 0, nnz = 0;
        for ( int i = 0; i < P; ++i )
        {
            int const p = rank[i];
            key[p] = i;
            rstart[p] = static_cast<int32_t>( row );
            rbase[p] = nnz;
            int64_t const n = problems[p].n_tgt;
            int const kb_p = query_bucket( kk[p] );
            if ( n > 0 )
            {
                if ( !kb.launches.empty() && kb.launches.back().end == row &&
                     query_bucket( kb.launches.back().kmax ) == kb_p )
                {
                    kb.launches.back().end = row + n;
                    kb.launches.back().kmax =
                        std::max( kb.launches.back().kmax, (int)kk[p] );
                }
                else
                    kb.launches.push_back( {row, row + n, kk[p]} );
                if ( !groups.empty() && groups.back().m == m[p] &&
                     groups.back().r1 == row )
                    groups.back().r1 = row + n;
                else
                    groups.push_back( {m[p], row, row + n, 0} );
            }
            row += n;
            nnz += n * kk[p];
        }
        b->nnz = nnz;

        // device copies of the bookkeeping
        auto upload = [&]( auto &buf, auto const &vec ) {
            using T = typename std::decay_t<decltype( vec )>::value_type;
            buf.resize( vec.size(), s );
            DTKX_HIP( hipMemcpyAsync( buf.get(), vec.data(),
                                      vec.size() * sizeof( T ),
                                      hipMemcpyHostToDevice, s ) );
        };
        upload( b->src_off, so );
        upload( b->tgt_off, to );
        upload( b->d_kk, kk );
        upload( b->d_key, key );
        upload( b->d_rstart, rstart );
        upload( b->d_rbase, rbase );
        upload( b->d_sp, sp );
        upload( b->d_tp, tp );
        b->src_prob.resize( ns, s );
        b->tgt_prob.resize( std::max<int64_t>( nt, 1 ), s );
        dtkx::problem_ids( b->src_off.get(), P, ns, b->src_prob.get(), s );
        dtkx::problem_ids( b->tgt_off.get(), P, nt, b->tgt_prob.get(), s );
        // the host vectors above must outlive the uploads
        DTKX_HIP( hipStreamSynchronize( s ) );

        b->counters.resize( 5, s );
        b->order.resize( std::max<int64_t>( nt, 1 ), s );
        b->perm.resize( ns, s );
        b->pos.resize( ns, s );
        b->row_off.resize( nt + 1, s );
        b->indices.resize( std::max<int64_t>( nnz, 1 ), s );
        if ( opts->precision == DTKX_PRECISION_FP32_BULK )
            b->coeffs32.resize( std::max<int64_t>( nnz, 1 ), s );
        else
            b->coeffs.resize( std::max<int64_t>( nnz, 1 ), s );
        // the FP32-bulk setup may take its radius from the search
        if ( DTKX_FP32_KNN_RADIUS &&
             opts->precision == DTKX_PRECISION_FP32_BULK && nt > 0 )
            b->row_max.resize( nt, s );
        kb.n_problems = P;
        kb.src_off = b->src_off.get();
        kb.src_prob = b->src_prob.get();
        kb.tgt_prob = b->tgt_prob.get();
        kb.row_key = b->d_key.get();
        kb.key_bits = 1;
        while ( ( int64_t( 1 ) << kb.key_bits ) < P )
            ++kb.key_bits;
        kb.kk_prob = b->d_kk.get();
        kb.row_base = b->d_rbase.get();
        kb.row_start = b->d_rstart.get();
        kb.nnz = nnz;

        StageEvents ev( s );
        enqueue_build( b.get(), &ev );
        read_build_counters( b.get() );
        b->stats.n_source = ns;
        b->stats.n_target = nt;
        b->stats.nnz = nnz;
        b->stats.basis_size = *std::max_element( m.begin(), m.end() );
        b->stats.k = *std::max_element( kk.begin(), kk.end() );
        b->stats.ms_knn = ev.ms( 0, 1 );
        b->stats.ms_setup = ev.ms( 1, 2 );
        b->stats.ms_knn_query = nt > 0 ? ev.ms( 3, 4 ) : 0.0;

        // apply resources, fixed for the captured graph
        if ( opts->precision == DTKX_PRECISION_FP32_BULK )
            b->xs32.resize( ns, s );
        else
            b->xs64.resize( ns, s );
        b->xptr.resize( P, s );
        b->yptr.resize( P, s );
        *out = b.release();
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_batch_rebuild( dtkx_mls_batch *b )
{
    return guarded( [&]() {
        REQ( b != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        hipStream_t const s = b->stream;
        if ( !b->build_exec )
        {
            // relaxed mode: the cache may allocate while capturing; every
            // block the captured work releases stays with the batch
            dtkx::KeepBlocks keep( &b->build_blocks );
            DTKX_HIP( hipStreamBeginCapture( s, hipStreamCaptureModeRelaxed ) );
            try
            {
                enqueue_build( b, nullptr );
            }
            catch ( ... )
            {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture( s, &g );
                if ( g )
                    (void)hipGraphDestroy( g );
                throw;
            }
            DTKX_HIP( hipStreamEndCapture( s, &b->build_graph ) );
            DTKX_HIP( hipGraphInstantiate( &b->build_exec, b->build_graph,
                                           nullptr, nullptr, 0 ) );
        }
        StageEvents ev( s );
        ev.mark( 0 );
        DTKX_HIP( hipGraphLaunch( b->build_exec, s ) );
        ev.mark( 1 );
        read_build_counters( b );
        // the stage times of a replay are not separated: the whole graph
        // is reported as the search, and the setup as 0
        b->stats.ms_knn = ev.ms( 0, 1 );
        b->stats.ms_setup = 0.0;
        b->stats.ms_knn_query = 0.0;
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_batch_apply( const dtkx_mls_batch *b,
                          const double *const *src_vals,
                          double *const *tgt_vals )
{
    return guarded( [&]() {
        REQ( b != nullptr && src_vals != nullptr && tgt_vals != nullptr,
             DTKX_ERR_INVALID_ARGUMENT );
        for ( int p = 0; p < b->n_prob; ++p )
            REQ( src_vals[p] != nullptr &&
                     ( tgt_vals[p] != nullptr || b->tgt_count[p] == 0 ),
                 DTKX_ERR_INVALID_ARGUMENT );
        // host arrays would fault the gather: the first problem's pair is
        // checked (a runtime query per pointer, too slow for all of them)
        REQ( dtkx::is_device_pointer( src_vals[0] ), DTKX_ERR_INVALID_ARGUMENT );
        REQ( b->tgt_count[0] == 0 || dtkx::is_device_pointer( tgt_vals[0] ),
             DTKX_ERR_INVALID_ARGUMENT );
        hipStream_t const s = b->stream;
        DTKX_HIP( hipMemcpyAsync( b->xptr.get(), src_vals,
                                  sizeof( double * ) * b->n_prob,
                                  hipMemcpyHostToDevice, s ) );
        DTKX_HIP( hipMemcpyAsync( b->yptr.get(), tgt_vals,
                                  sizeof( double * ) * b->n_prob,
                                  hipMemcpyHostToDevice, s ) );
        if ( !b->exec )
        {
            // gather + row sums of every problem, captured once and
            // replayed by each apply
            DTKX_HIP( hipStreamBeginCapture( s, hipStreamCaptureModeThreadLocal ) );
            try
            {
                b->launch_apply();
            }
            catch ( ... )
            {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture( s, &g );
                if ( g )
                    (void)hipGraphDestroy( g );
                throw;
            }
            DTKX_HIP( hipStreamEndCapture( s, &b->graph ) );
            DTKX_HIP( hipGraphInstantiate( &b->exec, b->graph, nullptr,
                                           nullptr, 0 ) );
        }
        DTKX_HIP( hipGraphLaunch( b->exec, s ) );
        DTKX_HIP( hipStreamSynchronize( s ) );
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_batch_get_stats( const dtkx_mls_batch *b, dtkx_mls_stats *stats )
{
    return guarded( [&]() {
        REQ( b != nullptr && stats != nullptr, DTKX_ERR_INVALID_ARGUMENT );
        *stats = b->stats;
        return DTKX_SUCCESS;
    } );
}

int dtkx_mls_batch_destroy( dtkx_mls_batch *b )
{
    return guarded( [&]() {
        delete b;
        return DTKX_SUCCESS;
    } );
}

} // extern "C"

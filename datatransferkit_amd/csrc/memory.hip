// Device block cache behind DeviceBuffer (see common.hpp).
//
// Free lists per (device, stream). A block freed on stream s is reused only
// by a later allocation on s: work on s that uses the new owner is queued
// after all work that used the old owner, so the reuse needs no event.
// Sizes are rounded (powers of two up to 1 MiB, 2 MiB multiples above) so
// the same sequence of allocations, as when operators of the same shape are
// built again, finds exact fits.
#include "common.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <unordered_map>
#include <utility>

namespace dtkx
{
namespace
{

struct BlockCache
{
    std::mutex mu;
    std::map<std::pair<int, hipStream_t>, std::multimap<std::size_t, void *>>
        free_lists;
    std::unordered_map<void *, std::pair<int, std::size_t>> live; // size, device
};

// Never destroyed: blocks must not be freed after the HIP runtime has shut
// down at process exit.
BlockCache &cache()
{
    static BlockCache *c = new BlockCache;
    return *c;
}

std::size_t round_size( std::size_t b )
{
    constexpr std::size_t kMiB = std::size_t( 1 ) << 20;
    if ( b > kMiB )
        return ( b + 2 * kMiB - 1 ) / ( 2 * kMiB ) * ( 2 * kMiB );
    std::size_t r = 512;
    while ( r < b )
        r <<= 1;
    return r;
}

// Frees every cached block of `dev` (caller holds the lock); pending work
// may still use them, so the device is synchronised first.
void release_locked( BlockCache &c, int dev )
{
    bool synced = false;
    for ( auto it = c.free_lists.begin(); it != c.free_lists.end(); )
    {
        if ( it->first.first != dev )
        {
            ++it;
            continue;
        }
        if ( !synced )
        {
            (void)hipDeviceSynchronize();
            synced = true;
        }
        for ( auto &b : it->second )
            (void)hipFree( b.second );
        it = c.free_lists.erase( it );
    }
}

} // namespace

void *cache_alloc( std::size_t bytes, hipStream_t s )
{
    std::size_t const r = round_size( bytes );
    int dev = 0;
    DTKX_HIP( hipGetDevice( &dev ) );
    BlockCache &c = cache();
    std::lock_guard<std::mutex> lock( c.mu );
    auto fl = c.free_lists.find( {dev, s} );
    if ( fl != c.free_lists.end() )
    {
        auto b = fl->second.lower_bound( r );
        // accept up to 1/8 slack
        if ( b != fl->second.end() && b->first <= r + r / 8 )
        {
            void *p = b->second;
            c.live[p] = {dev, b->first};
            fl->second.erase( b );
            return p;
        }
    }
    void *p = nullptr;
    hipError_t e = hipMalloc( &p, r );
    if ( e == hipErrorOutOfMemory )
    {
        (void)hipGetLastError();
        release_locked( c, dev );
        e = hipMalloc( &p, r );
    }
    if ( e != hipSuccess || !p )
    {
        (void)hipGetLastError();
        raise( DTKX_ERR_OUT_OF_MEMORY, "device allocation", __FILE__,
               __LINE__ );
    }
    c.live[p] = {dev, r};
    return p;
}

namespace
{
thread_local GraphBlocks *t_keep = nullptr;
}

void cache_free( void *p, std::size_t bytes, hipStream_t s )
{
    if ( !p )
        return;
    if ( t_keep )
    {
        t_keep->blocks.push_back( {p, bytes, s} );
        return;
    }
    BlockCache &c = cache();
    std::lock_guard<std::mutex> lock( c.mu );
    auto it = c.live.find( p );
    if ( it == c.live.end() )
        return;
    auto const [dev, size] = it->second;
    c.live.erase( it );
    c.free_lists[{dev, s}].emplace( size, p );
}

hipStream_t library_stream()
{
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    int dev = 0;
    DTKX_HIP( hipGetDevice( &dev ) );
    std::lock_guard<std::mutex> lock( mu );
    auto it = streams.find( dev );
    if ( it != streams.end() )
        return it->second;
    hipStream_t st = nullptr;
    DTKX_HIP( hipStreamCreate( &st ) );
    streams[dev] = st;
    return st;
}

hipStream_t side_stream()
{
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    int dev = 0;
    DTKX_HIP( hipGetDevice( &dev ) );
    std::lock_guard<std::mutex> lock( mu );
    auto it = streams.find( dev );
    if ( it != streams.end() )
        return it->second;
    hipStream_t st = nullptr;
    DTKX_HIP( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) );
    streams[dev] = st;
    return st;
}

namespace
{
// the forked work's own stream, apart from side_stream(), whose statistics
// reads wait for a create's last event only
hipStream_t fork_stream()
{
    static std::mutex mu;
    static std::map<int, hipStream_t> streams;
    int dev = 0;
    DTKX_HIP( hipGetDevice( &dev ) );
    std::lock_guard<std::mutex> lock( mu );
    auto it = streams.find( dev );
    if ( it != streams.end() )
        return it->second;
    hipStream_t st = nullptr;
    DTKX_HIP( hipStreamCreateWithFlags( &st, hipStreamNonBlocking ) );
    streams[dev] = st;
    return st;
}
} // namespace

Fork::Fork( hipStream_t s )
    : main_( s )
    , side_( fork_stream() )
{
    DTKX_HIP( hipEventCreateWithFlags( &fork_, hipEventDisableTiming ) );
    DTKX_HIP( hipEventCreateWithFlags( &join_, hipEventDisableTiming ) );
    DTKX_HIP( hipEventRecord( fork_, main_ ) );
    DTKX_HIP( hipStreamWaitEvent( side_, fork_, 0 ) );
}

void Fork::join()
{
    if ( joined_ )
        return;
    joined_ = true;
    DTKX_HIP( hipEventRecord( join_, side_ ) );
    DTKX_HIP( hipStreamWaitEvent( main_, join_, 0 ) );
}

Fork::~Fork()
{
    // a failed join leaves nothing better to do than wait for the side work
    if ( !joined_ && hipEventRecord( join_, side_ ) == hipSuccess )
        (void)hipStreamWaitEvent( main_, join_, 0 );
    (void)hipEventDestroy( fork_ );
    (void)hipEventDestroy( join_ );
}

void GraphBlocks::release()
{
    for ( auto const &b : blocks )
        cache_free( b.p, b.bytes, b.s );
    blocks.clear();
}

KeepBlocks::KeepBlocks( GraphBlocks *keep )
    : prev_( t_keep )
{
    t_keep = keep;
}

KeepBlocks::~KeepBlocks() { t_keep = prev_; }

namespace
{
__global__ void zero_kernel( uint32_t *p, std::size_t n )
{
    for ( std::size_t i = blockIdx.x * (std::size_t)blockDim.x + threadIdx.x;
          i < n; i += (std::size_t)gridDim.x * blockDim.x )
        p[i] = 0u;
}
} // namespace

void zero_async( void *p, std::size_t bytes, hipStream_t s )
{
    DTKX_REQUIRE( bytes % 4 == 0, DTKX_ERR_INVALID_ARGUMENT );
    std::size_t const n = bytes / 4;
    if ( n == 0 )
        return;
    std::size_t const blocks = std::min<std::size_t>( ( n + 255 ) / 256, 1024 );
    zero_kernel<<<static_cast<unsigned>( blocks ), 256, 0, s>>>(
        static_cast<uint32_t *>( p ), n );
    DTKX_HIP( hipGetLastError() );
}

void cache_release()
{
    int dev = 0;
    DTKX_HIP( hipGetDevice( &dev ) );
    BlockCache &c = cache();
    std::lock_guard<std::mutex> lock( c.mu );
    release_locked( c, dev );
}

} // namespace dtkx

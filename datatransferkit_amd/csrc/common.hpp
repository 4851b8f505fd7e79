// Shared host/device utilities of the dtkx HIP library: error reporting in the
// reference's DBC message shape, HIP call checking, device buffers, and the
// point formats used inside the library.
#ifndef DTKX_COMMON_HPP
#define DTKX_COMMON_HPP

#include "dtkx/mls.h"

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace dtkx
{

// Error raised inside the library; translated to a dtkx_status at the C ABI.
// The message follows the reference's DataTransferKitException output
// ("<cond>, failed in <file>, line <n>.", packages/Utils/src/DTK_DBC.cpp:40-48).
struct Error : std::runtime_error
{
    int status;
    Error( int s, std::string const &msg )
        : std::runtime_error( msg )
        , status( s )
    {
    }
};

[[noreturn]] void raise( int status, char const *cond, char const *file,
                         int line );

#define DTKX_REQUIRE( c, status )                                              \
    do                                                                         \
    {                                                                          \
        if ( !( c ) )                                                          \
            ::dtkx::raise( status, #c, __FILE__, __LINE__ );                   \
    } while ( 0 )

#define DTKX_HIP( call )                                                       \
    do                                                                         \
    {                                                                          \
        hipError_t e_ = ( call );                                              \
        if ( e_ != hipSuccess )                                                \
            ::dtkx::raise( DTKX_ERR_HIP, hipGetErrorString( e_ ), __FILE__,    \
                           __LINE__ );                                         \
    } while ( 0 )

// Device block cache (memory.hip). Blocks freed on a stream return to that
// stream's free list and are handed out again only to allocations on the
// same stream, whose later work is ordered after every earlier use, so no
// event or synchronisation is needed; a miss allocates with hipMalloc. The
// cache keeps blocks until dtkx_release_cached_memory() (or process end):
// building an operator then costs no allocator calls once warm, where the
// runtime's own stream-ordered allocator spends milliseconds per operator
// in hipFreeAsync at a million targets.
void *cache_alloc( std::size_t bytes, hipStream_t s );
void cache_free( void *p, std::size_t bytes, hipStream_t s );

// Blocks a captured graph uses. While a KeepBlocks installs a GraphBlocks
// on the current thread (around a stream capture), every block released to
// the cache is kept in it instead: the graph's nodes use those addresses on
// each replay, so they must not be handed out again. release() (or the
// destructor) returns them to the free lists of their streams; the graph's
// launches must all be queued on those streams, so later users of a block
// are ordered after them.
struct GraphBlocks
{
    struct Block
    {
        void *p;
        std::size_t bytes;
        hipStream_t s;
    };
    std::vector<Block> blocks;
    void release();
    GraphBlocks() = default;
    GraphBlocks( GraphBlocks const & ) = delete;
    GraphBlocks &operator=( GraphBlocks const & ) = delete;
    ~GraphBlocks() { release(); }
};

class KeepBlocks
{
  public:
    explicit KeepBlocks( GraphBlocks *keep );
    ~KeepBlocks();
    KeepBlocks( KeepBlocks const & ) = delete;
    KeepBlocks &operator=( KeepBlocks const & ) = delete;

  private:
    GraphBlocks *prev_;
};

// p[0, bytes) = 0 on stream s (bytes a multiple of 4) by a kernel, which,
// unlike a memset node, replays correctly from a captured graph.
void zero_async( void *p, std::size_t bytes, hipStream_t s );
// Returns the cached blocks of the current device to the runtime.
void cache_release();
// A stream of the library's own on the current device (default priority),
// created on first use and kept: for work that cannot run on the null
// stream, such as graph capture.
hipStream_t library_stream();
// A second stream of the library's own on the current device (default
// priority, non-blocking: it does not wait for the null stream), created
// on first use and kept: forked work, and reads that must wait for one
// event only, not for later work queued on a caller's stream.
hipStream_t side_stream();

// Work forked from stream s onto a stream of the library's own (default
// priority, non-blocking, one per device, not side_stream()): it first
// waits for everything queued on s so far, and join() (or the destructor)
// makes s wait for everything queued on the forked stream since. Buffers
// the forked work uses are allocated on s before the fork and released on
// s after the join, so the block cache's stream order holds for them.
class Fork
{
  public:
    explicit Fork( hipStream_t s );
    ~Fork();
    Fork( Fork const & ) = delete;
    Fork &operator=( Fork const & ) = delete;
    hipStream_t side() const { return side_; }
    void join();

  private:
    hipStream_t main_, side_;
    hipEvent_t fork_ = nullptr, join_ = nullptr;
    bool joined_ = false;
};

// Owning device buffer, move-only. With a stream it draws from the block
// cache of that stream; without one it uses hipMalloc / hipFree.
template <typename T>
struct DeviceBuffer
{
    T *ptr = nullptr;
    std::size_t n = 0;
    hipStream_t stream = nullptr;
    bool cached = false;
    DeviceBuffer() = default;
    explicit DeviceBuffer( std::size_t count ) { resize( count ); }
    DeviceBuffer( std::size_t count, hipStream_t s ) { resize( count, s ); }
    DeviceBuffer( DeviceBuffer const & ) = delete;
    DeviceBuffer &operator=( DeviceBuffer const & ) = delete;
    DeviceBuffer( DeviceBuffer &&o ) noexcept { steal( o ); }
    DeviceBuffer &operator=( DeviceBuffer &&o ) noexcept
    {
        if ( this != &o )
        {
            release();
            steal( o );
        }
        return *this;
    }
    ~DeviceBuffer() { release(); }
    void resize( std::size_t count )
    {
        release();
        if ( count > 0 )
        {
            hipError_t e = hipMalloc( &ptr, count * sizeof( T ) );
            if ( e != hipSuccess )
            {
                ptr = nullptr;
                ::dtkx::raise( DTKX_ERR_OUT_OF_MEMORY, "hipMalloc", __FILE__,
                               __LINE__ );
            }
        }
        n = count;
        cached = false;
    }
    void resize( std::size_t count, hipStream_t s )
    {
        release();
        if ( count > 0 )
            ptr = static_cast<T *>( cache_alloc( count * sizeof( T ), s ) );
        n = count;
        stream = s;
        cached = true;
    }
    void release()
    {
        if ( ptr )
        {
            if ( cached )
                cache_free( ptr, n * sizeof( T ), stream );
            else
                (void)hipFree( ptr );
        }
        ptr = nullptr;
        n = 0;
    }
    T *get() const { return ptr; }

  private:
    void steal( DeviceBuffer &o )
    {
        ptr = o.ptr;
        n = o.n;
        stream = o.stream;
        cached = o.cached;
        o.ptr = nullptr;
        o.n = 0;
    }
};

// True if p is device-accessible memory (hipMalloc, managed, or registered).
bool is_device_pointer( void const *p );

// Internal point formats: padded to 16/32 bytes for single-load gathers.
// float4 for float geometry (kNN), double4 for double geometry and setup.
void points_to_float4( double const *xyz, int64_t n, int layout, float4 *out,
                       hipStream_t s );
// outf (nullable): float copies in the same pass (the values
// double4_to_float4 gives)
void points_to_double4( double const *xyz, int64_t n, int layout,
                        double4 *out, hipStream_t s, float4 *outf = nullptr );

void double4_to_float4( double4 const *in, int64_t n, float4 *out,
                        hipStream_t s );

// out[i] = in[perm[i]]
void gather_points( double4 const *in, int32_t const *perm, int64_t n,
                    double4 *out, hipStream_t s );

// Batches: prob[i] = the problem p with off[p] <= i < off[p + 1]
void problem_ids( int32_t const *off, int n_prob, int64_t n, int32_t *prob,
                  hipStream_t s );
// out[i] = point i - off[p] of problem p = prob[i], read from the AoS
// device array ptrs[p]
void concat_points( double const *const *ptrs, int32_t const *off,
                    int32_t const *prob, int64_t n, double4 *out,
                    hipStream_t s );

// inv[perm[i]] = i
void invert_permutation( int32_t const *perm, int64_t n, int32_t *inv,
                         hipStream_t s );

// Work tile of the current block under the MI355X dispatch order, which
// hands consecutive workgroups to the 8 XCDs in turn (each XCD with its own
// L2). A bijection of [0, gridDim.x) in one of two shapes:
//  * kChunk = 0: each XCD takes one contiguous eighth of the tiles, so the
//    spatially adjacent work of Morton-ordered rows shares an L2. The kNN
//    queries use this: an XCD's L2 then holds the subtrees of its own region
//    only (2.03 ms for the 1M x 1M, k = 32 query, against 2.06-2.13 ms with
//    chunks of 32-512 tiles).
//  * kChunk > 0: the tiles are cut into chunks of kChunk consecutive tiles,
//    and chunk j goes to XCD j mod 8, each XCD taking its chunks in order:
//    adjacent work inside a chunk still shares an L2, and the eight XCDs
//    sweep the whole range side by side, so work whose cost changes along
//    the range stays balanced over them. The setup and apply use this: a
//    batch's rows are ordered by neighbour count, and with contiguous
//    eighths the XCDs holding its 64-neighbour rows had twice the average
//    work (1M rows in four k groups: setup 0.99 ms contiguous, 0.46 ms with
//    chunks of 512). Tiles past the last whole round of 8 chunks are split
//    into contiguous eighths, as with kChunk = 0 (in dispatch order they
//    would alternate between the XCDs tile by tile, and the 1M-row setup
//    has a fifth of its tiles there).
constexpr int64_t kBalancedChunk = 512;
template <int64_t kChunk = kBalancedChunk>
__device__ inline int64_t xcd_tile()
{
    constexpr int64_t kXcds = 8;
    int64_t const nb = gridDim.x, b = blockIdx.x;
    int64_t const x = b % kXcds, i = b / kXcds;
    if constexpr ( kChunk == 0 )
    {
        int64_t const per = nb / kXcds, rem = nb % kXcds;
        return x * per + ( x < rem ? x : rem ) + i;
    }
    else
    {
        constexpr int64_t kRound = kXcds * kChunk;
        int64_t const full = nb / kRound * kRound;
        if ( b < full )
            return i / kChunk * kRound + x * kChunk + i % kChunk;
        // (full is a multiple of 8: the tail's block b - full is on XCD x)
        int64_t const nt = nb - full, it = ( b - full ) / kXcds;
        int64_t const per = nt / kXcds, rem = nt % kXcds;
        return full + x * per + ( x < rem ? x : rem ) + it;
    }
}

inline unsigned ceil_div( int64_t a, int64_t b )
{
    return static_cast<unsigned>( ( a + b - 1 ) / b );
}

} // namespace dtkx

#endif

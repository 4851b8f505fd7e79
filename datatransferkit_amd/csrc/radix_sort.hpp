// Stable LSD radix sort of (key, index) pairs for the kNN build's Morton
// codes, written for gfx950 wave64.
//
// One launch per 8-bit digit. A block takes a tile of 8192 keys (a ticket
// from a counter fixes the tile order, so every tile it waits for is already
// running), ranks them in LDS, publishes the tile's 256 digit counts and
// reads the counts of the tiles before it from their published words
// (decoupled look-back: a word holds a flag and either the tile's own count
// or the inclusive count of every tile up to it), then writes the tile in
// digit order to its place. A pass thus reads and writes each pair once,
// with no separate histogram or scan launch.
//
// Ranking is stable: each of the block's 16 waves holds 512 consecutive
// keys as 8 slots of 64, and ranks them slot by slot in index order. For one slot the lanes with
// the same digit are found with one ballot per digit bit; the lowest of
// them adds the group's size to the wave's count for that digit.
//
// The digit counts of every pass over the whole array (the starting offset
// of each digit) come from the kernel that writes the keys; the look-back
// words of the first pass and the tickets start at zero (cleared by an
// earlier kernel of the chain). Pass p clears the words pass p + 1 uses.
// Two independent arrays (the sources' 63-bit codes and the targets' 30-bit
// ones) share the launches. The last pass can also write the inverse
// permutation and gather up to two per-index records into sorted order.
#ifndef DTKX_RADIX_SORT_HPP
#define DTKX_RADIX_SORT_HPP

#include "common.hpp"

namespace dtkx
{
namespace rsort
{

// 16 waves of 8 keys a lane, 8192 keys a tile (112 KB of LDS, one block
// per CU). Per pass at 2^20 source keys (with the targets' 2^20 in the
// first four passes), with 4096-key tiles of 16 waves: 18-19 us alone,
// 29-34 us with the targets; 4 waves of 16 keys took 29 / 38-46 us.
// Fewer, larger tiles shorten the look-back chains, which outweighs the
// CUs they leave idle: the kNN build (scene box to hierarchy, 1M x 1M)
// took 353 us with 8192-key tiles, 369 with 4096, 416 with 2048 (512
// threads) and 551 with 1024 (256 threads).
#ifndef DTKX_RSORT_THREADS
#define DTKX_RSORT_THREADS 1024
#endif
#ifndef DTKX_RSORT_TILE
#define DTKX_RSORT_TILE 8192
#endif
constexpr int kThreads = DTKX_RSORT_THREADS;
constexpr int kWaves = kThreads / 64;
constexpr int kItems = DTKX_RSORT_TILE / kThreads;
constexpr int kTile = kThreads * kItems;
static_assert( kThreads >= 256 && kThreads % 64 == 0 && kItems >= 1 &&
                   kTile == DTKX_RSORT_TILE,
               "a thread per digit, whole waves, whole slots" );
constexpr int kDigits = 256;
constexpr int kMaxPasses = 8;
// look-back words read per round trip (32 took 3-5 us longer a pass: its
// registers, not the round trips, cost more)
#ifndef DTKX_RSORT_WINDOW
#define DTKX_RSORT_WINDOW 8
#endif
constexpr int kWindow = DTKX_RSORT_WINDOW;
// look-back word: flag << 32 | count
constexpr uint64_t kAggregate = 1ull << 32;
constexpr uint64_t kInclusive = 2ull << 32;

struct Array
{
    void *keys[2] = {nullptr, nullptr}; // pass p reads [p & 1], writes the other
    int32_t *vals[2] = {nullptr, nullptr}; // likewise; pass 0 reads the identity
    int32_t *vals_out = nullptr;        // indices in sorted order (last pass)
    int32_t *pos_out = nullptr;         // nullable: pos_out[index] = position
    // nullable: gather_out[g][position] = gather_in[g][index] (last pass),
    // records of 4, 8, 16 or 32 bytes
    void const *gather_in[2] = {nullptr, nullptr};
    void *gather_out[2] = {nullptr, nullptr};
    int gather_bytes[2] = {0, 0};
    uint32_t const *hist = nullptr; // [passes][256] digit counts
    uint64_t *look[2] = {nullptr, nullptr}; // [tiles][256], by pass parity
    int64_t n = 0;
    int tiles = 0;
    int passes = 0;
    int key_bits = 0;
    int key_bytes = 8;
    bool keep_keys = false; // the last pass writes the sorted keys
};

struct PassArgs
{
    Array a[2];
    int pass = 0;
    uint32_t *tickets = nullptr; // [kMaxPasses], zero before the first pass
};

inline int passes_for( int key_bits ) { return ( key_bits + 7 ) / 8; }

// The device state of one sort of up to two arrays, in one buffer: digit
// counts (hist_words), tickets and the first pass's look-back words (the
// region [0, zero_bytes), to be cleared before the first pass), then the
// second pass's look-back words.
struct State
{
    uint32_t *hist = nullptr, *tickets = nullptr;
    uint64_t *look0 = nullptr, *look1 = nullptr; // [tiles0 + tiles1][256]
    void *zero = nullptr;
    std::size_t zero_bytes = 0;
};

inline std::size_t state_bytes( int hist_words, int tiles )
{
    std::size_t const head = ( (std::size_t)hist_words + 4 * kMaxPasses ) * 4;
    std::size_t const look = (std::size_t)tiles * kDigits * 8;
    return ( ( head + 15 ) & ~std::size_t( 15 ) ) + 2 * look;
}

// Carves base (state_bytes( hist_words, tiles ) bytes, 16-aligned).
inline State carve_state( void *base, int hist_words, int tiles )
{
    State st;
    char *const b = static_cast<char *>( base );
    std::size_t const head =
        ( ( (std::size_t)hist_words + 4 * kMaxPasses ) * 4 + 15 ) &
        ~std::size_t( 15 );
    std::size_t const look = (std::size_t)tiles * kDigits * 8;
    st.hist = reinterpret_cast<uint32_t *>( b );
    st.tickets = st.hist + hist_words;
    st.look0 = reinterpret_cast<uint64_t *>( b + head );
    st.look1 = reinterpret_cast<uint64_t *>( b + head + look );
    st.zero = b;
    st.zero_bytes = head + look;
    return st;
}
inline int tiles_for( int64_t n )
{
    return static_cast<int>( ( n + kTile - 1 ) / kTile );
}

struct Smem
{
    uint32_t whist[kWaves][kDigits]; // per-wave digit counts, then offsets
    uint32_t tpre[kDigits];          // tile-local start of each digit
    int32_t base[kDigits];           // output position of tile slot 0 per digit
    uint32_t scan[kWaves][2];
    uint32_t ticket;
    uint64_t keys[kTile]; // the tile in digit order (32-bit keys use half)
    int32_t vals[kTile];
};

// Number of set bits of m at lanes below this one.
__device__ inline uint32_t lanes_below( uint64_t m )
{
    return __builtin_amdgcn_mbcnt_hi(
        static_cast<uint32_t>( m >> 32 ),
        __builtin_amdgcn_mbcnt_lo( static_cast<uint32_t>( m ), 0u ) );
}

__device__ inline uint32_t wave_inclusive_scan( uint32_t v )
{
    int const lane = threadIdx.x & 63;
#pragma unroll
    for ( int o = 1; o < 64; o <<= 1 )
    {
        uint32_t const u = __shfl_up( v, o, 64 );
        if ( lane >= o )
            v += u;
    }
    return v;
}

// Exclusive prefix sums over the block's threads of a and b.
__device__ inline void block_exclusive_scan2( uint32_t a, uint32_t b,
                                              uint32_t &ea, uint32_t &eb,
                                              Smem &sm )
{
    int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t const ia = wave_inclusive_scan( a );
    uint32_t const ib = wave_inclusive_scan( b );
    if ( lane == 63 )
    {
        sm.scan[w][0] = ia;
        sm.scan[w][1] = ib;
    }
    __syncthreads();
    uint32_t oa = 0, ob = 0;
    for ( int i = 0; i < w; ++i )
    {
        oa += sm.scan[i][0];
        ob += sm.scan[i][1];
    }
    ea = oa + ia - a;
    eb = ob + ib - b;
}

__device__ inline void copy_record( void const *in, void *out, int bytes,
                                    int32_t from, int32_t to )
{
    uint4 const *src = static_cast<uint4 const *>( in );
    uint4 *dst = static_cast<uint4 *>( out );
    if ( bytes == 4 )
        static_cast<uint32_t *>( out )[to] =
            static_cast<uint32_t const *>( in )[from];
    else if ( bytes == 8 )
        static_cast<uint64_t *>( out )[to] =
            static_cast<uint64_t const *>( in )[from];
    else if ( bytes == 16 )
        dst[to] = src[from];
    else
    {
        uint4 const a = src[2 * (int64_t)from];
        uint4 const b = src[2 * (int64_t)from + 1];
        dst[2 * (int64_t)to] = a;
        dst[2 * (int64_t)to + 1] = b;
    }
}

template <typename Key>
__device__ void sort_tile( Array const &A, int pass, int tile, Smem &sm )
{
    int const lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int const t = threadIdx.x;
    int const shift = 8 * pass;
    int const nbits = min( 8, A.key_bits - shift );
    uint32_t const dmask = ( 1u << nbits ) - 1u;
    bool const odd = pass & 1;
    bool const last = pass + 1 == A.passes;
    Key const *kin = static_cast<Key const *>( odd ? A.keys[1] : A.keys[0] );
    Key *kout = static_cast<Key *>( odd ? A.keys[0] : A.keys[1] );
    int32_t const *vin = odd ? A.vals[1] : A.vals[0];
    int32_t *vout = last ? A.vals_out : ( odd ? A.vals[0] : A.vals[1] );
    uint64_t *look = odd ? A.look[1] : A.look[0];
    uint64_t *look_next = odd ? A.look[0] : A.look[1];
    int64_t const t0 = (int64_t)tile * kTile;
    int64_t const w0 = t0 + w * ( kTile / kWaves );

    for ( int i = lane; i < kDigits; i += 64 )
        sm.whist[w][i] = 0;

    Key key[kItems];
    int32_t val[kItems];
    uint32_t rank[kItems];
#pragma unroll
    for ( int s = 0; s < kItems; ++s )
    {
        int64_t const idx = w0 + s * 64 + lane;
        bool const ok = idx < A.n;
        key[s] = ok ? kin[idx] : Key( 0 );
        val[s] = ok ? ( pass == 0 ? static_cast<int32_t>( idx ) : vin[idx] )
                    : 0;
    }
    // rank slot by slot: the wave's earlier keys of a digit come first
#pragma unroll
    for ( int s = 0; s < kItems; ++s )
    {
        bool const ok = w0 + s * 64 + lane < A.n;
        uint32_t const d = static_cast<uint32_t>( key[s] >> shift ) & dmask;
        uint64_t m = __ballot( ok );
#pragma unroll
        for ( int b = 0; b < 8; ++b )
            if ( b < nbits )
            {
                bool const bit = ( d >> b ) & 1u;
                uint64_t const bal = __ballot( bit );
                m &= bit ? bal : ~bal;
            }
        uint32_t const below = lanes_below( m );
        uint32_t prior = 0;
        if ( ok )
            prior = sm.whist[w][d];
        if ( ok && below == 0 )
            sm.whist[w][d] = prior + static_cast<uint32_t>( __popcll( m ) );
        rank[s] = prior + below;
    }
    __syncthreads();

    // thread t < 256 owns digit t: offsets of the waves, the tile's count
    bool const digit = t < kDigits;
    uint32_t total = 0;
    uint64_t *mine = look + (int64_t)tile * kDigits + t;
    if ( digit )
    {
#pragma unroll
        for ( int v = 0; v < kWaves; ++v )
        {
            uint32_t const c = sm.whist[v][t];
            sm.whist[v][t] = total;
            total += c;
        }
        __hip_atomic_store( mine,
                            ( tile == 0 ? kInclusive : kAggregate ) | total,
                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
    }
    // (threads past the digits scan zeros after them)
    uint32_t tpre = 0, gbase = 0;
    block_exclusive_scan2( total, digit ? A.hist[pass * kDigits + t] : 0u,
                           tpre, gbase, sm );
    // look back over the earlier tiles of this digit, kWindow words per
    // round trip: add the counts from the nearest tile back to the first
    // inclusive word; a tile that has not published yet ends the round,
    // and the next round starts at it
    uint32_t excl = 0;
    if ( digit && tile > 0 )
    {
        int j = tile - 1;
        bool done = false;
        while ( !done )
        {
            uint64_t wd[kWindow];
#pragma unroll
            for ( int i = 0; i < kWindow; ++i )
                wd[i] = j - i >= 0
                            ? __hip_atomic_load( look + (int64_t)( j - i ) * kDigits + t,
                                                 __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT )
                            : kInclusive;
            int used = 0;
            bool stalled = false;
#pragma unroll
            for ( int i = 0; i < kWindow; ++i )
            {
                uint32_t const flag = static_cast<uint32_t>( wd[i] >> 32 );
                if ( done || stalled )
                    continue;
                if ( flag == 0 )
                {
                    stalled = true;
                    continue;
                }
                excl += static_cast<uint32_t>( wd[i] );
                ++used;
                done = flag == 2;
            }
            j -= used;
            if ( stalled )
                __builtin_amdgcn_s_sleep( 1 );
        }
        __hip_atomic_store( mine, kInclusive | ( excl + total ),
                            __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT );
    }
    if ( digit )
    {
        sm.tpre[t] = tpre;
        // gbase counts every key of a smaller digit, tpre only this tile's
        sm.base[t] = static_cast<int32_t>( gbase + excl - tpre );
        if ( !last )
            look_next[(int64_t)tile * kDigits + t] = 0;
    }
    __syncthreads();

    Key *skeys = reinterpret_cast<Key *>( sm.keys );
#pragma unroll
    for ( int s = 0; s < kItems; ++s )
        if ( w0 + s * 64 + lane < A.n )
        {
            uint32_t const d =
                static_cast<uint32_t>( key[s] >> shift ) & dmask;
            uint32_t const lp = sm.tpre[d] + sm.whist[w][d] + rank[s];
            skeys[lp] = key[s];
            sm.vals[lp] = val[s];
        }
    __syncthreads();

    int const cnt = static_cast<int>( min<int64_t>( kTile, A.n - t0 ) );
    bool const write_keys = !last || A.keep_keys;
#pragma unroll 4
    for ( int s = 0; s < kItems; ++s )
    {
        int const j = s * kThreads + t;
        if ( j >= cnt )
            break;
        Key const k = skeys[j];
        int32_t const v = sm.vals[j];
        uint32_t const d = static_cast<uint32_t>( k >> shift ) & dmask;
        int32_t const dest = sm.base[d] + j;
        if ( write_keys )
            kout[dest] = k;
        vout[dest] = v;
        if ( last )
        {
            if ( A.pos_out )
                A.pos_out[v] = dest;
            if ( A.gather_out[0] )
                copy_record( A.gather_in[0], A.gather_out[0],
                             A.gather_bytes[0], v, dest );
            if ( A.gather_out[1] )
                copy_record( A.gather_in[1], A.gather_out[1],
                             A.gather_bytes[1], v, dest );
        }
    }
}

// One digit pass over the tiles of both arrays (an array whose passes are
// done contributes no tiles). Grid: the active arrays' tile counts summed.
__global__ void __launch_bounds__( kThreads ) pass_kernel( PassArgs pa )
{
    __shared__ Smem sm;
    if ( threadIdx.x == 0 )
        sm.ticket = atomicAdd( pa.tickets + pa.pass, 1u );
    __syncthreads();
    int tile = static_cast<int>( sm.ticket );
    int const first = pa.pass < pa.a[0].passes ? pa.a[0].tiles : 0;
    bool const second = tile >= first;
    if ( second )
        tile -= first;
    Array const &A = second ? pa.a[1] : pa.a[0];
    if ( A.key_bytes == 8 )
        sort_tile<uint64_t>( A, pa.pass, tile, sm );
    else
        sort_tile<uint32_t>( A, pa.pass, tile, sm );
}

// Runs every pass of both arrays (a[1] may be empty: n = 0) on stream s.
inline void sort( PassArgs pa, hipStream_t s )
{
    int const passes = max( pa.a[0].passes, pa.a[1].passes );
    for ( int p = 0; p < passes; ++p )
    {
        int tiles = 0;
        for ( int i = 0; i < 2; ++i )
            if ( p < pa.a[i].passes )
                tiles += pa.a[i].tiles;
        if ( tiles == 0 )
            continue;
        pa.pass = p;
        pass_kernel<<<tiles, kThreads, 0, s>>>( pa );
        DTKX_HIP( hipGetLastError() );
    }
}

} // namespace rsort
} // namespace dtkx

#endif

// LBVH k-nearest-neighbour search for gfx950.
//
// Build: scene box -> 63-bit Morton codes (21 bits per axis) with every
// radix pass's digit counts -> the hand-written radix sort of (code, index)
// (radix_sort.hpp; the targets' 30-bit codes sort in the same launches, and
// the last pass writes both inverse permutations and the points in sorted
// order) -> Karras hierarchy (duplicate codes split by index), in which each
// internal node also writes both child boxes: exact for small children, the
// Morton cell of the child's common prefix for large ones (a conservative
// box needs no bottom-up pass, whose device-scope fences are costly across
// the XCDs' separate L2 caches).
// Query: one lane per target, targets processed in Morton order so a wave
// walks a compact region; a per-lane stack in LDS; the k best candidates
// stay sorted in registers as 64-bit keys (float distance bits << 32 |
// index), so the ordering is (distance, index) with one integer compare.
// Pruning uses squared distances against a threshold slightly above the
// square of the current k-th distance, which never drops a candidate that
// could tie or beat it; the kept set is therefore exactly the k smallest
// (distance, index) pairs, as brute force gives. For larger k (and double
// coordinates, whose key comparisons cost more) the list is replaced by two
// traversals: the exact k-th squared distance from a list of distances
// alone (one median-of-three per slot), then a fixed-radius pass that
// collects the row, sorted at the end by (distance, index).
#include "knn.hpp"
#include "radix_sort.hpp"


#include <cfloat>
#include <cstring>
#include <vector>

namespace dtkx
{
namespace
{

template <typename T>
struct Vec4;
template <>
struct Vec4<float>
{
    using type = float4;
};
template <>
struct Vec4<double>
{
    using type = double4;
};

// Child code c >= 0: internal node c; c < 0: leaf holding source index ~c.
template <typename T>
struct alignas( 16 ) Node
{
    T lo[2][3];
    T hi[2][3];
    int32_t child[2];
};

constexpr int kBlock = 256;

// ------------------------------------------------------------- scene box
// Also clears zero[0, n_zero): counters that later kernels of the same
// build accumulate into.
template <typename T, typename V>
__global__ void __launch_bounds__( kBlock )
    bbox_partial( V const *pts, int64_t n, T *partial, uint4 *zero,
                  int64_t n_zero )
{
    for ( int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n_zero;
          i += (int64_t)gridDim.x * kBlock )
        zero[i] = make_uint4( 0u, 0u, 0u, 0u );
    T lo[3] = {INFINITY, INFINITY, INFINITY};
    T hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for ( int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
          i += (int64_t)gridDim.x * kBlock )
    {
        V p = pts[i];
        lo[0] = fmin( lo[0], p.x );
        lo[1] = fmin( lo[1], p.y );
        lo[2] = fmin( lo[2], p.z );
        hi[0] = fmax( hi[0], p.x );
        hi[1] = fmax( hi[1], p.y );
        hi[2] = fmax( hi[2], p.z );
    }
    __shared__ T sh[6][kBlock];
    for ( int d = 0; d < 3; ++d )
    {
        sh[d][threadIdx.x] = lo[d];
        sh[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for ( int off = kBlock / 2; off > 0; off >>= 1 )
    {
        if ( threadIdx.x < off )
            for ( int d = 0; d < 3; ++d )
            {
                sh[d][threadIdx.x] =
                    fmin( sh[d][threadIdx.x], sh[d][threadIdx.x + off] );
                sh[3 + d][threadIdx.x] = fmax(
                    sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + off] );
            }
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        for ( int d = 0; d < 6; ++d )
            partial[blockIdx.x * 6 + d] = sh[d][0];
}

template <typename T>
__global__ void __launch_bounds__( kBlock )
    bbox_final( T const *partial, int nparts, T *box )
{
    __shared__ T sh[6][kBlock];
    T v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for ( int i = threadIdx.x; i < nparts; i += kBlock )
        for ( int d = 0; d < 3; ++d )
        {
            v[d] = fmin( v[d], partial[i * 6 + d] );
            v[3 + d] = fmax( v[3 + d], partial[i * 6 + 3 + d] );
        }
    for ( int d = 0; d < 6; ++d )
        sh[d][threadIdx.x] = v[d];
    __syncthreads();
    for ( int off = kBlock / 2; off > 0; off >>= 1 )
    {
        if ( threadIdx.x < off )
            for ( int d = 0; d < 3; ++d )
            {
                sh[d][threadIdx.x] =
                    fmin( sh[d][threadIdx.x], sh[d][threadIdx.x + off] );
                sh[3 + d][threadIdx.x] = fmax(
                    sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + off] );
            }
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        for ( int d = 0; d < 6; ++d )
            box[d] = sh[d][0];
}

// ---------------------------------------------------------- Morton codes
__device__ inline uint64_t expand21( uint32_t v )
{
    uint64_t x = v & 0x1fffffu;
    x = ( x | x << 32 ) & 0x1f00000000ffffull;
    x = ( x | x << 16 ) & 0x1f0000ff0000ffull;
    x = ( x | x << 8 ) & 0x100f00f00f00f00full;
    x = ( x | x << 4 ) & 0x10c30c30c30c30c3ull;
    x = ( x | x << 2 ) & 0x1249249249249249ull;
    return x;
}

// 63-bit Morton code of p (21 bits per axis) in the scene box
// box[0..2] = lower corner, box[3..5] = upper corner.
template <typename T, typename V>
__device__ inline uint64_t morton63( V p, T const *box )
{
    double c[3] = {(double)p.x, (double)p.y, (double)p.z};
    uint32_t q[3];
    for ( int d = 0; d < 3; ++d )
    {
        double lo = box[d], ext = (double)box[3 + d] - lo;
        double u = ext > 0 ? ( c[d] - lo ) / ext * 2097151.0 : 0.0;
        u = u < 0 ? 0 : ( u > 2097151.0 ? 2097151.0 : u );
        q[d] = (uint32_t)u;
    }
    return ( expand21( q[0] ) << 2 ) | ( expand21( q[1] ) << 1 ) |
           expand21( q[2] );
}

// Radix passes of the two key kinds: the sources' 63-bit codes and the
// targets' top 30 bits.
constexpr int kSrcKeyBits = 63, kTgtKeyBits = 30;
constexpr int kSrcPasses = ( kSrcKeyBits + 7 ) / 8;
constexpr int kTgtPasses = ( kTgtKeyBits + 7 ) / 8;
constexpr int kHistWords = ( kSrcPasses + kTgtPasses ) * rsort::kDigits;

// Morton codes of the sources (63-bit) and of the targets (top 30 bits) in
// one launch, one sort tile of points a block (blocks [0, src_blocks) take
// the sources), with the digit counts of every radix pass added into hist
// (sources' passes first, zero on entry). src_prob / tgt_prob (nullable, a batch):
// each point's problem, whose scene box is box + 6 problem.
template <typename T, typename V>
__global__ void __launch_bounds__( rsort::kThreads )
    morton_hist_kernel( V const *src, int64_t n_src, V const *tgt,
                        int64_t n_tgt, T const *box, int32_t const *src_prob,
                        int32_t const *tgt_prob, int src_blocks,
                        uint64_t *scodes, uint32_t *tcodes, uint32_t *hist )
{
    __shared__ uint32_t h[kSrcPasses * rsort::kDigits];
    bool const is_src = static_cast<int>( blockIdx.x ) < src_blocks;
    int const words = ( is_src ? kSrcPasses : kTgtPasses ) * rsort::kDigits;
    for ( int i = threadIdx.x; i < words; i += rsort::kThreads )
        h[i] = 0;
    __syncthreads();
    int64_t const c0 =
        (int64_t)( is_src ? blockIdx.x : blockIdx.x - src_blocks ) *
        rsort::kTile;
    int64_t const n = is_src ? n_src : n_tgt;
    V const *pts = is_src ? src : tgt;
    int32_t const *prob = is_src ? src_prob : tgt_prob;
    for ( int s = 0; s < rsort::kItems; ++s )
    {
        int64_t const i = c0 + s * rsort::kThreads + threadIdx.x;
        if ( i >= n )
            break;
        uint64_t const code =
            morton63<T, V>( pts[i], prob ? box + 6 * prob[i] : box );
        if ( is_src )
        {
            scodes[i] = code;
#pragma unroll
            for ( int p = 0; p < kSrcPasses; ++p )
                atomicAdd( &h[p * rsort::kDigits + ( ( code >> ( 8 * p ) ) & 255 )],
                           1u );
        }
        else
        {
            uint32_t const c = static_cast<uint32_t>( code >> 33 );
            tcodes[i] = c;
#pragma unroll
            for ( int p = 0; p < kTgtPasses; ++p )
                atomicAdd( &h[p * rsort::kDigits + ( ( c >> ( 8 * p ) ) & 255 )],
                           1u );
        }
    }
    __syncthreads();
    uint32_t *g = hist + ( is_src ? 0 : kSrcPasses * rsort::kDigits );
    for ( int i = threadIdx.x; i < words; i += rsort::kThreads )
    {
        uint32_t const v = h[i];
        if ( v )
            atomicAdd( g + i, v );
    }
}

// ------------------------------------------------------ Karras hierarchy
// delta(i, j): common prefix length of keys i and j, with the index used
// to break ties between equal codes; -1 outside [0, n).
__device__ inline int prefix_delta( uint64_t const *codes, int64_t n,
                                    int64_t i, int64_t j )
{
    if ( j < 0 || j >= n )
        return -1;
    uint64_t a = codes[i], b = codes[j];
    if ( a != b )
        return __clzll( a ^ b );
    return 64 + __clz( (uint32_t)i ^ (uint32_t)j );
}

__device__ inline uint32_t compact21( uint64_t x )
{
    x &= 0x1249249249249249ull;
    x = ( x | x >> 2 ) & 0x10c30c30c30c30c3ull;
    x = ( x | x >> 4 ) & 0x100f00f00f00f00full;
    x = ( x | x >> 8 ) & 0x1f0000ff0000ffull;
    x = ( x | x >> 16 ) & 0x1f00000000ffffull;
    x = ( x | x >> 32 ) & 0x1fffffull;
    return (uint32_t)x;
}

__device__ inline float round_down( double v ) { return __double2float_rd( v ); }
__device__ inline float round_up( double v ) { return __double2float_ru( v ); }
__device__ inline double round_down_d( double v ) { return v; }

// Conservative box of every point whose Morton code shares the first
// `delta` bits of `code` (bit 63 is always 0): the quantisation cell of the
// prefix, mapped back to world coordinates with a margin and outward
// rounding, so it contains the points whatever the rounding of the forward
// quantisation (morton63).
template <typename T>
__device__ inline void cell_box( uint64_t code, int delta, T const *scene,
                                 T *lo, T *hi )
{
    int const nb = delta >= 64 ? 63 : ( delta - 1 ); // known bits below 63
    uint64_t const all = ( 1ull << 63 ) - 1;
    uint64_t const mask = nb <= 0 ? 0ull : ( all & ~( ( 1ull << ( 63 - nb ) ) - 1 ) );
    uint64_t const pmin = code & mask;
    uint64_t const pmax = ( code & mask ) | ( all & ~mask );
    uint32_t const qmin[3] = {compact21( pmin >> 2 ), compact21( pmin >> 1 ),
                              compact21( pmin )};
    uint32_t const qmax[3] = {compact21( pmax >> 2 ), compact21( pmax >> 1 ),
                              compact21( pmax )};
    for ( int d = 0; d < 3; ++d )
    {
        double const s0 = scene[d], ext = (double)scene[3 + d] - s0;
        double const h = ext / 2097151.0;
        double const margin = 1e-9 * ext + 1e-300;
        double l = s0 + qmin[d] * h - margin;
        double u = s0 + ( qmax[d] + 1.0 ) * h + margin;
        if ( sizeof( T ) == 4 )
        {
            lo[d] = (T)round_down( l );
            hi[d] = (T)round_up( u );
        }
        else
        {
            lo[d] = (T)l;
            hi[d] = (T)u;
        }
    }
}

// Children of internal node i: ranges, codes, parents and child boxes.
// A child that is a leaf gets its point as a degenerate box and the code
// ~source_index; an internal child covering at most kTight points gets its
// exact box, a larger one the Morton cell of its common prefix. Query at
// 1M x 1M for kTight 8 / 16 / 32: k = 32 2.11 / 2.08 / 2.07 ms, k = 16
// 1.03 / 1.01 / 1.00, 1M clustered (k = 32) 2.06 / 2.03 / 2.21 (the build
// takes longer at 32).
#ifndef DTKX_KNN_TIGHT
#define DTKX_KNN_TIGHT 16
#endif
constexpr int kTight = DTKX_KNN_TIGHT;

// Problem owning internal node i of a batch: the largest p with
// src_off[p] - p <= i (problem p's internal nodes are
// [src_off[p] - p, src_off[p + 1] - p - 1), empty for one source).
__device__ inline int node_problem( int32_t const *src_off, int n_prob,
                                    int64_t i )
{
    int lo = 0, hi = n_prob - 1;
    while ( lo < hi )
    {
        int const mid = ( lo + hi + 1 ) >> 1;
        if ( src_off[mid] - mid <= i )
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

// One problem (src_off null): n points, internal nodes [0, n - 1).
// Batch: n_int internal nodes in all; node i belongs to problem p, whose
// sorted points are [src_off[p], src_off[p + 1]) with scene box scene + 6 p;
// the hierarchy of each problem is built over its own range (child node
// indices and leaf codes are global).
template <typename T, typename V>
__global__ void __launch_bounds__( kBlock )
    karras_kernel( uint64_t const *codes_all, int32_t const *sorted_all,
                   V const *pts, V const *sorted_pts_all, int64_t n_int,
                   T const *scenes, int32_t const *src_off, int n_prob,
                   Node<T> *nodes, int32_t *sizes )
{
    int64_t const gi = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( gi >= n_int )
        return;
    int64_t so = 0, io = 0, n = n_int + 1;
    T const *scene = scenes;
    if ( src_off )
    {
        int const pb = node_problem( src_off, n_prob, gi );
        so = src_off[pb];
        n = src_off[pb + 1] - so;
        io = so - pb;
        scene = scenes + 6 * pb;
    }
    uint64_t const *codes = codes_all + so;
    int32_t const *sorted_ids = sorted_all + so;
    // the points in sorted order: a copy when given, else through the ids
    V const *sorted_pts = sorted_pts_all ? sorted_pts_all + so : nullptr;
    auto point = [&]( int64_t q ) {
        return sorted_pts ? sorted_pts[q] : pts[sorted_ids[q]];
    };
    int64_t const i = gi - io;
    int d = ( prefix_delta( codes, n, i, i + 1 ) -
              prefix_delta( codes, n, i, i - 1 ) ) >= 0
                ? 1
                : -1;
    int dmin = prefix_delta( codes, n, i, i - d );
    int64_t lmax = 2;
    while ( prefix_delta( codes, n, i, i + lmax * d ) > dmin )
        lmax <<= 1;
    int64_t l = 0;
    for ( int64_t t = lmax >> 1; t >= 1; t >>= 1 )
        if ( prefix_delta( codes, n, i, i + ( l + t ) * d ) > dmin )
            l += t;
    int64_t j = i + l * d;
    int dnode = prefix_delta( codes, n, i, j );
    int64_t s = 0, t = l;
    do
    {
        t = ( t + 1 ) >> 1;
        if ( prefix_delta( codes, n, i, i + ( s + t ) * d ) > dnode )
            s += t;
    } while ( t > 1 );
    int64_t const gamma = i + s * d + ( d < 0 ? -1 : 0 );
    int64_t const first = i < j ? i : j, last = i < j ? j : i;
    sizes[gi] = (int32_t)( last - first + 1 );
    int64_t const rb[2] = {first, gamma + 1};
    int64_t const re[2] = {gamma, last};
    Node<T> nd;
    for ( int c = 0; c < 2; ++c )
    {
        int64_t const b = rb[c], e = re[c];
        if ( b == e )
        {
            int32_t const orig = sorted_ids[b];
            V const p = point( b );
            nd.lo[c][0] = nd.hi[c][0] = p.x;
            nd.lo[c][1] = nd.hi[c][1] = p.y;
            nd.lo[c][2] = nd.hi[c][2] = p.z;
            nd.child[c] = ~orig;
            continue;
        }
        int64_t const child = io + ( c == 0 ? gamma : gamma + 1 );
        nd.child[c] = (int32_t)child;
        if ( e - b + 1 <= kTight )
        {
            T lo[3] = {INFINITY, INFINITY, INFINITY};
            T hi[3] = {-INFINITY, -INFINITY, -INFINITY};
            for ( int64_t q = b; q <= e; ++q )
            {
                V const p = point( q );
                lo[0] = fmin( lo[0], p.x );
                lo[1] = fmin( lo[1], p.y );
                lo[2] = fmin( lo[2], p.z );
                hi[0] = fmax( hi[0], p.x );
                hi[1] = fmax( hi[1], p.y );
                hi[2] = fmax( hi[2], p.z );
            }
            for ( int q = 0; q < 3; ++q )
            {
                nd.lo[c][q] = lo[q];
                nd.hi[c][q] = hi[q];
            }
        }
        else
            cell_box<T>( codes[b], prefix_delta( codes, n, b, e ), scene,
                         nd.lo[c], nd.hi[c] );
    }
    nodes[gi] = nd;
}

// --------------------------------------------------------- k-best lists
template <int K>
struct KBestF
{
    uint64_t key[K];
    __device__ void init()
    {
#pragma unroll
        for ( int j = 0; j < K; ++j )
            key[j] = ~0ull;
    }
    // squared-distance threshold above which nothing can enter
    // every slot a candidate at distance w with an index above any source
    // (kept out of the results as long as kk sources lie within w)
    __device__ void seed( float w )
    {
        uint64_t const v = ( (uint64_t)__float_as_uint( w ) << 32 ) | 0xffffffffu;
#pragma unroll
        for ( int j = 0; j < K; ++j )
            key[j] = v;
    }
    __device__ float threshold() const
    {
        if ( key[K - 1] == ~0ull )
            return INFINITY;
        float w = __uint_as_float( (uint32_t)( key[K - 1] >> 32 ) );
        return w * w * 1.00000095367431640625f; // (1 + 2^-20)
    }
    __device__ void insert( float d, int32_t id )
    {
        uint64_t c = ( (uint64_t)__float_as_uint( d ) << 32 ) | (uint32_t)id;
        if ( c < key[K - 1] )
        {
#pragma unroll
            for ( int j = K - 1; j > 0; --j )
            {
                uint64_t prev = key[j - 1], cur = key[j];
                key[j] = c < prev ? prev : ( c < cur ? c : cur );
            }
            key[0] = c < key[0] ? c : key[0];
        }
    }
    __device__ int32_t index( int j ) const
    {
        return (int32_t)( key[j] & 0xffffffffu );
    }
    __device__ double distance( int j ) const
    {
        return (double)__uint_as_float( (uint32_t)( key[j] >> 32 ) );
    }
};

template <int K>
struct KBestD
{
    double dk[K];
    int32_t ik[K];
    __device__ void init()
    {
#pragma unroll
        for ( int j = 0; j < K; ++j )
        {
            dk[j] = INFINITY;
            ik[j] = INT32_MAX;
        }
    }
    __device__ void seed( double w )
    {
#pragma unroll
        for ( int j = 0; j < K; ++j )
        {
            dk[j] = w;
            ik[j] = INT32_MAX;
        }
    }
    __device__ double threshold() const
    {
        if ( ik[K - 1] == INT32_MAX && dk[K - 1] == INFINITY )
            return INFINITY;
        double w = dk[K - 1];
        return w * w * 1.0000000000000009; // (1 + 2^-50)
    }
    __device__ static bool less( double a, int32_t ia, double b, int32_t ib )
    {
        return a < b || ( a == b && ia < ib );
    }
    __device__ void insert( double d, int32_t id )
    {
        if ( less( d, id, dk[K - 1], ik[K - 1] ) )
        {
#pragma unroll
            for ( int j = K - 1; j > 0; --j )
            {
                bool lp = less( d, id, dk[j - 1], ik[j - 1] );
                bool lc = less( d, id, dk[j], ik[j] );
                double nd = lp ? dk[j - 1] : ( lc ? d : dk[j] );
                int32_t ni = lp ? ik[j - 1] : ( lc ? id : ik[j] );
                dk[j] = nd;
                ik[j] = ni;
            }
            if ( less( d, id, dk[0], ik[0] ) )
            {
                dk[0] = d;
                ik[0] = id;
            }
        }
    }
    __device__ int32_t index( int j ) const { return ik[j]; }
    __device__ double distance( int j ) const { return dk[j]; }
};

template <typename T, int K>
struct KBest;
template <int K>
struct KBest<float, K>
{
    using type = KBestF<K>;
};
template <int K>
struct KBest<double, K>
{
    using type = KBestD<K>;
};

// q clamped to [lo, hi] (lo <= hi): one median-of-three for float, a max
// and a min for double, the values of q < lo ? lo : (q > hi ? hi : q) (a
// signed-zero difference between the forms squares to the same zero).
__device__ inline float clamp_to( float q, float lo, float hi )
{
    return __builtin_amdgcn_fmed3f( q, lo, hi );
}
__device__ inline double clamp_to( double q, double lo, double hi )
{
    return fmin( fmax( q, lo ), hi );
}

// Squared distance from q to its projection onto [lo, hi]: the operation
// sequence of the point distance (tmp = b - a; s += tmp * tmp), so for a
// degenerate box it is that distance exactly and otherwise a lower bound.
template <typename T>
__device__ inline T box_sq( T const *q, T const *lo, T const *hi )
{
#pragma clang fp contract( off )
    T s = 0;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        T tmp = clamp_to( q[d], lo[d], hi[d] ) - q[d];
        s += tmp * tmp;
    }
    return s;
}

// Squared distance from q to the farthest corner of [lo, hi], in
// box_sq's operations: per axis the larger of the two rounded differences
// q - lo and hi - q. Rounding is monotonic, so for every point p of the
// box the rounded |p - q| is at most that axis term, and the sum bounds
// box_sq(q, p, p) from above.
template <typename T>
__device__ inline T far_sq( T const *q, T const *lo, T const *hi )
{
#pragma clang fp contract( off )
    T s = 0;
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        T const tmp = fmax( q[d] - lo[d], hi[d] - q[d] );
        s += tmp * tmp;
    }
    return s;
}

// Seed of a search for kk neighbours below internal node `root` (whose
// subtree holds at least kk sources): follow the query's nearer child down
// to the deepest node still holding kk sources and return far_sq of its
// box, an upper bound of the kk-th squared distance as the traversal
// computes it. sizes: sources under each internal node.
template <typename T>
__device__ inline T seed_far_sq( Node<T> const *nodes, int32_t const *sizes,
                                 int32_t root, T const *q, int kk )
{
    Node<T> nd = nodes[root];
    T lo[3], hi[3];
#pragma unroll
    for ( int d = 0; d < 3; ++d )
    {
        lo[d] = fmin( nd.lo[0][d], nd.lo[1][d] );
        hi[d] = fmax( nd.hi[0][d], nd.hi[1][d] );
    }
    while ( true )
    {
        int const c =
            box_sq( q, nd.lo[1], nd.hi[1] ) < box_sq( q, nd.lo[0], nd.hi[0] ) ? 1
                                                                              : 0;
        int32_t const ch = nd.child[c];
        if ( ch < 0 || sizes[ch] < kk )
            break;
#pragma unroll
        for ( int d = 0; d < 3; ++d )
        {
            lo[d] = nd.lo[c][d];
            hi[d] = nd.hi[c][d];
        }
        nd = nodes[ch];
    }
    return far_sq( q, lo, hi );
}

// Correctly rounded square roots: llvm.sqrt without fast-math flags (the
// library is built with -fhip-fp32-correctly-rounded-divide-sqrt). HIP's
// __fsqrt_rn maps to the native 1-ulp instruction and must not be used.
__device__ inline float sqrt_rn( float s ) { return __builtin_sqrtf( s ); }
__device__ inline double sqrt_rn( double s ) { return __builtin_sqrt( s ); }

constexpr int kQueryBlock = 64;

// Seed the two-pass search's k-th distance from the deepest node on the
// query's path holding k sources.
#ifndef DTKX_KNN_SEED
#define DTKX_KNN_SEED 1
#endif
constexpr bool kSeed = DTKX_KNN_SEED != 0;

// Queries run with an LDS stack of kLdsStack entries; the rare one whose
// descent needs more is re-run with a global stack of kMaxDepth entries.
// The LDS stack bounds the occupancy (8 bytes per entry and lane). With
// the seeded k-th distance (seed_far_sq) a descent pushes only far
// children inside the seed's bound, so shallow stacks suffice: 22 entries
// (14 waves per CU) against 24 at 1M x 1M: k = 32 2.11 / 2.12 ms, k = 16
// 1.03 / 1.13, k = 10 0.73 / 0.81; 20 entries (with an 8-entry second
// pass) 13 % slower at k = 32, and 82 clustered queries overflow. Before
// seeding, 24 entries were needed, and 40 % of the queries of 1M points in
// 200 Gaussian clusters (sigma 1e-3) overflowed them; none do now.
// Every level of the hierarchy lengthens the common prefix of the node's
// keys (63 Morton bits, then 32 index bits for equal codes), so no path
// from the root has more than 96 internal levels.
#ifndef DTKX_KNN_LDS_STACK
#define DTKX_KNN_LDS_STACK 22
#endif
constexpr int kLdsStack = DTKX_KNN_LDS_STACK;
// The one-pass search (K below two_pass_k) keeps no row in its stack's
// slots, so its stack could be shorter: 16 entries took 0.63 ms against
// 0.72 at k = 10 and 0.90 against 1.00 at k = 16 (1M x 1M uniform), but on
// 1M points in 200 clusters (sigma 1e-3) 15 and 65 queries overflowed and
// their re-run's serial tail made the search 0.86 against 0.73 and 1.23
// against 0.98 ms; 22 overflows none there.
#ifndef DTKX_KNN_LDS_STACK1
#define DTKX_KNN_LDS_STACK1 22
#endif
constexpr int kLdsStack1 = DTKX_KNN_LDS_STACK1;
// stack entries of the fixed-radius pass of a two-pass search when its row
// is collected in the LDS stack's other slots (its radius is final, so it
// descends far less deep than the first pass); the row needs
// K <= 2 kLdsStack - kPass2Stack slots for float coordinates
#ifndef DTKX_KNN_PASS2_STACK
#define DTKX_KNN_PASS2_STACK 12
#endif
constexpr int kPass2Stack = DTKX_KNN_PASS2_STACK;
constexpr int kMaxDepth = 97;
// the largest K whose two-pass rows are sorted inside the traversal kernel
#ifndef DTKX_KNN_FUSED_SORT_K
#define DTKX_KNN_FUSED_SORT_K 32
#endif
constexpr int kFusedSortK = DTKX_KNN_FUSED_SORT_K;
// collect the rows of K > kFusedSortK in LDS too (an extra array beside
// the dead stack slots), written out in one burst per row. Off: the extra
// LDS costs more occupancy than the scattered row stores cost traffic
// (1M x 1M query, k = 64: 5.30 against 4.09 ms; k = 48: 3.83 against 3.22).
#ifndef DTKX_KNN_BIG_ROW_LDS
#define DTKX_KNN_BIG_ROW_LDS 0
#endif
constexpr bool kBigRowLds = DTKX_KNN_BIG_ROW_LDS != 0;

// launch mode of the separate row sort
constexpr int kSortRows = -1;
// The smallest K searched in two passes. Measured for 1M uniform sources
// and targets with the 24-entry LDS stack: with float coordinates one pass
// is faster up to k = 24 (1.35 against 1.67 ms at k = 16, 1.93 against
// 2.09 ms at k = 24) and two passes from k = 32 (2.53 against 2.68 ms);
// with double coordinates, whose (distance, index) comparisons cost three
// operations, two passes win from k = 16 (with the 32-entry stack: 3.2
// against 4.2 ms at k = 16, 5.1 against 11.9 ms at k = 32). With the
// seeded k-th distance and the 22-entry stack the crossover is unchanged:
// one pass 1.54 against 1.75 ms at k = 24, two passes 2.12 against 2.21 ms
// at k = 32.
#ifndef DTKX_KNN_TWO_PASS_KF
#define DTKX_KNN_TWO_PASS_KF 32
#endif
template <typename T>
constexpr int two_pass_k()
{
    return sizeof( T ) == 4 ? DTKX_KNN_TWO_PASS_KF : 16;
}

// Arguments of a query launch. One problem: rows [row0, row0 + n_rows) of
// the Morton-ordered targets against n_src sources, kk results per row.
// Batch (src_off set): target t belongs to problem tgt_prob[t], whose
// sources are [src_off[p], src_off[p + 1]), whose hierarchy is rooted at
// internal node src_off[p] - p and whose rows hold kk_prob[p] results at
// CSR offset row_off[r].
template <typename T>
struct QueryArgs
{
    Node<T> const *nodes = nullptr;
    // nullable: points under each internal node (seeds the k-th distance)
    int32_t const *node_size = nullptr;
    typename Vec4<T>::type const *src = nullptr, *tgt = nullptr;
    int32_t const *order = nullptr; // row -> target
    int64_t row0 = 0, n_rows = 0;
    int64_t n_src = 0;
    int kk = 0;
    int32_t const *src_off = nullptr, *tgt_prob = nullptr,
                  *kk_prob = nullptr;
    int64_t const *row_off = nullptr;
    int32_t *out_idx = nullptr;
    double *out_dist = nullptr;
    // Morton layout: row r written at row r with indices as positions
    int32_t const *src_pos = nullptr;
    // nullable: row r's largest neighbour distance, written at row r
    float *row_max = nullptr;
    // LDS-stack launch: rows whose descent would push past the LDS stack
    // are listed here (and written by the re-run) instead of finished
    int32_t *overflow_rows = nullptr;
    int32_t *n_overflow = nullptr;
    // global-stack re-run: rows rows[i] for i < *n_rows_dev, taken by the
    // grid's n_rows lanes in turn, a stack of kMaxDepth entries per lane
    // laid out [level][lane] (stride gstride)
    int32_t const *rows = nullptr;
    int32_t const *n_rows_dev = nullptr;
    // nullable: the re-run adds the list's length here (one lane)
    int32_t *n_rows_total = nullptr;
    int2 *gstack = nullptr;
    int64_t gstride = 0;
};

// One row of a one-pass search: lane qi (its global-stack slot), row r;
// st_node and st_d are the block's LDS stack. STACK > 0: an LDS stack of
// that many entries per lane (a query that would need more is listed for
// the re-run); STACK == 0: a per-lane stack in global memory (gstack, laid
// out [level][lane]) deep enough for any tree.
template <typename T, int K, int STACK>
__device__ inline void query_row( QueryArgs<T> const &a, int64_t const qi,
                                 int64_t const r, int32_t *st_node, T *st_d )
{
    int32_t const t = a.order[r];
    int64_t so = 0, np = a.n_src;
    int32_t root = 0;
    int kk = a.kk;
    if ( a.src_off )
    {
        int const pb = a.tgt_prob[t];
        so = a.src_off[pb];
        np = a.src_off[pb + 1] - so;
        root = (int32_t)( so - pb );
        kk = a.kk_prob[pb];
    }
    typename Vec4<T>::type qp = a.tgt[t];
    T const q[3] = {qp.x, qp.y, qp.z};
    typename KBest<T, K>::type best;
    best.init();
    if ( a.node_size && np > 1 )
        best.seed(
            sqrt_rn( seed_far_sq( a.nodes, a.node_size, root, q, kk ) ) );
    T thr = best.threshold();

    if ( np == 1 )
    {
        // single source: the tree has no internal node
        typename Vec4<T>::type sp0 = a.src[so];
        T const b[3] = {sp0.x, sp0.y, sp0.z};
        best.insert( sqrt_rn( box_sq( q, b, b ) ), (int32_t)so );
    }
    else
    {
        int sp = 0;
        int32_t node = root;
        while ( node >= 0 )
        {
            Node<T> const nd = a.nodes[node];
            T s0 = box_sq( q, nd.lo[0], nd.hi[0] );
            T s1 = box_sq( q, nd.lo[1], nd.hi[1] );
            int32_t c0 = nd.child[0], c1 = nd.child[1];
            if ( c0 < 0 && s0 <= thr )
            {
                best.insert( sqrt_rn( s0 ), ~c0 );
                thr = best.threshold();
            }
            if ( c1 < 0 && s1 <= thr )
            {
                best.insert( sqrt_rn( s1 ), ~c1 );
                thr = best.threshold();
            }
            bool g0 = c0 >= 0 && s0 <= thr;
            bool g1 = c1 >= 0 && s1 <= thr;
            int32_t next = -1;
            if ( g0 && g1 )
            {
                bool near0 = s0 <= s1;
                int32_t far = near0 ? c1 : c0;
                T fs = near0 ? s1 : s0;
                next = near0 ? c0 : c1;
                if ( STACK > 0 )
                {
                    if ( sp == STACK )
                    {
                        // deeper than the LDS stack: left to the re-run
                        a.overflow_rows[atomicAdd( a.n_overflow, 1 )] =
                            (int32_t)r;
                        return;
                    }
                    st_node[sp * kQueryBlock + threadIdx.x] = far;
                    st_d[sp * kQueryBlock + threadIdx.x] = fs;
                }
                else
                {
                    int2 e;
                    e.x = far;
                    e.y = __float_as_int( (float)fs );
                    a.gstack[sp * a.gstride + qi] = e;
                }
                ++sp;
            }
            else if ( g0 )
                next = c0;
            else if ( g1 )
                next = c1;
            else
            {
                while ( sp > 0 )
                {
                    --sp;
                    int32_t cand;
                    T cs;
                    if ( STACK > 0 )
                    {
                        cand = st_node[sp * kQueryBlock + threadIdx.x];
                        cs = st_d[sp * kQueryBlock + threadIdx.x];
                    }
                    else
                    {
                        int2 e = a.gstack[sp * a.gstride + qi];
                        cand = e.x;
                        // the float copy of a double bound is rounded; a
                        // rounded-down copy could prune wrongly, so it is
                        // only used to order, never to prune, for double
                        cs = sizeof( T ) == 4 ? (T)__int_as_float( e.y ) : 0;
                    }
                    if ( cs <= thr )
                    {
                        next = cand;
                        break;
                    }
                }
            }
            node = next;
        }
    }
    // compile-time indices keep the k-best list in registers (a loop bounded
    // by the runtime kk would index it dynamically and move it to scratch)
    int64_t const base =
        a.row_off ? a.row_off[r] : ( a.src_pos ? r : (int64_t)t ) * kk;
#pragma unroll
    for ( int j = 0; j < K; ++j )
        if ( j < kk )
        {
            int32_t const id = best.index( j );
            a.out_idx[base + j] = a.src_pos ? a.src_pos[id] : id;
            if ( a.out_dist )
                a.out_dist[base + j] = best.distance( j );
            if ( a.row_max && j == kk - 1 )
                a.row_max[r] = (float)best.distance( j );
        }
}

template <typename T, int K, int STACK>
__global__ void __launch_bounds__( kQueryBlock )
    knn_query_kernel( QueryArgs<T> const a )
{
    constexpr int LS = STACK > 0 ? STACK : 1;
    __shared__ int32_t st_node[LS * kQueryBlock];
    __shared__ T st_d[LS * kQueryBlock];
    int64_t const qi = xcd_tile<0>() * kQueryBlock + threadIdx.x;
    if constexpr ( STACK == 0 )
        if ( a.rows )
        {
            // re-run of listed rows: the grid's n_rows lanes stride over
            // the list, whose length only the device knows
            int64_t const n = *a.n_rows_dev;
            if ( a.n_rows_total && qi == 0 )
                atomicAdd( a.n_rows_total, (int32_t)n );
            for ( int64_t i = qi; i < n; i += a.n_rows )
                query_row<T, K, STACK>( a, qi, a.rows[i], st_node, st_d );
            return;
        }
    // LDS-stack launches run rows [row0, row0 + n_rows), one per lane
    if ( qi >= a.n_rows )
        return;
    query_row<T, K, STACK>( a, qi, a.row0 + qi, st_node, st_d );
}

// ------------------------------------------------- two-pass k-best search
// For larger k the sorted (distance, index) list costs five operations per
// slot for every candidate, and a wave runs it whenever any of its lanes
// meets one. Two traversals do the same work with fewer operations:
// pass 1 keeps only the k smallest SQUARED distances, a sorted register
// array updated with one median-of-three per slot; its last entry s_k gives
// the k-th distance d_k = sqrt(s_k) (a correctly rounded square root is
// monotone, so it maps order statistics to order statistics). Exactly the
// pass-1 entries whose root is below d_k are sources at distance < d_k, say
// c_lt of them. Pass 2 is a fixed-radius traversal that writes those c_lt
// sources and the k - c_lt smallest indices among the sources at distance
// == d_k to the row; the row is then sorted by (distance, index) in
// registers. The result is the set and order the sorted list gives.
template <typename T>
__device__ inline T med3( T a, T b, T c );
template <>
__device__ inline float med3( float a, float b, float c )
{
    return __builtin_amdgcn_fmed3f( a, b, c );
}
template <>
__device__ inline double med3( double a, double b, double c )
{
    return fmin( c, fmax( a, b ) ); // a <= c
}

// The k smallest squared distances, ascending. Slots below K - kk hold -1,
// so the last slot is the kk-th smallest for any kk <= K.
template <typename T, int K>
struct KthList
{
    T d[K];
    __device__ void init( int kk )
    {
#pragma unroll
        for ( int j = 0; j < K; ++j )
            d[j] = j < K - kk ? T( -1 ) : T( INFINITY );
    }
    // new d[j] = clamp(s, d[j - 1], d[j]): the j-th of the sorted union
    __device__ void insert( T s )
    {
#pragma unroll
        for ( int j = K - 1; j > 0; --j )
            d[j] = med3( d[j - 1], s, d[j] );
        d[0] = fmin( d[0], s );
    }
    __device__ T last() const { return d[K - 1]; }
    // kk entries of value b (padding stays -1): b acts as kk candidates
    // at squared distance b, so the list ends at the true kk-th value
    // whenever at least kk sources lie at or below b
    __device__ void seed( int kk, T b )
    {
#pragma unroll
        for ( int j = 0; j < K; ++j )
            if ( j >= K - kk )
                d[j] = b;
    }
};

// A squared-distance bound above every s whose rounded root is <= w.
__device__ inline float root_bound( float w )
{
    return w * w * 1.00000095367431640625f; // (1 + 2^-20)
}
__device__ inline double root_bound( double w )
{
    return w * w * 1.0000000000000009; // (1 + 2^-50)
}

// (distance, index) keys of the final sort
template <typename T>
struct SortKey;
template <>
struct SortKey<float>
{
    using type = uint64_t;
    __device__ static type make( float d, int32_t id )
    {
        return ( (uint64_t)__float_as_uint( d ) << 32 ) | (uint32_t)id;
    }
    __device__ static type none() { return ~0ull; }
    __device__ static bool less( type a, type b ) { return a < b; }
    __device__ static int32_t index( type a ) { return (int32_t)( a & 0xffffffffu ); }
    __device__ static double distance( type a )
    {
        return (double)__uint_as_float( (uint32_t)( a >> 32 ) );
    }
};
struct DKey
{
    double d;
    int32_t id;
};
template <>
struct SortKey<double>
{
    using type = DKey;
    __device__ static type make( double d, int32_t id ) { return {d, id}; }
    __device__ static type none() { return {INFINITY, INT32_MAX}; }
    __device__ static bool less( type a, type b )
    {
        return a.d < b.d || ( a.d == b.d && a.id < b.id );
    }
    __device__ static int32_t index( type a ) { return a.id; }
    __device__ static double distance( type a ) { return a.d; }
};

constexpr int pow2_at_least( int k )
{
    int n = 1;
    while ( n < k )
        n <<= 1;
    return n;
}

// ascending bitonic sorting network, all indices compile-time constants
template <typename T, int N>
__device__ inline void bitonic_sort( typename SortKey<T>::type ( &a )[N] )
{
    using SK = SortKey<T>;
#pragma unroll
    for ( int k = 2; k <= N; k <<= 1 )
#pragma unroll
        for ( int j = k >> 1; j > 0; j >>= 1 )
#pragma unroll
            for ( int i = 0; i < N; ++i )
            {
                int const l = i ^ j;
                if ( l > i )
                {
                    bool const up = ( i & k ) == 0;
                    auto const x = a[i], y = a[l];
                    bool const swap = up ? SK::less( y, x ) : SK::less( x, y );
                    a[i] = swap ? y : x;
                    a[l] = swap ? x : y;
                }
            }
}

// A two-pass row (its unordered indices read through get(j)) sorted by
// (distance, index) into row, distances recomputed exactly as the
// traversals computed them, indices then mapped to positions for the Morton
// layout. (Fused: the LDS stack, not the network's registers, limits the
// kernel's occupancy.)
template <typename T, int K, typename Get>
__device__ inline void sort_row( QueryArgs<T> const &a, T const *q,
                                 int64_t r, int64_t base, int kk,
                                 int32_t *row, Get const &get )
{
    using SK = SortKey<T>;
    constexpr int N = pow2_at_least( K );
    typename SK::type key[N];
#pragma unroll
    for ( int j = 0; j < N; ++j )
    {
        key[j] = SK::none();
        if ( j < kk )
        {
            int32_t const id = get( j );
            typename Vec4<T>::type const p = a.src[id];
            T const b[3] = {p.x, p.y, p.z};
            key[j] = SK::make( sqrt_rn( box_sq( q, b, b ) ), id );
        }
    }
    bitonic_sort<T, N>( key );
#pragma unroll
    for ( int j = 0; j < K; ++j )
        if ( j < kk )
        {
            int32_t const id = SK::index( key[j] );
            row[j] = a.src_pos ? a.src_pos[id] : id;
            if ( a.out_dist )
                a.out_dist[base + j] = SK::distance( key[j] );
            if ( a.row_max && j == kk - 1 )
                a.row_max[r] = (float)SK::distance( key[j] );
        }
}

// One row of a two-pass search: lane qi (its global-stack slot), row r;
// st_node and st_d are the block's LDS stack.
template <typename T, int K, int STACK>
__device__ inline void query2_row( QueryArgs<T> const &a, int64_t const qi,
                                  int64_t const r, int32_t *st_node, T *st_d )
{
    using SK = SortKey<T>;
    int32_t const t = a.order[r];
    int64_t so = 0, np = a.n_src;
    int32_t root = 0;
    int kk = a.kk;
    if ( a.src_off )
    {
        int const pb = a.tgt_prob[t];
        so = a.src_off[pb];
        np = a.src_off[pb + 1] - so;
        root = (int32_t)( so - pb );
        kk = a.kk_prob[pb];
    }
    typename Vec4<T>::type qp = a.tgt[t];
    T const q[3] = {qp.x, qp.y, qp.z};
    int64_t const base =
        a.row_off ? a.row_off[r] : ( a.src_pos ? r : (int64_t)t ) * kk;
    int32_t *row = a.out_idx + base;

    auto overflow = [&]() {
        a.overflow_rows[atomicAdd( a.n_overflow, 1 )] = (int32_t)r;
    };

    // pass 1: the kk-th smallest squared distance
    KthList<T, K> list;
    list.init( kk );
    if ( a.node_size && np > 1 )
        // pruning starts at the seed, so the descent pushes no far child
        // beyond it
        list.seed( kk, root_bound( sqrt_rn(
                           seed_far_sq( a.nodes, a.node_size, root, q, kk ) ) ) );
    if ( np == 1 )
    {
        typename Vec4<T>::type sp0 = a.src[so];
        T const b[3] = {sp0.x, sp0.y, sp0.z};
        list.insert( box_sq( q, b, b ) );
    }
    else
    {
        T thr = list.last();
        int sp = 0;
        int32_t node = root;
        while ( node >= 0 )
        {
            Node<T> const nd = a.nodes[node];
            T const s0 = box_sq( q, nd.lo[0], nd.hi[0] );
            T const s1 = box_sq( q, nd.lo[1], nd.hi[1] );
            int32_t const c0 = nd.child[0], c1 = nd.child[1];
            // an entry equal to the bound leaves the list's values unchanged
            if ( c0 < 0 && s0 < thr )
            {
                list.insert( s0 );
                thr = list.last();
            }
            if ( c1 < 0 && s1 < thr )
            {
                list.insert( s1 );
                thr = list.last();
            }
            bool const g0 = c0 >= 0 && s0 < thr;
            bool const g1 = c1 >= 0 && s1 < thr;
            int32_t next = -1;
            if ( g0 && g1 )
            {
                bool const near0 = s0 <= s1;
                int32_t const far = near0 ? c1 : c0;
                T const fs = near0 ? s1 : s0;
                next = near0 ? c0 : c1;
                if ( STACK > 0 )
                {
                    if ( sp == STACK )
                    {
                        overflow();
                        return;
                    }
                    st_node[sp * kQueryBlock + threadIdx.x] = far;
                    st_d[sp * kQueryBlock + threadIdx.x] = fs;
                }
                else
                {
                    int2 e;
                    e.x = far;
                    e.y = __float_as_int( (float)fs );
                    a.gstack[sp * a.gstride + qi] = e;
                }
                ++sp;
            }
            else if ( g0 )
                next = c0;
            else if ( g1 )
                next = c1;
            else
            {
                while ( sp > 0 )
                {
                    --sp;
                    int32_t cand;
                    T cs;
                    if ( STACK > 0 )
                    {
                        cand = st_node[sp * kQueryBlock + threadIdx.x];
                        cs = st_d[sp * kQueryBlock + threadIdx.x];
                    }
                    else
                    {
                        int2 e = a.gstack[sp * a.gstride + qi];
                        cand = e.x;
                        // a rounded copy of a double bound only orders
                        cs = sizeof( T ) == 4 ? (T)__int_as_float( e.y ) : 0;
                    }
                    if ( cs < thr )
                    {
                        next = cand;
                        break;
                    }
                }
            }
            node = next;
        }
    }

    // pass 2: the sources at distance < d_k, and the smallest indices among
    // those at distance == d_k, written to the row unordered
    T const dk = sqrt_rn( list.last() );
    int c_lt = 0;
#pragma unroll
    for ( int j = 0; j < K; ++j )
        c_lt += ( list.d[j] >= 0 && sqrt_rn( list.d[j] ) < dk ) ? 1 : 0;
    T const s_hi = root_bound( dk );
    // The row is collected in LDS when it fits beside pass 2's stack (whose
    // pass-1 distances are dead): the stack keeps the first kPass2Stack node
    // slots, the row takes the distance slots (as ints) and the node slots
    // above. Scattered 4-byte stores to the row in global memory, spread
    // over the traversal, would leave partly written lines for the L2 to
    // evict (8 times the row's bytes, measured).
    constexpr int kDistSlots = STACK * (int)( sizeof( T ) / 4 );
    constexpr int kDeadSlots = kDistSlots + STACK - kPass2Stack;
    // rows of K > kFusedSortK: the slots beyond the dead ones in an LDS
    // array of their own, the row copied out at the end
    constexpr bool kBigRow = kBigRowLds && STACK > kPass2Stack &&
                             K > kFusedSortK;
    constexpr bool kRowLds = STACK > kPass2Stack &&
                             ( kBigRow || ( K <= kFusedSortK && K <= kDeadSlots ) );
    constexpr int kExtra = kBigRow && K > kDeadSlots ? K - kDeadSlots : 1;
    __shared__ int32_t st_extra[kBigRow ? kExtra * kQueryBlock : 1];
    constexpr int kStack2 = kRowLds ? kPass2Stack : STACK;
    int32_t *const ids_d = reinterpret_cast<int32_t *>( st_d );
    auto slot = [&]( int j ) -> int32_t & {
        if constexpr ( kRowLds )
        {
            if ( j < kDistSlots )
                return ids_d[j * kQueryBlock + threadIdx.x];
            if ( !kBigRow || j < kDeadSlots )
                return st_node[( kPass2Stack + j - kDistSlots ) * kQueryBlock +
                               threadIdx.x];
            return st_extra[( j - kDeadSlots ) * kQueryBlock + threadIdx.x];
        }
        else
            return row[j];
    };
    int n_lt = 0, n_eq = 0;
    auto take = [&]( T s, int32_t id ) {
        T const d = sqrt_rn( s );
        if ( d < dk )
        {
            if ( n_lt < c_lt )
                slot( n_lt++ ) = id;
        }
        else if ( d == dk )
        {
            if ( c_lt + n_eq < kk )
                slot( c_lt + n_eq++ ) = id;
            else
            {
                // all tie slots taken: replace the largest index if larger
                int jm = c_lt;
                for ( int j = c_lt + 1; j < kk; ++j )
                    jm = slot( j ) > slot( jm ) ? j : jm;
                if ( id < slot( jm ) )
                    slot( jm ) = id;
            }
        }
    };
    if ( np == 1 )
    {
        typename Vec4<T>::type sp0 = a.src[so];
        T const b[3] = {sp0.x, sp0.y, sp0.z};
        take( box_sq( q, b, b ), (int32_t)so );
    }
    else
    {
        int sp = 0;
        int32_t node = root;
        while ( node >= 0 )
        {
            Node<T> const nd = a.nodes[node];
            T const s0 = box_sq( q, nd.lo[0], nd.hi[0] );
            T const s1 = box_sq( q, nd.lo[1], nd.hi[1] );
            int32_t const c0 = nd.child[0], c1 = nd.child[1];
            if ( c0 < 0 && s0 <= s_hi )
                take( s0, ~c0 );
            if ( c1 < 0 && s1 <= s_hi )
                take( s1, ~c1 );
            bool const g0 = c0 >= 0 && s0 <= s_hi;
            bool const g1 = c1 >= 0 && s1 <= s_hi;
            int32_t next = -1;
            if ( g0 && g1 )
            {
                next = c0;
                if ( STACK > 0 )
                {
                    if ( sp == kStack2 )
                    {
                        overflow();
                        return;
                    }
                    st_node[sp * kQueryBlock + threadIdx.x] = c1;
                }
                else
                    a.gstack[sp * a.gstride + qi].x = c1;
                ++sp;
            }
            else if ( g0 )
                next = c0;
            else if ( g1 )
                next = c1;
            else if ( sp > 0 )
            {
                --sp;
                next = STACK > 0 ? st_node[sp * kQueryBlock + threadIdx.x]
                                 : a.gstack[sp * a.gstride + qi].x;
            }
            node = next;
        }
    }
    if constexpr ( K <= kFusedSortK )
        sort_row<T, K>( a, q, r, base, kk, row, slot );
    else if constexpr ( kBigRow )
        for ( int j = 0; j < kk; ++j )
            row[j] = slot( j );
}

// Registers for four waves per SIMD (128): the float K = 32 kernel then
// spills 52 bytes a lane, and its LDS stack admits 14 waves per CU against
// the 12 that 135 registers allowed: 1.92 ms against 2.02 for the 1M x 1M,
// k = 32 query. (Kernels whose LDS already allows fewer waves, the double
// ones, keep their registers.)
template <typename T, int K, int STACK>
__global__ void __launch_bounds__( kQueryBlock )
    __attribute__( ( amdgpu_waves_per_eu( 4 ) ) )
    knn_query2_kernel( QueryArgs<T> const a )
{
    constexpr int LS = STACK > 0 ? STACK : 1;
    __shared__ int32_t st_node[LS * kQueryBlock];
    __shared__ T st_d[LS * kQueryBlock];
    int64_t const qi = xcd_tile<0>() * kQueryBlock + threadIdx.x;
    if constexpr ( STACK == 0 )
        if ( a.rows )
        {
            // re-run of listed rows: the grid's n_rows lanes stride over
            // the list, whose length only the device knows
            int64_t const n = *a.n_rows_dev;
            if ( a.n_rows_total && qi == 0 )
                atomicAdd( a.n_rows_total, (int32_t)n );
            for ( int64_t i = qi; i < n; i += a.n_rows )
                query2_row<T, K, STACK>( a, qi, a.rows[i], st_node, st_d );
            return;
        }
    // LDS-stack launches run rows [row0, row0 + n_rows), one per lane
    if ( qi >= a.n_rows )
        return;
    query2_row<T, K, STACK>( a, qi, a.row0 + qi, st_node, st_d );
}

// The row sort of a two-pass launch for K > kFusedSortK, whose sorting
// network would take the traversal's registers. A row's 64 slots (entries
// past the row's length hold keys that sort last) are spread over a group
// of 64 / E lanes, lane g of the group holding slots g E .. g E + E - 1, so
// a wave sorts E rows. The bitonic network of sort_row then runs its
// stages of distance j < E inside each lane and the others across the
// group's lanes (ds_bpermute). The row's index loads and stores are
// contiguous per group. 1M rows of 64 (float, uniform points) take 0.40 ms
// with E = 4, against 0.42 / 0.50 / 0.93 ms with E = 2 / 8 / 16 (fewer
// VALU per row with more slots per lane, but more registers and fewer
// rows in flight), 0.59 ms with one row per wave, and 1.19 ms with one row
// per lane (64 keys in 268 registers, one wave per SIMD, each lane's loads
// and stores touching 64 cache lines).
#ifndef DTKX_SORT_ROWS_E
#define DTKX_SORT_ROWS_E 4
#endif
constexpr int kSortE = DTKX_SORT_ROWS_E;   // slots per lane
constexpr int kSortWaves = 4;              // waves per block
constexpr int kSortSlots = 64;             // slots per row

// the 32-bit value of the lane whose byte address is given (every lane of
// the wave is active)
__device__ inline int lane_read( int v, int addr )
{
    return __builtin_amdgcn_ds_bpermute( addr, v );
}

template <typename T>
__device__ inline typename SortKey<T>::type
lane_read_key( typename SortKey<T>::type k, int addr )
{
    if constexpr ( sizeof( T ) == 4 )
    {
        uint32_t const lo = (uint32_t)lane_read( (int)( k & 0xffffffffu ), addr );
        uint32_t const hi = (uint32_t)lane_read( (int)( k >> 32 ), addr );
        return ( (uint64_t)hi << 32 ) | lo;
    }
    else
    {
        int64_t const b = __double_as_longlong( k.d );
        int const lo = lane_read( (int)( b & 0xffffffff ), addr );
        int const hi = lane_read( (int)( b >> 32 ), addr );
        return {__longlong_as_double( ( (int64_t)hi << 32 ) | (uint32_t)lo ),
                lane_read( k.id, addr )};
    }
}

// a < b and (c ? a : b) field by field, without short-circuit branches
__device__ inline bool key_less( uint64_t a, uint64_t b ) { return a < b; }
__device__ inline bool key_less( DKey const &a, DKey const &b )
{
    return ( a.d < b.d ) | ( ( a.d == b.d ) & ( a.id < b.id ) );
}
__device__ inline uint64_t key_pick( bool c, uint64_t a, uint64_t b )
{
    return c ? a : b;
}
__device__ inline DKey key_pick( bool c, DKey const &a, DKey const &b )
{
    return {c ? a.d : b.d, c ? a.id : b.id};
}

template <typename T, int K>
__global__ void __launch_bounds__( 64 * kSortWaves )
    knn_sort_rows_kernel( QueryArgs<T> const a )
{
    constexpr int E = kSortE, G = kSortSlots / E;
    static_assert( K <= kSortSlots && G >= 1 && G * E == kSortSlots,
                   "a row's slots fill its lane group" );
    using SK = SortKey<T>;
    int const lane = threadIdx.x & 63;
    int const g = lane % G; // lane within the row's group
    int64_t const qi = ( xcd_tile<0>() * kSortWaves + threadIdx.x / 64 ) * E +
                       lane / G;
    // rows past the launch's end still run the network (every lane takes
    // part in the cross-lane reads) with keys that sort last
    bool const valid = qi < a.n_rows;
    int64_t const r = a.row0 + ( valid ? qi : 0 );
    int kk = 0;
    typename SK::type key[E];
#pragma unroll
    for ( int e = 0; e < E; ++e )
        key[e] = SK::none();
    int32_t *row = nullptr;
    int64_t base = 0;
    if ( valid )
    {
        int32_t const t = a.order[r];
        kk = a.kk;
        if ( a.src_off )
            kk = a.kk_prob[a.tgt_prob[t]];
        typename Vec4<T>::type qp = a.tgt[t];
        T const q[3] = {qp.x, qp.y, qp.z};
        base = a.row_off ? a.row_off[r] : ( a.src_pos ? r : (int64_t)t ) * kk;
        row = a.out_idx + base;
#pragma unroll
        for ( int e = 0; e < E; ++e )
        {
            int const j = g * E + e;
            if ( j < kk )
            {
                int32_t const id = row[j];
                typename Vec4<T>::type const p = a.src[id];
                T const b[3] = {p.x, p.y, p.z};
                key[e] = SK::make( sqrt_rn( box_sq( q, b, b ) ), id );
            }
        }
    }
    // slot i = g E + e; of a pair (i, i ^ j), the lower slot of an
    // ascending block (i & k == 0) keeps the smaller key
#pragma unroll
    for ( int k = 2; k <= kSortSlots; k <<= 1 )
#pragma unroll
        for ( int j = k >> 1; j > 0; j >>= 1 )
        {
            if ( j < E )
            {
#pragma unroll
                for ( int e = 0; e < E; ++e )
                {
                    int const f = e ^ j;
                    if ( f > e )
                    {
                        bool const up = ( ( g * E + e ) & k ) == 0;
                        auto const x = key[e], y = key[f];
                        bool const lt = key_less( y, x ), gt = key_less( x, y );
                        bool const swap = ( up & lt ) | ( !up & gt );
                        key[e] = key_pick( swap, y, x );
                        key[f] = key_pick( swap, x, y );
                    }
                }
            }
            else
            {
                int const addr = ( lane ^ ( j / E ) ) << 2;
                // (i & k) and (i & j) depend on the lane only: e < E <= j < k
                bool const keep_min =
                    ( ( ( g * E ) & k ) == 0 ) == ( ( ( g * E ) & j ) == 0 );
#pragma unroll
                for ( int e = 0; e < E; ++e )
                {
                    typename SK::type const other = lane_read_key<T>( key[e], addr );
                    bool const lt = key_less( other, key[e] );
                    bool const gt = key_less( key[e], other );
                    // both comparisons, combined without a branch per lane
                    bool const take = ( keep_min & lt ) | ( !keep_min & gt );
                    key[e] = key_pick( take, other, key[e] );
                }
            }
        }
    if ( valid )
    {
#pragma unroll
        for ( int e = 0; e < E; ++e )
        {
            int const j = g * E + e;
            if ( j < kk )
            {
                int32_t const id = SK::index( key[e] );
                row[j] = a.src_pos ? a.src_pos[id] : id;
                if ( a.out_dist )
                    a.out_dist[base + j] = SK::distance( key[e] );
                if ( a.row_max && j == kk - 1 )
                    a.row_max[r] = (float)SK::distance( key[e] );
            }
        }
    }
}

template <typename T>
struct QueryLauncher
{
    template <int K>
    static void run( int stack, QueryArgs<T> const &a, hipStream_t s )
    {
        dim3 grid( ceil_div( a.n_rows, kQueryBlock ) ), block( kQueryBlock );
        if constexpr ( K >= two_pass_k<T>() )
        {
            if ( stack == kLdsStack )
                knn_query2_kernel<T, K, kLdsStack><<<grid, block, 0, s>>>( a );
            else if ( stack == 0 )
                knn_query2_kernel<T, K, 0><<<grid, block, 0, s>>>( a );
            else if ( K > kFusedSortK && stack == kSortRows )
                knn_sort_rows_kernel<T, K>
                    <<<ceil_div( a.n_rows, kSortWaves * kSortE ),
                       64 * kSortWaves, 0, s>>>( a );
        }
        else if ( stack == kLdsStack )
            knn_query_kernel<T, K, kLdsStack1><<<grid, block, 0, s>>>( a );
        else if ( stack == 0 )
            knn_query_kernel<T, K, 0><<<grid, block, 0, s>>>( a );
        DTKX_HIP( hipGetLastError() );
    }
    // k: the largest number of results of a row in this launch
    static void dispatch( int k, int stack, QueryArgs<T> const &a,
                          hipStream_t s )
    {
        if ( a.n_rows == 0 )
            return;
#define DTKX_KNN_CASE( KV )                                                    \
    if ( k <= KV )                                                             \
    {                                                                          \
        run<KV>( stack, a, s );                                                \
        return;                                                                \
    }
        DTKX_KNN_CASE( 1 )
        DTKX_KNN_CASE( 4 )
        DTKX_KNN_CASE( 8 )
        DTKX_KNN_CASE( 10 )
        DTKX_KNN_CASE( 16 )
        DTKX_KNN_CASE( 24 )
        DTKX_KNN_CASE( 32 )
        DTKX_KNN_CASE( 48 )
        DTKX_KNN_CASE( 64 )
#undef DTKX_KNN_CASE
        DTKX_REQUIRE( k <= knn_max_k, DTKX_ERR_UNSUPPORTED );
    }
};

// the K of the query kernel instantiated for up to k results
inline int query_bucket( int k )
{
    int const ks[] = {1, 4, 8, 10, 16, 24, 32, 48, 64};
    for ( int v : ks )
        if ( k <= v )
            return v;
    return 64;
}

// blocks of the global-stack re-run (16k lanes, 12.7 MB of stack)
constexpr int64_t kRerunBlocks = 256;

// The query launch with the LDS stack, then the re-run of its overflow.
template <typename T>
void run_queries( int kmax, QueryArgs<T> qa, int64_t n_rows, KnnInfo *info,
                  hipStream_t s )
{
    if ( n_rows == 0 )
        return;
    DeviceBuffer<int32_t> rows( n_rows, s ), count( 1, s );
    zero_async( count.get(), sizeof( int32_t ), s );
    qa.overflow_rows = rows.get();
    qa.n_overflow = count.get();
    QueryLauncher<T>::dispatch( kmax, kLdsStack, qa, s );
    // The re-run is launched without reading the list's length back (a
    // stream synchronisation, 60 us of idle GPU at 1M x 1M): a bounded grid
    // whose lanes stride over the list, each with its own global stack.
    // Almost always the list is empty and the launch costs a few us.
    int64_t const lanes =
        std::min<int64_t>( ceil_div( n_rows, kQueryBlock ), kRerunBlocks ) *
        kQueryBlock;
    DeviceBuffer<int2> gstack( static_cast<std::size_t>( kMaxDepth ) * lanes,
                               s );
    QueryArgs<T> ra = qa;
    ra.overflow_rows = nullptr;
    ra.n_overflow = nullptr;
    ra.rows = rows.get();
    ra.n_rows_dev = count.get();
    ra.n_rows = lanes;
    ra.gstack = gstack.get();
    ra.gstride = lanes;
    if ( info && info->n_overflow_dev )
        ra.n_rows_total = info->n_overflow_dev;
    QueryLauncher<T>::dispatch( kmax, 0, ra, s );
    if ( info && !info->n_overflow_dev )
    {
        int32_t n_over = 0;
        DTKX_HIP( hipMemcpyAsync( &n_over, count.get(), sizeof( int32_t ),
                                  hipMemcpyDeviceToHost, s ) );
        DTKX_HIP( hipStreamSynchronize( s ) );
        info->n_overflow += n_over;
    }
    // two-pass rows of large K are written unordered: sort them (a no-op
    // launch mode for every other K)
    QueryLauncher<T>::dispatch( kmax, kSortRows, qa, s );
}

template <typename T>
void knn_impl( void const *src_v, int64_t n_src, void const *tgt_v,
               int64_t n_tgt, int k, KnnOutput const &out, KnnInfo *info,
               hipStream_t s )
{
    using V = typename Vec4<T>::type;
    V const *src = static_cast<V const *>( src_v );
    V const *tgt = static_cast<V const *>( tgt_v );
    int const kk = static_cast<int>( k < n_src ? k : n_src );

    // Sort state, one buffer: the digit counts of every pass, the tile
    // tickets and the look-back words of the first pass (all cleared by the
    // scene-box launch), then those of the second pass (cleared by the
    // first pass).
    int const src_tiles = rsort::tiles_for( n_src );
    int const tgt_tiles = rsort::tiles_for( n_tgt );
    std::size_t const head = ( kHistWords + 4 * rsort::kMaxPasses ) * 4;
    std::size_t const look_src = (std::size_t)src_tiles * rsort::kDigits * 8;
    std::size_t const look_tgt = (std::size_t)tgt_tiles * rsort::kDigits * 8;
    std::size_t const zero_bytes = head + look_src + look_tgt;
    static_assert( ( kHistWords + 4 * rsort::kMaxPasses ) * 4 % 16 == 0,
                   "uint4 clearing" );
    DeviceBuffer<uint4> state( ( zero_bytes + look_src + look_tgt ) / 16, s );
    char *const st = reinterpret_cast<char *>( state.get() );
    uint32_t *const hist = reinterpret_cast<uint32_t *>( st );
    uint32_t *const tickets = hist + kHistWords;
    uint64_t *const look0 = reinterpret_cast<uint64_t *>( st + head );
    uint64_t *const look1 = reinterpret_cast<uint64_t *>( st + zero_bytes );

    // scene box of the sources
    int nparts = static_cast<int>( std::min<int64_t>(
        1024, std::max<int64_t>( 1, ceil_div( n_src, kBlock ) ) ) );
    DeviceBuffer<T> partial( 6 * nparts, s ), box( 6, s );
    bbox_partial<T, V><<<nparts, kBlock, 0, s>>>(
        src, n_src, partial.get(), state.get(),
        static_cast<int64_t>( zero_bytes / 16 ) );
    bbox_final<T><<<1, kBlock, 0, s>>>( partial.get(), nparts, box.get() );
    DTKX_HIP( hipGetLastError() );

    // Morton codes of both point sets (the targets' order only has to be
    // local, so their top 30 bits do) with every pass's digit counts
    DeviceBuffer<uint64_t> scodes( n_src, s ), scodes1( n_src, s );
    DeviceBuffer<uint32_t> tcodes( n_tgt, s ), tcodes1( n_tgt, s );
    morton_hist_kernel<T, V><<<src_tiles + tgt_tiles, rsort::kThreads, 0, s>>>(
        src, n_src, tgt, n_tgt, box.get(), nullptr, nullptr, src_tiles,
        scodes.get(), tcodes.get(), hist );
    DTKX_HIP( hipGetLastError() );

    // outputs of the sort: the permutations, their inverses, the points in
    // sorted order (the hierarchy's copy, and the caller's double4 copy,
    // one buffer when they coincide)
    DeviceBuffer<int32_t> ids_sorted_buf, pos_buf, order_buf;
    int32_t *ids_sorted = out.src_perm;
    if ( !ids_sorted )
    {
        ids_sorted_buf.resize( n_src, s );
        ids_sorted = ids_sorted_buf.get();
    }
    int32_t *src_pos = out.src_pos;
    if ( out.morton_layout && !src_pos )
    {
        pos_buf.resize( n_src, s );
        src_pos = pos_buf.get();
    }
    int32_t *order = out.query_order;
    if ( !order && n_tgt > 0 )
    {
        order_buf.resize( n_tgt, s );
        order = order_buf.get();
    }
    bool const share = out.sorted_src &&
                       static_cast<void const *>( out.src4 ) == src_v &&
                       sizeof( V ) == sizeof( double4 );
    DeviceBuffer<V> spts_buf;
    V *spts = reinterpret_cast<V *>( out.sorted_src );
    if ( !share )
    {
        spts_buf.resize( n_src, s );
        spts = spts_buf.get();
    }
    DeviceBuffer<int32_t> svals( 2 * n_src, s ), tvals( 2 * n_tgt, s );

    rsort::PassArgs pa;
    pa.tickets = tickets;
    rsort::Array &a = pa.a[0];
    a.keys[0] = scodes.get();
    a.keys[1] = scodes1.get();
    a.vals[0] = svals.get();
    a.vals[1] = svals.get() + n_src;
    a.vals_out = ids_sorted;
    a.pos_out = src_pos;
    a.gather_in[0] = src;
    a.gather_out[0] = spts;
    a.gather_bytes[0] = sizeof( V );
    if ( out.sorted_src && !share )
    {
        a.gather_in[1] = out.src4;
        a.gather_out[1] = out.sorted_src;
        a.gather_bytes[1] = sizeof( double4 );
    }
    a.hist = hist;
    a.look[0] = look0;
    a.look[1] = look1;
    a.n = n_src;
    a.tiles = src_tiles;
    a.passes = kSrcPasses;
    a.key_bits = kSrcKeyBits;
    a.key_bytes = 8;
    a.keep_keys = true;
    rsort::Array &b = pa.a[1];
    if ( n_tgt > 0 )
    {
        b.keys[0] = tcodes.get();
        b.keys[1] = tcodes1.get();
        b.vals[0] = tvals.get();
        b.vals[1] = tvals.get() + n_tgt;
        b.vals_out = order;
        b.pos_out = out.query_pos;
        b.hist = hist + kSrcPasses * rsort::kDigits;
        b.look[0] = look0 + (std::size_t)src_tiles * rsort::kDigits;
        b.look[1] = look1 + (std::size_t)src_tiles * rsort::kDigits;
        b.n = n_tgt;
        b.tiles = tgt_tiles;
        b.passes = kTgtPasses;
        b.key_bits = kTgtKeyBits;
        b.key_bytes = 4;
    }
    rsort::sort( pa, s );
    // the sorted codes are back in scodes after an even number of passes
    static_assert( kSrcPasses % 2 == 0, "sorted codes in scodes" );

    // hierarchy
    int64_t const n_int = n_src - 1;
    DeviceBuffer<Node<T>> nodes( std::max<int64_t>( n_int, 1 ), s );
    DeviceBuffer<int32_t> sizes( std::max<int64_t>( n_int, 1 ), s );
    if ( n_int > 0 )
    {
        karras_kernel<T, V><<<ceil_div( n_int, kBlock ), kBlock, 0, s>>>(
            scodes.get(), ids_sorted, src, spts, n_int, box.get(), nullptr, 1,
            nodes.get(), sizes.get() );
        DTKX_HIP( hipGetLastError() );
    }
    if ( n_tgt == 0 )
        return;

    QueryArgs<T> qa;
    qa.nodes = nodes.get();
    qa.node_size = kSeed ? sizes.get() : nullptr;
    qa.src = src;
    qa.tgt = tgt;
    qa.order = order;
    qa.row0 = 0;
    qa.n_rows = n_tgt;
    qa.n_src = n_src;
    qa.kk = kk;
    qa.out_idx = out.indices;
    qa.out_dist = out.dist;
    qa.src_pos = out.morton_layout ? src_pos : nullptr;
    qa.row_max = out.morton_layout ? out.row_max : nullptr;
    if ( info && info->query_begin )
        DTKX_HIP( hipEventRecord( info->query_begin, s ) );
    run_queries<T>( k, qa, n_tgt, info, s );
    if ( info && info->query_end )
        DTKX_HIP( hipEventRecord( info->query_end, s ) );
    // buffers are released at scope exit, ordered on the stream
}


// ------------------------------------------------------------- batched
template <typename T, typename V>
__global__ void __launch_bounds__( kBlock )
    bbox_segments( V const *pts, int32_t const *src_off, T *boxes )
{
    int const pb = blockIdx.x;
    int64_t const b = src_off[pb], e = src_off[pb + 1];
    T lo[3] = {INFINITY, INFINITY, INFINITY};
    T hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for ( int64_t i = b + threadIdx.x; i < e; i += kBlock )
    {
        V p = pts[i];
        lo[0] = fmin( lo[0], p.x );
        lo[1] = fmin( lo[1], p.y );
        lo[2] = fmin( lo[2], p.z );
        hi[0] = fmax( hi[0], p.x );
        hi[1] = fmax( hi[1], p.y );
        hi[2] = fmax( hi[2], p.z );
    }
    __shared__ T sh[6][kBlock];
    for ( int d = 0; d < 3; ++d )
    {
        sh[d][threadIdx.x] = lo[d];
        sh[3 + d][threadIdx.x] = hi[d];
    }
    __syncthreads();
    for ( int off = kBlock / 2; off > 0; off >>= 1 )
    {
        if ( threadIdx.x < off )
            for ( int d = 0; d < 3; ++d )
            {
                sh[d][threadIdx.x] =
                    fmin( sh[d][threadIdx.x], sh[d][threadIdx.x + off] );
                sh[3 + d][threadIdx.x] = fmax(
                    sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + off] );
            }
        __syncthreads();
    }
    if ( threadIdx.x == 0 )
        for ( int d = 0; d < 6; ++d )
            boxes[pb * 6 + d] = sh[d][0];
}

// Clears [a, a + na) and [b, b + nb) (uint4 units).
__global__ void __launch_bounds__( kBlock )
    clear2_kernel( uint4 *a, int64_t na, uint4 *b, int64_t nb )
{
    for ( int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < na + nb;
          i += (int64_t)gridDim.x * kBlock )
        ( i < na ? a[i] : b[i - na] ) = make_uint4( 0u, 0u, 0u, 0u );
}

// Second-stage keys of a batch's sorts, read through the first stage's
// order: a source's problem, a target's row key (row_key[problem]), with
// the digit counts of their passes added into hist (sources' first, zero on
// entry); one sort tile of keys a block, blocks [0, src_blocks) take the
// sources.
__global__ void __launch_bounds__( rsort::kThreads )
    group_keys_kernel( int32_t const *src_ids, int32_t const *src_prob,
                       int64_t n_src, int src_passes, int32_t const *tgt_ids,
                       int32_t const *tgt_prob, int32_t const *row_key,
                       int64_t n_tgt, int tgt_passes, int src_blocks,
                       uint32_t *src_keys, uint32_t *tgt_keys, uint32_t *hist )
{
    __shared__ uint32_t h[4 * rsort::kDigits];
    bool const is_src = static_cast<int>( blockIdx.x ) < src_blocks;
    int const words = ( is_src ? src_passes : tgt_passes ) * rsort::kDigits;
    for ( int i = threadIdx.x; i < words; i += rsort::kThreads )
        h[i] = 0;
    __syncthreads();
    int64_t const c0 =
        (int64_t)( is_src ? blockIdx.x : blockIdx.x - src_blocks ) *
        rsort::kTile;
    int64_t const n = is_src ? n_src : n_tgt;
    int const passes = is_src ? src_passes : tgt_passes;
    for ( int s = 0; s < rsort::kItems; ++s )
    {
        int64_t const i = c0 + s * rsort::kThreads + threadIdx.x;
        if ( i >= n )
            break;
        uint32_t const key =
            is_src ? static_cast<uint32_t>( src_prob[src_ids[i]] )
                   : static_cast<uint32_t>( row_key[tgt_prob[tgt_ids[i]]] );
        ( is_src ? src_keys : tgt_keys )[i] = key;
        for ( int p = 0; p < passes; ++p )
            atomicAdd( &h[p * rsort::kDigits + ( ( key >> ( 8 * p ) ) & 255 )],
                       1u );
    }
    __syncthreads();
    uint32_t *g = hist + ( is_src ? 0 : src_passes * rsort::kDigits );
    for ( int i = threadIdx.x; i < words; i += rsort::kThreads )
    {
        uint32_t const v = h[i];
        if ( v )
            atomicAdd( g + i, v );
    }
}

__global__ void __launch_bounds__( kBlock )
    row_offsets_kernel( int32_t const *order, int32_t const *tgt_prob,
                        int64_t const *row_base, int32_t const *row_start,
                        int32_t const *kk_prob, int64_t n, int64_t nnz,
                        int64_t *row_off )
{
    int64_t r = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( r > n )
        return;
    if ( r == n )
    {
        row_off[n] = nnz;
        return;
    }
    int32_t const pb = tgt_prob[order[r]];
    row_off[r] = row_base[pb] + ( r - row_start[pb] ) * (int64_t)kk_prob[pb];
}

template <typename T>
void knn_batched_impl( void const *src_v, int64_t n_src, void const *tgt_v,
                       int64_t n_tgt, KnnBatch const &b,
                       KnnBatchOutput const &out, KnnInfo *info,
                       hipStream_t s )
{
    using V = typename Vec4<T>::type;
    V const *src = static_cast<V const *>( src_v );
    V const *tgt = static_cast<V const *>( tgt_v );
    int const P = b.n_problems;
    int prob_bits = 1;
    while ( ( int64_t( 1 ) << prob_bits ) < P )
        ++prob_bits;

    DeviceBuffer<T> boxes( 6 * (std::size_t)P, s );
    bbox_segments<T, V><<<P, kBlock, 0, s>>>( src, b.src_off, boxes.get() );
    DTKX_HIP( hipGetLastError() );

    // Two stable radix sorts of both point sets, in shared launches
    // (radix_sort.hpp): by Morton code in the problem's box (sources 63 bits,
    // targets 30, as one problem's search orders them), then by the
    // problem's key (sources: the problem; targets: its row key). The second
    // sorts positions of the first, and its last pass gathers the first's
    // indices (and the sources' codes) into the final order.
    int const src_tiles = rsort::tiles_for( n_src );
    int const tgt_tiles = rsort::tiles_for( n_tgt );
    int const src_passes2 = rsort::passes_for( prob_bits );
    int const tgt_passes2 = rsort::passes_for( b.key_bits );
    int const hist1 = kHistWords;
    int const hist2 = ( src_passes2 + tgt_passes2 ) * rsort::kDigits;
    std::size_t const sb1 = rsort::state_bytes( hist1, src_tiles + tgt_tiles );
    std::size_t const sb2 = rsort::state_bytes( hist2, src_tiles + tgt_tiles );
    DeviceBuffer<uint4> state( ( sb1 + sb2 ) / 16, s );
    rsort::State const st1 =
        rsort::carve_state( state.get(), hist1, src_tiles + tgt_tiles );
    rsort::State const st2 = rsort::carve_state(
        reinterpret_cast<char *>( state.get() ) + sb1, hist2,
        src_tiles + tgt_tiles );
    clear2_kernel<<<std::min<int64_t>( 1024, ceil_div( sb1 + sb2, 16 * kBlock ) ),
                    kBlock, 0, s>>>(
        static_cast<uint4 *>( st1.zero ), st1.zero_bytes / 16,
        static_cast<uint4 *>( st2.zero ), st2.zero_bytes / 16 );
    DTKX_HIP( hipGetLastError() );

    DeviceBuffer<uint64_t> scodes( n_src, s ), scodes1( n_src, s ),
        codes_sorted( n_src, s );
    DeviceBuffer<uint32_t> tcodes( n_tgt, s ), tcodes1( n_tgt, s );
    morton_hist_kernel<T, V><<<src_tiles + tgt_tiles, rsort::kThreads, 0, s>>>(
        src, n_src, tgt, n_tgt, boxes.get(), b.src_prob, b.tgt_prob,
        src_tiles, scodes.get(), tcodes.get(), st1.hist );
    DTKX_HIP( hipGetLastError() );
    DeviceBuffer<int32_t> svals( 2 * n_src, s ), tvals( 2 * n_tgt, s ),
        sids1( n_src, s ), tids1( n_tgt, s );
    auto first_stage = [&]( rsort::Array &a, rsort::Array &t ) {
        a.keys[0] = scodes.get();
        a.keys[1] = scodes1.get();
        a.vals[0] = svals.get();
        a.vals[1] = svals.get() + n_src;
        a.vals_out = sids1.get();
        a.hist = st1.hist;
        a.look[0] = st1.look0;
        a.look[1] = st1.look1;
        a.n = n_src;
        a.tiles = src_tiles;
        a.passes = kSrcPasses;
        a.key_bits = kSrcKeyBits;
        a.key_bytes = 8;
        a.keep_keys = true;
        if ( n_tgt == 0 )
            return;
        t.keys[0] = tcodes.get();
        t.keys[1] = tcodes1.get();
        t.vals[0] = tvals.get();
        t.vals[1] = tvals.get() + n_tgt;
        t.vals_out = tids1.get();
        t.hist = st1.hist + kSrcPasses * rsort::kDigits;
        t.look[0] = st1.look0 + (std::size_t)src_tiles * rsort::kDigits;
        t.look[1] = st1.look1 + (std::size_t)src_tiles * rsort::kDigits;
        t.n = n_tgt;
        t.tiles = tgt_tiles;
        t.passes = kTgtPasses;
        t.key_bits = kTgtKeyBits;
        t.key_bytes = 4;
    };
    rsort::PassArgs pa1;
    pa1.tickets = st1.tickets;
    first_stage( pa1.a[0], pa1.a[1] );
    rsort::sort( pa1, s );

    // second stage: keys through the first order; scodes holds the sorted
    // codes (an even number of passes), tcodes the targets' keys
    DeviceBuffer<uint32_t> skeys( n_src, s ), skeys1( n_src, s ),
        tkeys( n_tgt, s ), tkeys1( n_tgt, s );
    group_keys_kernel<<<src_tiles + tgt_tiles, rsort::kThreads, 0, s>>>(
        sids1.get(), b.src_prob, n_src, src_passes2, tids1.get(), b.tgt_prob,
        b.row_key, n_tgt, tgt_passes2, src_tiles, skeys.get(), tkeys.get(),
        st2.hist );
    DTKX_HIP( hipGetLastError() );
    DeviceBuffer<int32_t> spos1( n_src, s ), tpos1( n_tgt, s );
    rsort::PassArgs pa2;
    pa2.tickets = st2.tickets;
    {
        rsort::Array &a = pa2.a[0];
        a.keys[0] = skeys.get();
        a.keys[1] = skeys1.get();
        a.vals[0] = svals.get();
        a.vals[1] = svals.get() + n_src;
        a.vals_out = spos1.get();
        a.gather_in[0] = sids1.get();
        a.gather_out[0] = out.src_perm;
        a.gather_bytes[0] = 4;
        a.gather_in[1] = scodes.get();
        a.gather_out[1] = codes_sorted.get();
        a.gather_bytes[1] = 8;
        a.hist = st2.hist;
        a.look[0] = st2.look0;
        a.look[1] = st2.look1;
        a.n = n_src;
        a.tiles = src_tiles;
        a.passes = src_passes2;
        a.key_bits = prob_bits;
        a.key_bytes = 4;
        if ( n_tgt > 0 )
        {
            rsort::Array &t = pa2.a[1];
            t.keys[0] = tkeys.get();
            t.keys[1] = tkeys1.get();
            t.vals[0] = tvals.get();
            t.vals[1] = tvals.get() + n_tgt;
            t.vals_out = tpos1.get();
            t.gather_in[0] = tids1.get();
            t.gather_out[0] = out.order;
            t.gather_bytes[0] = 4;
            t.hist = st2.hist + src_passes2 * rsort::kDigits;
            t.look[0] = st2.look0 + (std::size_t)src_tiles * rsort::kDigits;
            t.look[1] = st2.look1 + (std::size_t)src_tiles * rsort::kDigits;
            t.n = n_tgt;
            t.tiles = tgt_tiles;
            t.passes = tgt_passes2;
            t.key_bits = b.key_bits;
            t.key_bytes = 4;
        }
    }
    rsort::sort( pa2, s );
    invert_permutation( out.src_perm, n_src, out.src_pos, s );

    // one hierarchy per problem, all in one launch
    int64_t const n_int = n_src - P;
    DeviceBuffer<Node<T>> nodes( std::max<int64_t>( n_int, 1 ), s );
    DeviceBuffer<int32_t> sizes( std::max<int64_t>( n_int, 1 ), s );
    if ( n_int > 0 )
    {
        karras_kernel<T, V><<<ceil_div( n_int, kBlock ), kBlock, 0, s>>>(
            codes_sorted.get(), out.src_perm, src, nullptr, n_int,
            boxes.get(), b.src_off, P, nodes.get(), sizes.get() );
        DTKX_HIP( hipGetLastError() );
    }
    codes_sorted.release();
    if ( n_tgt == 0 )
        return;

    // rows: ordered by (row key of the problem, Morton code)
    row_offsets_kernel<<<ceil_div( n_tgt + 1, kBlock ), kBlock, 0, s>>>(
        out.order, b.tgt_prob, b.row_base, b.row_start, b.kk_prob, n_tgt,
        b.nnz, out.row_off );
    DTKX_HIP( hipGetLastError() );

    if ( info && info->query_begin )
        DTKX_HIP( hipEventRecord( info->query_begin, s ) );
    for ( auto const &l : b.launches )
    {
        QueryArgs<T> qa;
        qa.nodes = nodes.get();
        qa.node_size = kSeed ? sizes.get() : nullptr;
        qa.src = src;
        qa.tgt = tgt;
        qa.order = out.order;
        qa.row0 = l.begin;
        qa.n_rows = l.end - l.begin;
        qa.src_off = b.src_off;
        qa.tgt_prob = b.tgt_prob;
        qa.kk_prob = b.kk_prob;
        qa.row_off = out.row_off;
        qa.out_idx = out.indices;
        qa.src_pos = out.src_pos;
        qa.row_max = out.row_max;
        run_queries<T>( l.kmax, qa, l.end - l.begin, info, s );
    }
    if ( info && info->query_end )
        DTKX_HIP( hipEventRecord( info->query_end, s ) );
}

} // namespace

void knn_search_batched( int geometry, void const *src, int64_t n_src,
                         void const *tgt, int64_t n_tgt, KnnBatch const &b,
                         KnnBatchOutput const &out, KnnInfo *info,
                         hipStream_t s )
{
    DTKX_REQUIRE( b.n_problems >= 1, DTKX_ERR_INVALID_ARGUMENT );
    DTKX_REQUIRE( n_src >= b.n_problems, DTKX_ERR_EMPTY_SOURCE );
    DTKX_REQUIRE( n_src < INT32_MAX && n_tgt < INT32_MAX,
                  DTKX_ERR_UNSUPPORTED );
    if ( geometry == DTKX_GEOMETRY_FLOAT )
        knn_batched_impl<float>( src, n_src, tgt, n_tgt, b, out, info, s );
    else
        knn_batched_impl<double>( src, n_src, tgt, n_tgt, b, out, info, s );
}

void knn_search( int geometry, void const *src, int64_t n_src,
                 void const *tgt, int64_t n_tgt, int k, KnnOutput const &out,
                 KnnInfo *info, hipStream_t s )
{
    DTKX_REQUIRE( n_src > 0, DTKX_ERR_EMPTY_SOURCE );
    DTKX_REQUIRE( k >= 1 && k <= knn_max_k, DTKX_ERR_UNSUPPORTED );
    DTKX_REQUIRE( n_src < INT32_MAX && n_tgt < INT32_MAX,
                  DTKX_ERR_INVALID_ARGUMENT );
    if ( geometry == DTKX_GEOMETRY_FLOAT )
        knn_impl<float>( src, n_src, tgt, n_tgt, k, out, info, s );
    else
        knn_impl<double>( src, n_src, tgt, n_tgt, k, out, info, s );
}

} // namespace dtkx

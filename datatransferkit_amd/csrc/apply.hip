// CRS apply kernels (HBM-bound gather-dot) and the transpose.
#include "apply.hpp"

#include <hipcub/hipcub.hpp>

namespace dtkx
{
namespace
{

constexpr int kBlock = 256;

// Accumulators: plain (FP64 path, the reference's sequential row sum) or
// compensated (FP32 path: the error of each float addition is carried in a
// second float, so a long row sums to float accuracy, not k times worse).
template <typename T>
struct PlainSum
{
    T s = 0;
    __device__ void add( T v ) { s += v; }
    __device__ T value() const { return s; }
};

struct KahanSum
{
    float s = 0.f, c = 0.f;
    __device__ void add( float v )
    {
#pragma clang fp contract( off )
        float const y = v - c;
        float const t = s + y;
        c = ( t - s ) - y;
        s = t;
    }
    __device__ float value() const { return s; }
};

template <typename Acc>
struct SumOf
{
    using type = PlainSum<Acc>;
};
template <>
struct SumOf<float>
{
    using type = KahanSum;
};

// The NF field values of one source position (or one row), contiguous.
template <typename T, int N>
struct alignas( sizeof( T ) * N > 16 ? 16 : sizeof( T ) * N ) Pack
{
    T v[N];
};

// xs[p] = (x[f][perm[p]] for f < nf, zero-padded to NF), converted to XS.
// Also used for the rows of the transpose (perm = order).
template <typename X, typename XS, int NF>
__global__ void __launch_bounds__( kBlock )
    gather_fields_kernel( X const *__restrict__ x, int64_t ld, int nf,
                          int32_t const *__restrict__ perm, int64_t n,
                          Pack<XS, NF> *__restrict__ xs )
{
    int64_t const p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( p >= n )
        return;
    int64_t const s = perm ? perm[p] : p;
    Pack<XS, NF> v;
#pragma unroll
    for ( int f = 0; f < NF; ++f )
        v.v[f] = f < nf ? (XS)x[f * ld + s] : XS( 0 );
    xs[p] = v;
}

// The same packing driven by the sources: each lane reads its source's
// fields (coalesced across lanes) and writes one packed record to the
// source's position, instead of reading fields at permuted addresses.
template <typename X, typename XS, int NF>
__global__ void __launch_bounds__( kBlock )
    scatter_fields_kernel( X const *__restrict__ x, int64_t ld, int nf,
                           int32_t const *__restrict__ spos, int64_t n,
                           Pack<XS, NF> *__restrict__ xs )
{
    int64_t const s = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( s >= n )
        return;
    Pack<XS, NF> v;
#pragma unroll
    for ( int f = 0; f < NF; ++f )
        v.v[f] = f < nf ? (XS)x[f * ld + s] : XS( 0 );
    xs[spos[s]] = v;
}

// One lane per stored row (consecutive lanes: spatially adjacent targets);
// every field of the row reuses the row's coefficients and indices.
// Sequential sum per row, j ascending. Each wave owns 64 consecutive rows
// and stages their indices and coefficients through LDS kChunk columns at
// a time, so the HBM reads are runs of whole rows rather than one strided
// element per lane. 16 columns (64-byte runs of a 4-byte array) measured
// best: the FP32 apply of 1M rows of 32 takes 0.23 ms, against 0.30 ms at
// 8 columns, 0.34 ms at 32 (LDS-bound occupancy) and 0.44 ms at 4.
#ifndef DTKX_APPLY_CHUNK
#define DTKX_APPLY_CHUNK 16
#endif
constexpr int kChunk = DTKX_APPLY_CHUNK;
constexpr int kPitch = kChunk + 1; // odd row pitch: no LDS bank conflicts

template <typename C, typename XS, typename X, typename Acc, int NF>
__global__ void __launch_bounds__( kBlock )
    crs_rows_kernel( int32_t const *__restrict__ idx,
                     C const *__restrict__ coeffs, int64_t n_tgt, int kk,
                     Pack<XS, NF> const *__restrict__ xs,
                     int32_t const *__restrict__ order, X *__restrict__ y,
                     int64_t ld_tgt, int nf, Pack<X, NF> *__restrict__ yrows )
{
#pragma clang fp contract( off )
    using Sum = typename SumOf<Acc>::type;
    constexpr int kWaves = kBlock / 64;
    __shared__ int32_t s_idx[kWaves][64 * kPitch];
    __shared__ C s_c[kWaves][64 * kPitch];
    int const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    int64_t const r0 = xcd_tile() * kBlock + wave * 64;
    if ( r0 >= n_tgt )
        return; // whole wave: no block-level barrier below
    int const nrows = n_tgt - r0 < 64 ? (int)( n_tgt - r0 ) : 64;
    int32_t *const si = s_idx[wave];
    C *const sc = s_c[wave];
    Sum acc[NF];
    for ( int j0 = 0; j0 < kk; j0 += kChunk )
    {
        int const ch = kk - j0 < kChunk ? kk - j0 : kChunk;
        // runs of ch contiguous entries per row
        for ( int e = lane; e < nrows * ch; e += 64 )
        {
            int const rr = e / ch, jj = e - rr * ch;
            int64_t const g = ( r0 + rr ) * kk + j0 + jj;
            si[rr * kPitch + jj] = idx[g];
            sc[rr * kPitch + jj] = coeffs[g];
        }
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
        if ( lane < nrows )
            for ( int jj = 0; jj < ch; ++jj )
            {
                int32_t const p = si[lane * kPitch + jj];
                Acc const c = (Acc)sc[lane * kPitch + jj];
                Pack<XS, NF> const v = xs[p];
#pragma unroll
                for ( int f = 0; f < NF; ++f )
                    acc[f].add( c * (Acc)v.v[f] );
            }
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
    }
    if ( lane >= nrows )
        return;
    int64_t const r = r0 + lane;
    if ( yrows )
    {
        // row order: one contiguous record per row
        Pack<X, NF> v;
#pragma unroll
        for ( int f = 0; f < NF; ++f )
            v.v[f] = (X)acc[f].value();
        yrows[r] = v;
        return;
    }
    int64_t const t = order ? order[r] : r;
#pragma unroll
    for ( int f = 0; f < NF; ++f )
        if ( f < nf )
            y[f * ld_tgt + t] = (X)acc[f].value();
}

// y[f][i] = rec[at[i]].v[f]: results stored one record per row (or
// column) moved to the caller's order with the lanes over the caller's
// items, so the field stores are coalesced and the random accesses are
// whole-record reads. Stored straight from the rows, targets in random
// caller order scatter one value per field and row: at 1M x 1M (k = 32,
// 4 float fields) the apply took 0.211 ms against 0.142 ms for the same
// targets given in Morton order, and 0.163 ms staged.
template <typename X, int NF>
__global__ void __launch_bounds__( kBlock )
    records_to_caller_kernel( Pack<X, NF> const *__restrict__ rec,
                              int32_t const *__restrict__ at, int64_t n,
                              X *__restrict__ y, int64_t ld, int nf )
{
    int64_t const i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i >= n )
        return;
    Pack<X, NF> const v = rec[at[i]];
#pragma unroll
    for ( int f = 0; f < NF; ++f )
        if ( f < nf )
            y[f * ld + i] = v.v[f];
}

// fields processed per pass: 1, 2, 4 or 8 (the last pass zero-padded)
inline int pass_width( int remaining )
{
    return remaining >= 5 ? 8 : ( remaining >= 3 ? 4 : remaining );
}

template <typename C, typename X, typename XS, typename Acc, int NF>
void apply_pass( RowLayout const &g, C const *coeffs, X const *x,
                 int64_t ld_src, X *y, int64_t ld_tgt, int nf, XS *scratch,
                 X *yscratch, hipStream_t s )
{
    auto *xs = reinterpret_cast<Pack<XS, NF> *>( scratch );
    if ( g.spos )
        scatter_fields_kernel<X, XS, NF>
            <<<ceil_div( g.n_src, kBlock ), kBlock, 0, s>>>(
                x, ld_src, nf, g.spos, g.n_src, xs );
    else
        gather_fields_kernel<X, XS, NF>
            <<<ceil_div( g.n_src, kBlock ), kBlock, 0, s>>>(
                x, ld_src, nf, g.perm, g.n_src, xs );
    auto *yrows = reinterpret_cast<Pack<X, NF> *>( yscratch );
    crs_rows_kernel<C, XS, X, Acc, NF>
        <<<ceil_div( g.n_tgt, kBlock ), kBlock, 0, s>>>(
            g.idx, coeffs, g.n_tgt, g.kk, xs, g.order, y, ld_tgt, nf, yrows );
    if ( yrows )
        records_to_caller_kernel<X, NF>
            <<<ceil_div( g.n_tgt, kBlock ), kBlock, 0, s>>>(
                yrows, g.trow, g.n_tgt, y, ld_tgt, nf );
    DTKX_HIP( hipGetLastError() );
}

inline std::size_t align256( std::size_t b ) { return ( b + 255 ) & ~std::size_t( 255 ); }

// The apply's two scratch arrays: the position-ordered gathered fields
// (n_src records of the widest pass) and, when rows are in Morton order,
// the row results staged before they move to the targets.
template <typename X, typename XS>
void apply_scratch_sizes( RowLayout const &g, int n_fields, std::size_t &xs_bytes,
                          std::size_t &ys_bytes )
{
    int const w = pass_width( n_fields );
    xs_bytes = align256( (std::size_t)g.n_src * w * sizeof( XS ) );
    bool const staged = g.order && g.trow;
    ys_bytes = staged ? align256( (std::size_t)g.n_tgt * w * sizeof( X ) ) : 0;
}

// ws: caller-provided scratch of apply_scratch_sizes' total (no allocation),
// or null to take it from the block cache
template <typename C, typename X, typename XS, typename Acc>
void launch_apply( RowLayout const &g, C const *coeffs, X const *x,
                   int64_t ld_src, X *y, int64_t ld_tgt, int n_fields,
                   hipStream_t s, void *ws = nullptr )
{
    if ( g.n_tgt == 0 || n_fields == 0 )
        return;
    std::size_t xs_bytes, ys_bytes;
    apply_scratch_sizes<X, XS>( g, n_fields, xs_bytes, ys_bytes );
    DeviceBuffer<char> owned( ws ? 0 : xs_bytes + ys_bytes, s );
    char *const base = ws ? static_cast<char *>( ws ) : owned.get();
    XS *const scratch = reinterpret_cast<XS *>( base );
    // rows in Morton order: results staged by row, then moved to targets
    X *const yscratch =
        ys_bytes ? reinterpret_cast<X *>( base + xs_bytes ) : nullptr;
    for ( int f0 = 0; f0 < n_fields; )
    {
        int const w = pass_width( n_fields - f0 );
        int const nf = std::min( w, n_fields - f0 );
        X const *xf = x + f0 * ld_src;
        X *yf = y + f0 * ld_tgt;
        switch ( w )
        {
        case 1:
            apply_pass<C, X, XS, Acc, 1>( g, coeffs, xf, ld_src, yf, ld_tgt,
                                          nf, scratch, yscratch, s );
            break;
        case 2:
            apply_pass<C, X, XS, Acc, 2>( g, coeffs, xf, ld_src, yf, ld_tgt,
                                          nf, scratch, yscratch, s );
            break;
        case 4:
            apply_pass<C, X, XS, Acc, 4>( g, coeffs, xf, ld_src, yf, ld_tgt,
                                          nf, scratch, yscratch, s );
            break;
        default:
            apply_pass<C, X, XS, Acc, 8>( g, coeffs, xf, ld_src, yf, ld_tgt,
                                          nf, scratch, yscratch, s );
        }
        f0 += nf;
    }
}

// ------------------------------------------------------------ transpose
// Entries enumerated in canonical order e = t * kk + j (target t, rank j):
// key = the entry's column (source position), value = its storage slot.
__global__ void __launch_bounds__( kBlock )
    canonical_keys_kernel( int32_t const *__restrict__ idx,
                           int32_t const *__restrict__ trow, int64_t nnz,
                           int kk, int32_t *__restrict__ keys,
                           int32_t *__restrict__ slots )
{
    int64_t const e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( e >= nnz )
        return;
    int64_t const t = e / kk, j = e - t * kk;
    int64_t const l = ( trow ? trow[t] : t ) * kk + j;
    keys[e] = idx[l];
    slots[e] = (int32_t)l;
}

// col_offsets[p] = first index q in the sorted keys with key[q] >= p
__global__ void __launch_bounds__( kBlock )
    col_offsets_kernel( int32_t const *keys, int64_t nnz, int64_t n_src,
                        int32_t *off )
{
    int64_t q = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( q > nnz )
        return;
    int64_t const prev = q == 0 ? -1 : keys[q - 1];
    int64_t const cur = q == nnz ? n_src : keys[q];
    for ( int64_t p = prev + 1; p <= cur; ++p )
        off[p] = (int32_t)q;
}

// One lane per source position: its column's entries in canonical order,
// each times the row's gathered ybar values.
// xbar[s] = sum over column s of coef * ybar[row], one lane per column in
// the CSC's canonical order. A wave's 64 columns are contiguous in the CSC
// arrays; each step copies the next kChunk entries of every column into
// LDS (coalesced), then each lane sums its column's part. The wave runs to
// its longest column.
template <int NF>
__global__ void __launch_bounds__( kBlock )
    transpose_kernel( int32_t const *__restrict__ off,
                      int32_t const *__restrict__ rows,
                      double const *__restrict__ coefs, int64_t n_src,
                      Pack<double, NF> const *__restrict__ ys,
                      int32_t const *__restrict__ perm,
                      double *__restrict__ xbar, int64_t ld_src, int nf,
                      Pack<double, NF> *__restrict__ xrows )
{
#pragma clang fp contract( off )
    constexpr int kWaves = kBlock / 64;
    __shared__ int32_t s_row[kWaves][64 * kPitch];
    __shared__ double s_c[kWaves][64 * kPitch];
    __shared__ int32_t s_beg[kWaves][64];
    __shared__ int32_t s_len[kWaves][64];
    int const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    int64_t const p0 = xcd_tile() * kBlock + wave * 64;
    if ( p0 >= n_src )
        return; // whole wave: no block-level barrier below
    int const ncols = n_src - p0 < 64 ? (int)( n_src - p0 ) : 64;
    int32_t *const sr = s_row[wave];
    double *const sc = s_c[wave];
    int len = 0;
    if ( lane < ncols )
    {
        int32_t const b = off[p0 + lane];
        len = off[p0 + lane + 1] - b;
        s_beg[wave][lane] = b;
        s_len[wave][lane] = len;
    }
    int maxlen = len;
    for ( int o = 32; o > 0; o >>= 1 )
        maxlen = max( maxlen, __shfl_xor( maxlen, o, 64 ) );
    __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
    double acc[NF];
#pragma unroll
    for ( int f = 0; f < NF; ++f )
        acc[f] = 0.;
    for ( int j0 = 0; j0 < maxlen; j0 += kChunk )
    {
        for ( int e = lane; e < ncols * kChunk; e += 64 )
        {
            int const cc = e / kChunk, jj = e - cc * kChunk;
            if ( j0 + jj < s_len[wave][cc] )
            {
                int32_t const q = s_beg[wave][cc] + j0 + jj;
                sr[cc * kPitch + jj] = rows[q];
                sc[cc * kPitch + jj] = coefs[q];
            }
        }
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
        int const ch = len - j0 < kChunk ? len - j0 : kChunk;
        for ( int jj = 0; jj < ch; ++jj )
        {
            double const c = sc[lane * kPitch + jj];
            Pack<double, NF> const v = ys[sr[lane * kPitch + jj]];
#pragma unroll
            for ( int f = 0; f < NF; ++f )
                acc[f] += c * v.v[f];
        }
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
    }
    if ( lane >= ncols )
        return;
    if ( xrows )
    {
        // position order: one contiguous record per column
        Pack<double, NF> v;
#pragma unroll
        for ( int f = 0; f < NF; ++f )
            v.v[f] = acc[f];
        xrows[p0 + lane] = v;
        return;
    }
    int64_t const s = perm ? perm[p0 + lane] : p0 + lane;
#pragma unroll
    for ( int f = 0; f < NF; ++f )
        if ( f < nf )
            xbar[f * ld_src + s] = acc[f];
}

// rows[q] = entries[q] / kk, coefs[q] = coeffs[entries[q]]
__global__ void __launch_bounds__( kBlock )
    csc_pack_kernel( int32_t const *__restrict__ entries,
                     double const *__restrict__ coeffs, int kk, int64_t nnz,
                     int32_t *__restrict__ rows, double *__restrict__ coefs )
{
    int64_t const q = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( q >= nnz )
        return;
    int32_t const l = entries[q];
    rows[q] = l / kk;
    coefs[q] = coeffs[l];
}

template <int NF>
void transpose_pass( RowLayout const &g, CscIndex const &csc,
                     double const *coeffs, double const *ybar, int64_t ld_tgt,
                     double *xbar, int64_t ld_src, int nf, double *scratch,
                     double *xscratch, hipStream_t s )
{
    auto *ys = reinterpret_cast<Pack<double, NF> *>( scratch );
    // no targets: every column is empty and the transpose kernel writes
    // zeros, so only the packing launch (it would have zero blocks) is
    // skipped
    if ( g.n_tgt > 0 && g.order && g.trow )
        // lanes over targets: coalesced field reads, whole-record stores
        scatter_fields_kernel<double, double, NF>
            <<<ceil_div( g.n_tgt, kBlock ), kBlock, 0, s>>>(
                ybar, ld_tgt, nf, g.trow, g.n_tgt, ys );
    else if ( g.n_tgt > 0 )
        gather_fields_kernel<double, double, NF>
            <<<ceil_div( g.n_tgt, kBlock ), kBlock, 0, s>>>(
                ybar, ld_tgt, nf, g.order, g.n_tgt, ys );
    (void)coeffs; // read through the CSC's column-ordered copy
    auto *xrows = reinterpret_cast<Pack<double, NF> *>( xscratch );
    transpose_kernel<NF><<<ceil_div( g.n_src, kBlock ), kBlock, 0, s>>>(
        csc.col_offsets.get(), csc.rows.get(), csc.coefs.get(), g.n_src, ys,
        g.perm, xbar, ld_src, nf, xrows );
    if ( xrows )
        records_to_caller_kernel<double, NF>
            <<<ceil_div( g.n_src, kBlock ), kBlock, 0, s>>>(
                xrows, g.spos, g.n_src, xbar, ld_src, nf );
    DTKX_HIP( hipGetLastError() );
}

// out[spos[s] * pitch + f] = x[f][s], one lane per source
__global__ void __launch_bounds__( kBlock )
    scatter_pitched_kernel( double const *__restrict__ x, int64_t ld, int nf,
                            int32_t const *__restrict__ spos, int64_t n,
                            double *__restrict__ out, int pitch )
{
    int64_t const s = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( s >= n )
        return;
    double *const o = out + (int64_t)spos[s] * pitch;
    for ( int f = 0; f < nf; ++f )
        o[f] = x[f * ld + s];
}

__global__ void __launch_bounds__( kBlock )
    gather_pitched_kernel( double const *__restrict__ x, int64_t ld, int nf,
                           int32_t const *__restrict__ perm, int64_t n,
                           double *__restrict__ out, int pitch )
{
    int64_t const p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( p >= n )
        return;
    int64_t const s = perm ? perm[p] : p;
    for ( int f = 0; f < nf; ++f )
        out[p * pitch + f] = x[f * ld + s];
}

template <typename C>
__global__ void __launch_bounds__( kBlock )
    canonical_graph_kernel( int32_t const *__restrict__ idx,
                            int32_t const *__restrict__ trow,
                            int32_t const *__restrict__ perm,
                            C const *__restrict__ coeffs, int64_t nnz, int kk,
                            int32_t *__restrict__ idx_out,
                            C *__restrict__ coeffs_out )
{
    int64_t const e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( e >= nnz )
        return;
    int64_t const t = e / kk, j = e - t * kk;
    int64_t const l = ( trow ? trow[t] : t ) * kk + j;
    int32_t const p = idx[l];
    idx_out[e] = perm ? perm[p] : p;
    coeffs_out[e] = coeffs[l];
}

// ------------------------------------------------------------- batched
template <typename XS>
__global__ void __launch_bounds__( kBlock )
    batch_gather_kernel( double const *const *__restrict__ x,
                         int32_t const *__restrict__ perm,
                         int32_t const *__restrict__ src_prob,
                         int32_t const *__restrict__ src_off, int64_t n,
                         XS *__restrict__ xs )
{
    int64_t const p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( p >= n )
        return;
    int32_t const s = perm[p];
    int32_t const q = src_prob[s];
    xs[p] = (XS)x[q][s - src_off[q]];
}

// One lane per row, the row's entries in order (sequential sum), staged
// as in crs_rows_kernel: each step copies the next kChunk entries of every
// row of the wave into LDS (rows are contiguous in the CSR arrays, so the
// copy runs in row-long coalesced pieces), then each lane sums its row's
// part. Rows differ in length across problems; the wave runs to its
// longest row.
template <typename C, typename XS, typename Acc>
__global__ void __launch_bounds__( kBlock )
    batch_rows_kernel( int32_t const *__restrict__ idx,
                       C const *__restrict__ coeffs,
                       int64_t const *__restrict__ row_off, int64_t n_rows,
                       XS const *__restrict__ xs,
                       int32_t const *__restrict__ order,
                       int32_t const *__restrict__ tgt_prob,
                       int32_t const *__restrict__ tgt_off,
                       double *const *__restrict__ y )
{
#pragma clang fp contract( off )
    using Sum = typename SumOf<Acc>::type;
    constexpr int kWaves = kBlock / 64;
    __shared__ int32_t s_idx[kWaves][64 * kPitch];
    __shared__ C s_c[kWaves][64 * kPitch];
    __shared__ int64_t s_beg[kWaves][64];
    __shared__ int32_t s_len[kWaves][64];
    int const wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    int64_t const r0 = xcd_tile() * kBlock + wave * 64;
    if ( r0 >= n_rows )
        return; // whole wave: no block-level barrier below
    int const nrows = n_rows - r0 < 64 ? (int)( n_rows - r0 ) : 64;
    int32_t *const si = s_idx[wave];
    C *const sc = s_c[wave];
    int len = 0;
    if ( lane < nrows )
    {
        int64_t const b = row_off[r0 + lane];
        len = (int)( row_off[r0 + lane + 1] - b );
        s_beg[wave][lane] = b;
        s_len[wave][lane] = len;
    }
    int maxlen = len;
    for ( int o = 32; o > 0; o >>= 1 )
        maxlen = max( maxlen, __shfl_xor( maxlen, o, 64 ) );
    __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
    Sum acc;
    for ( int j0 = 0; j0 < maxlen; j0 += kChunk )
    {
        for ( int e = lane; e < nrows * kChunk; e += 64 )
        {
            int const rr = e / kChunk, jj = e - rr * kChunk;
            if ( j0 + jj < s_len[wave][rr] )
            {
                int64_t const g = s_beg[wave][rr] + j0 + jj;
                si[rr * kPitch + jj] = idx[g];
                sc[rr * kPitch + jj] = coeffs[g];
            }
        }
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
        int const ch = len - j0 < kChunk ? len - j0 : kChunk;
        for ( int jj = 0; jj < ch; ++jj )
            acc.add( (Acc)sc[lane * kPitch + jj] *
                     (Acc)xs[si[lane * kPitch + jj]] );
        __builtin_amdgcn_fence( __ATOMIC_RELEASE, "wavefront" );
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence( __ATOMIC_ACQUIRE, "wavefront" );
    }
    if ( lane >= nrows )
        return;
    int32_t const t = order[r0 + lane];
    int32_t const q = tgt_prob[t];
    y[q][t - tgt_off[q]] = (double)acc.value();
}

template <typename C, typename XS, typename Acc>
void launch_batch( BatchLayout const &g, C const *coeffs,
                   double const *const *x, double *const *y, XS *xs,
                   hipStream_t s )
{
    if ( g.n_tgt == 0 )
        return;
    batch_gather_kernel<XS><<<ceil_div( g.n_src, kBlock ), kBlock, 0, s>>>(
        x, g.perm, g.src_prob, g.src_off, g.n_src, xs );
    batch_rows_kernel<C, XS, Acc>
        <<<ceil_div( g.n_tgt, kBlock ), kBlock, 0, s>>>(
            g.idx, coeffs, g.row_off, g.n_tgt, xs, g.order, g.tgt_prob,
            g.tgt_off, y );
    DTKX_HIP( hipGetLastError() );
}

} // namespace

void batch_apply_f64( BatchLayout const &g, double const *coeffs,
                      double const *const *x, double *const *y, double *xs,
                      hipStream_t s )
{
    launch_batch<double, double, double>( g, coeffs, x, y, xs, s );
}

void batch_apply_f32( BatchLayout const &g, float const *coeffs,
                      double const *const *x, double *const *y, float *xs,
                      hipStream_t s )
{
    launch_batch<float, float, float>( g, coeffs, x, y, xs, s );
}

void crs_apply_f64( RowLayout const &g, double const *coeffs, double const *x,
                    int64_t ld_src, double *y, int64_t ld_tgt, int n_fields,
                    hipStream_t s )
{
    launch_apply<double, double, double, double>( g, coeffs, x, ld_src, y,
                                                  ld_tgt, n_fields, s );
}

void crs_apply_f32( RowLayout const &g, float const *coeffs, float const *x,
                    int64_t ld_src, float *y, int64_t ld_tgt, int n_fields,
                    hipStream_t s )
{
    launch_apply<float, float, float, float>( g, coeffs, x, ld_src, y, ld_tgt,
                                              n_fields, s );
}

void crs_apply_f32_dvals( RowLayout const &g, float const *coeffs,
                          double const *x, int64_t ld_src, double *y,
                          int64_t ld_tgt, int n_fields, hipStream_t s )
{
    // values are rounded to float by the gather, products and sums in float
    launch_apply<float, double, float, float>( g, coeffs, x, ld_src, y,
                                               ld_tgt, n_fields, s );
}

std::size_t crs_apply_scratch_bytes( RowLayout const &g, int n_fields,
                                     bool fp32_coeffs )
{
    if ( g.n_tgt == 0 || n_fields == 0 )
        return 0;
    std::size_t xs, ys;
    if ( fp32_coeffs )
        apply_scratch_sizes<double, float>( g, n_fields, xs, ys );
    else
        apply_scratch_sizes<double, double>( g, n_fields, xs, ys );
    return xs + ys;
}

void crs_apply_f64_ws( RowLayout const &g, double const *coeffs,
                       double const *x, int64_t ld_src, double *y,
                       int64_t ld_tgt, int n_fields, void *ws, hipStream_t s )
{
    launch_apply<double, double, double, double>( g, coeffs, x, ld_src, y,
                                                  ld_tgt, n_fields, s, ws );
}

void crs_apply_f32_dvals_ws( RowLayout const &g, float const *coeffs,
                             double const *x, int64_t ld_src, double *y,
                             int64_t ld_tgt, int n_fields, void *ws,
                             hipStream_t s )
{
    launch_apply<float, double, float, float>( g, coeffs, x, ld_src, y,
                                               ld_tgt, n_fields, s, ws );
}

void build_csc( RowLayout const &g, double const *coeffs, CscIndex &csc,
                hipStream_t s )
{
    int64_t const nnz = g.n_tgt * g.kk;
    DTKX_REQUIRE( nnz < INT32_MAX, DTKX_ERR_UNSUPPORTED );
    csc.col_offsets.resize( g.n_src + 1, s );
    csc.entries.resize( std::max<int64_t>( nnz, 1 ), s );
    csc.rows.resize( std::max<int64_t>( nnz, 1 ), s );
    csc.coefs.resize( std::max<int64_t>( nnz, 1 ), s );
    if ( nnz == 0 )
    {
        DTKX_HIP( hipMemsetAsync( csc.col_offsets.get(), 0,
                                  sizeof( int32_t ) * ( g.n_src + 1 ), s ) );
        return;
    }
    DeviceBuffer<int32_t> keys( nnz, s ), slots( nnz, s ),
        keys_sorted( nnz, s );
    canonical_keys_kernel<<<ceil_div( nnz, kBlock ), kBlock, 0, s>>>(
        g.idx, g.trow, nnz, g.kk, keys.get(), slots.get() );
    DTKX_HIP( hipGetLastError() );
    int end_bit = 1;
    while ( ( int64_t( 1 ) << end_bit ) < g.n_src )
        ++end_bit;
    // the radix sort is stable: each column keeps the canonical order
    std::size_t temp_bytes = 0;
    DTKX_HIP( hipcub::DeviceRadixSort::SortPairs(
        nullptr, temp_bytes, keys.get(), keys_sorted.get(), slots.get(),
        csc.entries.get(), static_cast<int>( nnz ), 0, end_bit, s ) );
    DeviceBuffer<char> temp( temp_bytes, s );
    DTKX_HIP( hipcub::DeviceRadixSort::SortPairs(
        temp.get(), temp_bytes, keys.get(), keys_sorted.get(), slots.get(),
        csc.entries.get(), static_cast<int>( nnz ), 0, end_bit, s ) );
    col_offsets_kernel<<<ceil_div( nnz + 1, kBlock ), kBlock, 0, s>>>(
        keys_sorted.get(), nnz, g.n_src, csc.col_offsets.get() );
    csc_pack_kernel<<<ceil_div( nnz, kBlock ), kBlock, 0, s>>>(
        csc.entries.get(), coeffs, g.kk, nnz, csc.rows.get(),
        csc.coefs.get() );
    DTKX_HIP( hipGetLastError() );
}

void crs_apply_transpose_f64( RowLayout const &g, CscIndex const &csc,
                              double const *coeffs, double const *ybar,
                              int64_t ld_tgt, double *xbar, int64_t ld_src,
                              int n_fields, hipStream_t s )
{
    if ( g.n_src == 0 || n_fields == 0 )
        return;
    DeviceBuffer<double> scratch(
        (std::size_t)std::max<int64_t>( g.n_tgt, 1 ) *
            pass_width( n_fields ),
        s );
    // columns in Morton position order: results staged by position, then
    // moved to sources
    bool const staged = g.perm && g.spos;
    DeviceBuffer<double> xscratch(
        staged ? (std::size_t)g.n_src * pass_width( n_fields ) : 0, s );
    for ( int f0 = 0; f0 < n_fields; )
    {
        int const w = pass_width( n_fields - f0 );
        int const nf = std::min( w, n_fields - f0 );
        double const *yf = ybar + f0 * ld_tgt;
        double *xf = xbar + f0 * ld_src;
        switch ( w )
        {
        case 1:
            transpose_pass<1>( g, csc, coeffs, yf, ld_tgt, xf, ld_src, nf,
                               scratch.get(), xscratch.get(), s );
            break;
        case 2:
            transpose_pass<2>( g, csc, coeffs, yf, ld_tgt, xf, ld_src, nf,
                               scratch.get(), xscratch.get(), s );
            break;
        case 4:
            transpose_pass<4>( g, csc, coeffs, yf, ld_tgt, xf, ld_src, nf,
                               scratch.get(), xscratch.get(), s );
            break;
        default:
            transpose_pass<8>( g, csc, coeffs, yf, ld_tgt, xf, ld_src, nf,
                               scratch.get(), xscratch.get(), s );
        }
        f0 += nf;
    }
}

namespace
{
// one lane per target: its index is read once for all fields, the reads
// and writes of a field run along the targets
__global__ void __launch_bounds__( kBlock )
    nn_gather_kernel( int32_t const *__restrict__ idx, int64_t n_tgt,
                      double const *__restrict__ x, int64_t ld_src,
                      double *__restrict__ y, int64_t ld_tgt, int nf )
{
    int64_t const t = xcd_tile() * kBlock + threadIdx.x;
    if ( t >= n_tgt )
        return;
    int64_t const s = idx[t];
    for ( int f = 0; f < nf; ++f )
        y[f * ld_tgt + t] = x[f * ld_src + s];
}
} // namespace

void nn_gather_f64( int32_t const *idx, int64_t n_tgt, double const *x,
                    int64_t ld_src, double *y, int64_t ld_tgt, int n_fields,
                    hipStream_t s )
{
    if ( n_tgt == 0 || n_fields == 0 )
        return;
    nn_gather_kernel<<<ceil_div( n_tgt, kBlock ), kBlock, 0, s>>>(
        idx, n_tgt, x, ld_src, y, ld_tgt, n_fields );
    DTKX_HIP( hipGetLastError() );
}

void gather_fields_f64( double const *x, int64_t ld, int nf,
                        int32_t const *perm, int32_t const *spos, int64_t n,
                        double *out, int pitch, hipStream_t s )
{
    if ( n == 0 || nf == 0 )
        return;
    if ( spos )
        scatter_pitched_kernel<<<ceil_div( n, kBlock ), kBlock, 0, s>>>(
            x, ld, nf, spos, n, out, pitch );
    else
        gather_pitched_kernel<<<ceil_div( n, kBlock ), kBlock, 0, s>>>(
            x, ld, nf, perm, n, out, pitch );
    DTKX_HIP( hipGetLastError() );
}

template <typename C>
void canonical_graph( RowLayout const &g, C const *coeffs, int32_t *idx_out,
                      C *coeffs_out, hipStream_t s )
{
    int64_t const nnz = g.n_tgt * g.kk;
    if ( nnz == 0 )
        return;
    canonical_graph_kernel<C><<<ceil_div( nnz, kBlock ), kBlock, 0, s>>>(
        g.idx, g.trow, g.perm, coeffs, nnz, g.kk, idx_out, coeffs_out );
    DTKX_HIP( hipGetLastError() );
}

template void canonical_graph<double>( RowLayout const &, double const *,
                                       int32_t *, double *, hipStream_t );
template void canonical_graph<float>( RowLayout const &, float const *,
                                      int32_t *, float *, hipStream_t );

} // namespace dtkx

// Compactly supported RBF spline: basis weights, the saddle-point operator
// C = P + M + P^T, the target evaluation (Q + N) and a restarted GMRES
// whose Krylov basis lives in device memory.
//
// Layout: an extended vector is [n_src source entries | 4 polynomial
// entries], the order of the reference's prolongation range map
// (DTK_DetailsSplineProlongationOperator.hpp:47-61). The basis is
// (restart + 1) such vectors at a stride ld (a multiple of 64 doubles).
// Every reduction runs over a fixed grid and sums its per-block partials in
// block order, so a solve is deterministic.
#include "mls_device.hpp"
#include "spline.hpp"

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

namespace dtkx
{
namespace
{

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;
// Krylov kernels. Dots (gemv_t): each block takes kRowsT rows per thread
// and a group of kGroup basis vectors, so a thread has kGroup
// independent loads per row and two wave reductions per vector for kRowsT
// rows. Updates (gemv_n, dcgs2_update, residual): kRowsN rows per thread,
// many waves. Solve at 1M sources (91 iterations), rows x group:
// 8 x 16 21.6 ms, 4 x 16 22.6, 8 x 8 20.0, 4 x 8 20.5, 16 x 8 19.9,
// 32 x 8 20.4, 8 x 4 19.8, 16 x 4 19.8; kRowsN 1, 2, 4: 21.8, 21.6, 21.9.
#ifndef DTKX_SPLINE_ROWS_T
#define DTKX_SPLINE_ROWS_T 16
#endif
#ifndef DTKX_SPLINE_GROUP
#define DTKX_SPLINE_GROUP 8
#endif
#ifndef DTKX_SPLINE_ROWS_N
#define DTKX_SPLINE_ROWS_N 2
#endif
constexpr int kRowsT = DTKX_SPLINE_ROWS_T;
constexpr int kTileT = kBlock * kRowsT;
constexpr int kGroup = DTKX_SPLINE_GROUP;
constexpr int kRowsN = DTKX_SPLINE_ROWS_N;
constexpr int kTileN = kBlock * kRowsN;
// largest number of basis vectors one orthogonalisation pass handles
constexpr int kMaxVectors = 512;
// grid of the operator kernels (grid-stride over the rows)
constexpr int kRowBlocks = 2048;

__device__ inline double wave_sum( double a )
{
    for ( int o = 32; o > 0; o >>= 1 )
        a += __shfl_xor( a, o, 64 );
    return a;
}

__device__ inline double4 as_double4( float4 p )
{
    return make_double4( p.x, p.y, p.z, 0. );
}
__device__ inline double4 as_double4( double4 p ) { return p; }

// ------------------------------------------------------------ weights
// One thread per row point: transformed coordinates in double
// (Impl.hpp:92-95), distances as computeRadius / computeWeights form them
// (Impl.hpp:121-125, 162-165), R = 1.1 max(10 DBL_EPSILON, max_l d_l)
// (Impl.hpp:118-134), phi_l = rbf(d_l / R).
__global__ void __launch_bounds__( kBlock )
    basis_weights_kernel( double4 const *src, double4 const *pts, int64_t n,
                          int kk, int32_t const *idx, int rbf, int geometry,
                          double *phi )
{
    int64_t const i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i >= n )
        return;
    double4 const t = pts[i];
    double dist[spline_k];
    double dmax = 10. * DBL_EPSILON;
    for ( int l = 0; l < kk; ++l )
    {
        double4 const s = src[idx[i * kk + l]];
        double const xr[3] = {s.x - t.x, s.y - t.y, s.z - t.z};
        dist[l] = dev::rel_distance( geometry, xr );
        if ( dist[l] > dmax )
            dmax = dist[l];
    }
    double const radius = 1.1 * dmax;
    for ( int l = 0; l < kk; ++l )
        phi[i * kk + l] = dev::rbf_eval( rbf, dist[l] / radius );
}

// ------------------------------------------------------------ operators
// Rows of (V + B) z for a basis operator B (kk entries per row) and the
// Vandermonde rows V(i) = {1, x_i, y_i, z_i} of the row points:
// y[i] = V(i) zp + sum_l phi[i][l] zs[idx[i][l]], the order of
// Thyra::add(P, M) (def.hpp:230-231). TRANS (the sources' rows only) also
// forms the block partials of P^T zs = sum_i V(i) zs[i]
// (DTK_DetailsPolynomialMatrix.hpp:136-190) into partial[p][block].
template <typename Pt, bool TRANS>
__global__ void __launch_bounds__( kBlock )
    rows_kernel( Pt const *pts, int64_t n_rows, int kk,
                 int32_t const *idx, double const *phi, double const *zs,
                 double const *zp, double *y, double *partial )
{
    double const p0 = zp[0], p1 = zp[1], p2 = zp[2], p3 = zp[3];
    double acc[4] = {0., 0., 0., 0.};
    int64_t const stride = (int64_t)gridDim.x * kBlock;
    for ( int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n_rows;
          i += stride )
    {
        double4 const v = as_double4( pts[i] );
        double yi = p0 + v.x * p1 + v.y * p2 + v.z * p3;
        double b = 0.;
        if ( kk == spline_k )
        {
            int4 const c = reinterpret_cast<int4 const *>( idx )[i];
            double2 const f0 = reinterpret_cast<double2 const *>( phi )[2 * i];
            double2 const f1 =
                reinterpret_cast<double2 const *>( phi )[2 * i + 1];
            b += f0.x * zs[c.x];
            b += f0.y * zs[c.y];
            b += f1.x * zs[c.z];
            b += f1.y * zs[c.w];
        }
        else
            for ( int l = 0; l < kk; ++l )
                b += phi[i * kk + l] * zs[idx[i * kk + l]];
        y[i] = yi + b;
        if ( TRANS )
        {
            double const zi = zs[i];
            acc[0] += zi;
            acc[1] += v.x * zi;
            acc[2] += v.y * zi;
            acc[3] += v.z * zi;
        }
    }
    if ( !TRANS )
        return;
    __shared__ double lds[4][kWaves];
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for ( int p = 0; p < 4; ++p )
    {
        double const a = wave_sum( acc[p] );
        if ( lane == 0 )
            lds[p][wave] = a;
    }
    __syncthreads();
    if ( threadIdx.x < 4 )
    {
        double a = 0.;
        for ( int w = 0; w < kWaves; ++w )
            a += lds[threadIdx.x][w];
        partial[threadIdx.x * (int64_t)gridDim.x + blockIdx.x] = a;
    }
}

// out[v] = sum_b partial[v][b] over nb blocks, one wave per v.
__global__ void __launch_bounds__( kBlock )
    reduce_kernel( double const *partial, int nb, int nv, double *out )
{
    int const v = blockIdx.x * kWaves + ( threadIdx.x >> 6 );
    int const lane = threadIdx.x & 63;
    if ( v >= nv )
        return;
    double a = 0.;
    for ( int b = lane; b < nb; b += 64 )
        a += partial[(int64_t)v * nb + b];
    a = wave_sum( a );
    if ( lane == 0 )
        out[v] = a;
}

// partial[2 v + k][row block] = sum over the block's rows of q_v[i] r_k[i]
// for the two right-hand sides r_0, r_1 and the vectors v of the block's
// group (blockIdx.y), v < nv.
__global__ void __launch_bounds__( kBlock )
    gemv_t_kernel( double const *basis, int64_t ld, int nv, int64_t n,
                   double const *r0, double const *r1, double *partial )
{
    __shared__ double lds[kWaves][2 * kGroup];
    int const v0 = blockIdx.y * kGroup;
    int const cnt = min( kGroup, nv - v0 );
    int64_t const base = blockIdx.x * (int64_t)kTileT + threadIdx.x;
    double a0[kGroup], a1[kGroup];
#pragma unroll
    for ( int g = 0; g < kGroup; ++g )
    {
        a0[g] = 0.;
        a1[g] = 0.;
    }
    // loads are unconditional: rows past n read row n - 1 with weight 0 and
    // vectors past nv read vector nv - 1 into accumulators never stored
    double const *q[kGroup];
#pragma unroll
    for ( int g = 0; g < kGroup; ++g )
        q[g] = basis + min( v0 + g, nv - 1 ) * ld;
#pragma unroll 2
    for ( int r = 0; r < kRowsT; ++r )
    {
        int64_t const i = base + r * kBlock;
        int64_t const ic = min( i, n - 1 );
        double const x0 = i < n ? r0[ic] : 0.;
        double const x1 = i < n ? r1[ic] : 0.;
#pragma unroll
        for ( int g = 0; g < kGroup; ++g )
        {
            double const qv = q[g][ic];
            a0[g] += qv * x0;
            a1[g] += qv * x1;
        }
    }
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for ( int g = 0; g < kGroup; ++g )
    {
        if ( g < cnt )
        {
            double const b0 = wave_sum( a0[g] );
            double const b1 = wave_sum( a1[g] );
            if ( lane == 0 )
            {
                lds[wave][2 * g] = b0;
                lds[wave][2 * g + 1] = b1;
            }
        }
    }
    __syncthreads();
    if ( threadIdx.x < 2 * cnt )
    {
        double a = 0.;
        for ( int w2 = 0; w2 < kWaves; ++w2 )
            a += lds[w2][threadIdx.x];
        partial[(int64_t)( 2 * v0 + threadIdx.x ) * gridDim.x + blockIdx.x] = a;
    }
}

// One step of the delayed reorthogonalisation, rows of kRowsN per thread:
//   q_j = (u - sum_{v<j} q_v s_v) / rho      (u is basis vector j)
//   v   = (w - sum_{v<j} q_v t_v - q_j gamma) / rho   (w is vector j + 1)
// with s_v = st[2 v], t_v = st[2 v + 1]; q_j and v overwrite u and w.
__global__ void __launch_bounds__( kBlock )
    dcgs2_update_kernel( double *basis, int64_t ld, int j, int64_t n,
                         double const *st, double inv_rho, double gamma )
{
    __shared__ double ss[kMaxVectors], ts[kMaxVectors];
    for ( int v = threadIdx.x; v < j; v += kBlock )
    {
        ss[v] = st[2 * v];
        ts[v] = st[2 * v + 1];
    }
    __syncthreads();
    double *const u = basis + j * ld;
    double *const w = u + ld;
    int64_t const base = blockIdx.x * (int64_t)kTileN + threadIdx.x;
    int64_t ii[kRowsN];
    double au[kRowsN], aw[kRowsN];
#pragma unroll
    for ( int r = 0; r < kRowsN; ++r )
    {
        // rows past n read row n - 1 and are not written back
        ii[r] = min( base + r * kBlock, n - 1 );
        au[r] = u[ii[r]];
        aw[r] = w[ii[r]];
    }
#pragma unroll 8
    for ( int v = 0; v < j; ++v )
    {
        double const *q = basis + v * ld;
        double const sv = ss[v], tv = ts[v];
#pragma unroll
        for ( int r = 0; r < kRowsN; ++r )
        {
            double const qv = q[ii[r]];
            au[r] -= qv * sv;
            aw[r] -= qv * tv;
        }
    }
#pragma unroll
    for ( int r = 0; r < kRowsN; ++r )
    {
        int64_t const i = base + r * kBlock;
        if ( i < n )
        {
            double const qj = au[r] * inv_rho;
            u[i] = qj;
            w[i] = ( aw[r] - qj * gamma ) * inv_rho;
        }
    }
}

// w[i] -= sum_v q_v[i] h[v], v < nv; with norm_partial, the block's
// sum of the new w[i]^2 as well.
__global__ void __launch_bounds__( kBlock )
    gemv_n_kernel( double const *basis, int64_t ld, int nv, int64_t n,
                   double const *h, double *w, double *norm_partial )
{
    __shared__ double hs[kMaxVectors];
    __shared__ double lds[kWaves];
    for ( int v = threadIdx.x; v < nv; v += kBlock )
        hs[v] = h[v];
    __syncthreads();
    int64_t const base = blockIdx.x * (int64_t)kTileN + threadIdx.x;
    double acc[kRowsN];
    int64_t ii[kRowsN];
#pragma unroll
    for ( int r = 0; r < kRowsN; ++r )
    {
        // rows past n read row n - 1 and are not written back
        ii[r] = min( base + r * kBlock, n - 1 );
        acc[r] = w[ii[r]];
    }
#pragma unroll 8
    for ( int v = 0; v < nv; ++v )
    {
        double const *q = basis + v * ld;
        double const hv = hs[v];
#pragma unroll
        for ( int r = 0; r < kRowsN; ++r )
            acc[r] -= q[ii[r]] * hv;
    }
    double s2 = 0.;
#pragma unroll
    for ( int r = 0; r < kRowsN; ++r )
    {
        int64_t const i = base + r * kBlock;
        if ( i < n )
        {
            w[i] = acc[r];
            s2 += acc[r] * acc[r];
        }
    }
    if ( !norm_partial )
        return;
    s2 = wave_sum( s2 );
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if ( lane == 0 )
        lds[wave] = s2;
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        double a = 0.;
        for ( int w2 = 0; w2 < kWaves; ++w2 )
            a += lds[w2];
        norm_partial[blockIdx.x] = a;
    }
}

// out[i] = in[i] * alpha
__global__ void __launch_bounds__( kBlock )
    scale_kernel( double const *in, int64_t n, double alpha, double *out )
{
    int64_t const i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i < n )
        out[i] = in[i] * alpha;
}

// r[i] = b[i] - r[i] and the block's sum of r[i]^2 (gemv_n's tiling).
__global__ void __launch_bounds__( kBlock )
    residual_kernel( double const *b, int64_t n, double *r,
                     double *norm_partial )
{
    __shared__ double lds[kWaves];
    int64_t const base = blockIdx.x * (int64_t)kTileN + threadIdx.x;
    double s2 = 0.;
#pragma unroll
    for ( int k = 0; k < kRowsN; ++k )
    {
        int64_t const i = base + k * kBlock;
        if ( i < n )
        {
            double const v = b[i] - r[i];
            r[i] = v;
            s2 += v * v;
        }
    }
    s2 = wave_sum( s2 );
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if ( lane == 0 )
        lds[wave] = s2;
    __syncthreads();
    if ( threadIdx.x == 0 )
    {
        double a = 0.;
        for ( int w2 = 0; w2 < kWaves; ++w2 )
            a += lds[w2];
        norm_partial[blockIdx.x] = a;
    }
}

// ------------------------------------------------------ renumbering
// Sources renumbered by their Morton position (perm: position -> source,
// pos: source -> position): row p of the position-ordered M is row
// perm[p] with its indices mapped through pos, and the point rows follow.
template <typename Pt>
__global__ void __launch_bounds__( kBlock )
    permute_rows_kernel( int64_t n_src, int kk, int32_t const *perm,
                         int32_t const *pos, int32_t const *idx,
                         double const *phi, Pt const *pts, int32_t *idx_p,
                         double *phi_p, Pt *pts_p )
{
    int64_t const p = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( p >= n_src )
        return;
    int64_t const r = perm[p];
    for ( int l = 0; l < kk; ++l )
    {
        idx_p[p * kk + l] = pos[idx[r * kk + l]];
        phi_p[p * kk + l] = phi[r * kk + l];
    }
    pts_p[p] = pts[r];
}

// out[e] = pos[idx[e]]
__global__ void __launch_bounds__( kBlock )
    remap_kernel( int64_t count, int32_t const *pos, int32_t const *idx,
                  int32_t *out )
{
    int64_t const e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( e < count )
        out[e] = pos[idx[e]];
}

// Extended vectors between source order and position order: entries below
// n_src move through perm, the polynomial entries after them are copied.
template <bool TO_POSITIONS>
__global__ void __launch_bounds__( kBlock )
    permute_vector_kernel( double const *in, int32_t const *perm,
                           int64_t n_src, int64_t n, double *out )
{
    int64_t const i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i >= n )
        return;
    if ( i >= n_src )
        out[i] = in[i];
    else if ( TO_POSITIONS )
        out[i] = in[perm[i]];
    else
        out[perm[i]] = in[i];
}

int row_blocks( int64_t n )
{
    int64_t const b = ( n + kBlock - 1 ) / kBlock;
    return (int)std::max<int64_t>( 1, std::min<int64_t>( b, kRowBlocks ) );
}

template <bool TRANS>
void launch_rows( int geometry, void const *pts, int64_t n_rows, int kk,
                  int32_t const *idx, double const *phi, double const *zs,
                  double const *zp, double *y, double *partial, int nb,
                  hipStream_t s )
{
    if ( geometry == DTKX_GEOMETRY_FLOAT )
        rows_kernel<float4, TRANS><<<nb, kBlock, 0, s>>>(
            static_cast<float4 const *>( pts ), n_rows, kk, idx, phi, zs, zp,
            y, partial );
    else
        rows_kernel<double4, TRANS><<<nb, kBlock, 0, s>>>(
            static_cast<double4 const *>( pts ), n_rows, kk, idx, phi, zs, zp,
            y, partial );
    DTKX_HIP( hipGetLastError() );
}

// C z into y; partial holds 4 * row_blocks(n_src) doubles.
void system_apply( SplineSystem const &sys, double const *z, double *y,
                   double *partial, hipStream_t s )
{
    int const nb = row_blocks( sys.n_src );
    launch_rows<true>( sys.geometry, sys.src_pts, sys.n_src, sys.kk,
                       sys.m_idx, sys.m_phi, z, z + sys.n_src, y, partial, nb,
                       s );
    reduce_kernel<<<1, kBlock, 0, s>>>( partial, nb, spline_k,
                                        y + sys.n_src );
    DTKX_HIP( hipGetLastError() );
}

int tiles_t( int64_t n ) { return (int)( ( n + kTileT - 1 ) / kTileT ); }
int tiles_n( int64_t n ) { return (int)( ( n + kTileN - 1 ) / kTileN ); }
int groups( int nv ) { return ( nv + kGroup - 1 ) / kGroup; }

} // namespace

void spline_renumber( int geometry, int64_t n_src, int kk,
                      int32_t const *perm, int32_t const *pos,
                      int32_t const *m_idx, double const *m_phi,
                      void const *src_pts, int32_t *m_idx_p, double *m_phi_p,
                      void *src_pts_p, int64_t n_tgt, int32_t const *n_idx,
                      int32_t *n_idx_p, hipStream_t s )
{
    int64_t const nb = ( n_src + kBlock - 1 ) / kBlock;
    if ( geometry == DTKX_GEOMETRY_FLOAT )
        permute_rows_kernel<float4><<<nb, kBlock, 0, s>>>(
            n_src, kk, perm, pos, m_idx, m_phi,
            static_cast<float4 const *>( src_pts ), m_idx_p, m_phi_p,
            static_cast<float4 *>( src_pts_p ) );
    else
        permute_rows_kernel<double4><<<nb, kBlock, 0, s>>>(
            n_src, kk, perm, pos, m_idx, m_phi,
            static_cast<double4 const *>( src_pts ), m_idx_p, m_phi_p,
            static_cast<double4 *>( src_pts_p ) );
    DTKX_HIP( hipGetLastError() );
    int64_t const count = n_tgt * kk;
    if ( count > 0 )
    {
        remap_kernel<<<( count + kBlock - 1 ) / kBlock, kBlock, 0, s>>>(
            count, pos, n_idx, n_idx_p );
        DTKX_HIP( hipGetLastError() );
    }
}

void spline_to_positions( double const *in, int32_t const *perm,
                          int64_t n_src, int64_t n, double *out,
                          hipStream_t s )
{
    permute_vector_kernel<true><<<( n + kBlock - 1 ) / kBlock, kBlock, 0, s>>>(
        in, perm, n_src, n, out );
    DTKX_HIP( hipGetLastError() );
}

void spline_to_sources( double const *in, int32_t const *perm, int64_t n_src,
                        int64_t n, double *out, hipStream_t s )
{
    permute_vector_kernel<false><<<( n + kBlock - 1 ) / kBlock, kBlock, 0, s>>>(
        in, perm, n_src, n, out );
    DTKX_HIP( hipGetLastError() );
}

void spline_basis_weights( double4 const *src, double4 const *pts, int64_t n,
                           int kk, int32_t const *idx, int rbf, int geometry,
                           double *phi, hipStream_t s )
{
    if ( n == 0 )
        return;
    int64_t const nb = ( n + kBlock - 1 ) / kBlock;
    basis_weights_kernel<<<nb, kBlock, 0, s>>>( src, pts, n, kk, idx, rbf,
                                                geometry, phi );
    DTKX_HIP( hipGetLastError() );
}

void spline_system_apply( SplineSystem const &sys, double const *z, double *y,
                          hipStream_t s )
{
    DeviceBuffer<double> partial( 4 * (size_t)row_blocks( sys.n_src ), s );
    system_apply( sys, z, y, partial.get(), s );
}

void spline_evaluate( SplineSystem const &sys, double const *z, double *y,
                      hipStream_t s )
{
    if ( sys.n_tgt == 0 )
        return;
    launch_rows<false>( sys.geometry, sys.tgt_pts, sys.n_tgt, sys.kk,
                        sys.n_idx, sys.n_phi, z, z + sys.n_src, y, nullptr,
                        row_blocks( sys.n_tgt ), s );
}

SplineSolver::~SplineSolver()
{
    if ( host_ )
        (void)hipHostFree( host_ );
}

void SplineSolver::reserve( int64_t n, int restart, hipStream_t s )
{
    int64_t const ld = ( n + 63 ) / 64 * 64;
    if ( ld == ld_ && restart <= restart_ )
        return;
    restart = std::max( restart, restart_ );
    ld_ = ld;
    restart_ = restart;
    size_t const nv = (size_t)restart + 2;
    basis_.resize( ( nv - 1 ) * ld, s );
    rhs_.resize( ld, s );
    res_.resize( ld, s );
    size_t const rows = 4 * (size_t)row_blocks( n );
    // two right-hand sides per basis vector in the projections
    partial_.resize(
        std::max( { 2 * nv * tiles_t( n ), (size_t)tiles_n( n ), rows } ), s );
    dots_.resize( 3 * nv, s );
    if ( host_ )
        DTKX_HIP( hipHostFree( host_ ) );
    host_ = nullptr;
    DTKX_HIP( hipHostMalloc( (void **)&host_, 3 * nv * sizeof( double ),
                             hipHostMallocDefault ) );
}

GmresInfo SplineSolver::solve( SplineSystem const &sys, double const *x,
                               double *z, GmresOptions const &opt,
                               hipStream_t s )
{
    int64_t const n = sys.n_src + spline_k;
    int const m = opt.restart;
    reserve( n, m, s );
    int64_t const ld = ld_;
    int const nbt = tiles_t( n ), nbn = tiles_n( n );
    double *const Q = basis_.get();
    double *const b = rhs_.get();
    double *const r = res_.get();
    double *const part = partial_.get();
    double *const dots = dots_.get();
    auto fetch = [&]( double const *src, int count ) {
        DTKX_HIP( hipMemcpyAsync( host_, src, count * sizeof( double ),
                                  hipMemcpyDeviceToHost, s ) );
        DTKX_HIP( hipStreamSynchronize( s ) );
    };
    // dots[2 v + k] = Q[v]^T r_k, v < nv
    auto project = [&]( int nv, double const *r0, double const *r1 ) {
        gemv_t_kernel<<<dim3( nbt, groups( nv ) ), kBlock, 0, s>>>(
            Q, ld, nv, n, r0, r1, part );
        reduce_kernel<<<( 2 * nv + kWaves - 1 ) / kWaves, kBlock, 0, s>>>(
            part, nbt, 2 * nv, dots );
        DTKX_HIP( hipGetLastError() );
    };
    // b = S x = [x; 0], z = 0 (def.hpp:200-203, 286-291)
    DTKX_HIP( hipMemcpyAsync( b, x, sys.n_src * sizeof( double ),
                              hipMemcpyDeviceToDevice, s ) );
    DTKX_HIP( hipMemsetAsync( b + sys.n_src, 0, spline_k * sizeof( double ),
                              s ) );
    DTKX_HIP( hipMemsetAsync( z, 0, n * sizeof( double ), s ) );
    GmresInfo info;
    // |b|^2: b as the only basis vector of a projection
    gemv_t_kernel<<<dim3( nbt, 1 ), kBlock, 0, s>>>( b, ld, 1, n, b, b, part );
    reduce_kernel<<<1, kBlock, 0, s>>>( part, nbt, 1, dots );
    DTKX_HIP( hipGetLastError() );
    fetch( dots, 1 );
    double const beta = std::sqrt( host_[0] );
    if ( !( beta > 0. ) )
    {
        info.converged = std::isfinite( beta );
        return info;
    }
    double const target = opt.tolerance * beta;
    // Arnoldi with classical Gram-Schmidt and delayed reorthogonalisation
    // (DCGS2): iteration j applies C to the not yet reorthogonalised,
    // unnormalised u = Q[j]; one pass of dots gives s = Q_{j-1}^T u,
    // |u|^2, t = Q_{j-1}^T w and u^T w for w = C u; one update pass then
    // reorthogonalises u into q_j = (u - Q_{j-1} s) / rho,
    // rho^2 = |u|^2 - |s|^2, and forms the next u = (I - Q_j Q_j^T) w / rho.
    // Column j - 1 of the Hessenberg matrix is final once rho is known:
    // H(0:j, j-1) += s, H(j, j-1) = rho. Column j is
    // ([t; gamma] - H(0:j, 0:j-1) s) / rho with H(j+1, j) = 1 until the
    // next iteration, gamma = (u^T w - s^T t) / rho. The basis is read
    // twice per iteration.
    std::vector<double> Ha( (size_t)( m + 1 ) * m ), R( (size_t)( m + 1 ) * m ),
        cs( m ), sn( m ), g( m + 1 ), yv( m ), hs( m + 1 );
    auto at = [&]( std::vector<double> &M, int i, int j ) -> double & {
        return M[(size_t)j * ( m + 1 ) + i];
    };
    double const *rvec = b;
    double rnorm = beta;
    int total = 0;
    int const gblocks = (int)( ( n + kBlock - 1 ) / kBlock );
    int const ublocks = tiles_n( n );
    while ( true )
    {
        // u = r / |r|
        scale_kernel<<<gblocks, kBlock, 0, s>>>( rvec, n, 1. / rnorm, Q );
        DTKX_HIP( hipGetLastError() );
        std::fill( g.begin(), g.end(), 0. );
        int j = 0; // u = Q[j]
        int k = 0; // final columns
        while ( true )
        {
            // j == m: only column m - 1 is left to finalise
            bool const last = j == m;
            if ( j == 0 && total >= opt.max_iterations )
                break;
            double *const u = Q + j * ld;
            double *const w = u + ld;
            if ( !last )
                system_apply( sys, u, w, part, s );
            project( j + 1, u, last ? u : w );
            fetch( dots, 2 * ( j + 1 ) );
            double ss = 0., st = 0.;
            for ( int v = 0; v < j; ++v )
            {
                ss += host_[2 * v] * host_[2 * v];
                st += host_[2 * v] * host_[2 * v + 1];
            }
            double const sigma = host_[2 * j], tau = host_[2 * j + 1];
            double const rho2 = sigma - ss;
            double const rho = rho2 > 0. ? std::sqrt( rho2 ) : 0.;
            if ( j == 0 )
                g[0] = rnorm * rho; // r = rnorm u = rnorm rho q_0
            else
            {
                // column j - 1 is final
                int const c = j - 1;
                for ( int v = 0; v < j; ++v )
                    at( Ha, v, c ) += host_[2 * v];
                at( Ha, j, c ) = rho;
                for ( int v = 0; v <= j; ++v )
                    at( R, v, c ) = at( Ha, v, c );
                for ( int i = 0; i < c; ++i )
                {
                    double const t0 = cs[i] * at( R, i, c ) + sn[i] * at( R, i + 1, c );
                    at( R, i + 1, c ) =
                        -sn[i] * at( R, i, c ) + cs[i] * at( R, i + 1, c );
                    at( R, i, c ) = t0;
                }
                double const a = at( R, c, c ), bb = at( R, c + 1, c );
                double const rr = std::hypot( a, bb );
                cs[c] = rr > 0. ? a / rr : 1.;
                sn[c] = rr > 0. ? bb / rr : 0.;
                at( R, c, c ) = rr;
                at( R, c + 1, c ) = 0.;
                g[c + 1] = -sn[c] * g[c];
                g[c] = cs[c] * g[c];
                ++total;
                ++info.reorthogonalized;
                k = j;
                if ( !std::isfinite( rho ) || std::fabs( g[j] ) <= target ||
                     !( rho > 0. ) || last || total >= opt.max_iterations )
                    break; // converged, invariant Krylov space, or done
            }
            if ( !( rho > 0. ) || !std::isfinite( rho ) )
                break;
            // provisional column j: ([t; gamma] - H(0:j, 0:j-1) s) / rho
            double const gamma = ( tau - st ) / rho;
            for ( int v = 0; v <= j; ++v )
            {
                double a = 0.;
                for ( int c = 0; c < j && c + 1 >= v; ++c )
                    a += at( Ha, v, c ) * host_[2 * c];
                hs[v] = a;
            }
            for ( int v = 0; v < j; ++v )
                at( Ha, v, j ) = ( host_[2 * v + 1] - hs[v] ) / rho;
            at( Ha, j, j ) = ( gamma - hs[j] ) / rho;
            at( Ha, j + 1, j ) = 1.;
            dcgs2_update_kernel<<<ublocks, kBlock, 0, s>>>( Q, ld, j, n, dots,
                                                            1. / rho, gamma );
            DTKX_HIP( hipGetLastError() );
            ++j;
        }
        // z += Q_k y, R(0:k, 0:k) y = g(0:k)
        for ( int i = k - 1; i >= 0; --i )
        {
            double t = g[i];
            for ( int l = i + 1; l < k; ++l )
                t -= at( R, i, l ) * yv[l];
            yv[i] = at( R, i, i ) != 0. ? t / at( R, i, i ) : 0.;
        }
        if ( k > 0 )
        {
            for ( int i = 0; i < k; ++i )
                host_[i] = -yv[i];
            DTKX_HIP( hipMemcpyAsync( dots, host_, k * sizeof( double ),
                                      hipMemcpyHostToDevice, s ) );
            gemv_n_kernel<<<nbn, kBlock, 0, s>>>( Q, ld, k, n, dots, z,
                                                  nullptr );
            DTKX_HIP( hipGetLastError() );
        }
        // explicit residual r = b - C z
        system_apply( sys, z, r, part, s );
        residual_kernel<<<nbn, kBlock, 0, s>>>( b, n, r, part );
        reduce_kernel<<<1, kBlock, 0, s>>>( part, nbn, 1, dots );
        DTKX_HIP( hipGetLastError() );
        fetch( dots, 1 );
        rnorm = std::sqrt( host_[0] );
        info.iterations = total;
        info.residual = rnorm / beta;
        info.converged = rnorm <= target;
        if ( info.converged || total >= opt.max_iterations || k == 0 ||
             !std::isfinite( rnorm ) )
            return info;
        rvec = r;
        ++info.restarts;
    }
}

} // namespace dtkx

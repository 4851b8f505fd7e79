// Point-format conversion kernels and host utilities shared by all stages.
#include "common.hpp"

#include <cstdio>
#include <string>

namespace dtkx
{

[[noreturn]] void raise( int status, char const *cond, char const *file,
                         int line )
{
    // strip the directory part of the file name
    std::string f( file );
    auto pos = f.find_last_of( '/' );
    if ( pos != std::string::npos )
        f = f.substr( pos + 1 );
    throw Error( status, std::string( "dtkx: " ) + cond + ", failed in " + f +
                             ", line " + std::to_string( line ) + "." );
}

bool is_device_pointer( void const *p )
{
    if ( !p )
        return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes( &attr, p );
    if ( e != hipSuccess )
    {
        (void)hipGetLastError(); // plain host memory: clear the error
        return false;
    }
    return attr.type == hipMemoryTypeDevice ||
           attr.type == hipMemoryTypeManaged ||
           attr.type == hipMemoryTypeUnified ||
           ( attr.type == hipMemoryTypeHost && attr.devicePointer != nullptr );
}

namespace
{

constexpr int kBlock = 256;

// outf (nullable): the same points rounded to float as well
template <typename V, typename T>
__global__ void __launch_bounds__( kBlock )
    to_vec4_kernel( double const *xyz, int64_t n, int layout, V *out,
                    float4 *outf )
{
    int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i >= n )
        return;
    double x, y, z;
    if ( layout == DTKX_LAYOUT_SOA )
    {
        x = xyz[i];
        y = xyz[n + i];
        z = xyz[2 * n + i];
    }
    else
    {
        x = xyz[3 * i];
        y = xyz[3 * i + 1];
        z = xyz[3 * i + 2];
    }
    V v;
    v.x = static_cast<T>( x );
    v.y = static_cast<T>( y );
    v.z = static_cast<T>( z );
    v.w = 0;
    out[i] = v;
    if ( outf )
        outf[i] = make_float4( static_cast<float>( x ), static_cast<float>( y ),
                               static_cast<float>( z ), 0.f );
}

__global__ void __launch_bounds__( kBlock )
    d4_to_f4_kernel( double4 const *in, int64_t n, float4 *out )
{
    int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i >= n )
        return;
    double4 const v = in[i];
    out[i] = float4{(float)v.x, (float)v.y, (float)v.z, 0.f};
}

__global__ void __launch_bounds__( kBlock )
    gather_d4_kernel( double4 const *in, int32_t const *perm, int64_t n,
                      double4 *out )
{
    int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i < n )
        out[i] = in[perm[i]];
}

__global__ void __launch_bounds__( kBlock )
    invert_kernel( int32_t const *perm, int64_t n, int32_t *inv )
{
    int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i < n )
        inv[perm[i]] = (int32_t)i;
}

__global__ void __launch_bounds__( kBlock )
    problem_ids_kernel( int32_t const *off, int n_prob, int64_t n,
                        int32_t *prob )
{
    int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i >= n )
        return;
    int lo = 0, hi = n_prob - 1; // largest p with off[p] <= i
    while ( lo < hi )
    {
        int const mid = ( lo + hi + 1 ) >> 1;
        if ( off[mid] <= i )
            lo = mid;
        else
            hi = mid - 1;
    }
    prob[i] = lo;
}

__global__ void __launch_bounds__( kBlock )
    concat_points_kernel( double const *const *ptrs, int32_t const *off,
                          int32_t const *prob, int64_t n, double4 *out )
{
    int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x;
    if ( i >= n )
        return;
    int32_t const q = prob[i];
    double const *p = ptrs[q] + 3 * ( i - off[q] );
    out[i] = double4{p[0], p[1], p[2], 0.0};
}

} // namespace

void problem_ids( int32_t const *off, int n_prob, int64_t n, int32_t *prob,
                  hipStream_t s )
{
    if ( n == 0 )
        return;
    problem_ids_kernel<<<ceil_div( n, kBlock ), kBlock, 0, s>>>( off, n_prob,
                                                                 n, prob );
    DTKX_HIP( hipGetLastError() );
}

void concat_points( double const *const *ptrs, int32_t const *off,
                    int32_t const *prob, int64_t n, double4 *out,
                    hipStream_t s )
{
    if ( n == 0 )
        return;
    concat_points_kernel<<<ceil_div( n, kBlock ), kBlock, 0, s>>>(
        ptrs, off, prob, n, out );
    DTKX_HIP( hipGetLastError() );
}

void gather_points( double4 const *in, int32_t const *perm, int64_t n,
                    double4 *out, hipStream_t s )
{
    if ( n == 0 )
        return;
    gather_d4_kernel<<<ceil_div( n, kBlock ), kBlock, 0, s>>>( in, perm, n,
                                                               out );
    DTKX_HIP( hipGetLastError() );
}

void invert_permutation( int32_t const *perm, int64_t n, int32_t *inv,
                         hipStream_t s )
{
    if ( n == 0 )
        return;
    invert_kernel<<<ceil_div( n, kBlock ), kBlock, 0, s>>>( perm, n, inv );
    DTKX_HIP( hipGetLastError() );
}

void double4_to_float4( double4 const *in, int64_t n, float4 *out,
                        hipStream_t s )
{
    if ( n == 0 )
        return;
    d4_to_f4_kernel<<<ceil_div( n, kBlock ), kBlock, 0, s>>>( in, n, out );
    DTKX_HIP( hipGetLastError() );
}

void points_to_float4( double const *xyz, int64_t n, int layout, float4 *out,
                       hipStream_t s )
{
    if ( n == 0 )
        return;
    to_vec4_kernel<float4, float><<<ceil_div( n, kBlock ), kBlock, 0, s>>>(
        xyz, n, layout, out, nullptr );
    DTKX_HIP( hipGetLastError() );
}

void points_to_double4( double const *xyz, int64_t n, int layout,
                        double4 *out, hipStream_t s, float4 *outf )
{
    if ( n == 0 )
        return;
    to_vec4_kernel<double4, double><<<ceil_div( n, kBlock ), kBlock, 0, s>>>(
        xyz, n, layout, out, outf );
    DTKX_HIP( hipGetLastError() );
}

} // namespace dtkx

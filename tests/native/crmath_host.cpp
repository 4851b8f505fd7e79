// Host build of datatransferkit_amd/csrc/crmath.hpp for the CPU tests that
// pin the correctly rounded atan2/sin/cos/log against mpmath and glibc.
// Build: g++ -O2 -std=c++17 -ffp-contract=off -shared -fPIC
#include "../../datatransferkit_amd/csrc/crmath.hpp"
#include <cmath>
#include <cstdint>

extern "C" {
void crm_atan2( const double *y, const double *x, double *out, int64_t n )
{
    for ( int64_t i = 0; i < n; ++i )
        out[i] = dtkx::crm::atan2( y[i], x[i] );
}
void crm_sin( const double *x, double *out, int64_t n )
{
    for ( int64_t i = 0; i < n; ++i )
        out[i] = dtkx::crm::sin( x[i] );
}
void crm_cos( const double *x, double *out, int64_t n )
{
    for ( int64_t i = 0; i < n; ++i )
        out[i] = dtkx::crm::cos( x[i] );
}
void crm_sincos( const double *x, double *s, double *c, int64_t n )
{
    for ( int64_t i = 0; i < n; ++i )
        dtkx::crm::sincos( x[i], s + i, c + i );
}
void crm_log( const double *x, double *out, int64_t n )
{
    for ( int64_t i = 0; i < n; ++i )
        out[i] = dtkx::crm::log( x[i] );
}
void libm_atan2( const double *y, const double *x, double *out, int64_t n )
{
    for ( int64_t i = 0; i < n; ++i )
        out[i] = std::atan2( y[i], x[i] );
}
void libm_sin( const double *x, double *out, int64_t n )
{
    for ( int64_t i = 0; i < n; ++i )
        out[i] = std::sin( x[i] );
}
void libm_cos( const double *x, double *out, int64_t n )
{
    for ( int64_t i = 0; i < n; ++i )
        out[i] = std::cos( x[i] );
}
}

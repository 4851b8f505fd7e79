// Adjoint of the MLS operator with respect to the target coordinates.
//
// The reference has no adjoint (its only transposed apply is the spline
// PolynomialMatrix, DTK_DetailsPolynomialMatrix.hpp:137-195). For target t
// with neighbours j (frozen), relative coordinates u_j = x_j - t, basis rows
// p_j = p(u_j), weights phi_j = W(|u_j| / R), R = 1.1 max_j |u_j|, moments
// A = sum_j phi_j p_j p_j^T and a0 = A^+ e_0, the operator row is
// c_j = phi_j (a0 . p_j) and y_f = sum_j c_j x_f(j) = e_0^T A^+ b_f with
// b_f = sum_j phi_j p_j x_f(j). Given the upstream values ybar_f of target t,
// L = sum_f ybar_f y_f depends on t through the combined field
// xbar_j = sum_f ybar_f x_f(j) only: with gamma = A^+ (sum_j phi_j p_j xbar_j)
// (the local fit of xbar), w_j = a0 . p_j and residuals
// r_j = xbar_j - gamma . p_j,
//   dL/dt = sum_j [ phi_j r_j (a0 . dp_j/dt) + w_j r_j dphi_j/dt
//                   - w_j phi_j (gamma . dp_j/dt) ],
// dp_j/dt = -grad_u p(u_j), dphi_j/dt = W'(rho_j) drho_j/dt,
// rho_j = |u_j| / R, drho_j/dt = -u_j / (|u_j| R) + (|u_j| / R^2) dR/dt and
// dR/dt = -1.1 u_* / |u_*| for the farthest neighbour u_*. This is the
// derivative of y = e_0^T A^{-1} b (dy = a0^T (db - dA A^{-1} b)), exact
// for systems with no truncated eigenvalue.
//
// A^+ is recomputed per target (no storage) with the register-resident
// cyclic Jacobi or, for Cholesky handles, the certified Cholesky inverse
// (the systems it cannot certify go to the cyclic solver), or is read back
// from the packed triangle the setup stored (store_pinv). Double geometry only: float-rounded coordinates make
// the operator piecewise constant at the float spacing.
#include "cyclic_device.hpp"
#include "gradient.hpp"

#include <tuple>

namespace dtkx
{
namespace
{

using namespace cyc;
constexpr int kBlock = 64;
constexpr int kMaxFields = 8;

// derivative of the RBFs of rbf_eval with respect to their argument
__device__ inline double rbf_deriv( int rbf, double x )
{
    double const a = 1.0 - x;
    double const a2 = a * a, a3 = a2 * a, a5 = a3 * a2, a7 = a5 * a2;
    switch ( rbf )
    {
    case DTKX_RBF_WENDLAND0:
        return -2.0 * a;
    case DTKX_RBF_WENDLAND2:
        return -20.0 * x * a3;
    case DTKX_RBF_WENDLAND4:
        return -56.0 * x * a5 * ( 5.0 * x + 1.0 );
    case DTKX_RBF_WENDLAND6:
        return -22.0 * x * a7 * ( ( 16.0 * x + 7.0 ) * x + 1.0 );
    case DTKX_RBF_WU2:
        return -7.0 * x * a3 * ( ( 3.0 * x + 9.0 ) * x + 8.0 );
    case DTKX_RBF_WU4:
        return -11.0 * x * a5 *
               ( ( ( ( 5.0 * x + 25.0 ) * x + 48.0 ) * x + 40.0 ) * x + 8.0 );
    case DTKX_RBF_BUHMANN2:
    {
        double const x2 = x * x, x3 = x2 * x;
        return x > 0 ? 8.0 * x3 * crm::log( x ) - 12.0 * x3 + 16.0 * x2 - 4.0 * x
                     : 0.0;
    }
    case DTKX_RBF_BUHMANN3:
    {
        double const x2 = x * x, x3 = x2 * x, x5 = x3 * x2, x7 = x5 * x2;
        double const sx = sqrt( x );
        return 8.0 * x7 - 504.0 / 5.0 * x5 + 4608.0 / 5.0 * x3 * sx -
               1512.0 * x3 + 3584.0 / 5.0 * x2 * sx - 168.0 / 5.0 * x;
    }
    default: // DTKX_RBF_BUHMANN4
    {
        double const x2 = x * x, x3 = x2 * x, x4 = x2 * x2, x5 = x4 * x,
                     x7 = x5 * x2;
        double const sx = sqrt( x );
        return 792.0 / 35.0 * x7 - 792.0 * x5 + 50688.0 / 35.0 * x4 * sx -
               50688.0 / 35.0 * x3 * sx + 792.0 * x3 - 792.0 / 5.0 * x;
    }
    }
}

// g[d] = sum_i a_i d p_i / d u_d (gradient of a . p(u) in u)
template <int M>
__device__ inline void basis_grad_dot( double const *a, double x, double y,
                                       double z, double *g )
{
    g[0] = g[1] = g[2] = 0.0;
    if ( M >= 4 )
    {
        g[0] = a[1];
        g[1] = a[2];
        g[2] = a[3];
    }
    if ( M >= 10 )
    {
        g[0] += 2.0 * x * a[4] + y * a[5] + z * a[6];
        g[1] += x * a[5] + 2.0 * y * a[7] + z * a[8];
        g[2] += x * a[6] + y * a[8] + 2.0 * z * a[9];
    }
}

// sum of a lane-group value over the H lanes of a system
template <int H>
__device__ inline double group_sum( double v )
{
    if ( H > 1 )
        v += __shfl_xor( v, 1, kBlock );
    return v;
}

// SOLVE: the source of A^+ -- stored (the packed triangle), cyclic Jacobi,
// or certified Cholesky. rows (nullable): the rows rows[i], i < n_list, in
// place of [0, n_tgt). A Cholesky system that cannot be certified is listed
// in fallback for a cyclic launch over the list and written there.
enum Solve
{
    kStored,
    kCyclic,
    kCholesky
};

template <int M, int H, int SOLVE>
__global__ void __launch_bounds__( kBlock )
    target_gradient_kernel( double4 const *__restrict__ src,
                            double4 const *__restrict__ tgt, int64_t n_tgt,
                            int32_t const *__restrict__ indices,
                            int32_t const *__restrict__ order, int kk,
                            int rbf, int truncation, double rcond,
                            int max_sweeps,
                            double const *__restrict__ pinv,
                            double const *__restrict__ xs, int pitch, int nf,
                            double const *__restrict__ ybar, int64_t ld_tgt,
                            double *__restrict__ d_tgt, bool accumulate,
                            int32_t const *__restrict__ rows, int64_t n_list,
                            int32_t *__restrict__ fallback,
                            int32_t *__restrict__ n_fallback )
{
    using S = SymIdx<M>;
    constexpr int R = M / H;
    constexpr int SPB = kBlock / H;
    int const h = threadIdx.x % H;
    int64_t const li = xcd_tile() * SPB + threadIdx.x / H;
    bool valid = rows ? li < n_list : li < n_tgt;
    int64_t const slot = rows ? ( valid ? rows[li] : 0 ) : li;
    int64_t const t = valid ? ( order ? order[slot] : slot ) : 0;
    int32_t const *nbr = indices + slot * kk;
    double4 const tp = valid ? tgt[t] : double4{0, 0, 0, 0};

    // radius (Impl.hpp:118-134)
    double dmax = 10. * DBL_EPSILON;
    if ( valid )
        for ( int l = 0; l < kk; ++l )
        {
            double4 const sp = src[nbr[l]];
            double const u[3] = {sp.x - tp.x, sp.y - tp.y, sp.z - tp.z};
            double const d = dev::rel_distance( DTKX_GEOMETRY_DOUBLE, u );
            dmax = d > dmax ? d : dmax;
        }
    double const radius = 1.1 * dmax;
    // dR/dt from the farthest neighbour (the first of equal maxima; zero
    // when the radius is the floor); found after the eigen-solve so that it
    // is not live across it
    auto radius_derivative = [&]( double *dr ) {
        dr[0] = dr[1] = dr[2] = 0.0;
        if ( valid )
            for ( int l = 0; l < kk; ++l )
            {
                double4 const sp = src[nbr[l]];
                double const u[3] = {sp.x - tp.x, sp.y - tp.y, sp.z - tp.z};
                if ( dev::rel_distance( DTKX_GEOMETRY_DOUBLE, u ) == dmax )
                {
#pragma unroll
                    for ( int c = 0; c < 3; ++c )
                        dr[c] = -1.1 * u[c] / dmax;
                    break;
                }
            }
    };

    // a0 = A^+ e_0 and gamma = A^+ bbar; the upstream values are loaded
    // after the eigen-solve so they are not live across it
    double a0[M], gamma[M], bbar[M];
    double yb[kMaxFields];
    auto load_upstream = [&]() {
#pragma unroll
        for ( int f = 0; f < kMaxFields; ++f )
            yb[f] = ( valid && f < nf ) ? ybar[f * ld_tgt + t] : 0.0;
    };
    // bbar = sum_j phi_j p_j xbar_j, neighbours split over the H lanes
    auto form_bbar = [&]() {
#pragma unroll
        for ( int i = 0; i < M; ++i )
            bbar[i] = 0.0;
        if ( valid )
            for ( int l = h; l < kk; l += H )
            {
                int32_t const pj = nbr[l];
                double P[M];
                double const phi = nbr_row<M, false>(
                    DTKX_GEOMETRY_DOUBLE, rbf, radius, src[pj], tp, P );
                double xb = 0.0;
#pragma unroll
                for ( int f = 0; f < kMaxFields; ++f )
                    if ( f < nf )
                        xb = fma( yb[f], xs[(int64_t)pj * pitch + f], xb );
#pragma unroll
                for ( int i = 0; i < M; ++i )
                    bbar[i] = fma( P[i] * phi, xb, bbar[i] );
            }
#pragma unroll
        for ( int i = 0; i < M; ++i )
            bbar[i] = group_sum<H>( bbar[i] );
    };

    if constexpr ( SOLVE == kStored )
    {
        load_upstream();
        form_bbar();
        double const *pt = pinv + ( valid ? slot : 0 ) * S::size;
#pragma unroll
        for ( int i = 0; i < M; ++i )
        {
            double g = 0.0;
#pragma unroll
            for ( int j = 0; j < M; ++j )
                g = fma( pt[S::at( i, j )], bbar[j], g );
            gamma[i] = g;
            a0[i] = pt[S::at( 0, i )];
        }
    }
    else if constexpr ( SOLVE == kCholesky )
    {
        double A[S::size];
#pragma unroll
        for ( int e = 0; e < S::size; ++e )
            A[e] = 0.0;
        if ( valid )
            for ( int l = 0; l < kk; ++l )
            {
                double P[M];
                double const phi = nbr_row<M, false>(
                    DTKX_GEOMETRY_DOUBLE, rbf, radius, src[nbr[l]], tp, P );
#pragma unroll
                for ( int r = 0; r < M; ++r )
                {
                    double const pr = P[r] * phi;
#pragma unroll
                    for ( int c = r; c < M; ++c )
                        A[S::at( r, c )] = fma( pr, P[c], A[S::at( r, c )] );
                }
            }
        double trace = 0.0;
#pragma unroll
        for ( int i = 0; i < M; ++i )
            trace += A[S::at( i, i )];
        if ( !valid )
#pragma unroll
            for ( int i = 0; i < M; ++i )
                A[S::at( i, i )] = 1.0; // keeps the factorisation finite
        bool const ok = cholesky_inverse<M>( A, trace, truncation, rcond );
        if ( valid && !ok )
        {
            // left to the cyclic launch over the list
            fallback[atomicAdd( n_fallback, 1 )] = (int32_t)slot;
            valid = false;
        }
        load_upstream();
        form_bbar();
        double e0[M];
#pragma unroll
        for ( int i = 0; i < M; ++i )
            e0[i] = i == 0 ? 1.0 : 0.0;
        apply_vvt<M>( A, e0, a0 );
        apply_vvt<M>( A, bbar, gamma );
    }
    else
    {
        double A[S::size];
#pragma unroll
        for ( int e = 0; e < S::size; ++e )
            A[e] = 0.0;
        if ( valid )
            for ( int l = 0; l < kk; ++l )
            {
                double P[M];
                double const phi = nbr_row<M, false>(
                    DTKX_GEOMETRY_DOUBLE, rbf, radius, src[nbr[l]], tp, P );
#pragma unroll
                for ( int r = 0; r < M; ++r )
                {
                    double const pr = P[r] * phi;
#pragma unroll
                    for ( int c = r; c < M; ++c )
                        A[S::at( r, c )] = fma( pr, P[c], A[S::at( r, c )] );
                }
            }
        double V[R * M];
#pragma unroll
        for ( int e = 0; e < R * M; ++e )
            V[e] = ( h * R + e / M == e % M ) ? 1.0 : 0.0;
        // each system stops at its own convergence, so its result does not
        // depend on which systems share its wave
        int sweeps = 0;
        bool done = !valid || converged<M>( A );
        for ( int pass = 0; !__all( done ) && pass < max_sweeps; ++pass )
        {
            sweep_all<M, R>( A, V, done );
            sweeps += done ? 0 : 1;
            done = done || converged<M>( A );
        }
        double thr = DBL_EPSILON;
        if ( truncation == DTKX_TRUNCATION_RELATIVE )
        {
            double lmax = 0.;
#pragma unroll
            for ( int k = 0; k < M; ++k )
                lmax = fmax( lmax, fabs( A[S::at( k, k )] ) );
            thr = rcond * lmax;
        }
        // q0_k = w_k V(0, k), qb_k = w_k (V^T bbar)_k
        double wk[M], q0[M], qb[M];
#pragma unroll
        for ( int k = 0; k < M; ++k )
        {
            double const lk = A[S::at( k, k )];
            wk[k] = fabs( lk ) >= thr ? 1.0 / lk : 0.0;
            double v0k = V[k];
            if ( H > 1 )
                v0k = __shfl( v0k, ( threadIdx.x / H ) * H, kBlock );
            q0[k] = wk[k] * v0k;
        }
        load_upstream();
        form_bbar();
#pragma unroll
        for ( int k = 0; k < M; ++k )
        {
            double vb = 0.0;
#pragma unroll
            for ( int r = 0; r < R; ++r )
                vb = fma( V[r * M + k], bbar[h * R + r], vb );
            qb[k] = wk[k] * group_sum<H>( vb );
        }
        // own rows of a0 = V q0 and gamma = V qb, then exchange halves
        double a0o[R], go[R];
#pragma unroll
        for ( int r = 0; r < R; ++r )
        {
            double x0 = 0.0, xg = 0.0;
#pragma unroll
            for ( int k = 0; k < M; ++k )
            {
                x0 = fma( V[r * M + k], q0[k], x0 );
                xg = fma( V[r * M + k], qb[k], xg );
            }
            a0o[r] = x0;
            go[r] = xg;
        }
#pragma unroll
        for ( int r = 0; r < R; ++r )
        {
            double const oa = H > 1 ? __shfl_xor( a0o[r], 1, kBlock ) : a0o[r];
            double const og = H > 1 ? __shfl_xor( go[r], 1, kBlock ) : go[r];
            a0[h * R + r] = a0o[r];
            gamma[h * R + r] = go[r];
            if ( H > 1 )
            {
                a0[( 1 - h ) * R + r] = oa;
                gamma[( 1 - h ) * R + r] = og;
            }
        }
    }

    // dL/dt, neighbours split over the H lanes
    double dr[3];
    radius_derivative( dr );
    double grad[3] = {0.0, 0.0, 0.0};
    if ( valid )
        for ( int l = h; l < kk; l += H )
        {
            int32_t const pj = nbr[l];
            double4 const sp = src[pj];
            double const u[3] = {sp.x - tp.x, sp.y - tp.y, sp.z - tp.z};
            double const d = dev::rel_distance( DTKX_GEOMETRY_DOUBLE, u );
            double const rho = d / radius;
            double P[M];
            basis_fast<M, double>( u[0], u[1], u[2], P );
            double const phi = dev::rbf_eval( rbf, rho );
            double const dphi = rbf_deriv( rbf, rho );
            double xb = 0.0;
#pragma unroll
            for ( int f = 0; f < kMaxFields; ++f )
                if ( f < nf )
                    xb = fma( yb[f], xs[(int64_t)pj * pitch + f], xb );
            double w = 0.0, gp = 0.0;
#pragma unroll
            for ( int i = 0; i < M; ++i )
            {
                w = fma( a0[i], P[i], w );
                gp = fma( gamma[i], P[i], gp );
            }
            double const res = xb - gp;
            double ga[3], gg[3];
            basis_grad_dot<M>( a0, u[0], u[1], u[2], ga );
            basis_grad_dot<M>( gamma, u[0], u[1], u[2], gg );
            double const inv_dr = d > 0 ? 1.0 / ( d * radius ) : 0.0;
            double const k2 = d / ( radius * radius );
#pragma unroll
            for ( int c = 0; c < 3; ++c )
            {
                double const drho = -u[c] * inv_dr - k2 * dr[c];
                grad[c] += -phi * res * ga[c] + w * res * dphi * drho +
                           w * phi * gg[c];
            }
        }
#pragma unroll
    for ( int c = 0; c < 3; ++c )
        grad[c] = group_sum<H>( grad[c] );
    if ( valid && h == 0 )
#pragma unroll
        for ( int c = 0; c < 3; ++c )
        {
            double const prev = accumulate ? d_tgt[t * 3 + c] : 0.0;
            d_tgt[t * 3 + c] = prev + grad[c];
        }
}

template <int M, int H>
void launch_gradient( GradientArgs const &a, double const *xs, int pitch,
                      int nf, double const *ybar, bool accumulate,
                      hipStream_t s )
{
    int const sweeps = a.max_sweeps > 0 ? a.max_sweeps : 30;
    auto args = [&]( int32_t const *rows, int64_t n_list, int32_t *fb,
                     int32_t *n_fb ) {
        return std::make_tuple( a.src, a.tgt, a.n_tgt, a.indices, a.order,
                                a.kk, a.rbf, a.truncation, a.rcond, sweeps,
                                a.pinv, xs, pitch, nf, ybar, a.ld_tgt, a.d_tgt,
                                accumulate, rows, n_list, fb, n_fb );
    };
    auto launch = [&]( auto kernel, int64_t n, int spb, auto tup ) {
        if ( n == 0 )
            return;
        dim3 const grid( ceil_div( n, spb ) ), block( kBlock );
        std::apply(
            [&]( auto... xs_ ) { kernel<<<grid, block, 0, s>>>( xs_... ); },
            tup );
        DTKX_HIP( hipGetLastError() );
    };
    if ( a.pinv )
    {
        launch( target_gradient_kernel<M, 1, kStored>, a.n_tgt, kBlock,
                args( nullptr, 0, nullptr, nullptr ) );
        return;
    }
    if ( a.solver != DTKX_SOLVER_CHOLESKY )
    {
        launch( target_gradient_kernel<M, H, kCyclic>, a.n_tgt, kBlock / H,
                args( nullptr, 0, nullptr, nullptr ) );
        return;
    }
    // certified Cholesky, then the cyclic solver on the rows it lists (the
    // count is read back: the list is usually empty)
    DeviceBuffer<int32_t> list( a.n_tgt, s ), count( 1, s );
    DTKX_HIP( hipMemsetAsync( count.get(), 0, sizeof( int32_t ), s ) );
    launch( target_gradient_kernel<M, 1, kCholesky>, a.n_tgt, kBlock,
            args( nullptr, 0, list.get(), count.get() ) );
    int32_t n_list = 0;
    DTKX_HIP( hipMemcpyAsync( &n_list, count.get(), sizeof( int32_t ),
                              hipMemcpyDeviceToHost, s ) );
    DTKX_HIP( hipStreamSynchronize( s ) );
    launch( target_gradient_kernel<M, H, kCyclic>, n_list, kBlock / H,
            args( list.get(), n_list, nullptr, nullptr ) );
}

} // namespace

void target_gradient( GradientArgs const &a, double const *x, int64_t ld_src,
                      int n_fields, double const *ybar, hipStream_t s )
{
    if ( a.n_tgt == 0 )
        return;
    // field chunks of up to kMaxFields: the gradient is linear in the
    // combined field, so the chunks' contributions add
    int const pitch = std::min( n_fields, kMaxFields );
    DeviceBuffer<double> xs( (std::size_t)a.n_src * pitch, s );
    for ( int f0 = 0; f0 < n_fields; f0 += kMaxFields )
    {
        int const nf = std::min( kMaxFields, n_fields - f0 );
        gather_fields_f64( x + f0 * ld_src, ld_src, nf, a.perm, a.spos,
                           a.n_src, xs.get(), pitch, s );
        double const *yf = ybar + f0 * a.ld_tgt;
        bool const acc = f0 > 0;
        switch ( a.m )
        {
        case 1:
            launch_gradient<1, 1>( a, xs.get(), pitch, nf, yf, acc, s );
            break;
        case 4:
            launch_gradient<4, 1>( a, xs.get(), pitch, nf, yf, acc, s );
            break;
        default:
            launch_gradient<10, 2>( a, xs.get(), pitch, nf, yf, acc, s );
        }
    }
}

} // namespace dtkx

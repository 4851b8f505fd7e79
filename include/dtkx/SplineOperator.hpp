/*
 * dtkx — C++ drop-in for DataTransferKit's SplineOperator.
 *
 * Header-only, above the C ABI in dtkx/mls.h (dtkx_spline_*); no HIP or
 * torch types appear here.
 *
 * Mirrors (DataTransferKit, packages/Meshfree/src):
 *   DTK_SplineOperator_decl.hpp:40-113
 *       template <DeviceType, RBF = Wendland<0>,
 *                 PolynomialBasis = MultivariatePolynomialBasis<Linear, 3>>
 *       class SplineOperator : public PointCloudOperator<DeviceType>;
 *       static_assert: linear basis only (:47-49)
 *       ctor(comm, source_points, target_points); apply(...) const
 *   DTK_SplineOperator_def.hpp:170-295 (A = (Q + N) (P + M + P^T)^-1 S,
 *       Belos pseudo-block GMRES, tolerance 1e-10)
 *
 * Differences: one GPU and a one-rank communicator, so the constructor
 * takes no communicator; points are [n][3] doubles (host or device
 * pointers); the solve runs on the GPU (restarted GMRES, same tolerance);
 * apply also takes several fields at once.
 */
#ifndef DTKX_SPLINE_OPERATOR_HPP
#define DTKX_SPLINE_OPERATOR_HPP

#include "dtkx/MovingLeastSquaresOperator.hpp"
#include "dtkx/mls.h"

#include <cstdint>
#include <type_traits>

namespace dtkx
{

// Solver and geometry options beyond the template policies; the defaults
// are the reference's (float points, GMRES to 1e-10).
inline dtkx_spline_options default_spline_options()
{
    dtkx_spline_options o;
    check( dtkx_spline_options_init( &o ) );
    return o;
}

template <typename CompactlySupportedRadialBasisFunction = Wendland<0>,
          typename PolynomialBasis = MultivariatePolynomialBasis<Linear, 3>>
class SplineOperator : public PointCloudOperator
{
    static_assert(
        std::is_same<PolynomialBasis, MultivariatePolynomialBasis<Linear, 3>>::value,
        "Only implemented for linear basis functions!" );

  public:
    using polynomial_basis = PolynomialBasis;

    // source_points: [n_source][3], target_points: [n_target][3] doubles
    SplineOperator( double const *source_points, std::int64_t n_source,
                    double const *target_points, std::int64_t n_target,
                    dtkx_spline_options options = default_spline_options() )
        : _n_source( n_source )
        , _n_target( n_target )
    {
        options.rbf = Details::RbfId<CompactlySupportedRadialBasisFunction>::value;
        check( dtkx_spline_create( source_points, n_source, target_points,
                                   n_target, &options, &_handle ) );
    }

    SplineOperator( SplineOperator const & ) = delete;
    SplineOperator &operator=( SplineOperator const & ) = delete;

    ~SplineOperator() override
    {
        if ( _handle )
            dtkx_spline_destroy( _handle );
    }

    // target_values = A source_values
    void apply( double const *source_values,
                double *target_values ) const override
    {
        apply( source_values, target_values, 1 );
    }

    // n_fields contiguous fields ([n_fields][n_source] -> [n_fields][n_target])
    void apply( double const *source_values, double *target_values,
                int n_fields ) const
    {
        check( dtkx_spline_apply( _handle, source_values, target_values,
                                  n_fields, _n_source, _n_target ) );
    }

    // the last apply's solver report
    dtkx_spline_info info() const
    {
        dtkx_spline_info i;
        check( dtkx_spline_get_info( _handle, &i ) );
        return i;
    }

    std::int64_t n_source() const { return _n_source; }
    std::int64_t n_target() const { return _n_target; }

  private:
    std::int64_t _n_source;
    std::int64_t _n_target;
    dtkx_spline *_handle = nullptr;
};

} // namespace dtkx

#endif

/*
 * dtkx — C++ drop-in for DataTransferKit's MovingLeastSquaresOperator.
 *
 * Header-only, above the C ABI in dtkx/mls.h; no HIP or torch types appear
 * here, so a caller compiles it with any C++17 compiler and links libdtkx.so.
 *
 * Mirrors (DataTransferKit, packages/Meshfree/src):
 *   DTK_PointCloudOperator.hpp:24-37          PointCloudOperator::apply
 *   DTK_MovingLeastSquaresOperator_decl.hpp:36-63
 *       template <DeviceType, RBF = Wendland<0>,
 *                 Basis = MultivariatePolynomialBasis<Linear, 3>>
 *       class MovingLeastSquaresOperator; typedefs polynomial_basis,
 *       radial_basis_function; ctor(comm, source_points, target_points);
 *       apply(source_values, target_values) const
 *   DTK_CompactlySupportedRadialBasisFunctions.hpp  Wendland<k>, Wu<k>,
 *                                                   Buhmann<k> policies
 *   DTK_MultivariatePolynomialBasis.hpp             Constant/Linear/Quadratic,
 *                                                   MultivariatePolynomialBasis
 *   packages/Utils/src/DTK_DBC.hpp:36               DataTransferKitException
 *
 * Differences: the device is one GPU with a one-rank communicator, so the
 * constructor takes no communicator; points are [n][3] doubles (host or
 * device pointers); apply also takes several fields at once.
 */
#ifndef DTKX_MOVING_LEAST_SQUARES_OPERATOR_HPP
#define DTKX_MOVING_LEAST_SQUARES_OPERATOR_HPP

#include "dtkx/mls.h"

#include <cstdint>
#include <stdexcept>
#include <string>

namespace dtkx
{

// ------------------------------------------------------------ exceptions
class DataTransferKitException : public std::logic_error
{
  public:
    DataTransferKitException( int status, std::string const &what )
        : std::logic_error( what )
        , _status( status )
    {
    }
    int status() const { return _status; }

  private:
    int _status;
};

inline void check( int status )
{
    if ( status != DTKX_SUCCESS )
        throw DataTransferKitException( status, dtkx_last_error() );
}

// ------------------------------------------------------------- policies
struct Constant
{
};
struct Linear
{
};
struct Quadratic
{
};

namespace Details
{
template <typename Basis, int DIM>
struct Size;
template <int DIM>
struct Size<Constant, DIM>
{
    static constexpr int value = 1;
};
template <>
struct Size<Linear, 3>
{
    static constexpr int value = 4;
};
template <>
struct Size<Quadratic, 3>
{
    static constexpr int value = 10;
};
} // namespace Details

template <typename Basis, int DIM>
struct MultivariatePolynomialBasis
{
    static_assert( DIM == 3, "the MLS operator is three-dimensional" );
    static constexpr int size = Details::Size<Basis, DIM>::value;
};

template <int k>
struct Wendland;
template <int k>
struct Wu;
template <int k>
struct Buhmann;

namespace Details
{
template <typename RBF>
struct RbfId;
template <>
struct RbfId<Wendland<0>>
{
    static constexpr int value = DTKX_RBF_WENDLAND0;
};
template <>
struct RbfId<Wendland<2>>
{
    static constexpr int value = DTKX_RBF_WENDLAND2;
};
template <>
struct RbfId<Wendland<4>>
{
    static constexpr int value = DTKX_RBF_WENDLAND4;
};
template <>
struct RbfId<Wendland<6>>
{
    static constexpr int value = DTKX_RBF_WENDLAND6;
};
template <>
struct RbfId<Wu<2>>
{
    static constexpr int value = DTKX_RBF_WU2;
};
template <>
struct RbfId<Wu<4>>
{
    static constexpr int value = DTKX_RBF_WU4;
};
template <>
struct RbfId<Buhmann<2>>
{
    static constexpr int value = DTKX_RBF_BUHMANN2;
};
template <>
struct RbfId<Buhmann<3>>
{
    static constexpr int value = DTKX_RBF_BUHMANN3;
};
template <>
struct RbfId<Buhmann<4>>
{
    static constexpr int value = DTKX_RBF_BUHMANN4;
};
} // namespace Details

// ------------------------------------------------------------ operators
class PointCloudOperator
{
  public:
    virtual ~PointCloudOperator() = default;
    // target_values[i] = sum over neighbours of coeff * source_values[...]
    virtual void apply( double const *source_values,
                        double *target_values ) const = 0;
};

// Options beyond the template policies; defaults are the reference's
// semantics (k = basis size, float geometry, classical Jacobi, absolute
// DBL_EPSILON truncation).
inline dtkx_mls_options default_options()
{
    dtkx_mls_options o;
    check( dtkx_mls_options_init( &o ) );
    return o;
}

template <typename CompactlySupportedRadialBasisFunction = Wendland<0>,
          typename PolynomialBasis = MultivariatePolynomialBasis<Linear, 3>>
class MovingLeastSquaresOperator : public PointCloudOperator
{
  public:
    using polynomial_basis = PolynomialBasis;
    using radial_basis_function = CompactlySupportedRadialBasisFunction;

    // source_points: [n_source][3], target_points: [n_target][3] doubles
    MovingLeastSquaresOperator( double const *source_points,
                                std::int64_t n_source,
                                double const *target_points,
                                std::int64_t n_target,
                                dtkx_mls_options options = default_options() )
        : _n_source( n_source )
        , _n_target( n_target )
    {
        options.basis = PolynomialBasis::size;
        options.rbf = Details::RbfId<CompactlySupportedRadialBasisFunction>::value;
        check( dtkx_mls_create( source_points, n_source, target_points,
                                n_target, &options, &_handle ) );
    }

    MovingLeastSquaresOperator( MovingLeastSquaresOperator const & ) = delete;
    MovingLeastSquaresOperator &
    operator=( MovingLeastSquaresOperator const & ) = delete;

    ~MovingLeastSquaresOperator() override
    {
        if ( _handle )
            dtkx_mls_destroy( _handle );
    }

    // the reference's apply: one field of n_source values -> n_target values
    void apply( double const *source_values,
                double *target_values ) const override
    {
        apply( source_values, target_values, 1 );
    }

    // n_fields contiguous fields ([n_fields][n_source] -> [n_fields][n_target])
    void apply( double const *source_values, double *target_values,
                int n_fields ) const
    {
        check( dtkx_mls_apply( _handle, source_values, target_values,
                               n_fields, _n_source, _n_target ) );
    }

    void apply_transpose( double const *target_bar, double *source_bar,
                          int n_fields = 1 ) const
    {
        check( dtkx_mls_apply_transpose( _handle, target_bar, source_bar,
                                         n_fields, _n_target, _n_source ) );
    }

    dtkx_mls_stats stats() const
    {
        dtkx_mls_stats s;
        check( dtkx_mls_get_stats( _handle, &s ) );
        return s;
    }

    dtkx_mls const *handle() const { return _handle; }
    std::int64_t n_source() const { return _n_source; }
    std::int64_t n_target() const { return _n_target; }

  private:
    std::int64_t _n_source;
    std::int64_t _n_target;
    dtkx_mls *_handle = nullptr;
};

} // namespace dtkx

#endif

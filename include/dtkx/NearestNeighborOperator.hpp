/*
 * dtkx — C++ drop-in for DataTransferKit's NearestNeighborOperator.
 *
 * Header-only, above the C ABI in dtkx/mls.h (dtkx_nn_*); no HIP or torch
 * types appear here.
 *
 * Mirrors (DataTransferKit, packages/Meshfree/src):
 *   DTK_NearestNeighborOperator_decl.hpp:24-50
 *       template <DeviceType> class NearestNeighborOperator
 *           : public PointCloudOperator<DeviceType>;
 *       ctor(comm, source_points, target_points);
 *       apply(source_values, target_values) const
 *   DTK_NearestNeighborOperator_def.hpp:23-86 (k = 1 query, fetch)
 *
 * Differences: one GPU and a one-rank communicator, so the constructor
 * takes no communicator; points are [n][3] doubles (host or device
 * pointers); apply also takes several fields at once.
 */
#ifndef DTKX_NEAREST_NEIGHBOR_OPERATOR_HPP
#define DTKX_NEAREST_NEIGHBOR_OPERATOR_HPP

#include "dtkx/MovingLeastSquaresOperator.hpp"
#include "dtkx/mls.h"

#include <cstdint>

namespace dtkx
{

class NearestNeighborOperator : public PointCloudOperator
{
  public:
    // source_points: [n_source][3], target_points: [n_target][3] doubles;
    // the search uses the reference's float points unless geometry says
    // DTKX_GEOMETRY_DOUBLE
    NearestNeighborOperator( double const *source_points,
                             std::int64_t n_source,
                             double const *target_points,
                             std::int64_t n_target,
                             int geometry = DTKX_GEOMETRY_FLOAT,
                             void *stream = nullptr )
        : _n_source( n_source )
        , _n_target( n_target )
    {
        check( dtkx_nn_create( source_points, n_source, target_points,
                               n_target, geometry, DTKX_LAYOUT_AOS, stream,
                               &_handle ) );
    }

    NearestNeighborOperator( NearestNeighborOperator const & ) = delete;
    NearestNeighborOperator &
    operator=( NearestNeighborOperator const & ) = delete;

    ~NearestNeighborOperator() override
    {
        if ( _handle )
            dtkx_nn_destroy( _handle );
    }

    // target_values[i] = source_values[nearest source of target i]
    void apply( double const *source_values,
                double *target_values ) const override
    {
        apply( source_values, target_values, 1 );
    }

    // n_fields contiguous fields ([n_fields][n_source] -> [n_fields][n_target])
    void apply( double const *source_values, double *target_values,
                int n_fields ) const
    {
        check( dtkx_nn_apply( _handle, source_values, target_values, n_fields,
                              _n_source, _n_target ) );
    }

    std::int64_t n_source() const { return _n_source; }
    std::int64_t n_target() const { return _n_target; }

  private:
    std::int64_t _n_source;
    std::int64_t _n_target;
    dtkx_nn *_handle = nullptr;
};

} // namespace dtkx

#endif

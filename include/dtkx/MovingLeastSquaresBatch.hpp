/*
 * dtkx — many independent MLS operators built and applied together.
 *
 * Header-only, above the C ABI in dtkx/mls.h (dtkx_mls_batch_*); no HIP or
 * torch types appear here. No counterpart in the reference, where each
 * point-cloud pair gets its own MovingLeastSquaresOperator
 * (DTK_MovingLeastSquaresOperator_decl.hpp:36-63): here one segmented tree
 * build and one query launch per neighbour-list size serve every problem,
 * the setup runs one launch per basis, the apply replays a captured HIP
 * graph, and rebuild() builds every operator again from moved points as one
 * graph replay. Each problem's result has the bits of a
 * MovingLeastSquaresOperator with the same options.
 *
 * Points and values are DEVICE pointers (the single operator also takes host
 * arrays; the batch does not stage them). Two forms of the same batch:
 *   - one pointer pair per problem (dtkx_mls_problem records), or
 *   - every problem's points, and values, concatenated in problem order in
 *     one array each, with per-problem counts.
 */
#ifndef DTKX_MOVING_LEAST_SQUARES_BATCH_HPP
#define DTKX_MOVING_LEAST_SQUARES_BATCH_HPP

#include "dtkx/MovingLeastSquaresOperator.hpp"

#include <cstddef>
#include <cstdint>
#include <vector>

namespace dtkx
{

template <typename CompactlySupportedRadialBasisFunction = Wendland<0>>
class MovingLeastSquaresBatch
{
  public:
    using radial_basis_function = CompactlySupportedRadialBasisFunction;

    // One record per problem: device points [n][3], basis size
    // (DTKX_BASIS_CONSTANT, _LINEAR or _QUADRATIC) and k (0 = basis size).
    explicit MovingLeastSquaresBatch(
        std::vector<dtkx_mls_problem> const &problems,
        dtkx_mls_options options = default_options() )
    {
        create( problems, options );
    }

    // Concatenated points: source_points [sum n_source][3] and
    // target_points [sum n_target][3] in problem order; problem p owns
    // n_source[p] sources and n_target[p] targets, with basis[p] and k[p].
    MovingLeastSquaresBatch( double const *source_points,
                             std::vector<std::int64_t> const &n_source,
                             double const *target_points,
                             std::vector<std::int64_t> const &n_target,
                             std::vector<int> const &basis,
                             std::vector<int> const &k,
                             dtkx_mls_options options = default_options() )
    {
        std::size_t const n = n_source.size();
        if ( n_target.size() != n || basis.size() != n || k.size() != n )
            throw DataTransferKitException(
                DTKX_ERR_SIZE_MISMATCH,
                "n_source, n_target, basis and k: one entry per problem" );
        std::vector<dtkx_mls_problem> problems( n );
        std::int64_t s = 0, t = 0;
        for ( std::size_t p = 0; p < n; ++p )
        {
            if ( n_source[p] < 0 || n_target[p] < 0 )
                throw DataTransferKitException( DTKX_ERR_INVALID_ARGUMENT,
                                                "negative point count" );
            problems[p] = {source_points + 3 * s, n_source[p],
                           target_points + 3 * t, n_target[p], basis[p], k[p]};
            s += n_source[p];
            t += n_target[p];
        }
        create( problems, options );
    }

    MovingLeastSquaresBatch( MovingLeastSquaresBatch const & ) = delete;
    MovingLeastSquaresBatch &operator=( MovingLeastSquaresBatch const & ) = delete;

    ~MovingLeastSquaresBatch()
    {
        if ( _handle )
            dtkx_mls_batch_destroy( _handle );
    }

    // One field per problem: target_values[p] = C_p source_values[p]
    // (device arrays of the problem's sizes).
    void apply( std::vector<double const *> const &source_values,
                std::vector<double *> const &target_values ) const
    {
        if ( source_values.size() != size() || target_values.size() != size() )
            throw DataTransferKitException( DTKX_ERR_SIZE_MISMATCH,
                                            "one array per problem" );
        check( dtkx_mls_batch_apply( _handle, source_values.data(),
                                     target_values.data() ) );
    }

    // The values of every problem concatenated in problem order:
    // source_values [sum n_source], target_values [sum n_target].
    void apply( double const *source_values, double *target_values ) const
    {
        std::vector<double const *> x( size() );
        std::vector<double *> y( size() );
        for ( std::size_t p = 0; p < size(); ++p )
        {
            x[p] = source_values + _src_off[p];
            y[p] = target_values + _tgt_off[p];
        }
        check( dtkx_mls_batch_apply( _handle, x.data(), y.data() ) );
    }

    // Builds every operator again from the current contents of the point
    // arrays given at construction (same sizes, bases and k).
    void rebuild() { check( dtkx_mls_batch_rebuild( _handle ) ); }

    dtkx_mls_stats stats() const
    {
        dtkx_mls_stats s;
        check( dtkx_mls_batch_get_stats( _handle, &s ) );
        return s;
    }

    std::size_t size() const { return _n_source.size(); }
    std::int64_t n_source( std::size_t p ) const { return _n_source[p]; }
    std::int64_t n_target( std::size_t p ) const { return _n_target[p]; }
    dtkx_mls_batch const *handle() const { return _handle; }

  private:
    void create( std::vector<dtkx_mls_problem> const &problems,
                 dtkx_mls_options options )
    {
        options.rbf = Details::RbfId<CompactlySupportedRadialBasisFunction>::value;
        std::int64_t s = 0, t = 0;
        for ( auto const &q : problems )
        {
            _n_source.push_back( q.n_src );
            _n_target.push_back( q.n_tgt );
            _src_off.push_back( s );
            _tgt_off.push_back( t );
            s += q.n_src;
            t += q.n_tgt;
        }
        check( dtkx_mls_batch_create( static_cast<std::int32_t>( problems.size() ),
                                      problems.data(), &options, &_handle ) );
    }

    std::vector<std::int64_t> _n_source, _n_target, _src_off, _tgt_off;
    dtkx_mls_batch *_handle = nullptr;
};

} // namespace dtkx

#endif

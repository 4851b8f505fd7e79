// Compactly supported RBF spline (one GPU).
//
// Replaces DataTransferKit's SplineOperator
// (DTK_SplineOperator_def.hpp:39-295) on a one-rank communicator:
// A = (Q + N) (P + M + P^T)^-1 S, where M and N are basis operators (kNN
// rows of rbf values, buildBasisOperator, def.hpp:46-149), P and Q
// Vandermonde operators of the linear basis {1, x, y, z}
// (DTK_DetailsPolynomialMatrix.hpp:71-196) and S the prolongation that
// appends four zero polynomial entries
// (DTK_DetailsSplineProlongationOperator.hpp:41-108). The saddle-point system
// is solved by restarted GMRES on the device, as the reference does with
// Belos pseudo-block GMRES (def.hpp:248-263).
#ifndef DTKX_SPLINE_HPP
#define DTKX_SPLINE_HPP

#include "common.hpp"

namespace dtkx
{

// Neighbours per row: PolynomialBasis::size of the linear basis
// (def.hpp:188).
constexpr int spline_k = 4;

// Device views of the four blocks. Point rows are float4 (geometry FLOAT,
// the reference's ArborX::Point rounding) or double4 (geometry DOUBLE);
// their x, y, z are the Vandermonde row {1, x, y, z}. m_* rows are the
// sources' kNN among the sources, n_* rows the targets'.
struct SplineSystem
{
    int64_t n_src = 0, n_tgt = 0;
    int kk = 0; // min(spline_k, n_src)
    int geometry = 0;
    void const *src_pts = nullptr;
    void const *tgt_pts = nullptr;
    int32_t const *m_idx = nullptr;
    double const *m_phi = nullptr;
    int32_t const *n_idx = nullptr;
    double const *n_phi = nullptr;
};

// phi[i][l] = rbf(|src[idx[i][l]] - pts[i]| / R_i), R_i = 1.1 max_l of the
// distances (at least 10 DBL_EPSILON): computeRadius / computeWeights on
// the transformed coordinates (Impl.hpp:92-95, 118-134, 161-165).
void spline_basis_weights( double4 const *src, double4 const *pts, int64_t n,
                           int kk, int32_t const *idx, int rbf, int geometry,
                           double *phi, hipStream_t s );

// The sources renumbered by their Morton position (perm: position ->
// source, pos: its inverse), so that the operator's gathers stay local:
// M's rows and the source points are reordered and their indices mapped,
// N's indices are mapped. Every buffer is device memory.
void spline_renumber( int geometry, int64_t n_src, int kk,
                      int32_t const *perm, int32_t const *pos,
                      int32_t const *m_idx, double const *m_phi,
                      void const *src_pts, int32_t *m_idx_p, double *m_phi_p,
                      void *src_pts_p, int64_t n_tgt, int32_t const *n_idx,
                      int32_t *n_idx_p, hipStream_t s );

// Vectors of n entries, the first n_src of which are source entries,
// between source order and position order (the rest are copied).
void spline_to_positions( double const *in, int32_t const *perm,
                          int64_t n_src, int64_t n, double *out,
                          hipStream_t s );
void spline_to_sources( double const *in, int32_t const *perm, int64_t n_src,
                        int64_t n, double *out, hipStream_t s );

// y = C z = P z + M z + P^T z over the extended vector (n_src + 4 entries).
void spline_system_apply( SplineSystem const &sys, double const *z, double *y,
                          hipStream_t s );

// y[i] = V_t(i) z_poly + sum_l phi_n[i][l] z[n_idx[i][l]] for the targets.
void spline_evaluate( SplineSystem const &sys, double const *z, double *y,
                      hipStream_t s );

struct GmresOptions
{
    double tolerance = 1e-10; // relative residual |b - C z| / |b|
    int restart = 300;        // Krylov vectors per cycle
    int max_iterations = 1000;
};

struct GmresInfo
{
    int iterations = 0;
    int reorthogonalized = 0; // basis vectors reorthogonalised
    int restarts = 0;
    double residual = 0.0; // explicit relative residual of the result
    bool converged = false;
};

// Restarted GMRES for C z = [x; 0] from z = 0. Keeps its Krylov basis and
// scratch between calls (sized for the largest restart seen).
class SplineSolver
{
  public:
    SplineSolver() = default;
    ~SplineSolver();
    SplineSolver( SplineSolver const & ) = delete;
    SplineSolver &operator=( SplineSolver const & ) = delete;
    // x: device [n_src]; z: device [n_src + 4].
    GmresInfo solve( SplineSystem const &sys, double const *x, double *z,
                     GmresOptions const &opt, hipStream_t s );

  private:
    void reserve( int64_t n, int restart, hipStream_t s );
    int64_t ld_ = 0;
    int restart_ = 0;
    DeviceBuffer<double> basis_, rhs_, res_, partial_, dots_;
    double *host_ = nullptr; // pinned: projections and norms
};

} // namespace dtkx

#endif

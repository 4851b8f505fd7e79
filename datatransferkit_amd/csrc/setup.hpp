// Fused MLS setup: for each target, gather its neighbours, form relative
// coordinates, Vandermonde rows, radius, Wendland weights, the moment matrix
// P^T Phi P, its pseudo-inverse (row 0) and the CRS coefficients, in one
// kernel. Replaces the chain of separate kernels of the reference's
// constructor (DTK_MovingLeastSquaresOperator_def.hpp:66-118), whose
// intermediates (transformed coordinates, P, radius, phi, A, A^+, aux) all
// go through global memory.
#ifndef DTKX_SETUP_HPP
#define DTKX_SETUP_HPP

#include "common.hpp"

namespace dtkx
{

struct SetupParams
{
    int m = 4;
    int kk = 4;
    int rbf = DTKX_RBF_WENDLAND0;
    int geometry = DTKX_GEOMETRY_FLOAT;
    int truncation = DTKX_TRUNCATION_ABSOLUTE;
    double rcond = 0.0;
    long max_rot = 0;
    int group_size = 0; // classical solver: lanes per system, 0 = choose
    // rows [row0, n_tgt) are set up; row_off (nullable) gives CSR offsets
    // of the rows' neighbour lists (kk per row when null)
    int64_t row0 = 0;
    int64_t const *row_off = nullptr;
    // row list (nullable, cyclic solver): set up rows rows[i] for
    // i < *n_rows (device count, at most n_tgt - row0) instead of a range
    int32_t const *rows = nullptr;
    int32_t const *n_rows = nullptr;
    // FP32 bulk (nullable): each row's largest neighbour distance from the
    // kNN, which then gives the radius without a pass over the neighbours
    float const *row_max = nullptr;
    // cyclic solver over a row list (nullable): rows whose eigenvalues all
    // pass the truncation test are appended here (device count n_refine)
    int32_t *refine_rows = nullptr;
    int32_t *n_refine = nullptr;
    // cyclic solver over a list that setup_screen made: the sweeps and
    // unconverged systems are counted for truncated rows only (the screen
    // counted the others)
    bool screened = false;
};

// counters[0] = n_underdetermined, [1] = n_unconverged, [2] = rotations
// (classical) or sweeps (cyclic), [3] = systems the Cholesky path left to
// the eigen-solver; accumulated atomically.
// Row r of indices and coeffs belongs to target order[r] (identity when
// order is null); indices address src, which may be the sources in a
// permuted (Morton) order.

// Classical max-pivot two-sided Jacobi, reference operation order.
int setup_classical( double4 const *src, double4 const *tgt, int64_t n_tgt,
                     int32_t const *indices, int32_t const *order,
                     SetupParams const &p, double *coeffs,
                     unsigned long long *counters, hipStream_t s );

// Lane group size the classical solver uses for (m, kk) when not forced.
int setup_classical_group( int m, int kk );

// Cyclic symmetric Jacobi eigen-solver, registers only (fast path);
// p.max_rot caps the sweeps (0 = 30). With coeffs32 != nullptr the FP32-bulk
// form runs: float geometry, weights, Vandermonde and coefficients (written
// to coeffs32), FP64 moments, Jacobi and truncation. pinv (nullable)
// receives each row's pseudo-inverse A^+ as a packed upper triangle
// [n][m(m+1)/2] (entry (i, j), i <= j, at i*m - i(i-1)/2 + j - i).
void setup_cyclic( double4 const *src, double4 const *tgt, int64_t n_tgt,
                   int32_t const *indices, int32_t const *order,
                   SetupParams const &p, double *coeffs, float *coeffs32,
                   double *pinv, unsigned long long *counters, hipStream_t s );

// Eigenvalue screen of the listed rows p.rows[0 .. *p.n_rows): the cyclic
// Jacobi of setup_cyclic without eigenvectors (the same rotations of A, so
// the same eigenvalues, sweeps and convergence). Rows whose eigenvalues all
// pass the truncation test are appended to p.refine_rows, the others to
// full_rows; counters[1] and [2] are taken for the first kind. f32: the
// FP32-bulk form's geometry and weights.
void setup_screen( double4 const *src, double4 const *tgt, int64_t n_tgt,
                   int32_t const *indices, int32_t const *order,
                   SetupParams const &p, bool f32, int32_t *full_rows,
                   int32_t *n_full, unsigned long long *counters,
                   hipStream_t s );

// Certified Cholesky fast path (DTKX_SOLVER_CHOLESKY): each system is
// factored A = U^T U in FP64; when the bound lambda_min(A) >= 1 / |U^-1|_F^2
// proves that no eigenvalue reaches the truncation threshold (with a margin
// for the eigen-solver's rounding), A^+ = A^-1 and a0 = U^-1 U^-T e_0 is
// used. The other systems are listed; those of them with no eigenvalue
// below the threshold (setup_screen) have A^+ = A^-1 as well and are solved
// again by Cholesky, whose rounding error follows the equilibrated matrix,
// not the eigen-decomposition's eps |A| / lambda (measured on 10-neighbour
// quadratic systems: 5e-10 against 2e-4 from the operator in extended
// precision); the rest are set up by the cyclic eigen-solver.
// Same outputs and counters as setup_cyclic.
void setup_fast( double4 const *src, double4 const *tgt, int64_t n_tgt,
                 int32_t const *indices, int32_t const *order,
                 SetupParams const &p, double *coeffs, float *coeffs32,
                 double *pinv, unsigned long long *counters, hipStream_t s );

// Batched pseudo-inverse of general n x n matrices (n <= 64), one matrix
// per workgroup (tstSVD path).
void svd_pinv_batched( int n, int64_t n_mat, double const *a, double *pinv,
                       long max_rot, unsigned long long *counters,
                       hipStream_t s );

} // namespace dtkx

#endif

/*
 * dtkx — MI355X-native meshfree moving-least-squares (MLS) transfer.
 *
 * C ABI of the MLS transfer path. Every entry point takes plain pointers and
 * sizes; no C++ or torch types cross this boundary. Point and value pointers
 * may be device pointers (hipMalloc / torch CUDA tensors, used in place) or
 * ordinary host pointers (staged through device memory by the library).
 *
 * Reference interface this replaces (DataTransferKit, read-only reference):
 *   packages/Meshfree/src/DTK_MovingLeastSquaresOperator_decl.hpp:36-63
 *     template <DeviceType, RBF = Wendland<0>, Basis = Linear3>
 *     class MovingLeastSquaresOperator : public PointCloudOperator<DeviceType>
 *       ctor(MPI_Comm, View<Coordinate const**> src, View<Coordinate const**> tgt)
 *       void apply(View<double const*> src_vals, View<double*> tgt_vals) const
 *   packages/Meshfree/src/DTK_PointCloudOperator.hpp:24-37 (abstract apply)
 * The communicator is always a one-rank communicator here (one GPU); the
 * reference's MPI gather (`fetch`) degenerates to an in-kernel gather.
 *
 * Error convention: every function returns a dtkx_status; on failure a
 * message shaped like the reference's DBC output
 *   "dtkx: <condition>, failed in <file>, line <n>."
 * (packages/Utils/src/DTK_DBC.cpp:40-48) is available from
 * dtkx_last_error() (thread-local). Checks are always on, unlike the
 * reference's DTK_REQUIRE (active only with HAVE_DTK_DBC).
 */
#ifndef DTKX_MLS_H
#define DTKX_MLS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DTKX_VERSION_MAJOR 0
#define DTKX_VERSION_MINOR 1

/* ---------------------------------------------------------------- status */
typedef enum dtkx_status {
    DTKX_SUCCESS = 0,
    DTKX_ERR_INVALID_ARGUMENT = 1, /* null pointer, bad enum, negative size */
    DTKX_ERR_SIZE_MISMATCH = 2,    /* ..._def.hpp:129-130 */
    DTKX_ERR_DIMENSION = 3,        /* ..._def.hpp:39-42 (only 3D) */
    DTKX_ERR_EMPTY_SOURCE = 4,     /* ..._def.hpp:49 !search_tree.empty() */
    DTKX_ERR_HIP = 5,              /* HIP runtime error */
    DTKX_ERR_NOT_CONVERGED = 6,    /* Jacobi rotation cap reached */
    DTKX_ERR_UNSUPPORTED = 7,      /* option combination not built */
    DTKX_ERR_OUT_OF_MEMORY = 8
} dtkx_status;

/* ------------------------------------------------------------- policies */
/* Polynomial basis; the value is the basis size m in 3D
 * (DTK_MultivariatePolynomialBasis.hpp:42-75). */
typedef enum dtkx_basis {
    DTKX_BASIS_CONSTANT = 1,  /* {1} */
    DTKX_BASIS_LINEAR = 4,    /* {1,x,y,z} */
    DTKX_BASIS_QUADRATIC = 10 /* {1,x,y,z,x^2,xy,xz,y^2,yz,z^2} */
} dtkx_basis;

/* Compactly supported radial basis functions
 * (DTK_CompactlySupportedRadialBasisFunctions.hpp:45-147). */
typedef enum dtkx_rbf {
    DTKX_RBF_WENDLAND0 = 0,
    DTKX_RBF_WENDLAND2 = 1,
    DTKX_RBF_WENDLAND4 = 2,
    DTKX_RBF_WENDLAND6 = 3,
    DTKX_RBF_WU2 = 4,
    DTKX_RBF_WU4 = 5,
    DTKX_RBF_BUHMANN2 = 6,
    DTKX_RBF_BUHMANN3 = 7,
    /* refused by every create call with DTKX_ERR_UNSUPPORTED: as the
     * reference writes it (:139-146) it is -67.9 at x = 1, so it is not
     * compactly supported and no test or identity pins its values */
    DTKX_RBF_BUHMANN4 = 8
} dtkx_rbf;

typedef enum dtkx_precision {
    /* everything in FP64 (geometry per dtkx_geometry) */
    DTKX_PRECISION_FP64 = 0,
    /* FP32 relative coordinates, radius, weights, Vandermonde,
     * coefficients and apply (compensated FP32 sums); FP64 moments
     * P^T W P, solve and singular-value truncation */
    DTKX_PRECISION_FP32_BULK = 1
} dtkx_precision;

typedef enum dtkx_geometry {
    /* Reference semantics: coordinates pass through a float point type
     * before kNN distances, radius, weights and monomials are formed
     * (DTK_DetailsMovingLeastSquaresOperatorImpl.hpp:40-43,121-125,162-165,
     * 184-185). */
    DTKX_GEOMETRY_FLOAT = 0,
    /* all geometry in double */
    DTKX_GEOMETRY_DOUBLE = 1
} dtkx_geometry;

typedef enum dtkx_solver {
    /* Reference order: classical max-pivot two-sided Jacobi with trig 2x2
     * SVDs (DTK_DetailsSVDImpl.hpp:119-302). One system per lane group. */
    DTKX_SOLVER_CLASSICAL_JACOBI = 0,
    /* Cyclic symmetric Jacobi eigen-solver, one system per lane, registers
     * only. Same pseudo-inverse up to rounding (amplified by cond(A), which
     * at k = m reaches 1e12 and more); much faster. */
    DTKX_SOLVER_CYCLIC_JACOBI = 1,
    /* Cholesky solve for every system whose smallest eigenvalue is proved
     * (lambda_min >= 1/|U^-1|_F^2) to lie above the truncation threshold,
     * where the pseudo-inverse is the inverse; the cyclic eigen-solver for
     * the others. Same pseudo-inverse up to rounding; fastest. */
    DTKX_SOLVER_CHOLESKY = 2
} dtkx_solver;

typedef enum dtkx_truncation {
    /* keep singular values with |s| >= DBL_EPSILON (DTK_DetailsSVDImpl.hpp:291) */
    DTKX_TRUNCATION_ABSOLUTE = 0,
    /* keep singular values with |s| >= rcond * max|s| */
    DTKX_TRUNCATION_RELATIVE = 1
} dtkx_truncation;

typedef enum dtkx_layout {
    DTKX_LAYOUT_AOS = 0, /* [n][3] row-major, Kokkos LayoutRight */
    DTKX_LAYOUT_SOA = 1  /* [3][n], Kokkos LayoutLeft */
} dtkx_layout;

typedef struct dtkx_mls_options {
    int32_t dim;        /* spatial dimension; must be 3 */
    int32_t basis;      /* dtkx_basis; default LINEAR (the reference default) */
    int32_t rbf;        /* dtkx_rbf; default WENDLAND0 */
    int32_t k;          /* neighbours per target; 0 => basis size
                           (..._def.hpp:53-55 hard-wires k = basis size) */
    int32_t precision;  /* dtkx_precision */
    int32_t geometry;   /* dtkx_geometry */
    int32_t solver;     /* dtkx_solver */
    int32_t truncation; /* dtkx_truncation */
    double rcond;       /* for DTKX_TRUNCATION_RELATIVE */
    int32_t max_rotations; /* Jacobi cap: classical = rotations, cyclic =
                              sweeps; 0 => library default */
    int32_t layout;     /* dtkx_layout of the point arrays */
    int32_t group_size; /* lanes per system for the classical solver
                           (1,2,4,8,16); 0 => autotuned */
    int32_t store_pinv; /* keep each target's pseudo-inverse A^+ (packed
                           upper triangle, m(m+1)/2 doubles) for
                           dtkx_mls_target_gradient (1), or let the adjoint
                           recompute it (0); needs the cyclic solver */
    void *stream;       /* hipStream_t; NULL = default stream */
} dtkx_mls_options;

typedef struct dtkx_mls_stats {
    int64_t n_source;
    int64_t n_target;
    int64_t nnz;
    int32_t basis_size;
    int32_t k;                  /* neighbours per row = min(k, n_source) */
    int64_t n_underdetermined;  /* systems with a truncated singular value
                                   (the reference's parallel_reduce count,
                                   Impl.hpp:284-288) */
    int64_t n_unconverged;      /* systems that hit the Jacobi cap */
    int64_t total_rotations;    /* classical: rotations; cyclic: sweeps */
    double ms_knn;              /* stage times of the last create, measured */
    double ms_setup;            /* with HIP events on the handle's stream   */
    int32_t group_size_used;
    int32_t n_stack_overflow;   /* kNN queries deeper than the LDS traversal
                                   stack, re-run with a global one */
    double ms_knn_query;        /* the kNN query kernels alone */
    int64_t n_eigen_fallback;   /* Cholesky solver: systems not certified,
                                   set up by the cyclic eigen-solver */
} dtkx_mls_stats;

typedef struct dtkx_mls dtkx_mls; /* opaque operator handle */

/* ------------------------------------------------------------ generic */
const char *dtkx_last_error(void);
const char *dtkx_version(void);
/* Device memory freed by the library stays cached per (device, stream) for
 * reuse by later allocations on the same stream. This returns the current
 * device's cached blocks to the HIP runtime (after synchronising the
 * device). No counterpart in the reference (Kokkos manages its own pool). */
int dtkx_release_cached_memory(void);
/* Fill *opts with the defaults (reference semantics: Linear basis,
 * Wendland<0>, k = basis size, FP64, float geometry, classical Jacobi,
 * absolute DBL_EPSILON truncation, AoS points). */
int dtkx_mls_options_init(dtkx_mls_options *opts);

/* ---------------------------------------------------- operator (drop-in) */
/* MovingLeastSquaresOperator ctor (..._def.hpp:24-119): kNN over the source
 * points, then the fused setup (relative coordinates, Vandermonde, radius,
 * weights, moments, pseudo-inverse, coefficients). src/tgt are [n][3]
 * (AoS) or [3][n] (SoA) doubles. n_tgt may be 0; n_src must be > 0.
 * Returns once the work is queued on opts->stream (host point arrays are
 * copied before it returns); dtkx_mls_get_stats waits for it. */
int dtkx_mls_create(const double *src_xyz, int64_t n_src,
                    const double *tgt_xyz, int64_t n_tgt,
                    const dtkx_mls_options *opts, dtkx_mls **out);

/* Same as dtkx_mls_create but with caller-supplied neighbour lists:
 * row i holds indices[i*k_row .. i*k_row+k_row-1] (0 <= index < n_src).
 * No kNN is run; the order of a row is the summation order. */
int dtkx_mls_create_with_neighbors(const double *src_xyz, int64_t n_src,
                                   const double *tgt_xyz, int64_t n_tgt,
                                   const int32_t *indices, int32_t k_row,
                                   const dtkx_mls_options *opts,
                                   dtkx_mls **out);

/* MovingLeastSquaresOperator::apply (..._def.hpp:121-141):
 * tgt_vals[f*ld_tgt + i] = sum_j coeff(i,j) * src_vals[f*ld_src + idx(i,j)]
 * for f < n_fields. n_fields = 1, ld = n reproduces the reference call.
 * In FP32_BULK precision the values are still double at the boundary.
 * With device arrays it returns once the work is queued on the handle's
 * stream (later work on that stream sees the result; synchronise it before
 * reading from the host or another stream), as the reference's Kokkos
 * kernels do; with host arrays the result is complete on return. */
int dtkx_mls_apply(const dtkx_mls *h, const double *src_vals,
                   double *tgt_vals, int32_t n_fields, int64_t ld_src,
                   int64_t ld_tgt);

/* dtkx_mls_apply without any allocation. No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as dtkx_mls_apply (bit for bit), with these
 * limits:
 *   - src_vals and tgt_vals must be device memory (DTKX_ERR_INVALID_ARGUMENT
 *     otherwise: host arrays would need staging buffers);
 *   - the apply's scratch (the gathered fields, n_src records of up to 8
 *     fields, and the staged row results, n_tgt records) must fit in
 *     DTKX_NOALLOC_SCRATCH_BYTES (DTKX_ERR_UNSUPPORTED otherwise): FP64 with
 *     4 fields fits n_src + n_tgt <= 2,097,152;
 *   - the scratch is one fixed device array per GPU, shared by every call
 *     on that GPU and held until the process ends: calls that overlap on
 *     one GPU (other threads, other handles' streams) race; serialise them.
 * Call it with the handle's device current. */
#define DTKX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int dtkx_mls_apply_noalloc(const dtkx_mls *h, const double *src_vals,
                           double *tgt_vals, int32_t n_fields, int64_t ld_src,
                           int64_t ld_tgt);

/* FP32 apply: float values, float coefficients, float accumulation.
 * Requires a handle created with DTKX_PRECISION_FP32_BULK. Returns as
 * dtkx_mls_apply does. */
int dtkx_mls_apply_fp32(const dtkx_mls *h, const float *src_vals,
                        float *tgt_vals, int32_t n_fields, int64_t ld_src,
                        int64_t ld_tgt);

/* Transpose apply (no counterpart in the reference's MLS path):
 * src_bar[f*ld_src + s] = sum_{(i,j): idx(i,j)=s} coeff(i,j) *
 *                          tgt_bar[f*ld_tgt + i]  (overwrites src_bar). */
int dtkx_mls_apply_transpose(const dtkx_mls *h, const double *tgt_bar,
                             double *src_bar, int32_t n_fields,
                             int64_t ld_tgt, int64_t ld_src);

/* Gradient of L = sum_f <tgt_bar_f, C(t) src_vals_f> with respect to the
 * target coordinates t, neighbour sets frozen (derivatives of the weights,
 * the radius and the basis rows through the moment matrix). src_vals is
 * [n_fields][ld_src], tgt_bar [n_fields][ld_tgt]; d_tgt is [n_tgt][3] (AoS),
 * overwritten. Needs DTKX_GEOMETRY_DOUBLE and FP64. The pseudo-inverse is
 * read back when the handle stored it (store_pinv), else recomputed per
 * target with the cyclic eigen-solver. No counterpart in the reference,
 * whose MLS operator has no adjoint. */
int dtkx_mls_target_gradient(const dtkx_mls *h, const double *src_vals,
                             const double *tgt_bar, int32_t n_fields,
                             int64_t ld_src, int64_t ld_tgt, double *d_tgt);

/* CSR view of the operator, device pointers owned by the handle:
 * offsets[n_tgt+1], indices[nnz], coeffs[nnz] (double, or float for
 * FP32_BULK). Mirrors the reference's _offset/_indices/_coeffs members
 * (..._decl.hpp:57-62). */
int dtkx_mls_get_graph(const dtkx_mls *h, const int32_t **offsets,
                       const int32_t **indices, const void **coeffs,
                       int64_t *nnz);
/* The operator as a CSR file for downstream assembly (a Tpetra CrsMatrix,
 * C^T, another process): no counterpart in the reference, whose operator
 * keeps its graph in memory (..._decl.hpp:57-62; SplineOperator builds a
 * CrsMatrix from such a graph, DTK_SplineOperator_def.hpp:108-149).
 * Little-endian, in order:
 *   char    magic[8]        "DTKXCSR1"
 *   int64_t n_rows, n_cols, nnz, col_base
 *   int32_t coeff_bytes     8 (double) or 4 (float, FP32_BULK)
 *   int32_t reserved        0
 *   int64_t offsets[n_rows + 1]
 *   int64_t columns[nnz]    global ids: col_base + source index
 *   coeffs[nnz]             double or float
 * Rows are targets in caller order, each ascending by (distance, index). */
int dtkx_mls_write_graph(const dtkx_mls *h, const char *path,
                         int64_t col_base);
int dtkx_mls_get_stats(const dtkx_mls *h, dtkx_mls_stats *stats);
int dtkx_mls_destroy(dtkx_mls *h);

/* ------------------------------------------------------- stage kernels */
/* k nearest neighbours (ArborX Nearest<Point>{t, k} query,
 * ..._def.hpp:47-60 + Impl.hpp:28-47). indices is [n_tgt][min(k,n_src)],
 * each row ascending by (distance, index); dist (optional, may be NULL)
 * receives the matching distances (float for FLOAT geometry, written as
 * double). */
int dtkx_knn(const double *src_xyz, int64_t n_src, const double *tgt_xyz,
             int64_t n_tgt, int32_t k, int32_t geometry, int32_t layout,
             int32_t *indices, double *dist, void *stream);

/* Batched pseudo-inverse of n x n general square matrices
 * (SVDFunctor, DTK_DetailsSVDImpl.hpp:39-309, as driven by tstSVD.cpp):
 * pinv[m] = V^T E^+ U^T, singular values below DBL_EPSILON dropped.
 * 1 <= n <= 64. *n_underdetermined receives the reference's reduction. */
int dtkx_svd_pseudo_inverse(int32_t n, int64_t n_matrices, const double *a,
                            double *pinv, int64_t *n_underdetermined,
                            int64_t *n_unconverged, int32_t max_rotations,
                            void *stream);

/* ------------------------------------------- batched uneven operators */
typedef struct dtkx_mls_problem {
    const double *src_xyz; /* device pointers, AoS [n][3] */
    int64_t n_src;
    const double *tgt_xyz;
    int64_t n_tgt;
    int32_t basis; /* dtkx_basis */
    int32_t k;     /* 0 => basis size */
} dtkx_mls_problem;

typedef struct dtkx_mls_batch dtkx_mls_batch;

/* Build many independent operators, each with its own basis and k (other
 * options from opts; store_pinv is not supported). Every problem needs at
 * least one source; points are device pointers, copied here and again by
 * each dtkx_mls_batch_rebuild. All trees
 * are built by one set of kernels over the concatenated, problem-sorted
 * sources; queries run in one launch per neighbour-list size and the setup
 * in one launch per basis. With the classical solver and group_size 0 the
 * lanes per system are chosen by timing a sample of each basis' rows.
 * Each problem's result is bit-identical to a dtkx_mls_create of that
 * problem with the same options. No counterpart in the reference (one
 * operator per MovingLeastSquaresOperator construction). */
int dtkx_mls_batch_create(int32_t n_problems, const dtkx_mls_problem *problems,
                          const dtkx_mls_options *opts, dtkx_mls_batch **out);
/* One field per problem: src_vals[p] (n_src_p doubles) and tgt_vals[p]
 * (n_tgt_p doubles) are device pointers in host arrays of n_problems
 * entries (the first problem's pair is checked, DTKX_ERR_INVALID_ARGUMENT
 * for host memory; the others are not). The gather and row sums of all problems are captured in a HIP
 * graph on the first call and replayed; synchronous at return. */
int dtkx_mls_batch_apply(const dtkx_mls_batch *b,
                         const double *const *src_vals,
                         double *const *tgt_vals);
/* Builds every operator of the batch again from the problems' point arrays
 * at the addresses given to dtkx_mls_batch_create, whose contents may have
 * changed since (moving points; sizes, bases and k stay as created, and the
 * arrays must still be valid). The search and setup are captured in a HIP
 * graph on the first call and replayed: no host bookkeeping, uploads or
 * allocations after it. Each problem's result is bit-identical to a new
 * dtkx_mls_batch_create on the current points (the classical solver keeps
 * the lanes per system chosen at create; every choice gives the same bits).
 * Statistics: counters of this build; ms_knn is the whole replay, ms_setup
 * and ms_knn_query 0. Synchronous at return. No counterpart in the
 * reference (a new MovingLeastSquaresOperator is constructed instead). */
int dtkx_mls_batch_rebuild(dtkx_mls_batch *b);
int dtkx_mls_batch_get_stats(const dtkx_mls_batch *b, dtkx_mls_stats *stats);
int dtkx_mls_batch_destroy(dtkx_mls_batch *b);

/* ------------------------------------------ nearest-neighbour operator */
/* NearestNeighborOperator (DTK_NearestNeighborOperator_decl.hpp:24-50,
 * _def.hpp:23-86): each target takes the value of its nearest source. The
 * search is dtkx_knn with k = 1 (ArborX Nearest<Point> with k = 1 in the
 * reference, _def.hpp:46-53; ties go to the smallest source index); apply
 * copies x[f][nearest(t)] into y[f][t], bit for bit (the reference's fetch
 * and deep_copy, _def.hpp:78-82). One rank. */
typedef struct dtkx_nn dtkx_nn;

/* geometry: dtkx_geometry (FLOAT: the reference's float points); layout:
 * dtkx_layout; stream: hipStream_t or NULL. Fails with
 * DTKX_ERR_EMPTY_SOURCE when n_src == 0 (the reference's
 * DTK_CHECK(!search_tree.empty()), _def.hpp:43). */
int dtkx_nn_create(const double *src_xyz, int64_t n_src,
                   const double *tgt_xyz, int64_t n_tgt, int32_t geometry,
                   int32_t layout, void *stream, dtkx_nn **out);
/* Fields are [n_fields][ld]; ld_src >= n_src, ld_tgt >= n_tgt
 * (the reference's size preconditions, _def.hpp:73-76). Synchronous. */
int dtkx_nn_apply(const dtkx_nn *h, const double *src_vals, double *tgt_vals,
                  int32_t n_fields, int64_t ld_src, int64_t ld_tgt);
/* The nearest source of every target (device array of n_tgt, owned by the
 * handle): the reference's _indices member (_decl.hpp:46). */
int dtkx_nn_get_indices(const dtkx_nn *h, const int32_t **indices,
                        int64_t *n_tgt);
int dtkx_nn_destroy(dtkx_nn *h);

/* ----------------------------------------------- RBF spline operator */
/* SplineOperator<Wendland<0>, Linear3> (DTK_SplineOperator_decl.hpp:40-113,
 * _def.hpp:170-295): A = (Q + N) (P + M + P^T)^-1 S. M (sources x sources)
 * and N (targets x sources) hold, per row, the 4 nearest sources
 * (k = PolynomialBasis::size, _def.hpp:188) with the values
 * rbf(|x_l - t| / R), R = 1.1 max_l |x_l - t| (buildBasisOperator,
 * _def.hpp:46-149); P and Q are the Vandermonde rows {1, x, y, z} of the
 * sources and targets (buildPolynomialOperator, _def.hpp:151-167, and
 * DTK_DetailsPolynomialMatrix.hpp:71-196); S appends four zero polynomial
 * entries to a source vector (DTK_DetailsSplineProlongationOperator.hpp
 * :41-108). The saddle-point system C z = S x, C = [[M, P], [P^T, 0]], is
 * solved on the GPU by restarted GMRES from z = 0, as the reference does
 * with Belos pseudo-block GMRES (_def.hpp:248-263). One rank. */
typedef struct dtkx_spline dtkx_spline;

typedef struct dtkx_spline_options {
  int32_t rbf;            /* dtkx_rbf; default Wendland<0> (_def.hpp:319) */
  int32_t geometry;       /* dtkx_geometry; FLOAT = the reference's points */
  int32_t layout;         /* dtkx_layout of the point arrays */
  int32_t restart;        /* Krylov vectors per GMRES cycle, 1..512 (300);
                           * the basis takes (restart + 1) * 8 * (n_src + 4)
                           * bytes of device memory (2.4 GB at 1M sources):
                           * lower it for very large clouds */
  int32_t max_iterations; /* GMRES iterations per solve (1000) */
  int32_t reserved;
  double tolerance;       /* relative residual |S x - C z| / |S x|: 1e-10
                           * ("Convergence Tolerance", _def.hpp:258) */
  void *stream;           /* hipStream_t or NULL */
} dtkx_spline_options;

/* Report of a handle's last solve (one per field of the last apply/solve:
 * the largest iteration count and residual, and whether all converged). */
typedef struct dtkx_spline_info {
  int32_t iterations;
  int32_t converged;
  int32_t reorthogonalized; /* basis vectors reorthogonalised (the delayed
                             * classical Gram-Schmidt does every one) */
  int32_t restarts;         /* GMRES cycles after the first */
  double residual;
} dtkx_spline_info;

int dtkx_spline_options_init(dtkx_spline_options *opts);
/* Fails with DTKX_ERR_EMPTY_SOURCE when n_src == 0 (the reference's
 * DTK_CHECK(!distributed_tree.empty()), _def.hpp:70). */
int dtkx_spline_create(const double *src_xyz, int64_t n_src,
                       const double *tgt_xyz, int64_t n_tgt,
                       const dtkx_spline_options *opts, dtkx_spline **out);
/* y = A x for n_fields fields stored [n_fields][ld] (ld_src >= n_src,
 * ld_tgt >= n_tgt): one GMRES solve per field (apply, _def.hpp:266-295).
 * Synchronous. Not safe to call concurrently on one handle. */
int dtkx_spline_apply(const dtkx_spline *h, const double *src_vals,
                      double *tgt_vals, int32_t n_fields, int64_t ld_src,
                      int64_t ld_tgt);
/* The spline coefficients z = C^-1 S x (n_src + 4 doubles: the source
 * entries, then the four polynomial entries) for one field. */
int dtkx_spline_solve(const dtkx_spline *h, const double *src_vals,
                      double *coeffs);
/* y = (Q + N) z for coefficients z of n_src + 4 doubles. */
int dtkx_spline_evaluate(const dtkx_spline *h, const double *coeffs,
                         double *tgt_vals);
/* y = C z = (P + M + P^T) z, both of n_src + 4 doubles. */
int dtkx_spline_system_apply(const dtkx_spline *h, const double *z,
                             double *y);
/* The basis operators (device arrays owned by the handle): which = 0 for M
 * (n_rows = n_src), 1 for N (n_rows = n_tgt); row i holds kk = min(4, n_src)
 * source indices ascending by (distance, index) and their rbf values. */
int dtkx_spline_get_basis(const dtkx_spline *h, int32_t which,
                          const int32_t **indices, const double **values,
                          int64_t *n_rows, int32_t *kk);
int dtkx_spline_get_info(const dtkx_spline *h, dtkx_spline_info *info);
int dtkx_spline_destroy(dtkx_spline *h);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* DTKX_MLS_H */

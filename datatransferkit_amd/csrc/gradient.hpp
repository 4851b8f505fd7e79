// Target-coordinate adjoint of the MLS operator (gradient.hip).
#ifndef DTKX_GRADIENT_HPP
#define DTKX_GRADIENT_HPP

#include "apply.hpp"
#include "common.hpp"

namespace dtkx
{

struct GradientArgs
{
    double4 const *src = nullptr;  // sources in position order
    double4 const *tgt = nullptr;  // targets in caller order
    int32_t const *indices = nullptr; // [n_tgt * kk] positions, by row
    int32_t const *order = nullptr;   // row -> target (nullable)
    int32_t const *perm = nullptr;    // position -> source (nullable)
    int32_t const *spos = nullptr;    // source -> position (nullable)
    double const *pinv = nullptr;     // stored packed A^+ per row, or null
    int64_t n_src = 0, n_tgt = 0;
    int m = 4, kk = 4, rbf = 0, truncation = 0, max_sweeps = 0;
    int solver = 0;                   // DTKX_SOLVER_* of the handle
    double rcond = 0.0;
    int64_t ld_tgt = 0;               // leading dimension of ybar
    double *d_tgt = nullptr;          // [n_tgt][3], caller order
};

// d_tgt = dL/dt for L = sum_f <ybar_f, C(t) x_f>; x is [n_fields][ld_src]
// in caller source order, ybar [n_fields][a.ld_tgt] in caller target order.
void target_gradient( GradientArgs const &a, double const *x, int64_t ld_src,
                      int n_fields, double const *ybar, hipStream_t s );

} // namespace dtkx

#endif

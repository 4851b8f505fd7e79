"""Extended-precision MLS target values -- TEST INFRASTRUCTURE ONLY.

The MLS operator of DataTransferKit (DTK_DetailsMovingLeastSquaresOperator
Impl.hpp, double geometry) evaluated in numpy's x87 extended precision
(np.longdouble, 64-bit significand) for given, frozen neighbour lists:
relative coordinates u_j = x_j - t (computeTransformedCoordinates,
Impl.hpp:72-99), radius R = 1.1 max_j |u_j| (computeRadius, :101-138),
weights phi_j = W(|u_j| / R) (computeWeights, :143-169), basis rows
p(u_j) (MultivariatePolynomialBasis, DTK_MultivariatePolynomialBasis.hpp
:113-124), moments A = sum_j phi_j p_j p_j^T (computeMoments, :213-258)
and y_f = e_0^T A^{-1} sum_j phi_j p_j x_f(j), i.e. sum_j c_j x_f(j) with
the coefficients of computePolynomialCoefficients (:293-336) for a
full-rank A. Used to take central differences in target coordinates at
h = 1e-6 on clouds whose moment matrices are too ill-conditioned for a
double-precision difference quotient (the monomials of relative
coordinates of size R make cond(A) grow like R^-4): with 11 more bits the
quotient's rounding noise is 2^-11 of the double one's.

Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np

LD = np.longdouble


def _rbf(name, x):
    a = 1 - x
    if name == "wendland0":
        return a ** 2
    if name == "wendland2":
        return a ** 4 * (4 * x + 1)
    if name == "wendland4":
        return a ** 6 * (35 * x * x + 18 * x + 3)
    if name == "wendland6":
        return a ** 8 * (32 * x ** 3 + 25 * x * x + 8 * x + 1)
    raise ValueError(f"rbf {name} not restated in extended precision")


def _basis(m, u):
    x, y, z = u[..., 0], u[..., 1], u[..., 2]
    one = np.ones_like(x)
    cols = [one]
    if m >= 4:
        cols += [x, y, z]
    if m >= 10:
        cols += [x * x, x * y, x * z, y * y, y * z, z * z]
    return np.stack(cols, axis=-1)


def _solve(a, b):
    """Batched Gaussian elimination with partial pivoting, a: (n, m, m),
    b: (n, m, f), in the arrays' dtype."""
    a = a.copy()
    b = b.copy()
    n, m, _ = a.shape
    rows = np.arange(n)
    for c in range(m):
        piv = c + np.argmax(np.abs(a[:, c:, c]), axis=1)
        for arr in (a, b):
            tmp = arr[rows, c].copy()
            arr[rows, c] = arr[rows, piv]
            arr[rows, piv] = tmp
        f = a[:, c + 1:, c] / a[:, c:c + 1, c]
        a[:, c + 1:, :] -= f[:, :, None] * a[:, None, c, :]
        b[:, c + 1:, :] -= f[:, :, None] * b[:, None, c, :]
    x = np.zeros_like(b)
    for c in range(m - 1, -1, -1):
        s = b[:, c, :] - np.einsum("nk,nkf->nf", a[:, c, c + 1:], x[:, c + 1:, :])
        x[:, c, :] = s / a[:, c, c][:, None]
    return x


def target_values(src, tgt, nbr, x, m, rbf="wendland2"):
    """y[f, t] of the operator with neighbour lists nbr [n_t, k] for fields
    x [F, n_s], all evaluated in extended precision."""
    src = np.asarray(src, dtype=LD)
    tgt = np.asarray(tgt, dtype=LD)
    x = np.asarray(x, dtype=LD)
    nbr = np.asarray(nbr)
    u = src[nbr] - tgt[:, None, :]
    d = np.sqrt((u * u).sum(-1))
    r = LD(1.1) * np.maximum(d.max(axis=1), LD(10) * LD(np.finfo(np.float64).eps))
    phi = _rbf(rbf, d / r[:, None])
    p = _basis(m, u)
    a = np.einsum("tki,tk,tkj->tij", p, phi, p)
    b = np.einsum("tki,tk,ftk->tif", p, phi, x[:, nbr])
    beta = _solve(a, b)
    return beta[:, 0, :].T


def target_gradient_fd(src, tgt, nbr, x, ybar, m, rbf="wendland2", h=1e-6):
    """Central differences of L_t = sum_f ybar[f, t] y[f, t] in each target
    coordinate, in extended precision."""
    tgt = np.asarray(tgt, dtype=LD)
    ybar = np.asarray(ybar, dtype=LD)
    g = np.zeros((len(tgt), 3), dtype=LD)
    for c in range(3):
        tp, tm = tgt.copy(), tgt.copy()
        tp[:, c] += LD(h)
        tm[:, c] -= LD(h)
        lp = (target_values(src, tp, nbr, x, m, rbf) * ybar).sum(0)
        lm = (target_values(src, tm, nbr, x, m, rbf) * ybar).sum(0)
        g[:, c] = (lp - lm) / (2 * LD(h))
    return g.astype(np.float64)


def float_geometry_operator(src, tgt, nbr, m, rbf="wendland2"):
    """The operator's coefficients [n_t, k] for ArborXFloat geometry (the
    library's default and the reference's): relative coordinates rounded to
    float, float monomial products and float distances exactly as the
    oracle forms them (numpy float32 arithmetic, no contraction), radius
    R = 1.1 max(10 eps, max_j d_j) in double, then the moments and the solve
    A^-1 e_0 in extended precision (no truncation). Returns (coeffs,
    lambda_min), the smallest eigenvalue of the double moment matrix, by
    which a caller can leave out the systems whose truncation is in doubt."""
    src = np.asarray(src, dtype=np.float64)
    tgt = np.asarray(tgt, dtype=np.float64)
    nbr = np.asarray(nbr)
    u = (src[nbr] - tgt[:, None, :]).astype(np.float32)
    d = np.sqrt((u[..., 0] * u[..., 0] + u[..., 1] * u[..., 1]) + u[..., 2] * u[..., 2])
    dmax = np.maximum(d.max(axis=1).astype(np.float64), 10 * np.finfo(np.float64).eps)
    r = 1.1 * dmax
    x = d.astype(np.float64) / r[:, None]
    phi = _rbf(rbf, x.astype(LD))
    p = _basis(m, u).astype(np.float64)  # float products, exact in double
    a = np.einsum("tki,tk,tkj->tij", p.astype(LD), phi, p.astype(LD))
    e0 = np.zeros((len(tgt), m, 1), dtype=LD)
    e0[:, 0, 0] = 1
    a0 = _solve(a, e0)[:, :, 0]
    c = np.einsum("ti,tki->tk", a0, p.astype(LD)) * phi
    lam = np.linalg.eigvalsh(a.astype(np.float64))[:, 0]
    return c.astype(np.float64), lam

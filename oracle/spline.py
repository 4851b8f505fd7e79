"""CPU oracle for the compactly supported RBF spline -- TEST INFRASTRUCTURE ONLY.

A restatement of DataTransferKit's SplineOperator
(packages/Meshfree/src/DTK_SplineOperator_def.hpp:39-295) on one rank:

* M (def.hpp:211-212) and N (def.hpp:215-216) are basis operators
  (buildBasisOperator, def.hpp:46-149): row i holds the kNN of point i among
  the sources, k = PolynomialBasis::size = 4 (def.hpp:188), with the values
  rbf(|x_l - t_i| / R_i), R_i = 1.1 max_l |x_l - t_i|. Columns are source
  ids. M's rows are the sources, N's the targets.
* P and Q are PolynomialMatrix operators (DTK_DetailsPolynomialMatrix.hpp
  :71-196) over the Vandermonde rows {1, x, y, z} of the float-rounded
  absolute coordinates (computeVandermonde2, Impl.hpp:193-211):
  NO_TRANS adds V z_poly to the point rows; TRANS writes V^T z_src into the
  four polynomial entries.
* S (DTK_DetailsSplineProlongationOperator.hpp:41-108) extends a source
  vector with four zero polynomial entries (range map: the source ids, then
  max_id + 1 .. max_id + 4, appended on rank 0).
* The operator is A = (Q + N) (P + M + P^T)^-1 S (def.hpp:222-246): with
  C = [[M, V_s], [V_s^T, 0]], z = C^-1 [x; 0] and y = V_t z_poly + N z_src.

The reference solves C z = S x with Belos pseudo-block GMRES to a relative
residual of 1e-10 (def.hpp:248-258). The oracle solves it directly
(scipy's sparse LU of the system reduced to the row space of V_s, see
``SplineOracle``), so its answer is the minimum-norm solution of the same
system up to the LU's rounding; an iterative solution differs from it by
at most cond(C) times its residual.

The kNN rows come from the C++ oracle's k-d tree (ascending (distance,
index), pinned against brute force) and the weights from its
``dtkx_oracle_basis_weights``. Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import GEOMETRY_FLOAT, basis_weights, knn

K_SPLINE = 4  # PolynomialBasis::size of MultivariatePolynomialBasis<Linear, 3>


def vandermonde(points, geometry=GEOMETRY_FLOAT):
    """computeVandermonde2 (Impl.hpp:193-211) with the Linear basis
    (DTK_MultivariatePolynomialBasis.hpp:108-114): {1, x, y, z} of the point
    as ArborX::Point holds it (float), stored as double."""
    p = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    if geometry == GEOMETRY_FLOAT:
        p = p.astype(np.float32).astype(np.float64)
    return np.concatenate([np.ones((p.shape[0], 1)), p], axis=1)


def basis_operator(src, pts, rbf_name="wendland0", geometry=GEOMETRY_FLOAT):
    """buildBasisOperator (def.hpp:46-149): (indices [n, kk], phi [n, kk],
    scipy CSR [n, n_src])."""
    src = np.asarray(src, dtype=np.float64).reshape(-1, 3)
    pts = np.asarray(pts, dtype=np.float64).reshape(-1, 3)
    ns, n = src.shape[0], pts.shape[0]
    idx, _ = knn(src, pts, K_SPLINE, geometry=geometry)
    phi = basis_weights(src, pts, idx, rbf_name=rbf_name, geometry=geometry)
    kk = idx.shape[1]
    rows = np.repeat(np.arange(n), kk)
    mat = sp.csr_matrix((phi.ravel(), (rows, idx.ravel())), shape=(n, ns))
    return idx, phi, mat


class SplineOracle:
    """The four blocks of def.hpp:203-217 and the direct solve of the
    saddle-point system."""

    def __init__(self, src, tgt, rbf_name="wendland0", geometry=GEOMETRY_FLOAT):
        self.src = np.asarray(src, dtype=np.float64).reshape(-1, 3)
        self.tgt = np.asarray(tgt, dtype=np.float64).reshape(-1, 3)
        self.ns, self.nt = self.src.shape[0], self.tgt.shape[0]
        self.m_idx, self.m_phi, self.M = basis_operator(self.src, self.src,
                                                        rbf_name, geometry)
        self.n_idx, self.n_phi, self.N = basis_operator(self.src, self.tgt,
                                                        rbf_name, geometry)
        self.Vs = vandermonde(self.src, geometry)
        self.Vt = vandermonde(self.tgt, geometry)
        # Planar or collinear sources make V_s rank-deficient and C singular,
        # with null space {(0, w) : V_s w = 0}, which is also C^T's. GMRES
        # from a zero start keeps its iterates orthogonal to it, so the
        # reference's solution is the minimum-norm one: z_poly in the row
        # space of V_s. B spans that row space; the reduced system in
        # (z_src, u), z_poly = B u, is nonsingular.
        _, sv, vt = np.linalg.svd(self.Vs, full_matrices=False)
        rank = int(np.sum(sv > sv[0] * max(self.Vs.shape) * np.finfo(float).eps))
        self.B = vt[:rank].T
        vb = sp.csr_matrix(self.Vs @ self.B)
        self.C_reduced = sp.bmat([[self.M, vb], [vb.T, None]], format="csc")
        self._lu = None

    def saddle_apply(self, z):
        """C z = P z + M z + P^T z (def.hpp:230-234)."""
        z = np.asarray(z, dtype=np.float64)
        zs, zp = z[:self.ns], z[self.ns:]
        return np.concatenate([self.Vs @ zp + self.M @ zs, self.Vs.T @ zs])

    def solve(self, x):
        """z = C^-1 S x for one source field x [ns]."""
        if self._lu is None:
            self._lu = spla.splu(self.C_reduced)
        b = np.concatenate([np.asarray(x, dtype=np.float64),
                            np.zeros(self.B.shape[1])])
        w = self._lu.solve(b)
        return np.concatenate([w[:self.ns], self.B @ w[self.ns:]])

    def apply(self, x):
        """y = (Q + N) z (def.hpp:240-246, 278-283)."""
        z = self.solve(x)
        return self.Vt @ z[self.ns:] + self.N @ z[:self.ns]

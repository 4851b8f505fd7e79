"""Pins of the spline oracle (oracle/spline.py) on the CPU.

* The reference's own SplineOperator tests (tstMeshfreeOperators.cpp
  :130-448, Spline instantiations :504-516) with their tolerances:
  same_npoints_and_basis 2e-7, grid 1e-9, line 2e-9, single_point_in_radius
  finite.
* Closed forms: basis weights of a unit-spaced line (R = 1.1 * the 4th
  neighbour distance), interpolation at the sources, reproduction of linear
  functions, the side condition V_s^T z_src = 0.
* The minimum-norm reduction for planar / collinear sources against
  numpy's dense pseudo-inverse of the full singular C.
"""
import numpy as np
import pytest

from conftest import grid_points, load_golden, poly_for_basis


@pytest.fixture(scope="module")
def spl(oracle_lib):
    from oracle import spline

    return spline


F = poly_for_basis(4)


def rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(a), np.abs(b)))


def test_reference_spline_tests(spl):
    src = load_golden("mls_points_m4_src", (-1, 3))
    tgt = load_golden("mls_points_m4_tgt", (-1, 3))
    assert rel(spl.SplineOracle(src, tgt).apply(F(src)), F(tgt)) <= 2e-7
    g = grid_points((40, 40, 1), (0, 0, 0))
    t = grid_points((1, 1, 1), (19, 19, 0))
    assert rel(spl.SplineOracle(g, t).apply(F(g)), F(t)) <= 1e-9
    line = grid_points((40, 1, 1), (0, 0, 0))
    lt = grid_points((39, 1, 1), (0.5, 0, 0))
    assert rel(spl.SplineOracle(line, lt).apply(F(line)), F(lt)) <= 2e-9
    s1 = load_golden("mls_points_m1_src", (-1, 3))
    t1 = load_golden("mls_points_m1_tgt", (-1, 3))
    assert np.all(np.isfinite(spl.SplineOracle(s1, t1).apply(F(s1))))


def test_line_weights_closed_form(spl):
    line = grid_points((10, 1, 1), (0, 0, 0))
    idx, phi, _ = spl.basis_operator(line, line)
    # interior point 4: itself, 3 and 5 at 1, then 2 (smaller index) at 2
    assert list(idx[4]) == [4, 3, 5, 2]
    r = 1.1 * 2.0
    np.testing.assert_array_equal(phi[4], [1.0, (1 - 1 / r) ** 2, (1 - 1 / r) ** 2,
                                           (1 - 2 / r) ** 2])
    # end point 0: 1, 2, 3 at 1, 2, 3
    r = 1.1 * 3.0
    np.testing.assert_allclose(phi[0], [1.0, (1 - 1 / r) ** 2, (1 - 2 / r) ** 2,
                                        (1 - 3 / r) ** 2], rtol=1e-15)


def test_weights_match_float_distances(spl, oracle_lib):
    """phi from the reference's float distance of the float-rounded relative
    coordinate (Impl.hpp:121-125), formed here with numpy float32."""
    rng = np.random.default_rng(5)
    src = rng.random((300, 3))
    tgt = rng.random((200, 3))
    idx, phi, _ = spl.basis_operator(src, tgt)
    xr = (src[idx] - tgt[:, None, :]).astype(np.float32)
    s = np.zeros(idx.shape, dtype=np.float32)
    for d in range(3):
        tmp = np.float32(0) - xr[..., d]
        s = s + tmp * tmp
    dist = np.sqrt(s).astype(np.float64)
    radius = 1.1 * np.maximum(dist.max(axis=1), 10 * np.finfo(float).eps)
    np.testing.assert_array_equal(phi, (1 - dist / radius[:, None]) ** 2)


def test_interpolates_sources_and_reproduces_linear(spl):
    rng = np.random.default_rng(11)
    src = rng.random((500, 3))
    tgt = rng.random((400, 3))
    x = np.sin(3 * src[:, 0]) + src[:, 1] * src[:, 2]
    # targets = sources: (Q + N) z = (P + M) z, the source rows of C z = [x; 0]
    np.testing.assert_allclose(spl.SplineOracle(src, src).apply(x), x,
                               rtol=0, atol=1e-12)
    o = spl.SplineOracle(src, tgt)
    # linear data: z_src = 0 and z_poly = the coefficients, exactly
    # representable up to the float rounding of V's coordinates
    y = o.apply(F(src))
    np.testing.assert_allclose(y, F(tgt), rtol=2e-7)
    z = o.solve(x)
    assert np.max(np.abs(o.Vs.T @ z[:500])) <= 1e-12 * np.max(np.abs(z[:500]))
    np.testing.assert_allclose(o.saddle_apply(z), np.concatenate([x, np.zeros(4)]),
                               rtol=0, atol=1e-12)


@pytest.mark.parametrize("shape", [(12, 1, 1), (6, 5, 1)])
def test_min_norm_solution_of_singular_system(spl, shape):
    pts = grid_points(shape, (0, 0, 0)) * 0.37 + 0.1
    tgt = grid_points((3, 1, 1), (0.25, 0, 0))
    o = spl.SplineOracle(pts, tgt)
    x = np.cos(pts[:, 0]) + 0.5 * pts[:, 1]
    vs = o.Vs
    c = np.block([[o.M.toarray(), vs], [vs.T, np.zeros((4, 4))]])
    assert np.linalg.matrix_rank(c) < c.shape[0]
    b = np.concatenate([x, np.zeros(4)])
    np.testing.assert_allclose(o.solve(x), np.linalg.pinv(c) @ b, rtol=0,
                               atol=1e-10)

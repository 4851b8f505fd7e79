"""GPU parity of the RBF spline operator (dtkx_spline_*, the reference's
SplineOperator, packages/Meshfree/src/DTK_SplineOperator_def.hpp:39-295)
against the CPU oracle (oracle/spline.py).

Bars:
* basis operators M and N: kNN indices and rbf values bit-exact against the
  oracle (float and double coordinates, two RBFs);
* C z and (Q + N) z: within 1e-14 of the oracle's sums (summation order);
* the GMRES solution: explicit relative residual |[x; 0] - C z| / |x| at
  most 1e-10 (the reference's Belos tolerance, def.hpp:258) measured with
  the oracle's C, and the transferred values within 1e-7 (relative to the
  largest value) of the oracle's direct solve: an iterative solution
  differs from the exact one by at most cond(C) times its residual;
* the reference's own Spline tests (tstMeshfreeOperators.cpp:130-448) with
  their tolerances.
"""
import numpy as np
import pytest

from conftest import grid_points, load_golden, poly_for_basis

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

F = poly_for_basis(4)


@pytest.fixture(scope="module")
def dtk():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import datatransferkit_amd as d

    return d


@pytest.fixture(scope="module")
def spl(oracle_lib):
    from oracle import spline

    return spline


def cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


def rel(a, b):
    return np.max(np.abs(a - b) / np.maximum(np.abs(a), np.abs(b)))


def _apply(op, x, nt):
    y = torch.zeros(nt, dtype=torch.float64, device="cuda")
    op.apply(cuda(x), y)
    return y.cpu().numpy()


def cloud(n_src, n_tgt, seed):
    rng = np.random.default_rng(seed)
    src = rng.random((n_src, 3))
    tgt = rng.random((n_tgt, 3)) * 1.1 - 0.05
    x = np.sin(2 * np.pi * src[:, 0]) * np.cos(2 * np.pi * src[:, 1]) + src[:, 2] ** 2
    return src, tgt, x


@pytest.mark.parametrize("geometry,rbf", [("float", "wendland0"), ("double", "wendland0"),
                                          ("float", "wendland2")])
def test_basis_operators_bit_exact(dtk, spl, geometry, rbf):
    src, tgt, _ = cloud(20_000, 15_000, 3)
    print(f"spline basis {geometry} {rbf}", flush=True)
    g = 0 if geometry == "float" else 1
    op = dtk.SplineOperator(cuda(src), cuda(tgt), geometry=geometry, rbf=rbf)
    for which, pts in (("M", src), ("N", tgt)):
        gi, gp, _ = spl.basis_operator(src, pts, rbf, g)
        idx, val = op.basis(which)
        assert np.array_equal(idx.cpu().numpy(), gi), which
        assert np.array_equal(bits(val.cpu().numpy()), bits(gp)), which


def test_system_apply_and_evaluate(dtk, spl):
    src, tgt, _ = cloud(5_000, 4_000, 7)
    o = spl.SplineOracle(src, tgt)
    op = dtk.SplineOperator(cuda(src), cuda(tgt))
    z = np.random.default_rng(1).standard_normal(len(src) + 4)
    y = torch.empty(len(src) + 4, dtype=torch.float64, device="cuda")
    op.system_apply(cuda(z), y)
    ref = o.saddle_apply(z)
    np.testing.assert_allclose(y.cpu().numpy(), ref, rtol=0,
                               atol=1e-14 * np.max(np.abs(ref)) * 8)
    yt = torch.empty(len(tgt), dtype=torch.float64, device="cuda")
    op.evaluate(cuda(z), yt)
    ref = o.Vt @ z[len(src):] + o.N @ z[:len(src)]
    np.testing.assert_allclose(yt.cpu().numpy(), ref, rtol=0,
                               atol=1e-14 * np.max(np.abs(ref)) * 8)


@pytest.mark.parametrize("n,geometry", [(2_000, "float"), (30_000, "float"),
                                        (10_000, "double")])
def test_solve_residual_and_transfer_match_oracle(dtk, spl, n, geometry):
    src, tgt, x = cloud(n, n, 9)
    print(f"spline solve n={n} {geometry}", flush=True)
    o = spl.SplineOracle(src, tgt, geometry=0 if geometry == "float" else 1)
    op = dtk.SplineOperator(cuda(src), cuda(tgt), geometry=geometry)
    z = torch.empty(n + 4, dtype=torch.float64, device="cuda")
    op.solve(cuda(x), z)
    info = op.info()
    assert info["converged"] and info["residual"] <= 1e-10, info
    zc = z.cpu().numpy()
    b = np.concatenate([x, np.zeros(4)])
    assert np.linalg.norm(o.saddle_apply(zc) - b) <= 1.05e-10 * np.linalg.norm(b)
    y = _apply(op, x, n)
    yo = o.apply(x)
    assert np.max(np.abs(y - yo)) <= 1e-7 * np.max(np.abs(yo))
    assert op.info()["iterations"] == info["iterations"]


def test_reference_spline_tests(dtk):
    # same_npoints_and_basis (tstMeshfreeOperators.cpp:130-224)
    src = load_golden("mls_points_m4_src", (-1, 3))
    tgt = load_golden("mls_points_m4_tgt", (-1, 3))
    op = dtk.SplineOperator(cuda(src), cuda(tgt))
    assert rel(_apply(op, F(src), len(tgt)), F(tgt)) <= 2e-7
    # grid (:226-300) and line (:302-376): planar / collinear sources, C
    # singular, the solve's minimum-norm solution
    g = grid_points((40, 40, 1), (0, 0, 0))
    t = grid_points((1, 1, 1), (19, 19, 0))
    op = dtk.SplineOperator(cuda(g), cuda(t))
    assert rel(_apply(op, F(g), 1), F(t)) <= 1e-9
    line = grid_points((40, 1, 1), (0, 0, 0))
    lt = grid_points((39, 1, 1), (0.5, 0, 0))
    op = dtk.SplineOperator(cuda(line), cuda(lt))
    assert rel(_apply(op, F(line), 39), F(lt)) <= 2e-9
    # single_point_in_radius (:378-448)
    s1 = load_golden("mls_points_m1_src", (-1, 3))
    t1 = load_golden("mls_points_m1_tgt", (-1, 3))
    op = dtk.SplineOperator(cuda(s1), cuda(t1))
    assert np.all(np.isfinite(_apply(op, F(s1), 10)))


def test_singular_systems_match_min_norm_oracle(dtk, spl):
    for shape in [(40, 1, 1), (30, 20, 1)]:
        pts = grid_points(shape, (0, 0, 0)) * 0.37 + 0.1
        tgt = np.random.default_rng(2).random((50, 3)) * pts.max(axis=0)
        x = np.cos(pts[:, 0]) + 0.5 * pts[:, 1]
        print(f"spline singular {shape}", flush=True)
        o = spl.SplineOracle(pts, tgt)
        op = dtk.SplineOperator(cuda(pts), cuda(tgt))
        z = torch.empty(len(pts) + 4, dtype=torch.float64, device="cuda")
        op.solve(cuda(x), z)
        assert op.info()["converged"]
        np.testing.assert_allclose(z.cpu().numpy(), o.solve(x), rtol=0, atol=1e-8)


def test_fields_host_arrays_and_determinism(dtk, spl):
    src, tgt, x = cloud(3_000, 2_000, 13)
    op = dtk.SplineOperator(cuda(src), cuda(tgt))
    xs = np.stack([x, -2 * x + 1, np.zeros_like(x)])
    y = torch.zeros((3, len(tgt)), dtype=torch.float64, device="cuda")
    op.apply(cuda(xs), y)
    y = y.cpu().numpy()
    for f in range(3):
        assert np.array_equal(bits(y[f]), bits(_apply(op, xs[f], len(tgt))))
    assert np.all(y[2] == 0)
    # host arrays in and out, SoA points
    yh = np.zeros(len(tgt))
    op2 = dtk.SplineOperator(np.ascontiguousarray(src.T), np.ascontiguousarray(tgt.T),
                             layout="soa")
    op2.apply(x, yh)
    assert np.array_equal(bits(yh), bits(y[0]))


def test_corner_cases(dtk, spl):
    # fewer sources than neighbours (kk = 2) and no targets
    src = np.array([[1.0, 1, 1], [2, 2, 2]])
    tgt = np.array([[1.5, 1.5, 1.5], [1.0, 1.0, 1.0]])
    op = dtk.SplineOperator(cuda(src), cuda(tgt))
    y = _apply(op, np.array([255.0, 255.0]), 2)
    np.testing.assert_allclose(y, spl.SplineOracle(src, tgt).apply(np.array([255.0, 255.0])),
                               rtol=1e-9)
    # one source: kk = 1, V_s of rank 1, the minimum-norm solution
    one = np.array([[0.3, 0.4, 0.5]])
    y = _apply(op := dtk.SplineOperator(cuda(one), cuda(tgt)), np.array([7.0]), 2)
    assert op.info()["converged"]
    np.testing.assert_allclose(y, spl.SplineOracle(one, tgt).apply(np.array([7.0])),
                               rtol=1e-9)
    op = dtk.SplineOperator(cuda(src), cuda(np.zeros((0, 3))))
    op.apply(cuda(np.array([1.0, 2.0])), torch.zeros(0, dtype=torch.float64, device="cuda"))
    with pytest.raises(dtk.DtkxError):
        dtk.SplineOperator(cuda(np.zeros((0, 3))), cuda(tgt))
    with pytest.raises(dtk.DtkxError):
        dtk.SplineOperator(cuda(src), cuda(tgt), restart=513)


def test_restarts_and_iteration_cap(dtk, spl):
    """Restarted cycles (restart 10) still reach the tolerance and the same
    transfer; an iteration cap below what the tolerance needs ends the solve
    unconverged with its explicit residual reported."""
    src, tgt, x = cloud(8_000, 5_000, 17)
    o = spl.SplineOracle(src, tgt)
    yo = o.apply(x)
    op = dtk.SplineOperator(cuda(src), cuda(tgt), restart=10)
    y = _apply(op, x, len(tgt))
    info = op.info()
    assert info["converged"] and info["restarts"] >= 3 and info["residual"] <= 1e-10, info
    assert np.max(np.abs(y - yo)) <= 1e-7 * np.max(np.abs(yo))
    op = dtk.SplineOperator(cuda(src), cuda(tgt), max_iterations=5)
    y = _apply(op, x, len(tgt))
    info = op.info()
    assert not info["converged"] and info["iterations"] == 5, info
    assert 1e-10 < info["residual"] < 1.0
    z = torch.empty(len(src) + 4, dtype=torch.float64, device="cuda")
    op.solve(cuda(x), z)
    b = np.concatenate([x, np.zeros(4)])
    r = np.linalg.norm(o.saddle_apply(z.cpu().numpy()) - b) / np.linalg.norm(b)
    assert abs(r - op.info()["residual"]) <= 1e-12 + 1e-6 * r

"""FP32-bulk form: float geometry, weights, Vandermonde, coefficients and
apply (FP32 accumulation); FP64 moments, Jacobi and truncation. Its values
must stay within 1e-6 relative (infinity norm) of the FP64 oracle on
well-posed systems (overdetermined, k > basis size)."""
import numpy as np
import pytest

from conftest import c1_inputs, grid_points, load_golden, poly_for_basis

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dtk():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import datatransferkit_amd as d

    return d


def cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


@pytest.mark.parametrize("solver", ["cyclic", "cholesky"])
@pytest.mark.parametrize("m,basis,k,rbf", [
    (10, "quadratic", 32, "wendland2"),
    (10, "quadratic", 16, "wendland2"),
    (4, "linear", 16, "wendland4"),
    (10, "quadratic", 64, "wendland4"),
])
def test_fp32_bulk_against_fp64_oracle(dtk, oracle_lib, m, basis, k, rbf, solver):
    src, tgt, f = c1_inputs(20_000, seed=80 + k)
    tgt = tgt[:5000] * 0.9 + 0.05
    fields = np.stack([f, np.cos(3 * src[:, 0]) * src[:, 1], src[:, 2] - 0.5])
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis=basis, k=k, rbf=rbf,
                                        precision="fp32_bulk", solver=solver)
    _, idx, coeffs = op.graph()
    assert coeffs.dtype == torch.float32
    gi, gc, _ = oracle_lib.mls_create(src, tgt, m, k=k, rbf_name=rbf)
    assert np.array_equal(idx.cpu().numpy(), gi)
    ref = oracle_lib.apply(gi, gc, fields)
    # float values in, float values out (the halved-traffic path)
    y32 = torch.zeros((3, len(tgt)), dtype=torch.float32, device="cuda")
    op.apply(cuda(fields, torch.float32), y32)
    # double values at the boundary
    y64 = torch.zeros((3, len(tgt)), dtype=torch.float64, device="cuda")
    op.apply(cuda(fields), y64)
    for y in (y32.double().cpu().numpy(), y64.cpu().numpy()):
        rel = np.abs(y - ref).max(axis=1) / np.abs(ref).max(axis=1)
        print(f"fp32 {solver} m={m} k={k}: per-field rel_inf {rel}")
        assert np.all(rel < 1e-6)


@pytest.mark.parametrize("m,basis", [(4, "linear"), (10, "quadratic")])
def test_fp32_bulk_reference_tests(dtk, m, basis):
    f = poly_for_basis(m)
    src = load_golden(f"mls_points_m{m}_src", (-1, 3))
    tgt = load_golden(f"mls_points_m{m}_tgt", (-1, 3))
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis=basis,
                                        precision="fp32_bulk")
    y = torch.zeros(len(tgt), dtype=torch.float64, device="cuda")
    op.apply(cuda(f(src)), y)
    # float weights and coefficients: reproduction to single precision
    np.testing.assert_allclose(y.cpu().numpy(), f(tgt), rtol=2e-5)
    g = grid_points((40, 40, 1), (0, 0, 0))
    t = grid_points((1, 1, 1), (19, 19, 0))
    op = dtk.MovingLeastSquaresOperator(cuda(g), cuda(t), basis=basis, precision="fp32_bulk")
    y = torch.zeros(1, dtype=torch.float64, device="cuda")
    op.apply(cuda(f(g)), y)
    np.testing.assert_allclose(y.cpu().numpy(), f(t), rtol=1e-6)


def test_fp32_requires_fp32_handle(dtk):
    src, tgt, f = c1_inputs(1000, seed=3)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt))
    y = torch.zeros(len(tgt), dtype=torch.float32, device="cuda")
    with pytest.raises(dtk.DtkxError) as e:
        op.apply(cuda(f, torch.float32), y)
    assert e.value.status == 7

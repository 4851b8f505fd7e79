"""The fast FP64 setups on the GPU: the cyclic symmetric Jacobi
eigen-solver, and the certified Cholesky solve with its cyclic fallback.

It computes the same pseudo-inverse as the reference's two-sided Jacobi up
to rounding, so it is checked (a) on the reference's own MLS tests at their
tolerances, (b) against the oracle on random clouds with an error bound
scaled by the moment matrices' conditioning, and (c) for the same
neighbour graph and the same truncation decisions on well-posed systems.
"""
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


def _apply(op, vals, nt):
    out = torch.zeros(nt, dtype=torch.float64, device="cuda")
    op.apply(cuda(vals), out)
    return out.cpu().numpy()


SOLVERS = ["cyclic", "cholesky"]


@pytest.mark.parametrize("solver", SOLVERS)
@pytest.mark.parametrize("m,basis", [(1, "constant"), (4, "linear"), (10, "quadratic")])
def test_reference_mls_tests_cyclic(dtk, m, basis, solver):
    f = poly_for_basis(m)
    src = load_golden(f"mls_points_m{m}_src", (-1, 3))
    tgt = load_golden(f"mls_points_m{m}_tgt", (-1, 3))
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis=basis, solver=solver)
    np.testing.assert_allclose(_apply(op, f(src), len(tgt)), f(tgt), rtol=1e-7)
    g = grid_points((40, 40, 1), (0, 0, 0))
    t = grid_points((1, 1, 1), (19, 19, 0))
    op = dtk.MovingLeastSquaresOperator(cuda(g), cuda(t), basis=basis, solver=solver)
    np.testing.assert_allclose(_apply(op, f(g), 1), f(t), rtol=1e-14)
    line = grid_points((40, 1, 1), (0, 0, 0))
    lt = grid_points((39, 1, 1), (0.5, 0, 0))
    op = dtk.MovingLeastSquaresOperator(cuda(line), cuda(lt), basis=basis, solver=solver)
    np.testing.assert_allclose(_apply(op, f(line), 39), f(lt), rtol=1e-14)
    s1 = load_golden("mls_points_m1_src", (-1, 3))
    t1 = load_golden("mls_points_m1_tgt", (-1, 3))
    op = dtk.MovingLeastSquaresOperator(cuda(s1), cuda(t1), basis=basis, solver=solver)
    assert np.all(np.isfinite(_apply(op, f(s1), 10)))


@pytest.mark.parametrize("m,basis,k,rbf,bound", [
    (4, "linear", 4, "wendland0", 1e-6),
    (4, "linear", 16, "wendland2", 1e-12),
    # square quadratic interpolation: moment matrices with condition numbers
    # up to ~1e12, so any Jacobi ordering other than the reference's moves
    # values by ~eps * cond; the classical solver is the bit-exact path
    (10, "quadratic", 10, "wendland2", 1e-3),
    # overdetermined, but the unscaled quadratic basis (monomials ~h^2 next
    # to 1) gives cond(A) ~ 1e7: ~eps * cond
    (10, "quadratic", 32, "wendland2", 1e-10),
    (10, "quadratic", 64, "wendland4", 1e-10),
])
@pytest.mark.parametrize("solver", SOLVERS)
def test_cyclic_against_oracle(dtk, oracle_lib, m, basis, k, rbf, bound, solver):
    src, tgt, f = c1_inputs(20_000, seed=60 + k)
    tgt = tgt[:5000] * 0.9 + 0.05
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis=basis, k=k, rbf=rbf,
                                        solver=solver)
    _, idx, coeffs = op.graph()
    with oracle_lib.libm_mode(1):
        gi, gc, cnt = oracle_lib.mls_create(src, tgt, m, k=k, rbf_name=rbf)
    assert np.array_equal(idx.cpu().numpy(), gi)
    y = _apply(op, f, len(tgt))
    yo = oracle_lib.apply(gi, gc, f)
    rel_inf = np.abs(y - yo).max() / np.abs(yo).max()
    # normalised by sum |c_j f_j|, the scale of the cancellation in each row
    scale = (np.abs(gc) * np.abs(f[gi])).sum(1)
    norm_err = np.abs(y - yo) / scale
    st = op.stats()
    print(f"cyclic m={m} k={k}: rel_inf {rel_inf:.3e} normalised max {norm_err.max():.3e} "
          f"p99 {np.quantile(norm_err, 0.99):.3e} sweeps/target {st['total_rotations'] / len(tgt):.2f} "
          f"under gpu {st['n_underdetermined']} oracle {cnt.n_underdetermined}")
    assert st["n_unconverged"] == 0
    assert rel_inf < bound


def test_cyclic_store_pinv_is_consistent(dtk):
    src, tgt, f = c1_inputs(5000, seed=70)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic", k=16,
                                        rbf="wendland2", solver="cyclic", store_pinv=True)
    op2 = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic", k=16,
                                         rbf="wendland2", solver="cyclic")
    _, i1, c1 = op.graph()
    _, i2, c2 = op2.graph()
    assert torch.equal(i1, i2) and torch.equal(c1, c2)


@pytest.mark.parametrize("k", [10, 16, 32])
def test_cholesky_fallback_is_the_cyclic_solver(dtk, k):
    """Systems the Cholesky path cannot certify go to the cyclic solver, so
    the truncation and convergence counts equal the cyclic solver's, and the
    rows the certificate rejects are the cyclic solver's bit for bit."""
    src, tgt, f = c1_inputs(20_000, seed=70 + k)
    tgt = tgt[:6000] * 0.9 + 0.05
    kw = dict(basis="quadratic", k=k, rbf="wendland2")
    a = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), solver="cyclic", **kw)
    b = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), solver="cholesky", **kw)
    sa, sb = a.stats(), b.stats()
    assert sa["n_underdetermined"] == sb["n_underdetermined"]
    assert sa["n_unconverged"] == sb["n_unconverged"]
    _, ia, ca = a.graph()
    _, ib, cb = b.graph()
    assert torch.equal(ia, ib)
    same = (ca == cb).all(dim=1)
    # every truncated system was set up by the eigen-solver in both handles
    print(f"k={k}: rows identical to the cyclic solver {int(same.sum())} of {len(tgt)}, "
          f"underdetermined {sb['n_underdetermined']}")
    assert int(same.sum()) >= sb["n_underdetermined"]
    # both solvers are backward stable: ~eps * cond(A), cond(A) ~ 1e7 here
    rel = ((ca - cb).abs().max() / ca.abs().max()).item()
    assert rel < (1e-3 if k == 10 else 1e-7)

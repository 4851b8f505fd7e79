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


def _shell_inputs(n, seed=21):
    """The thin-shell cloud: sources on a spherical shell of radius 1 and
    thickness 1e-3 (uniform directions), targets on the unit sphere jittered
    +-2e-4 radially, four smooth fields."""
    rng = np.random.default_rng(seed)

    def sphere(m):
        v = rng.standard_normal((m, 3))
        return v / np.linalg.norm(v, axis=1, keepdims=True)

    src = sphere(n) * (1 + rng.uniform(-5e-4, 5e-4, (n, 1)))
    tgt = sphere(n) * (1 + rng.uniform(-2e-4, 2e-4, (n, 1)))
    x, y, z = src[:, 0], src[:, 1], src[:, 2]
    vals = np.stack([x, y * z, x * x - y * y, np.sin(3 * x) * np.cos(2 * y) + z])
    return src, tgt, vals


@pytest.mark.parametrize("solver", ["cholesky", "cyclic"])
def test_thin_shell_fp32_bulk_against_fp64(dtk, oracle_lib, solver):
    """Mixed precision on the ill-conditioned cloud (100k -> 100k, the 2M
    shell's density scaled down): quadratic, k = 32, Wendland<4>, relative
    truncation at rcond = 1e-12 (DTK_DetailsSVDImpl.hpp:281-297 with the
    threshold rcond * sigma_max). Every system is truncated there: the
    points lie on a surface, so the quadratic moments are near-singular.
    FP32 bulk must stay within 1e-6 per field of the oracle and of the
    all-FP64 GPU form, which itself is bit-exact to the oracle and truncates
    the same systems."""
    n = 100_000
    src, tgt, vals = _shell_inputs(n)
    print(f"thin shell n={n} solver={solver}", flush=True)
    kw = dict(basis="quadratic", rbf="wendland4", k=32, truncation="relative", rcond=1e-12)
    with oracle_lib.libm_mode(1):
        gi, gc, cnt = oracle_lib.mls_create(src, tgt, 10, k=32, rbf_name="wendland4",
                                            truncation=oracle_lib.TRUNC_RELATIVE,
                                            rcond=1e-12)
    ref = oracle_lib.apply(gi, gc, vals)
    assert cnt.n_underdetermined == n
    op64 = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), precision="fp64",
                                          solver="classical", **kw)
    _, idx, c64 = op64.graph()
    assert np.array_equal(idx.cpu().numpy(), gi)
    assert np.array_equal(c64.cpu().numpy().view(np.uint64), gc.view(np.uint64))
    assert op64.stats()["n_underdetermined"] == cnt.n_underdetermined
    y64 = torch.zeros((4, n), dtype=torch.float64, device="cuda")
    op64.apply(cuda(vals), y64)
    y64 = y64.cpu().numpy()
    op32 = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), precision="fp32_bulk",
                                          solver=solver, **kw)
    y32 = torch.zeros((4, n), dtype=torch.float32, device="cuda")
    op32.apply(cuda(vals, torch.float32), y32)
    y32 = y32.double().cpu().numpy()
    st = op32.stats()
    vs_oracle = np.abs(y32 - ref).max(axis=1) / np.abs(ref).max(axis=1)
    vs_fp64 = np.abs(y32 - y64).max(axis=1) / np.abs(y64).max(axis=1)
    print(f"  per-field rel_inf vs oracle {vs_oracle}, vs all-FP64 GPU {vs_fp64}, "
          f"n_underdetermined {st['n_underdetermined']} (oracle {cnt.n_underdetermined})")
    assert np.all(vs_oracle <= 1e-6) and np.all(vs_fp64 <= 1e-6)
    assert st["n_underdetermined"] == cnt.n_underdetermined


def test_k_equals_m_against_the_extended_precision_operator(dtk, oracle_lib):
    """1M sources, quadratic basis with k = m = 10 (the 1M x 1M, k = 10
    workload's density; 2,000 targets). Each system is square, cond(A)
    reaches 1e13, and the truth is the operator evaluated in extended
    precision from the same float-rounded geometry. Systems whose smallest
    moment eigenvalue is within 1000 eps of the reference's truncation
    threshold are left out (whether they are truncated depends on rounding).
    On the rest the Cholesky forms stay within 1e-8 (FP64) of the truth,
    while the reference's own two-sided Jacobi (the oracle, which the
    classical form matches bit for bit) is off by about 1e-4: a 1e-6 bar
    against the FP64 oracle cannot hold at k = m for any other solver."""
    from oracle import extended

    rng = np.random.default_rng(19)
    src = rng.random((1_000_000, 3))
    tgt = 0.05 + 0.9 * rng.random((2000, 3))
    f = np.stack([np.sin(2 * np.pi * src[:, 0]) * np.cos(2 * np.pi * src[:, 1]) + src[:, 2] ** 2,
                  src[:, 0] * src[:, 1] - src[:, 2]])
    print("k = m = 10 against extended precision", flush=True)
    with oracle_lib.libm_mode(1):
        gi, gc, cnt = oracle_lib.mls_create(src, tgt, 10, k=10, rbf_name="wendland2")
    ce, lam = extended.float_geometry_operator(src, tgt, gi, 10)
    clear = lam > 1e3 * np.finfo(np.float64).eps
    truth = np.einsum("tk,ftk->ft", ce, f[:, gi])

    def err(y):
        return (np.abs(y - truth)[:, clear].max(axis=1) / np.abs(truth[:, clear]).max(axis=1))

    e_oracle = err(oracle_lib.apply(gi, gc, f))
    out = {}
    for name, kw, vdt in (("fp64_cholesky", dict(precision="fp64", solver="cholesky"), torch.float64),
                          ("fp32_bulk", dict(precision="fp32_bulk", solver="cholesky"), torch.float32),
                          ("fp64_classical", dict(precision="fp64", solver="classical"), torch.float64)):
        op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic", k=10,
                                            rbf="wendland2", **kw)
        _, idx, _ = op.graph()
        assert np.array_equal(idx.cpu().numpy(), gi)
        y = torch.zeros((2, len(tgt)), dtype=vdt, device="cuda")
        op.apply(cuda(f, vdt), y)
        out[name] = err(y.double().cpu().numpy())
    print(f"  {clear.sum()} of {len(tgt)} systems clear of the threshold "
          f"(oracle truncates {cnt.n_underdetermined}); rel error vs extended: "
          f"oracle {e_oracle}, " + ", ".join(f"{k} {v}" for k, v in out.items()))
    assert np.all(out["fp64_cholesky"] <= 1e-8)
    assert np.array_equal(out["fp64_classical"], e_oracle)
    # measured: FP64 Cholesky 1.3e-12 to 1.7e-12, FP32 bulk 6.4e-7, oracle
    # 2.0e-7 and 1.3e-8 here; at the 1M x 1M, k = 10 bench: oracle up to
    # 1.9e-6, FP32 bulk up to 1.2e-6, FP64 Cholesky 9e-13
    assert np.all(out["fp32_bulk"] <= 3e-6)

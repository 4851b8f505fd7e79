"""Adjoint of the MLS transfer (pattern 3): dL/d(source values) must equal the
explicit transpose of the operator, and dL/d(target coordinates) must agree
with central finite differences (h = 1e-6, neighbour sets frozen) to 1e-6
relative (infinity norm over the targets). The radius R = 1.1 max_j |x_j - t|
is a maximum, so L is not differentiable where the two farthest neighbours
are equidistant; targets whose two largest distances are closer than 10 h
are left out of the finite-difference comparison (about 1 % of them): a
shift of +-h moves each distance by at most h, so beyond 4 h the farthest
neighbour cannot change, while a central difference across the kink
averages two one-sided slopes. The
stored and the recomputed pseudo-inverse must give the same gradient."""
import numpy as np
import pytest

from conftest import c1_inputs

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


def fields(src, nf, seed):
    rng = np.random.default_rng(seed)
    x, y, z = src[:, 0], src[:, 1], src[:, 2]
    base = [np.sin(6 * x) * np.cos(5 * y) + z * z, x * y - z, np.cos(7 * z) + x,
            np.exp(-x * y) * z]
    return np.stack([base[i % 4] * (1 + 0.1 * i) for i in range(nf)]) + \
        1e-3 * rng.standard_normal((nf, len(src)))


def per_target_loss(dtk, src, tgt, nbr, x, ybar, opts):
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), neighbors=nbr,
                                        geometry="double", **opts)
    y = torch.zeros((x.shape[0], len(tgt)), dtype=torch.float64, device="cuda")
    op.apply(x, y)
    return (y * ybar).sum(dim=0).cpu().numpy()


def radius_gap(src, tgt, nbr):
    """Gap between the two largest neighbour distances of each target."""
    d = np.sqrt(((src[nbr] - tgt[:, None, :]) ** 2).sum(-1))
    d.sort(axis=1)
    return d[:, -1] - d[:, -2] if d.shape[1] > 1 else np.full(len(tgt), np.inf)


def fd_gradient(dtk, src, tgt, nbr, x, ybar, opts, h=1e-6):
    """Central differences of L_t = sum_f ybar_f(t) y_f(t): target t's value
    depends on its own coordinates only, so one shifted operator per
    coordinate and sign differentiates every target at once."""
    g = np.zeros((len(tgt), 3))
    for c in range(3):
        tp, tm = tgt.copy(), tgt.copy()
        tp[:, c] += h
        tm[:, c] -= h
        lp = per_target_loss(dtk, src, tp, nbr, x, ybar, opts)
        lm = per_target_loss(dtk, src, tm, nbr, x, ybar, opts)
        g[:, c] = (lp - lm) / (2 * h)
    return g


CASES = [
    ("quadratic", "wendland2", 16),
    ("quadratic", "wendland4", 24),
    ("linear", "wendland0", 8),
    ("linear", "wu2", 10),
    ("quadratic", "buhmann3", 20),
    ("constant", "wendland2", 6),
]


@pytest.mark.parametrize("basis,rbf,k", CASES)
def test_target_gradient_matches_central_differences(dtk, basis, rbf, k):
    src, tgt, _ = c1_inputs(20_000, seed=90 + k)
    tgt = tgt[:3000] * 0.9 + 0.05
    nf = 8
    x = cuda(fields(src, nf, seed=k))
    rng = np.random.default_rng(k)
    ybar = cuda(rng.standard_normal((nf, len(tgt))))
    opts = dict(basis=basis, rbf=rbf, k=k, solver="cyclic")
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), geometry="double", **opts)
    g = op.target_gradient(x, ybar).cpu().numpy()
    _, nbr, _ = op.graph()
    h = 1e-6
    gfd = fd_gradient(dtk, src, tgt, nbr, x, ybar, opts, h=h)
    smooth = radius_gap(src, tgt, nbr.cpu().numpy()) > 10 * h
    err = np.abs(g - gfd)[smooth].max() / np.abs(gfd)[smooth].max()
    print(f"{basis} {rbf} k={k}: |g - g_fd|_inf / |g_fd|_inf = {err:.3e} over "
          f"{smooth.sum()} of {len(tgt)} targets")
    assert np.all(np.isfinite(g))
    assert smooth.mean() > 0.95
    assert err < 1e-6


@pytest.mark.parametrize("solver", ["cyclic", "cholesky"])
def test_store_and_recompute_agree(dtk, solver):
    src, tgt, _ = c1_inputs(20_000, seed=5)
    tgt = tgt[:4000] * 0.9 + 0.05
    x = cuda(fields(src, 3, seed=1))
    ybar = cuda(np.random.default_rng(2).standard_normal((3, len(tgt))))
    kw = dict(basis="quadratic", rbf="wendland2", k=16, solver=solver, geometry="double")
    g1 = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), **kw).target_gradient(x, ybar)
    op2 = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), store_pinv=True, **kw)
    g2 = op2.target_gradient(x, ybar)
    rel = (g1 - g2).abs().max().item() / g1.abs().max().item()
    print(f"store vs recompute: {rel:.3e}")
    # same eigen-decomposition; A^+ b formed as (V W V^T) b or V (W (V^T b))
    assert rel < 1e-8


def test_cholesky_recompute_matches_cyclic(dtk):
    # the certified Cholesky inverse and the cyclic Jacobi pseudo-inverse
    # agree to eps * cond(A) on full-rank systems
    src, tgt, _ = c1_inputs(20_000, seed=7)
    tgt = tgt[:4000] * 0.9 + 0.05
    x = cuda(fields(src, 3, seed=3))
    ybar = cuda(np.random.default_rng(4).standard_normal((3, len(tgt))))
    kw = dict(basis="quadratic", rbf="wendland2", k=16, geometry="double")
    gc = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), solver="cyclic",
                                        **kw).target_gradient(x, ybar)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), solver="cholesky", **kw)
    gk = op.target_gradient(x, ybar)
    rel = (gc - gk).abs().max().item() / gc.abs().max().item()
    print(f"cholesky vs cyclic recompute: {rel:.3e}")
    assert rel < 1e-8


def test_cholesky_gradient_matches_central_differences(dtk):
    src, tgt, _ = c1_inputs(20_000, seed=96)
    tgt = tgt[:3000] * 0.9 + 0.05
    x = cuda(fields(src, 4, seed=6))
    ybar = cuda(np.random.default_rng(6).standard_normal((4, len(tgt))))
    opts = dict(basis="quadratic", rbf="wendland2", k=16, solver="cholesky")
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), geometry="double", **opts)
    g = op.target_gradient(x, ybar).cpu().numpy()
    _, nbr, _ = op.graph()
    h = 1e-6
    gfd = fd_gradient(dtk, src, tgt, nbr, x, ybar, opts, h=h)
    smooth = radius_gap(src, tgt, nbr.cpu().numpy()) > 10 * h
    err = np.abs(g - gfd)[smooth].max() / np.abs(gfd)[smooth].max()
    print(f"cholesky: |g - g_fd|_inf / |g_fd|_inf = {err:.3e}")
    assert smooth.mean() > 0.95
    assert err < 1e-6


def test_cholesky_fallback_rows_equal_cyclic(dtk):
    # points on a plane: every quadratic moment matrix is singular, so no
    # system is certified and every row takes the cyclic path, bit for bit
    rng = np.random.default_rng(8)
    src = np.concatenate([rng.random((3000, 2)), np.full((3000, 1), 0.5)], axis=1)
    tgt = np.concatenate([0.1 + 0.8 * rng.random((500, 2)), np.full((500, 1), 0.5)],
                         axis=1)
    x = cuda(fields(src, 2, seed=9))
    ybar = cuda(rng.standard_normal((2, len(tgt))))
    kw = dict(basis="quadratic", rbf="wendland2", k=16, geometry="double")
    gc = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), solver="cyclic",
                                        **kw).target_gradient(x, ybar)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), solver="cholesky", **kw)
    assert op.stats()["n_eigen_fallback"] == len(tgt)
    gk = op.target_gradient(x, ybar)
    assert torch.equal(gc, gk)


def test_more_than_eight_fields_accumulate(dtk):
    # fields go through the adjoint in chunks of eight; linear in the fields
    src, tgt, _ = c1_inputs(8000, seed=6)
    tgt = tgt[:1000] * 0.9 + 0.05
    x = cuda(fields(src, 11, seed=3))
    ybar = cuda(np.random.default_rng(4).standard_normal((11, len(tgt))))
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic", k=16,
                                        rbf="wendland2", solver="cyclic", geometry="double")
    g = op.target_gradient(x, ybar)
    parts = sum(op.target_gradient(x[i:i + 1].contiguous(), ybar[i:i + 1].contiguous())
                for i in range(11))
    rel = (g - parts).abs().max().item() / g.abs().max().item()
    print(f"chunked vs per-field sum: {rel:.3e}")
    assert rel < 1e-9


def test_source_gradient_is_the_explicit_transpose(dtk):
    from datatransferkit_amd.autograd import mls_transfer

    src, tgt, _ = c1_inputs(10_000, seed=8)
    tgt = tgt[:4000] * 0.9 + 0.05
    x = cuda(fields(src, 8, seed=5)).requires_grad_(True)
    t = cuda(tgt).requires_grad_(True)
    y = mls_transfer(cuda(src), t, x, basis="quadratic", rbf="wendland2", k=16,
                     solver="cyclic")
    ybar = torch.randn_like(y)
    y.backward(ybar)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic",
                                        rbf="wendland2", k=16, solver="cyclic",
                                        geometry="double")
    _, idx, c = op.graph()
    idx, c = idx.cpu().numpy(), c.cpu().numpy()
    yb = ybar.cpu().numpy()
    ref = np.zeros((8, len(src)))
    for f in range(8):
        np.add.at(ref[f], idx.ravel(), (c * yb[f][:, None]).ravel())
    gx = x.grad.cpu().numpy()
    rel = np.abs(gx - ref).max() / np.abs(ref).max()
    print(f"dL/dx vs explicit C^T: {rel:.3e}")
    assert rel < 1e-12
    gt = op.target_gradient(x.detach(), ybar)
    assert torch.allclose(t.grad, gt, rtol=0, atol=0)


def test_gradcheck_small(dtk):
    from datatransferkit_amd.autograd import mls_transfer

    rng = np.random.default_rng(11)
    src = rng.random((200, 3))
    tgt = 0.2 + 0.6 * rng.random((12, 3))
    opts = dict(basis="quadratic", rbf="wendland2", k=14, solver="cyclic")
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), geometry="double", **opts)
    _, nbr, _ = op.graph()
    s = cuda(src)

    def f(x, t):
        return mls_transfer(s, t, x, neighbors=nbr, **opts)

    x = cuda(rng.standard_normal((2, 200))).requires_grad_(True)
    t = cuda(tgt).requires_grad_(True)
    assert torch.autograd.gradcheck(f, (x, t), eps=1e-6, atol=1e-6, rtol=1e-4)


def test_adjoint_preconditions(dtk):
    src, tgt, f = c1_inputs(1000, seed=3)
    x = cuda(f[None])
    yb = torch.ones((1, len(tgt)), dtype=torch.float64, device="cuda")
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="linear")
    with pytest.raises(dtk.DtkxError) as e:  # float geometry
        op.target_gradient(x, yb)
    assert e.value.status == 7
    with pytest.raises(dtk.DtkxError) as e:  # stored pinv needs the cyclic solver
        dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="linear",
                                       store_pinv=True)
    assert e.value.status == 7


def test_target_gradient_at_scale_against_extended_differences(dtk):
    """1M sources (the adjoint workload's density): the moment matrices of
    monomials of relative coordinates are ill-conditioned enough that a
    double-precision difference quotient at h = 1e-6 carries ~1e-5 noise, so
    the central differences are taken in extended precision (the oracle's
    restatement of the operator) on a sample of targets."""
    from oracle import extended

    rng = np.random.default_rng(31)
    src = rng.random((1_000_000, 3))
    tgt = 0.05 + 0.9 * rng.random((2000, 3))
    nf = 8
    x = fields(src, nf, seed=2)
    ybar = rng.standard_normal((nf, len(tgt)))
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic",
                                        rbf="wendland2", k=16, solver="cyclic",
                                        geometry="double")
    g = op.target_gradient(cuda(x), cuda(ybar)).cpu().numpy()
    _, nbr, _ = op.graph()
    nbr = nbr.cpu().numpy()
    h = 1e-6
    gfd = extended.target_gradient_fd(src, tgt, nbr, x, ybar, 10, "wendland2", h=h)
    smooth = radius_gap(src, tgt, nbr) > 10 * h
    err = np.abs(g - gfd)[smooth].max() / np.abs(gfd)[smooth].max()
    print(f"1M sources: |g - g_fd|_inf / |g_fd|_inf = {err:.3e} over {smooth.sum()} targets")
    assert smooth.mean() > 0.95
    assert err < 1e-6

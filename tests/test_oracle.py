"""Pins the CPU oracle (oracle/dtkx_oracle.cpp) before it is trusted.

Each oracle function is checked against the reference's own known answers
and test expectations, and against what mathematics fixes independently of
the oracle's formulas:

  rbf_eval          tstCompactlySupportedRadialBasisFunctions.cpp:52-108
                    (polynomial products at r = i/10, 1e-8; radius wrapper)
  basis_eval        tstMultivariatePolynomialBasis.cpp:159-209 known answers;
                    float rounding of ArborX points
  SVD pinv          tstSVD.cpp:67-150 inputs regenerated bit-for-bit (golden
                    fixtures) with A+ A = I at 1e-12 and n_underdetermined;
                    Moore-Penrose conditions vs numpy; diagonal/permutation;
                    2x2 closed form (orthogonal factors, diagonalisation)
  knn               brute force (numpy float32 restatement) incl. ties;
                    k-d tree == brute force
  mls setup         tstMeshfreeOperators.cpp:130-448 and
                    tstMeshfreeOperatorsSimpleProblem.cpp:113-147; constant
                    and polynomial reproduction; k = m interpolation; an
                    independent numpy weighted least squares for k > m
  apply/transpose   numpy CSR product; <y, C x> = <C^T y, x>
  Buhmann<2>, <3>   no reference test; pinned by what the mathematics fixes:
                    Buhmann<2> is C^3 and Buhmann<3> C^5 across the edge of
                    their support, so phi(1) = 0 and phi(1 - h) has the
                    leading term phi^(p)(1)/p! h^p (p = 4: 16/24, p = 6:
                    6048/720, derived by hand); Buhmann<2>(0+) = 1/6 and
                    Buhmann<3>(0) = 1
  Buhmann<4>        as the reference writes it (RBF.hpp:139-146) phi(1) =
                    -67.886: not compactly supported and nothing pins it; the
                    library rejects it (DTKX_ERR_UNSUPPORTED)
  libm modes        correctly rounded mode vs mpmath (atan2, sin, cos, log);
                    the two modes on C1 differ by the measured bound that the
                    GPU test asserts
"""
import numpy as np
import pytest

from conftest import c1_inputs, grid_points, load_golden, poly_for_basis


# ------------------------------------------------------------------- RBFs
@pytest.mark.parametrize("name,factors", [
    ("wendland0", ([1.0, -1.0], 2, [1.0])),
    ("wendland2", ([1.0, -1.0], 4, [1.0, 4.0])),
    ("wendland4", ([1.0, -1.0], 6, [3.0, 18.0, 35.0])),
    ("wendland6", ([1.0, -1.0], 8, [1.0, 8.0, 25.0, 32.0])),
    ("wu2", ([1.0, -1.0], 4, [4.0, 16.0, 12.0, 3.0])),
    ("wu4", ([1.0, -1.0], 6, [6.0, 36.0, 82.0, 72.0, 30.0, 5.0])),
])
def test_rbf_polynomial_forms(oracle_lib, name, factors):
    base, power, tail = factors
    poly = np.polynomial.Polynomial(base) ** power * np.polynomial.Polynomial(tail)
    r = np.arange(10) / 10.0
    ref = poly(r)
    got = oracle_lib.rbf(name, 1.0, r)
    np.testing.assert_allclose(got, ref, rtol=1e-8, atol=0)


@pytest.mark.parametrize("name,order,lead", [
    ("buhmann2", 4, 16.0 / 24.0),     # 4th derivative at 1: 48 log 1 + 16
    ("buhmann3", 6, 6048.0 / 720.0),  # 6th derivative at 1: 6048
])
@pytest.mark.parametrize("mode", [0, 1])
def test_buhmann_vanish_at_the_support_edge(oracle_lib, name, order, lead, mode):
    # Buhmann<2> is C^3 and Buhmann<3> C^5 across x = 1: phi(1) = 0 to
    # rounding (terms up to 378 cancel) and phi(1 - h) = lead h^order + O(h^(order+1)),
    # lead = phi^(order)(1) / order! from differentiating the closed forms by
    # hand. Richardson's 2 q(h/2) - q(h), q(h) = phi(1 - h) / h^order, removes
    # the O(h) term. A wrong coefficient breaks phi(1) = 0 or the order.
    with oracle_lib.libm_mode(mode):
        assert abs(oracle_lib.rbf(name, 1.0, np.array([1.0]))[0]) <= 1e-12
        h = np.array([0.04, 0.02])
        v = oracle_lib.rbf(name, 1.0, 1.0 - h)
    q = v / h ** order
    assert np.all(v > 0)
    assert abs((2 * q[1] - q[0]) / lead - 1.0) < 2e-3


def test_buhmann_values_at_zero(oracle_lib):
    # Buhmann<3>(0) = 1; Buhmann<2>(x) -> 1/6 as x -> 0 (x^4 log x -> 0), and at
    # x = 0 exactly the reference's 0 * log(0) is NaN
    assert oracle_lib.rbf("buhmann3", 1.0, np.array([0.0]))[0] == 1.0
    v = oracle_lib.rbf("buhmann2", 1.0, np.array([1e-4, 0.0]))
    assert abs(v[0] - 1.0 / 6.0) < 1e-7 and np.isnan(v[1])


def test_buhmann4_as_written_is_not_compactly_supported(oracle_lib):
    # the reason the library refuses Buhmann<4>
    v = oracle_lib.rbf("buhmann4", 1.0, np.array([1.0]))[0]
    assert abs(v - (99 / 35 - 132 + 9216 / 35 - 11264 / 35 + 198 - 396 / 5 + 1)) < 1e-12
    assert v < -67.8


def test_rbf_radius_wrapper(oracle_lib):
    # RadialBasisFunction divides by the radius and does not clamp x > 1
    # (tstCompactlySupportedRadialBasisFunctions.cpp:97-108 with an identity
    # RBF gives 0.5, 1, 2; Wendland<0> of those is exact below)
    got = oracle_lib.rbf("wendland0", 2.0, np.array([1.0, 2.0, 4.0]))
    assert got.tolist() == [0.25, 0.0, 1.0]


# ------------------------------------------------------------------ basis
def test_basis_known_answers_3d(oracle_lib):
    p = np.array([[0.0, 1.0, 2.0]])
    assert oracle_lib.basis(1, p).tolist() == [[1.0]]
    assert oracle_lib.basis(4, p).tolist() == [[1.0, 0.0, 1.0, 2.0]]
    assert oracle_lib.basis(10, p).tolist() == [[1.0, 0.0, 1.0, 2.0, 0.0, 0.0, 0.0, 1.0, 2.0, 4.0]]


def test_basis_known_answers_2d(oracle_lib):
    assert oracle_lib.basis(1, np.array([[0.0, 0.0]]), dim=2, float_point=False).tolist() == [[1.0]]
    assert oracle_lib.basis(3, np.array([[0.0, 1.0]]), dim=2, float_point=False).tolist() == [[1.0, 0.0, 1.0]]
    assert oracle_lib.basis(6, np.array([[1.0, 2.0]]), dim=2, float_point=False).tolist() == [[1.0, 1.0, 2.0, 1.0, 2.0, 4.0]]


def test_basis_float_point_rounding(oracle_lib):
    p = np.array([[0.1, -0.3, 0.7]])
    got = oracle_lib.basis(10, p, float_point=True)[0]
    f = p[0].astype(np.float32)
    exp = [1.0, f[0], f[1], f[2], f[0] * f[0], f[0] * f[1], f[0] * f[2],
           f[1] * f[1], f[1] * f[2], f[2] * f[2]]
    assert got.tolist() == [float(np.float32(v)) for v in exp]
    assert got[4] != 0.1 * 0.1  # the product is a float product
    gd = oracle_lib.basis(10, p, float_point=False)[0]
    assert gd[4] == 0.1 * 0.1


# -------------------------------------------------------------------- SVD
def _check_identity(mats, pinv, deficient=()):
    # tstSVD.cpp:22-65: result = pinv * A; diag 1 (rel 1e-12) except the
    # deficient indices; off-diagonal: result + 1 == 1 at rel 1e-12
    for a, p in zip(mats, pinv):
        res = p @ a
        n = a.shape[0]
        for i in range(n):
            for j in range(n):
                if i == j and i not in deficient:
                    assert abs(res[i, i] - 1.0) <= 1e-12
                else:
                    assert abs(res[i, j] + 1.0 - 1.0) <= 1e-12


@pytest.mark.parametrize("mode", [0, 1])
def test_svd_full_rank_tstSVD(oracle_lib, mode):
    mats = load_golden("svd_full_rank", (10, 32, 32))
    with oracle_lib.libm_mode(mode):
        pinv, c = oracle_lib.svd_pseudo_inverse(mats)
    _check_identity(mats, pinv)
    assert c.n_underdetermined == 0 and c.n_unconverged == 0


@pytest.mark.parametrize("mode", [0, 1])
def test_svd_rank_deficient_tstSVD(oracle_lib, mode):
    mats = load_golden("svd_rank_deficient", (10, 40, 40))
    assert np.all(mats[:, :, 3] == 0.0)
    with oracle_lib.libm_mode(mode):
        pinv, c = oracle_lib.svd_pseudo_inverse(mats)
    assert c.n_underdetermined == 10  # tstSVD.cpp:146
    _check_identity(mats, pinv, deficient=(3,))


def test_svd_moore_penrose_spd(oracle_lib):
    rng = np.random.default_rng(3)
    for n in (4, 10):
        b = rng.standard_normal((20, n, n + 3))
        mats = b @ b.transpose(0, 2, 1)
        pinv, c = oracle_lib.svd_pseudo_inverse(mats)
        for a, p in zip(mats, pinv):
            ref = np.linalg.pinv(a)
            scale = np.linalg.norm(ref) * np.linalg.cond(a)
            assert np.linalg.norm(p - ref) <= 1e-13 * scale
            assert np.allclose(a @ p @ a, a, rtol=0, atol=1e-10 * np.linalg.norm(a))
            assert np.allclose(p @ a @ p, p, rtol=0, atol=1e-10 * np.linalg.norm(p))
            assert np.allclose(a @ p, (a @ p).T, atol=1e-10)


def test_svd_rank_deficient_spd(oracle_lib):
    # a PSD matrix with an exactly zero row/column: the Jacobi never rotates
    # it, E stays 0 there and the singular value is dropped
    rng = np.random.default_rng(4)
    b = rng.standard_normal((6, 6))
    a = b @ b.T
    a[2, :] = 0.0
    a[:, 2] = 0.0
    pinv, c = oracle_lib.svd_pseudo_inverse(a[None])
    assert c.n_underdetermined == 1
    np.testing.assert_allclose(pinv[0], np.linalg.pinv(a), atol=1e-12 * np.linalg.norm(np.linalg.pinv(a)))
    assert np.all(pinv[0][2, :] == 0.0) and np.all(pinv[0][:, 2] == 0.0)


def test_svd_diagonal_and_permutation(oracle_lib):
    d = np.diag([4.0, 0.5, 2.0, 8.0])
    pinv, c = oracle_lib.svd_pseudo_inverse(d[None])
    assert c.total_rotations == 0
    assert np.array_equal(pinv[0], np.diag([0.25, 2.0, 0.5, 0.125]))
    # A permutation has a zero diagonal: the reference's 2x2 SVD of the
    # block [[0, 1], [1, 0]] gives atan2(0, 0) = 0 for both angles, i.e.
    # identity rotations, so its uncapped loop (SVD.hpp:239) never ends.
    # Here the cap stops it and the system is reported unconverged.
    perm = np.eye(6)[[1, 0, 3, 2, 5, 4]]
    pinv, c = oracle_lib.svd_pseudo_inverse(perm[None], max_rot=500)
    assert c.n_unconverged == 1 and c.total_rotations == 500
    L, R = oracle_lib.svd_2x2(np.array([[0.0, 1.0], [1.0, 0.0]]))
    assert np.array_equal(L, np.eye(2)) and np.array_equal(R, np.eye(2))


def test_svd_2x2_closed_form(oracle_lib):
    rng = np.random.default_rng(5)
    for _ in range(200):
        a = rng.standard_normal((2, 2))
        L, R = oracle_lib.svd_2x2(a)
        # orthogonal factors; L^T A R is diagonal (svd_2x2: U^T A V = E)
        np.testing.assert_allclose(L.T @ L, np.eye(2), atol=1e-15)
        np.testing.assert_allclose(R.T @ R, np.eye(2), atol=1e-15)
        d = L.T @ a @ R
        assert abs(d[0, 1]) <= 1e-14 * np.abs(a).max() and abs(d[1, 0]) <= 1e-14 * np.abs(a).max()
        # singular values up to order
        np.testing.assert_allclose(np.sort(np.abs(np.diag(d))),
                                   np.sort(np.linalg.svd(a, compute_uv=False)), rtol=1e-13)


# -------------------------------------------------------------------- kNN
def _numpy_brute_knn(src, tgt, k):
    s32 = src.astype(np.float32)
    t32 = tgt.astype(np.float32)
    out = []
    for q in t32:
        tmp = s32 - q  # tmp = b - a
        s = np.float32(0) + tmp[:, 0] * tmp[:, 0]
        s = s + tmp[:, 1] * tmp[:, 1]
        s = s + tmp[:, 2] * tmp[:, 2]
        d = np.sqrt(s)
        order = np.lexsort((np.arange(len(src)), d))
        out.append(order[:min(k, len(src))])
    return np.array(out, dtype=np.int32)


def test_knn_brute_matches_numpy(oracle_lib):
    rng = np.random.default_rng(6)
    src = rng.random((500, 3))
    tgt = rng.random((60, 3))
    for k in (1, 4, 10, 32):
        idx, _ = oracle_lib.knn(src, tgt, k, brute=True)
        assert np.array_equal(idx, _numpy_brute_knn(src, tgt, k))


def test_knn_ties_go_to_lower_index(oracle_lib):
    src = grid_points((5, 5, 1), (0.0, 0.0, 0.0))
    tgt = np.array([[2.0, 2.0, 0.0], [2.5, 2.5, 0.0]])
    idx, d = oracle_lib.knn(src, tgt, 5)
    # (2,2) itself, then four at distance 1 in index order
    assert idx[0].tolist() == [12, 7, 11, 13, 17]
    assert idx[1].tolist()[:4] == [12, 13, 17, 18]
    assert np.array_equal(idx, _numpy_brute_knn(src, tgt, 5))


@pytest.mark.parametrize("geometry", [0, 1])
def test_knn_tree_equals_brute(oracle_lib, geometry):
    rng = np.random.default_rng(7)
    src = rng.random((3000, 3))
    # duplicated sources and a lattice to force ties
    src = np.concatenate([src, src[:200], grid_points((6, 6, 6), (0, 0, 0)) / 5.0])
    tgt = np.concatenate([rng.random((400, 3)), src[:50]])
    for k in (1, 4, 10, 16, 64):
        a, da = oracle_lib.knn(src, tgt, k, geometry=geometry)
        b, db = oracle_lib.knn(src, tgt, k, geometry=geometry, brute=True)
        assert np.array_equal(a, b)
        assert np.array_equal(da, db)


def test_knn_fewer_sources_than_k(oracle_lib):
    src = np.array([[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]])
    idx, _ = oracle_lib.knn(src, np.array([[1.5, 1.5, 1.5]]), 4)
    assert idx.tolist() == [[0, 1]]


# ------------------------------------------------------------- MLS setup
def _run(oracle_lib, src, tgt, m, **kw):
    idx, c, cnt = oracle_lib.mls_create(src, tgt, m, **kw)
    return idx, c, cnt


@pytest.mark.parametrize("m", [1, 4, 10])
def test_mls_same_npoints_and_basis(oracle_lib, m):
    src = load_golden(f"mls_points_m{m}_src", (-1, 3))
    tgt = load_golden(f"mls_points_m{m}_tgt", (-1, 3))
    f = poly_for_basis(m)
    idx, c, _ = _run(oracle_lib, src, tgt, m)
    y = oracle_lib.apply(idx, c, f(src))
    np.testing.assert_allclose(y, f(tgt), rtol=1e-7)  # tstMeshfreeOperators.cpp:215-223


@pytest.mark.parametrize("m", [1, 4, 10])
def test_mls_grid(oracle_lib, m):
    src = grid_points((40, 40, 1), (0.0, 0.0, 0.0))
    tgt = grid_points((1, 1, 1), (19.0, 19.0, 0.0))
    f = poly_for_basis(m)
    idx, c, _ = _run(oracle_lib, src, tgt, m)
    np.testing.assert_allclose(oracle_lib.apply(idx, c, f(src)), f(tgt), rtol=1e-14)


@pytest.mark.parametrize("m", [1, 4, 10])
def test_mls_line(oracle_lib, m):
    src = grid_points((40, 1, 1), (0.0, 0.0, 0.0))
    tgt = grid_points((39, 1, 1), (0.5, 0.0, 0.0))
    f = poly_for_basis(m)
    idx, c, cnt = _run(oracle_lib, src, tgt, m)
    np.testing.assert_allclose(oracle_lib.apply(idx, c, f(src)), f(tgt), rtol=1e-14)
    if m > 1:
        assert cnt.n_underdetermined == 39  # collinear: rank deficient


@pytest.mark.parametrize("m", [1, 4, 10])
def test_mls_single_point_in_radius(oracle_lib, m):
    src = load_golden("mls_points_m1_src", (-1, 3))
    tgt = load_golden("mls_points_m1_tgt", (-1, 3))
    idx, c, _ = _run(oracle_lib, src, tgt, m)
    assert np.all(np.isfinite(oracle_lib.apply(idx, c, poly_for_basis(m)(src))))


def test_mls_corner_cases(oracle_lib):
    idx, c, _ = _run(oracle_lib, np.array([[1.0, 1, 1], [2, 2, 2]]), np.array([[1.5, 1.5, 1.5]]), 4)
    np.testing.assert_allclose(oracle_lib.apply(idx, c, np.array([255.0, 255.0])), [255.0], rtol=1e-12)
    idx, c, _ = _run(oracle_lib, np.array([[1.0, 1, 1]]), np.zeros((0, 3)), 4)
    assert idx.shape == (0, 1) and c.shape == (0, 1)


def _reproduction_errors(src, tgt, idx, c, m):
    # the basis sees the float-rounded relative coordinates (float geometry)
    rel = (src[idx] - tgt[:, None, :]).astype(np.float32).astype(np.float64)
    cons = np.abs(c.sum(1) - 1.0)
    lin = np.abs(np.einsum("tk,tkd->td", c, rel)).max(1) if m >= 4 else np.zeros(len(tgt))
    return cons, lin


@pytest.mark.parametrize("m,k,rbf", [(4, 4, "wendland0"), (4, 16, "wendland2"),
                                     (10, 10, "wendland2"), (10, 32, "wendland4")])
def test_mls_polynomial_reproduction(oracle_lib, m, k, rbf):
    src, tgt, _ = c1_inputs(3000, seed=11)
    idx, c, cnt = _run(oracle_lib, src, tgt, m, k=k, rbf_name=rbf)
    cons, lin = _reproduction_errors(src, tgt, idx, c, m)
    h = np.linalg.norm(src[idx] - tgt[:, None, :], axis=2).max(1)
    ok = np.ones(len(tgt), bool)
    if cnt.n_underdetermined:
        ok = cons < 1e-3  # truncated systems lose exactness; skip them
    # square systems (k = m) are interpolation with condition numbers up to
    # ~1e10 for the quadratic basis; overdetermined ones are well posed
    med, p99 = (1e-10, 1e-6) if k == m else (1e-13, 1e-10)
    assert np.median(cons[ok]) < med
    assert np.quantile(cons[ok], 0.99) < p99
    if m >= 4:
        assert np.quantile(lin[ok] / h[ok], 0.99) < p99


def test_mls_k_equals_m_is_interpolation(oracle_lib):
    # k = m: the weights cancel and c = e1^T P^{-1} (textbook interpolation)
    src, tgt, _ = c1_inputs(2000, seed=12)
    idx, c, _ = _run(oracle_lib, src, tgt, 4, k=4)
    rel = (src[idx] - tgt[:, None, :]).astype(np.float32).astype(np.float64)
    P = np.concatenate([np.ones(rel.shape[:2] + (1,)), rel], axis=2)
    e1 = np.tile([1.0, 0, 0, 0], (len(tgt), 1))[..., None]
    ref = np.linalg.solve(P.transpose(0, 2, 1), e1)[..., 0]
    cond = np.linalg.cond(P)
    err = np.abs(c - ref).max(1) / np.abs(ref).max(1)
    assert np.all(err <= 1e-14 * cond**2 + 1e-9)


def test_mls_weighted_least_squares_independent(oracle_lib):
    # overdetermined k > m: c = Phi P (P^T Phi P)^{-1} e1 with the radius
    # 1.1 * max distance and Wendland<2>, formed independently in numpy
    src, tgt, _ = c1_inputs(2000, seed=13)
    m, k = 10, 32
    idx, c, _ = _run(oracle_lib, src, tgt, m, k=k, rbf_name="wendland2", geometry=1)
    rel = src[idx] - tgt[:, None, :]
    d = np.linalg.norm(rel, axis=2)
    R = 1.1 * np.maximum(d.max(1), 10 * np.finfo(float).eps)
    x = d / R[:, None]
    phi = (1 - x) ** 4 * (4 * x + 1)
    X, Y, Z = rel[..., 0], rel[..., 1], rel[..., 2]
    P = np.stack([np.ones_like(X), X, Y, Z, X * X, X * Y, X * Z, Y * Y, Y * Z, Z * Z], axis=2)
    A = np.einsum("tkr,tk,tkc->trc", P, phi, P)
    a0 = np.linalg.solve(A, np.tile(np.eye(m)[0], (len(tgt), 1))[..., None])[..., 0]
    ref = phi * np.einsum("tkr,tr->tk", P, a0)
    err = np.abs(c - ref).max(1) / np.abs(ref).max(1)
    assert np.quantile(err, 0.99) < 1e-8 and err.max() < 1e-6


# ---------------------------------------------------------- apply / transpose
def test_apply_and_transpose(oracle_lib):
    rng = np.random.default_rng(8)
    nt, ns, kk = 300, 200, 7
    idx = rng.integers(0, ns, (nt, kk)).astype(np.int32)
    c = rng.standard_normal((nt, kk))
    x = rng.standard_normal((3, ns))
    y = oracle_lib.apply(idx, c, x)
    dense = np.zeros((nt, ns))
    np.add.at(dense, (np.repeat(np.arange(nt), kk), idx.ravel()), c.ravel())
    np.testing.assert_allclose(y, x @ dense.T, rtol=1e-13, atol=1e-13)
    ybar = rng.standard_normal((3, nt))
    xbar = oracle_lib.apply_transpose(idx, c, ybar, ns)
    np.testing.assert_allclose(xbar, ybar @ dense, rtol=1e-12, atol=1e-12)
    for f in range(3):
        assert abs(np.dot(ybar[f], y[f]) - np.dot(xbar[f], x[f])) < 1e-10


# --------------------------------------------------------------- libm modes
def test_correctly_rounded_mode_matches_mpmath(oracle_lib):
    mp = pytest.importorskip("mpmath")
    mp.mp.prec = 200
    rng = np.random.default_rng(9)
    x = np.concatenate([rng.uniform(-np.pi / 2, np.pi / 2, 1500), [np.pi / 2, -np.pi / 2, 1e-9]])
    y = rng.standard_normal(x.size)
    with oracle_lib.libm_mode(1):
        s = oracle_lib.elementary("sin", x)
        co = oracle_lib.elementary("cos", x)
        at = oracle_lib.elementary("atan2", x, y)
    for i in range(x.size):
        assert s[i] == float(mp.sin(mp.mpf(x[i])))
        assert co[i] == float(mp.cos(mp.mpf(x[i])))
        assert at[i] == float(mp.atan2(mp.mpf(y[i]), mp.mpf(x[i])))
    u = np.concatenate([rng.uniform(0.0, 1.0, 1500), 10.0 ** rng.uniform(-300, 300, 200),
                        [1.0, 2.0, 0.5, 1 - 2 ** -53, 1 + 2 ** -52, 5e-324]])
    with oracle_lib.libm_mode(1):
        lg = oracle_lib.elementary("log", u)
    for i in range(u.size):
        assert lg[i] == float(mp.log(mp.mpf(u[i])))


def test_c1_libm_modes_differ_by_the_measured_bound(oracle_lib):
    # The GPU is bit-exact to the correctly rounded mode; this is the gap to
    # what glibc gives the reference's CPU build on C1 (asserted on the GPU
    # in test_gpu_parity.test_c1_against_platform_libm)
    src, tgt, f = c1_inputs()
    y = {}
    for mode in (0, 1):
        with oracle_lib.libm_mode(mode):
            idx, c, _ = oracle_lib.mls_create(src, tgt, 4, k=0, rbf_name="wendland0")
        y[mode] = oracle_lib.apply(idx, c, f)
    rel_inf = np.abs(y[1] - y[0]).max() / np.abs(y[0]).max()
    pw = np.abs(y[1] - y[0]) / np.abs(y[0])
    assert 1e-12 < rel_inf <= 2e-12
    assert (pw > 1e-12).sum() == 5 and pw.max() <= 1.3e-11


# ------------------------------------------------ extended-precision values
def test_extended_precision_is_extended():
    assert np.finfo(np.longdouble).eps < 1e-18


def test_extended_values_match_double_oracle(oracle_lib):
    from oracle import extended

    rng = np.random.default_rng(4)
    src = rng.random((400, 3))
    tgt = 0.2 + 0.6 * rng.random((50, 3))
    x = np.stack([np.sin(3 * src[:, 0]) + src[:, 1], src[:, 2] ** 2])
    idx, c, _ = oracle_lib.mls_create(src, tgt, 10, k=24, rbf_name="wendland2",
                                      geometry=oracle_lib.GEOMETRY_DOUBLE)
    ref = oracle_lib.apply(idx, c, x)
    y = extended.target_values(src, tgt, idx, x, 10, "wendland2").astype(np.float64)
    np.testing.assert_allclose(y, ref, rtol=1e-11, atol=1e-12)


def test_extended_reproduces_polynomials_and_their_gradient():
    from oracle import extended

    rng = np.random.default_rng(5)
    src = rng.random((300, 3))
    tgt = 0.3 + 0.4 * rng.random((20, 3))
    nbr = np.argsort(((src[None] - tgt[:, None]) ** 2).sum(-1), axis=1)[:, :20]
    a = np.array([0.7, -1.3, 2.1])
    lin = src @ a + 0.25
    quad = src[:, 0] * src[:, 1] - 2 * src[:, 2] ** 2
    y = extended.target_values(src, tgt, nbr, np.stack([lin, quad]), 10, "wendland4")
    np.testing.assert_allclose(y[0].astype(float), tgt @ a + 0.25, rtol=0, atol=1e-14)
    np.testing.assert_allclose(y[1].astype(float),
                               tgt[:, 0] * tgt[:, 1] - 2 * tgt[:, 2] ** 2, atol=1e-14)
    # d/dt of a reproduced linear field is its gradient, whatever the weights
    g = extended.target_gradient_fd(src, tgt, nbr, lin[None], np.ones((1, 20)), 4,
                                    "wendland2")
    np.testing.assert_allclose(g, np.tile(a, (20, 1)), rtol=1e-9)

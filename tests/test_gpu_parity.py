"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars (integer/index work bit-exact; floating point per the plain FP64 form):
  * kNN indices and distances: bit-exact (ties by index).
  * SVD pseudo-inverse, MLS coefficients and applied values with the
    classical solver: bit-exact against the oracle's correctly rounded libm
    mode (the GPU uses correctly rounded atan2/sin/cos), and within 1e-12
    relative (|dy| / max|y|) of the platform-libm (glibc) oracle except on
    systems whose condition makes one-ulp trig differences visible, which
    are counted and bounded.
  * the reference's own MLS tests at their tolerances.
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


def bits(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a.view(np.uint64)


# --------------------------------------------------------------------- kNN
@pytest.mark.parametrize("geometry", ["float", "double"])
@pytest.mark.parametrize("k", [1, 4, 10, 16, 32, 64])
def test_knn_bit_exact(dtk, oracle_lib, geometry, k):
    src, tgt, _ = c1_inputs(10_000, seed=100 + k)
    print(f"knn geometry={geometry} k={k} n=10000", flush=True)
    idx, dist = dtk.knn(cuda(src), cuda(tgt), k, geometry=geometry, with_distances=True)
    gi, gd = oracle_lib.knn(src, tgt, k, geometry=0 if geometry == "float" else 1)
    assert np.array_equal(idx.cpu().numpy(), gi)
    assert np.array_equal(bits(dist.cpu().numpy()), bits(gd))


def test_knn_ties_duplicates_and_layouts(dtk, oracle_lib):
    rng = np.random.default_rng(1)
    lattice = grid_points((12, 12, 12), (0, 0, 0)) / 11.0
    src = np.concatenate([lattice, lattice[:300], rng.random((2000, 3))])
    tgt = np.concatenate([lattice[::7], rng.random((1000, 3)), src[:40]])
    for k in (1, 6, 27, 40):
        gi, gd = oracle_lib.knn(src, tgt, k)
        idx = dtk.knn(cuda(src), cuda(tgt), k)
        assert np.array_equal(idx.cpu().numpy(), gi)
        gi_d, _ = oracle_lib.knn(src, tgt, k, geometry=1)
        idx = dtk.knn(cuda(src), cuda(tgt), k, geometry="double")
        assert np.array_equal(idx.cpu().numpy(), gi_d)
        idx = dtk.knn(cuda(src.T.copy()), cuda(tgt.T.copy()), k, layout="soa")
        assert np.array_equal(idx.cpu().numpy(), gi)
        idx = dtk.knn(src, tgt, k)  # host pointers are staged
        assert np.array_equal(idx.cpu().numpy(), gi)


def test_knn_small_source_sets(dtk, oracle_lib):
    rng = np.random.default_rng(2)
    tgt = rng.random((100, 3))
    for ns in (1, 2, 3, 5):
        src = rng.random((ns, 3))
        idx, dist = dtk.knn(cuda(src), cuda(tgt), 4, with_distances=True)
        gi, gd = oracle_lib.knn(src, tgt, 4)
        assert idx.shape == (100, min(4, ns))
        assert np.array_equal(idx.cpu().numpy(), gi)
        assert np.array_equal(bits(dist.cpu().numpy()), bits(gd))


@pytest.mark.parametrize("geometry", ["float", "double"])
def test_knn_rounded_distance_ties(dtk, oracle_lib, geometry):
    """Sources on a sphere of radius 1/4 around a target, 1e-7 apart in
    radius: in float arithmetic thousands of them share a handful of
    rounded distances, many from different squared distances. The k-th
    result sits inside such a group, so the selection must take the
    smallest indices among every source whose rounded distance ties the
    k-th, whatever its squared distance (k from 16 up covers both the
    one-pass list and the two-pass search with its separate row sort)."""
    rng = np.random.default_rng(7)
    centre = np.array([0.5, 0.5, 0.5])
    d = rng.normal(size=(4000, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    src = (centre + 0.25 * d * (1 + rng.integers(-3, 4, size=(4000, 1)) * 1e-7))
    src = src.astype(np.float32).astype(np.float64)
    tgt = np.concatenate([centre[None, :], rng.random((300, 3))])
    g = 0 if geometry == "float" else 1
    for k in (16, 24, 32, 40, 64):
        gi, gd = oracle_lib.knn(src, tgt, k, geometry=g)
        idx, dist = dtk.knn(cuda(src), cuda(tgt), k, geometry=geometry, with_distances=True)
        assert np.array_equal(idx.cpu().numpy(), gi), k
        assert np.array_equal(bits(dist.cpu().numpy()), bits(gd)), k


@pytest.mark.parametrize("geometry", ["float", "double"])
def test_knn_fewer_sources_than_k(dtk, oracle_lib, geometry):
    rng = np.random.default_rng(5)
    tgt = rng.random((200, 3))
    g = 0 if geometry == "float" else 1
    for ns in (1, 2, 7, 33):
        src = rng.random((ns, 3))
        for k in (20, 40, 64):
            gi, gd = oracle_lib.knn(src, tgt, k, geometry=g)
            idx, dist = dtk.knn(cuda(src), cuda(tgt), k, geometry=geometry, with_distances=True)
            assert idx.shape == (200, min(k, ns))
            assert np.array_equal(idx.cpu().numpy(), gi)
            assert np.array_equal(bits(dist.cpu().numpy()), bits(gd))


def test_knn_large(dtk, oracle_lib):
    # 200k sources: a deeper tree; sampled targets checked against the oracle
    rng = np.random.default_rng(3)
    src = rng.random((200_000, 3))
    tgt = rng.random((20_000, 3)) * 0.9 + 0.05
    idx = dtk.knn(cuda(src), cuda(tgt), 32).cpu().numpy()
    gi, _ = oracle_lib.knn(src, tgt, 32)
    assert np.array_equal(idx, gi)


@pytest.mark.parametrize("geometry", ["float", "double"])
def test_knn_seeded_bound_edge_geometry(dtk, oracle_lib, geometry):
    """The search seeds its k-th distance with the farthest corner of the
    deepest node on the query's path holding k sources. Geometry where that
    bound is loose, zero or tight: 50 Gaussian clusters (sigma 1e-4) with
    exact duplicates, targets on duplicates, inside clusters, between
    clusters and far outside the sources' box; k across the one-pass and
    two-pass kernels. Indices and distances must equal brute force."""
    rng = np.random.default_rng(21)
    centres = rng.random((50, 3))
    src = centres[rng.integers(0, 50, 20_000)] + rng.normal(scale=1e-4, size=(20_000, 3))
    src[:2000] = src[rng.integers(2000, 20_000, 2000)]  # duplicates
    tgt = np.concatenate([src[rng.integers(0, 20_000, 300)],
                          centres[rng.integers(0, 50, 300)]
                          + rng.normal(scale=2e-4, size=(300, 3)),
                          rng.random((300, 3)),
                          rng.random((100, 3)) * 20 - 10])
    g = 0 if geometry == "float" else 1
    for k in (1, 4, 16, 32, 48):
        print(f"seeded knn {geometry} k={k}", flush=True)
        gi, gd = oracle_lib.knn(src, tgt, k, geometry=g)
        idx, dist = dtk.knn(cuda(src), cuda(tgt), k, geometry=geometry,
                            with_distances=True)
        assert np.array_equal(idx.cpu().numpy(), gi), k
        assert np.array_equal(bits(dist.cpu().numpy()), bits(gd)), k


@pytest.mark.parametrize("geometry", ["float", "double"])
def test_knn_deep_tree_overflow_rerun(dtk, oracle_lib, geometry):
    """A hierarchy deeper than the LDS traversal stack: pairs of points
    halving their distance to the origin (one Morton level each, and a pair
    is an internal node, so the traversal pushes it) ending in 65536 copies
    of the origin (16 more levels, split by index bits). For queries at the
    origin the seeded bound is 0 and every duplicate's node lies at
    distance 0, so both children of each pass and the descent pushes past
    the LDS stack; those queries are re-run with the global stack and must
    still equal brute force."""
    rng = np.random.default_rng(4)
    chain = np.repeat(np.array([[2.0 ** -i] * 3 for i in range(1, 22)]), 2, axis=0)
    dup = np.zeros((65536, 3))
    src = np.concatenate([rng.random((3000, 3)), chain, dup])
    tgt = np.concatenate([np.zeros((64, 3)), rng.random((500, 3)) * 1e-3,
                          rng.random((2000, 3))])
    g = 0 if geometry == "float" else 1
    for k in (16, 32):
        gi, gd = oracle_lib.knn(src, tgt, k, geometry=g)
        idx, dist = dtk.knn(cuda(src), cuda(tgt), k, geometry=geometry, with_distances=True)
        assert np.array_equal(idx.cpu().numpy(), gi)
        assert np.array_equal(bits(dist.cpu().numpy()), bits(gd))
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="constant", k=16,
                                        geometry=geometry)
    assert op.stats()["n_stack_overflow"] > 0
    _, idx, _ = op.graph()
    gi, _ = oracle_lib.knn(src, tgt, 16, geometry=g)
    assert np.array_equal(idx.cpu().numpy(), gi)


@pytest.mark.parametrize("geometry", ["float", "double"])
def test_knn_overflow_rerun_strides_over_many_rows(dtk, oracle_lib, geometry):
    """More overflowing queries (20,000) than the re-run has lanes (16,384),
    so lanes take several listed rows and reuse their global stack.

    Sources: pairs at (2^-i, 2^-i, 2^-i), i = 1..21, and 16 copies of the
    origin; k = all 58 sources; every query at the origin. The hierarchy is
    a chain: each level splits off the pair i from a node holding the
    origin. The query lies in that node's box (distance 0) and the k-th
    distance covers every pair, so each of the 21 levels pushes its pair,
    and the 16 duplicates (split by index, 4 levels at distance 0) push 3
    more: 24 entries, past the 22-entry LDS stack, for every query. So one
    query gives a count of 1 and 20,000 give 20,000; the rows must equal
    the oracle's."""
    chain = np.repeat(np.array([[2.0 ** -i] * 3 for i in range(1, 22)]), 2, axis=0)
    src = np.concatenate([chain, np.zeros((16, 3))])
    k = len(src)
    g = 0 if geometry == "float" else 1
    n = 20_000
    print(f"overflow rerun geometry={geometry} n_tgt={n} k={k}", flush=True)

    def make(t):
        op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(t), basis="constant", k=k,
                                            geometry=geometry)
        return op, op.stats()["n_stack_overflow"]

    assert make(np.zeros((1, 3)))[1] == 1
    op, n_all = make(np.zeros((n, 3)))
    assert n_all == n
    gi, _ = oracle_lib.knn(src, np.zeros((1, 3)), k, geometry=g)
    _, idx, _ = op.graph()
    assert np.array_equal(idx.cpu().numpy(), np.repeat(gi, n, axis=0))


# --------------------------------------------------------------------- SVD
@pytest.mark.parametrize("case,shape,deficient", [
    ("svd_full_rank", (10, 32, 32), ()),
    ("svd_rank_deficient", (10, 40, 40), (3,)),
])
def test_svd_tstSVD_bit_exact(dtk, oracle_lib, case, shape, deficient):
    mats = load_golden(case, shape)
    print(f"svd {case}", flush=True)
    pinv, under, unconv = dtk.svd_pseudo_inverse(cuda(mats))
    with oracle_lib.libm_mode(1):
        ref, c = oracle_lib.svd_pseudo_inverse(mats)
    got = pinv.cpu().numpy()
    assert np.array_equal(bits(got), bits(ref))
    assert under == (10 if deficient else 0) and unconv == 0
    for a, p in zip(mats, got):  # tstSVD.cpp:22-65
        res = p @ a
        for i in range(shape[1]):
            assert abs(res[i, i] - (0.0 if i in deficient else 1.0)) <= 1e-12 or i in deficient


# ---------------------------------------------------- reference MLS tests
def _apply(op, vals, nt):
    out = torch.zeros(nt, dtype=torch.float64, device="cuda")
    op.apply(cuda(vals), out)
    return out.cpu().numpy()


@pytest.mark.parametrize("m,basis", [(1, "constant"), (4, "linear"), (10, "quadratic")])
def test_reference_mls_tests(dtk, m, basis):
    f = poly_for_basis(m)
    # same_npoints_and_basis (tstMeshfreeOperators.cpp:130-224)
    src = load_golden(f"mls_points_m{m}_src", (-1, 3))
    tgt = load_golden(f"mls_points_m{m}_tgt", (-1, 3))
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis=basis)
    np.testing.assert_allclose(_apply(op, f(src), len(tgt)), f(tgt), rtol=1e-7)
    # grid (:226-300) and line (:302-376)
    g = grid_points((40, 40, 1), (0, 0, 0))
    t = grid_points((1, 1, 1), (19, 19, 0))
    op = dtk.MovingLeastSquaresOperator(cuda(g), cuda(t), basis=basis)
    np.testing.assert_allclose(_apply(op, f(g), 1), f(t), rtol=1e-14)
    line = grid_points((40, 1, 1), (0, 0, 0))
    lt = grid_points((39, 1, 1), (0.5, 0, 0))
    op = dtk.MovingLeastSquaresOperator(cuda(line), cuda(lt), basis=basis)
    np.testing.assert_allclose(_apply(op, f(line), 39), f(lt), rtol=1e-14)
    # single_point_in_radius (:378-448)
    s1 = load_golden("mls_points_m1_src", (-1, 3))
    t1 = load_golden("mls_points_m1_tgt", (-1, 3))
    op = dtk.MovingLeastSquaresOperator(cuda(s1), cuda(t1), basis=basis)
    assert np.all(np.isfinite(_apply(op, f(s1), 10)))


def test_reference_corner_cases(dtk):
    # tstMeshfreeOperatorsSimpleProblem.cpp:113-147
    op = dtk.MovingLeastSquaresOperator(np.array([[1.0, 1, 1], [2, 2, 2]]),
                                        np.array([[1.5, 1.5, 1.5]]))
    y = np.zeros(1)
    op.apply(np.array([255.0, 255.0]), y)
    np.testing.assert_allclose(y, [255.0], rtol=1e-12)
    op = dtk.MovingLeastSquaresOperator(np.array([[1.0, 1, 1]]), np.zeros((0, 3)))
    y = np.zeros(0)
    op.apply(np.array([255.0]), y)
    assert y.size == 0


def test_transpose_with_no_targets_zero_fills(dtk):
    # C^T of an operator with no targets is the zero map: every column of
    # src_bar (with its leading dimension) is zeroed, the gaps are left alone
    import ctypes

    from datatransferkit_amd import _lib

    src, _, _ = c1_inputs(500, seed=3)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(np.zeros((0, 3))), basis="linear")
    lds = 512
    xb = torch.full((3, lds), 9.0, dtype=torch.float64, device="cuda")
    _lib.check(_lib.load().dtkx_mls_apply_transpose(
        op._h, None, ctypes.c_void_p(xb.data_ptr()), 3, 0, lds))
    got = xb.cpu().numpy()
    assert np.all(got[:, :500] == 0.0) and np.all(got[:, 500:] == 9.0)


# ---------------------------------------------------------- C1 parity
def _oracle_full(oracle_lib, src, tgt, m, k, rbf, mode, geometry=0):
    with oracle_lib.libm_mode(mode):
        return oracle_lib.mls_create(src, tgt, m, k=k, rbf_name=rbf, geometry=geometry)


@pytest.mark.parametrize("group", [0, 1, 2, 4])
def test_c1_plain_fp64_bit_exact(dtk, oracle_lib, group):
    """10k/10k uniform, Linear basis, k = 4, Wendland<0>, FP64."""
    src, tgt, f = c1_inputs()
    print(f"c1 group={group}", flush=True)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="linear",
                                        rbf="wendland0", group_size=group)
    _, idx, coeffs = op.graph()
    gi, gc, cnt = _oracle_full(oracle_lib, src, tgt, 4, 0, "wendland0", 1)
    assert np.array_equal(idx.cpu().numpy(), gi)
    assert np.array_equal(bits(coeffs.cpu().numpy()), bits(gc))
    y = _apply(op, f, len(tgt))
    yo = oracle_lib.apply(gi, gc, f)
    assert np.array_equal(bits(y), bits(yo))
    st = op.stats()
    assert st["n_underdetermined"] == cnt.n_underdetermined
    assert st["total_rotations"] == cnt.total_rotations
    assert st["n_unconverged"] == 0


def _moment_cond(src, tgt, idx, t):
    """cond(A) of target t's linear moment matrix A = P^T W P (Wendland<0>,
    float relative coordinates, as the setup forms it)."""
    x = (src[idx[t]] - tgt[t]).astype(np.float32).astype(np.float64)
    d = np.sqrt((x ** 2).sum(1))
    w = (1.0 - d / (1.1 * d.max())) ** 2
    p = np.c_[np.ones(len(d)), x]
    return np.linalg.cond(p.T @ (w[:, None] * p))


def test_c1_against_platform_libm(dtk, oracle_lib):
    """Against the oracle with glibc's atan2/sin/cos (what the reference's CPU
    build computes here). The GPU rounds these functions correctly, glibc
    misses by an ulp on about 0.1 % of arguments, and on C1's most
    ill-conditioned systems that ulp shows above 1e-12. Measured on the CPU
    between the oracle's two modes (the GPU equals the correctly rounded
    one): infinity-norm 1.10e-12 relative; pointwise, 5 of 10,000 targets
    above 1e-12, at most 1.29e-11, all with cond(A) between 3.7e6 and
    1.1e9."""
    src, tgt, f = c1_inputs()
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt))
    y = _apply(op, f, len(tgt))
    gi, gc, _ = _oracle_full(oracle_lib, src, tgt, 4, 0, "wendland0", 0)
    yo = oracle_lib.apply(gi, gc, f)
    rel_inf = np.abs(y - yo).max() / np.abs(yo).max()
    pointwise = np.abs(y - yo) / np.abs(yo)
    over = np.nonzero(pointwise > 1e-12)[0]
    print(f"C1 vs glibc oracle: inf-norm rel {rel_inf:.3e}, "
          f"{over.size} of {len(tgt)} targets above 1e-12, max {pointwise.max():.3e}")
    for t in over:
        print(f"  target {t}: rel {pointwise[t]:.3e}, "
              f"cond(A) {_moment_cond(src, tgt, gi, t):.3e}")
    assert rel_inf <= 2e-12
    assert over.size <= 5
    assert pointwise.max() <= 1.3e-11


@pytest.mark.parametrize("m,basis,k,rbf", [
    (10, "quadratic", 10, "wendland2"),
    (10, "quadratic", 32, "wendland2"),
    (4, "linear", 16, "wendland4"),
    (10, "quadratic", 64, "wendland6"),
    (1, "constant", 8, "wu2"),
    (4, "linear", 12, "wu4"),
    (10, "quadratic", 16, "buhmann2"),
    (4, "linear", 12, "buhmann3"),
    (10, "quadratic", 32, "buhmann3"),
])
def test_mls_variants_bit_exact(dtk, oracle_lib, m, basis, k, rbf):
    src, tgt, f = c1_inputs(20_000, seed=7 + k)
    tgt = tgt[:4000] * 0.9 + 0.05
    print(f"variant m={m} k={k} rbf={rbf}", flush=True)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis=basis, k=k, rbf=rbf)
    _, idx, coeffs = op.graph()
    gi, gc, cnt = _oracle_full(oracle_lib, src, tgt, m, k, rbf, 1)
    assert np.array_equal(idx.cpu().numpy(), gi)
    assert np.array_equal(bits(coeffs.cpu().numpy()), bits(gc))
    assert op.stats()["total_rotations"] == cnt.total_rotations


def test_double_geometry_bit_exact(dtk, oracle_lib):
    src, tgt, f = c1_inputs(20_000, seed=31)
    tgt = tgt[:3000]
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic", k=16,
                                        rbf="wendland2", geometry="double")
    _, idx, coeffs = op.graph()
    gi, gc, _ = _oracle_full(oracle_lib, src, tgt, 10, 16, "wendland2", 1, geometry=1)
    assert np.array_equal(idx.cpu().numpy(), gi)
    assert np.array_equal(bits(coeffs.cpu().numpy()), bits(gc))


def test_with_neighbors_and_relative_truncation(dtk, oracle_lib):
    src, tgt, f = c1_inputs(5000, seed=41)
    gi, _ = oracle_lib.knn(src, tgt, 12)
    perm = np.random.default_rng(0).permutation(12)
    nb = gi[:, perm].copy()  # row order is the summation order
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic",
                                        rbf="wendland2", neighbors=cuda(nb, torch.int32),
                                        truncation="relative", rcond=1e-12)
    _, idx, coeffs = op.graph()
    with oracle_lib.libm_mode(1):
        gc, _ = oracle_lib.mls_setup(src, tgt, nb, 10, "wendland2", truncation=1, rcond=1e-12)
    assert np.array_equal(idx.cpu().numpy(), nb)
    assert np.array_equal(bits(coeffs.cpu().numpy()), bits(gc))


# -------------------------------------------------------- apply / transpose
def test_apply_multifield_and_transpose(dtk, oracle_lib):
    src, tgt, f = c1_inputs(10_000, seed=51)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic", k=16,
                                        rbf="wendland2")
    _, idx, coeffs = op.graph()
    gi, gc = idx.cpu().numpy(), coeffs.cpu().numpy()
    rng = np.random.default_rng(5)
    for nf in (1, 3, 4, 8):
        x = rng.standard_normal((nf, len(src)))
        y = torch.zeros((nf, len(tgt)), dtype=torch.float64, device="cuda")
        op.apply(cuda(x), y)
        assert np.array_equal(bits(y.cpu().numpy()), bits(oracle_lib.apply(gi, gc, x)))
        ybar = rng.standard_normal((nf, len(tgt)))
        xbar = torch.zeros((nf, len(src)), dtype=torch.float64, device="cuda")
        op.apply_transpose(cuda(ybar), xbar)
        ref = oracle_lib.apply_transpose(gi, gc, ybar, len(src))
        assert np.array_equal(bits(xbar.cpu().numpy()), bits(ref))


def test_apply_padded_leading_dimensions(dtk, oracle_lib):
    """Fields with leading dimensions beyond n (the C ABI's ld_src, ld_tgt)
    and rows staged in Morton order: each value lands at f * ld + i, bit for
    bit the oracle's, and the gaps keep their contents, forward and
    transpose."""
    import ctypes

    from datatransferkit_amd import _lib

    src, tgt, _ = c1_inputs(5_000, seed=52)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic", k=16,
                                        rbf="wendland2")
    _, idx, coeffs = op.graph()
    gi, gc = idx.cpu().numpy(), coeffs.cpu().numpy()
    ns, nt, nf = len(src), len(tgt), 3
    lds, ldt = ns + 37, nt + 11
    rng = np.random.default_rng(6)
    L = _lib.load()
    vp = ctypes.c_void_p
    x = rng.standard_normal((nf, ns))
    xp = torch.full((nf, lds), 7.0, dtype=torch.float64, device="cuda")
    xp[:, :ns] = cuda(x)
    yp = torch.full((nf, ldt), -3.0, dtype=torch.float64, device="cuda")
    _lib.check(L.dtkx_mls_apply(op._h, vp(xp.data_ptr()), vp(yp.data_ptr()), nf, lds, ldt))
    y = yp.cpu().numpy()
    assert np.array_equal(bits(y[:, :nt]), bits(oracle_lib.apply(gi, gc, x)))
    assert np.all(y[:, nt:] == -3.0)
    ybar = rng.standard_normal((nf, nt))
    yb = torch.full((nf, ldt), 5.0, dtype=torch.float64, device="cuda")
    yb[:, :nt] = cuda(ybar)
    xb = torch.full((nf, lds), -1.0, dtype=torch.float64, device="cuda")
    _lib.check(L.dtkx_mls_apply_transpose(op._h, vp(yb.data_ptr()), vp(xb.data_ptr()), nf,
                                          ldt, lds))
    xbar = xb.cpu().numpy()
    assert np.array_equal(bits(xbar[:, :ns]), bits(oracle_lib.apply_transpose(gi, gc, ybar, ns)))
    assert np.all(xbar[:, ns:] == -1.0)


def test_errors_are_reported(dtk):
    src, tgt, f = c1_inputs(100, seed=1)
    op = dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt))
    with pytest.raises(dtk.DtkxError) as e:
        op.apply(cuda(f[:50]), torch.zeros(100, dtype=torch.float64, device="cuda"))
    assert e.value.status == 2
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresOperator(cuda(np.zeros((0, 3))), cuda(tgt))
    assert e.value.status == 4
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresOperator(cuda(src), cuda(tgt), k=65)
    assert e.value.status == 7

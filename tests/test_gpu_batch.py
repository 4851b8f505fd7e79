"""Batched uneven operators (pattern 2): many independent problems built by
one segmented kNN and one setup launch per basis, applied through a
captured HIP graph. Every problem's result must be bit-identical to the
single-operator result for the same solver (same neighbours, same
per-target arithmetic, same row sums), which the single-operator tests pin
to the oracle; a few problems are also checked against the oracle
directly."""
import numpy as np
import pytest

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


BASES = ["constant", "linear", "quadratic"]
M = {"constant": 1, "linear": 4, "quadratic": 10}


def make_problems(n, seed, max_src=3000):
    """Uneven problems: log-uniform source counts (with a few tiny ones),
    0.5-2x as many targets (some problems without targets), random k and
    basis, points in random affine-distorted unit cubes."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        ns = int(round(np.exp(rng.uniform(np.log(8), np.log(max_src)))))
        if i % 11 == 3:
            ns = int(rng.integers(1, 4))
        nt = int(round(ns * rng.uniform(0.5, 2.0)))
        if i % 13 == 5:
            nt = 0
        basis = BASES[int(rng.integers(0, 3))]
        k = int(rng.choice([0, 4, 10, 16, 32, 64]))
        a = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
        b = rng.standard_normal(3)
        src = rng.random((ns, 3)) @ a.T + b
        tgt = (0.05 + 0.9 * rng.random((nt, 3))) @ a.T + b
        vals = np.sin(2 * src[:, 0]) + src[:, 1] * src[:, 2]
        out.append((src, tgt, basis, k, vals))
    return out


FORMS = [
    dict(solver="classical", precision="fp64"),
    dict(solver="cyclic", precision="fp64"),
    dict(solver="cholesky", precision="fp64"),
    dict(solver="cyclic", precision="fp32_bulk"),
    dict(solver="cholesky", precision="fp32_bulk"),
]


@pytest.mark.parametrize("form", FORMS, ids=lambda f: f"{f['solver']}-{f['precision']}")
@pytest.mark.parametrize("geometry", ["float", "double"])
def test_batch_equals_single_operators(dtk, form, geometry):
    probs = make_problems(48, seed=3)
    dev = [(cuda(s), cuda(t), b, k) for s, t, b, k, _ in probs]
    opts = dict(rbf="wendland2", geometry=geometry, **form)
    batch = dtk.MovingLeastSquaresBatch(dev, **opts)
    xs = [cuda(v) for *_, v in probs]
    ys = [torch.full((len(t),), np.nan, dtype=torch.float64, device="cuda")
          for _, t, *_ in probs]
    batch.apply(xs, ys)
    st = batch.stats()
    assert st["n_target"] == sum(len(t) for _, t, *_ in probs)
    for i, ((s, t, b, k), x, y) in enumerate(zip(dev, xs, ys)):
        if len(t) == 0:
            continue
        op = dtk.MovingLeastSquaresOperator(s, t, basis=b, k=k, **opts)
        ref = torch.empty(len(t), dtype=torch.float64, device="cuda")
        op.apply(x, ref)
        assert torch.equal(y, ref), f"problem {i}: {b} k={k} n_src={len(s)}"


def test_batch_classical_matches_oracle(dtk, oracle_lib):
    probs = make_problems(12, seed=5, max_src=800)
    dev = [(cuda(s), cuda(t), b, k) for s, t, b, k, _ in probs]
    batch = dtk.MovingLeastSquaresBatch(dev, rbf="wendland2")
    xs = [cuda(v) for *_, v in probs]
    ys = [torch.zeros(len(t), dtype=torch.float64, device="cuda") for _, t, *_ in probs]
    batch.apply(xs, ys)
    with oracle_lib.libm_mode(oracle_lib.LIBM_CORRECTLY_ROUNDED):
        for (s, t, b, k, v), y in zip(probs, ys):
            if len(t) == 0:
                continue
            idx, c, _ = oracle_lib.mls_create(s, t, M[b], k=k, rbf_name="wendland2")
            ref = oracle_lib.apply(idx, c, v[None])[0]
            assert np.array_equal(y.cpu().numpy(), ref)


def test_graph_replay_reads_new_values(dtk):
    probs = make_problems(20, seed=9)
    dev = [(cuda(s), cuda(t), b, k) for s, t, b, k, _ in probs]
    batch = dtk.MovingLeastSquaresBatch(dev, rbf="wendland2", solver="cyclic")
    singles = [dtk.MovingLeastSquaresOperator(s, t, basis=b, k=k, rbf="wendland2",
                                              solver="cyclic") if len(t) else None
               for s, t, b, k in dev]
    for rep in range(3):
        xs = [cuda(v * (rep + 1) + rep) for *_, v in probs]
        # fresh output buffers on some replays, reused on others
        ys = [torch.zeros(len(t), dtype=torch.float64, device="cuda") for _, t, *_ in probs]
        batch.apply(xs, ys)
        torch.cuda.synchronize()
        for op, x, y in zip(singles, xs, ys):
            if op is None:
                continue
            ref = torch.empty_like(y)
            op.apply(x, ref)
            assert torch.equal(y, ref)


def test_apply_concatenated_equals_apply(dtk):
    """Values concatenated over the problems give the bits of the per-problem
    apply, with problems without targets inside and at the end of the batch
    (their pointer lands on the next problem's values or one past the end,
    and nothing is written there)."""
    probs = make_problems(20, seed=19)
    probs.append((probs[0][0], np.zeros((0, 3)), "linear", 4, probs[0][4]))
    assert sum(len(t) == 0 for _, t, *_ in probs) >= 2 and len(probs[-1][1]) == 0
    dev = [(cuda(s), cuda(t), b, k) for s, t, b, k, _ in probs]
    batch = dtk.MovingLeastSquaresBatch(dev, rbf="wendland2", solver="cholesky",
                                        precision="fp32_bulk")
    xs = [cuda(v) for *_, v in probs]
    ys = [torch.zeros(len(t), dtype=torch.float64, device="cuda") for _, t, *_ in probs]
    batch.apply(xs, ys)
    x_flat = torch.cat(xs)
    # one sentinel past the targets: the last problem has none, so its
    # pointer is one past y_flat's end, where nothing may be written
    y_buf = torch.full((sum(len(y) for y in ys) + 1,), 7.0, dtype=torch.float64,
                       device="cuda")
    y_flat = y_buf[:-1]
    for _ in range(2):
        batch.apply_concatenated(x_flat, y_flat)
        torch.cuda.synchronize()
        assert torch.equal(y_flat, torch.cat(ys))
        assert y_buf[-1].item() == 7.0
    with pytest.raises(dtk.DtkxError) as e:
        batch.apply_concatenated(x_flat[:-1], y_flat)
    assert e.value.status == 2
    with pytest.raises(dtk.DtkxError) as e:
        batch.apply_concatenated(x_flat, y_flat.float())
    assert e.value.status == 1
    with pytest.raises(dtk.DtkxError) as e:
        batch.apply_concatenated(torch.stack([x_flat, x_flat], 1)[:, 0], y_flat)
    assert e.value.status == 1
    # host values are refused before any work is queued
    with pytest.raises(dtk.DtkxError) as e:
        batch.apply_concatenated(x_flat.cpu(), y_flat)
    assert e.value.status == 1
    with pytest.raises(dtk.DtkxError) as e:
        batch.apply([x.cpu() for x in xs], ys)
    assert e.value.status == 1
    # the C ABI checks the first problem's pair itself: a CPU tensor, the
    # memory the create's point check refuses in test_batch_errors
    from datatransferkit_amd import _lib

    host = torch.zeros(len(xs[0]), dtype=torch.float64)
    xp = np.array([host.data_ptr()] + [x.data_ptr() for x in xs[1:]], dtype=np.uint64)
    yp = np.array([y.data_ptr() for y in ys], dtype=np.uint64)
    assert _lib.load().dtkx_mls_batch_apply(batch._h, xp.ctypes.data, yp.ctypes.data) == 1


def test_from_concatenated_equals_constructor(dtk):
    """The batch built from concatenated points is the constructor's batch
    over the per-problem slices, bit for bit: per-problem basis and k, one
    shared basis and k, problems without targets inside and at the end, and
    a rebuild after the flat points move in place."""
    probs = make_problems(24, seed=23)
    probs.append((probs[1][0], np.zeros((0, 3)), "quadratic", 16, probs[1][4]))
    src = cuda(np.concatenate([s for s, *_ in probs]))
    tgt = cuda(np.concatenate([t for _, t, *_ in probs]))
    ns = [len(s) for s, *_ in probs]
    nt = [len(t) for _, t, *_ in probs]
    views = list(zip(torch.split(src, ns), torch.split(tgt, nt)))
    x_flat = cuda(np.concatenate([v for *_, v in probs]))
    opts = dict(rbf="wendland2", solver="cholesky", precision="fp32_bulk")

    def run(batch):
        y = torch.full((sum(nt),), np.nan, dtype=torch.float64, device="cuda")
        return batch.apply_concatenated(x_flat, y)

    for bases, ks in (([b for *_, b, _, _ in probs], [k for *_, k, _ in probs]),
                      ("linear", 16)):
        each = [(s, t, bases if isinstance(bases, str) else bases[i],
                 ks if isinstance(ks, int) else ks[i]) for i, (s, t) in enumerate(views)]
        ref = dtk.MovingLeastSquaresBatch(each, **opts)
        flat = dtk.MovingLeastSquaresBatch.from_concatenated(src, tgt, ns, nt, bases, ks,
                                                             **opts)
        assert torch.equal(run(flat), run(ref))
        assert flat.stats()["nnz"] == ref.stats()["nnz"]
    # move the points in place: a rebuild of the flat batch reads them again
    src.mul_(1.5).add_(0.25)
    tgt.mul_(1.5).add_(0.25)
    torch.cuda.synchronize()
    flat.rebuild()
    assert torch.equal(run(flat), run(dtk.MovingLeastSquaresBatch(each, **opts)))
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch.from_concatenated(src, tgt, ns[:-1], nt, "linear", 4)
    assert e.value.status == 2
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch.from_concatenated(src, tgt[:-1], ns, nt, "linear", 4)
    assert e.value.status == 2
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch.from_concatenated(src[:, :2].contiguous(), tgt, ns, nt,
                                                      "linear", 4)
    assert e.value.status == 3
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch([(views[0][0][:, :2].contiguous(), views[0][1],
                                      "linear", 4)])
    assert e.value.status == 3


def deep_chain_problem():
    """Sources whose hierarchy is deeper than the query's LDS stack (pairs
    halving their distance to the origin, then 65,536 copies of it), with
    targets at and within 1e-3 of the origin: the chain pairs inside a near
    target's k-th distance and the duplicates push past the LDS stack, and
    the search re-runs those targets with the global stack."""
    rng = np.random.default_rng(4)
    chain = np.repeat(np.array([[2.0 ** -i] * 3 for i in range(1, 22)]), 2, axis=0)
    src = np.concatenate([rng.random((3000, 3)), chain, np.zeros((65536, 3))])
    tgt = np.concatenate([np.zeros((64, 3)), rng.random((500, 3)) * 1e-3,
                          rng.random((300, 3))])
    return src, tgt, "constant", 16, np.cos(3 * src[:, 0]) + src[:, 1]


REBUILD_FORMS = [FORMS[0], FORMS[2], FORMS[4]]


@pytest.mark.parametrize("form", REBUILD_FORMS,
                         ids=lambda f: f"{f['solver']}-{f['precision']}")
def test_batch_rebuild_equals_new_batch(dtk, form):
    """rebuild() after the points move in place: the captured search and
    setup, replayed, give bit for bit the operators (and counters) of a new
    batch over the moved points. Replays alternate with eager work (new
    batches and applies), and the last move returns to the first points,
    which must give the first results again. The deep problem makes the
    search's overflow re-run part of the graph."""
    probs = make_problems(24, seed=41) + [deep_chain_problem()]
    dev = [(cuda(s), cuda(t), b, k) for s, t, b, k, _ in probs]
    orig = [(s.clone(), t.clone()) for s, t, _, _ in dev]
    opts = dict(rbf="wendland2", geometry="float", **form)
    batch = dtk.MovingLeastSquaresBatch(dev, **opts)
    xs = [cuda(v) for *_, v in probs]

    def run(b):
        ys = [torch.full((len(t),), np.nan, dtype=torch.float64, device="cuda")
              for _, t, _, _ in dev]
        b.apply(xs, ys)
        torch.cuda.synchronize()
        return ys

    first = run(batch)
    assert batch.stats()["n_stack_overflow"] > 0
    keys = ("n_underdetermined", "n_unconverged", "total_rotations",
            "n_eigen_fallback", "n_stack_overflow")
    rng = np.random.default_rng(5)
    for rep in range(4):
        print(f"rebuild {rep} {form}", flush=True)
        for (s, t, _, _), (s0, t0) in zip(dev, orig):
            if rep == 3:
                s.copy_(s0)
                t.copy_(t0)
            else:
                # scaled, shifted and jittered: the neighbour lists change
                s.copy_(s0 * (1 + 0.1 * rep) + 0.3 * rep
                        + cuda(1e-3 * rng.standard_normal(tuple(s.shape))))
                t.copy_(t0 * (1 + 0.1 * rep) + 0.3 * rep
                        + cuda(1e-3 * rng.standard_normal(tuple(t.shape))))
        batch.rebuild()
        got = run(batch)
        fresh = dtk.MovingLeastSquaresBatch(dev, **opts)
        ref = run(fresh)
        for i, (a, r) in enumerate(zip(got, ref)):
            assert torch.equal(a, r), f"rebuild {rep}, problem {i}"
        st, sf = batch.stats(), fresh.stats()
        for key in keys:
            assert st[key] == sf[key], (rep, key, st[key], sf[key])
        if rep == 3:
            for a, r in zip(got, first):
                assert torch.equal(a, r)


def test_autotuned_group_gives_the_same_bits(dtk):
    probs = make_problems(16, seed=12)
    dev = [(cuda(s), cuda(t), b, k) for s, t, b, k, _ in probs]
    xs = [cuda(v) for *_, v in probs]
    outs = []
    for g in (0, 4):
        batch = dtk.MovingLeastSquaresBatch(dev, rbf="wendland2", group_size=g)
        ys = [torch.zeros(len(t), dtype=torch.float64, device="cuda") for _, t, *_ in probs]
        batch.apply(xs, ys)
        outs.append(ys)
        assert batch.stats()["group_size_used"] in (1, 2, 4, 8, 16)
    for a, b in zip(*outs):
        assert torch.equal(a, b)


def test_batch_errors(dtk):
    s = cuda(np.random.default_rng(0).random((10, 3)))
    t = cuda(np.random.default_rng(1).random((5, 3)))
    empty = torch.zeros((0, 3), dtype=torch.float64, device="cuda")
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch([(s, t, "linear", 4), (empty, t, "linear", 4)])
    assert e.value.status == 4
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch([(s.cpu(), t, "linear", 4)])
    assert e.value.status == 1
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch([(s, t, "linear", 99)])
    assert e.value.status == 7


@pytest.mark.parametrize("form", FORMS[1:], ids=lambda f: f"{f['solver']}-{f['precision']}")
def test_batch_fast_solvers_match_oracle(dtk, oracle_lib, form):
    """The fast solvers in the batch against the FP64 oracle, problem by
    problem, on overdetermined systems (at least 1.6 m neighbours, at least
    2 m sources): within 1e-6 relative (infinity norm) for the FP32-bulk
    form and 1e-9 for the FP64 forms (the oracle's own rounding on these
    systems is below 1e-10)."""
    rng = np.random.default_rng(31)
    probs = []
    for s, t, b, k, v in make_problems(40, seed=21, max_src=3000):
        m = M[b]
        if len(s) < 2 * m or len(t) == 0:
            continue
        k = int(rng.choice({1: [4, 8], 4: [8, 16, 32], 10: [16, 32, 64]}[m]))
        probs.append((s, t, b, k, v))
    dev = [(cuda(s), cuda(t), b, k) for s, t, b, k, _ in probs]
    batch = dtk.MovingLeastSquaresBatch(dev, rbf="wendland2", **form)
    xs = [cuda(v) for *_, v in probs]
    ys = [torch.zeros(len(t), dtype=torch.float64, device="cuda") for _, t, *_ in probs]
    batch.apply(xs, ys)
    bar = 1e-6 if form["precision"] == "fp32_bulk" else 1e-9
    worst = 0.0
    for i, ((s, t, b, k, v), y) in enumerate(zip(probs, ys)):
        idx, c, _ = oracle_lib.mls_create(s, t, M[b], k=k, rbf_name="wendland2")
        ref = oracle_lib.apply(idx, c, v[None])[0]
        rel = np.abs(y.cpu().numpy() - ref).max() / np.abs(ref).max()
        worst = max(worst, rel)
        assert rel <= bar, f"problem {i}: {b} k={k} n_src={len(s)} rel {rel:.2e}"
    print(f"{form}: {len(probs)} problems, worst rel error vs oracle {worst:.2e}")


def test_batch_c3_full_size(dtk, oracle_lib):
    """C3 at its full size, as the batched bench runs it: 4,096 uneven
    problems (sources log-uniform in [64, 65536], 0.5-2x as many targets,
    k in {10, 16, 32, 64}, linear or quadratic basis; ~47M targets, ~1.4G
    entries), FP32 bulk with the certified Cholesky solver. Checks whose
    meaning does not depend on the size:
      * 48 random problems bit-identical to single operators of the same
        form (which the single-operator tests pin to the oracle);
      * MLS reproduces its basis on every row of the overdetermined problems
        (linear basis, or quadratic with k >= 32): a constant field maps to
        the constant, a linear field to its values at the targets. Square
        quadratic systems (k = 10) are left out: there the reference's
        absolute DBL_EPSILON rule truncates the ill-conditioned ones, and a
        truncated pseudo-inverse does not reproduce its basis;
      * a sample of problems (k >= 16) against the FP64 oracle at the form's
        1e-6 bar."""
    rng = np.random.default_rng(7)
    probs = []
    for _ in range(4096):
        ns = int(round(np.exp(rng.uniform(np.log(64), np.log(65536)))))
        nt = int(round(ns * rng.uniform(0.5, 2.0)))
        k = int(rng.choice([10, 16, 32, 64]))
        basis = "linear" if rng.random() < 0.5 else "quadratic"
        a = np.eye(3) + 0.25 * rng.standard_normal((3, 3))
        b = rng.standard_normal(3)
        probs.append((rng.random((ns, 3)) @ a.T + b, rng.random((nt, 3)) @ a.T + b,
                      basis, k))
    n_tgt = sum(len(t) for _, t, _, _ in probs)
    print(f"C3 full size: {len(probs)} problems, {n_tgt} targets", flush=True)
    dev = [(cuda(s), cuda(t), bs, k) for s, t, bs, k in probs]
    form = dict(rbf="wendland2", precision="fp32_bulk", solver="cholesky")
    batch = dtk.MovingLeastSquaresBatch(dev, **form)
    st = batch.stats()
    assert st["n_target"] == n_tgt
    coef = rng.standard_normal((len(probs), 4))
    xs = [c[0] + s[:, 0] * c[1] + s[:, 1] * c[2] + s[:, 2] * c[3]
          for (s, _, _, _), c in zip(dev, coef)]
    ys = [torch.zeros(len(t), dtype=torch.float64, device="cuda") for _, t, _, _ in dev]
    batch.apply(xs, ys)
    for i in rng.choice(len(probs), size=48, replace=False):
        s, t, bs, k = dev[i]
        op = dtk.MovingLeastSquaresOperator(s, t, basis=bs, k=k, **form)
        ref = torch.empty_like(ys[i])
        op.apply(xs[i], ref)
        assert torch.equal(ys[i], ref), f"problem {i}: {bs} k={k} n_src={len(s)}"
    over = [i for i, (_, _, bs, k) in enumerate(probs) if bs == "linear" or k >= 32]
    worst_lin = 0.0
    for i in over:
        s, t, _, _ = dev[i]
        c = coef[i]
        f = c[0] + t[:, 0] * c[1] + t[:, 1] * c[2] + t[:, 2] * c[3]
        worst_lin = max(worst_lin, float((ys[i] - f).abs().max() / xs[i].abs().max()))
    ones = [torch.full((len(s),), 3.5, dtype=torch.float64, device="cuda") for s, _, _, _ in dev]
    batch.apply(ones, ys)
    worst_const = max(float((ys[i] - 3.5).abs().max()) / 3.5 for i in over)
    print(f"  {len(over)} overdetermined problems: linear reproduction {worst_lin:.2e}, "
          f"constants {worst_const:.2e}", flush=True)
    # measured: 6.3e-7 and 1.0e-6
    assert worst_lin <= 5e-6
    assert worst_const <= 5e-6
    # oracle on a sample (sizes the CPU oracle does in seconds)
    batch.apply(xs, ys)
    small = [i for i, (s, t, _, k) in enumerate(probs) if len(s) <= 4000 and k >= 16]
    worst = 0.0
    for i in rng.choice(small, size=6, replace=False):
        s, t, bs, k = probs[i]
        idx, co, _ = oracle_lib.mls_create(s, t, M[bs], k=k, rbf_name="wendland2")
        ref = oracle_lib.apply(idx, co, xs[i].cpu().numpy()[None])[0]
        worst = max(worst, float(np.abs(ys[i].cpu().numpy() - ref).max() / np.abs(ref).max()))
    print(f"  sampled problems vs oracle {worst:.2e}", flush=True)
    assert worst <= 1e-6

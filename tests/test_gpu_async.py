"""The create and the device-array apply return once their work is queued
on the handle's stream; the statistics wait for the create that made them.
Results must not depend on it: operators built and applied back to back
without any wait, on the null stream and on a stream of the caller's, equal
one built and applied with a wait after each call, bit for bit, with the
same statistics."""
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


KEYS = ("n_underdetermined", "n_unconverged", "total_rotations", "nnz",
        "n_stack_overflow", "n_eigen_fallback")


def _waited(dtk, src, tgt, f, **kw):
    op = dtk.MovingLeastSquaresOperator(src, tgt, **kw)
    torch.cuda.synchronize()
    y = torch.zeros(tgt.shape[0], dtype=torch.float64, device="cuda")
    op.apply(f, y)
    torch.cuda.synchronize()
    return y.cpu().numpy(), {k: op.stats()[k] for k in KEYS}


@pytest.mark.parametrize("kw", [
    dict(basis="quadratic", k=32, rbf="wendland2", precision="fp64", solver="classical"),
    dict(basis="quadratic", k=32, rbf="wendland2", precision="fp32_bulk", solver="cholesky"),
], ids=["classical", "fp32_bulk"])
def test_queued_create_and_apply_equal_waited(dtk, kw):
    src, tgt, f = c1_inputs(40_000, seed=3)
    d_src = torch.as_tensor(src).cuda()
    d_tgt = torch.as_tensor(tgt).cuda()
    d_f = torch.as_tensor(f).cuda()
    print(f"queued create/apply {kw['solver']} {kw['precision']}", flush=True)
    ref_y, ref_st = _waited(dtk, d_src, d_tgt, d_f, **kw)

    # three operators queued back to back; statistics read last, in reverse
    ops, ys = [], []
    for _ in range(3):
        op = dtk.MovingLeastSquaresOperator(d_src, d_tgt, **kw)
        y = torch.zeros(len(tgt), dtype=torch.float64, device="cuda")
        op.apply(d_f, y)
        ops.append(op)
        ys.append(y)
    for op in reversed(ops):
        assert {k: op.stats()[k] for k in KEYS} == ref_st
    for y in ys:
        assert np.array_equal(y.cpu().numpy(), ref_y)

    # on a stream of the caller's (default priority)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        op = dtk.MovingLeastSquaresOperator(d_src, d_tgt, stream=s.cuda_stream, **kw)
        y = torch.zeros(len(tgt), dtype=torch.float64, device="cuda")
        op.apply(d_f, y)
    s.synchronize()
    assert np.array_equal(y.cpu().numpy(), ref_y)
    assert {k: op.stats()[k] for k in KEYS} == ref_st

"""dtkx_mls_apply_noalloc: the apply with its scratch in a fixed device
array instead of allocated buffers. Through the C ABI, against
dtkx_mls_apply on the same handle, bit for bit, on calls in a row with
different (smaller) operators, and its refusals (host arrays, a scratch
need above DTKX_NOALLOC_SCRATCH_BYTES)."""
import ctypes

import numpy as np
import pytest

from conftest import c1_inputs

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

_vp = ctypes.c_void_p
SCRATCH_BYTES = 64 << 20


@pytest.fixture(scope="module")
def lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import datatransferkit_amd as d
    from datatransferkit_amd import _lib

    return d, _lib, _lib.load()


def cuda(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).cuda()


def _both(lib, op, x, ns, nt, nf, pad_s, pad_t):
    """(apply, apply_noalloc) results of nf fields at padded leading
    dimensions; the gaps must come back untouched."""
    d, _lib, L = lib
    lds, ldt = ns + pad_s, nt + pad_t
    xp = torch.full((nf, lds), 7.0, dtype=torch.float64, device="cuda")
    xp[:, :ns] = cuda(x)
    out = []
    for fn in (L.dtkx_mls_apply, L.dtkx_mls_apply_noalloc):
        yp = torch.full((nf, ldt), -3.0, dtype=torch.float64, device="cuda")
        _lib.check(fn(op._h, _vp(xp.data_ptr()), _vp(yp.data_ptr()), nf, lds, ldt))
        y = yp.cpu().numpy()
        assert np.all(y[:, nt:] == -3.0)
        out.append(y[:, :nt])
    return out


@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_noalloc_equals_apply_on_calls_in_a_row(lib, precision):
    d = lib[0]
    rng = np.random.default_rng(12)
    # first call: 40k -> 30k, quadratic, 5 fields (passes of 8 and padding)
    src, tgt, _ = c1_inputs(40_000, seed=12)
    tgt = tgt[:30_000] * 0.9 + 0.05
    print(f"noalloc {precision}: 40000 -> 30000, 5 fields", flush=True)
    op = d.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="quadratic", k=16,
                                      rbf="wendland2", precision=precision)
    x = rng.standard_normal((5, len(src)))
    a, b = _both(lib, op, x, len(src), len(tgt), 5, 13, 5)
    assert np.array_equal(a.view(np.uint64), b.view(np.uint64))
    # second call, smaller: 6k -> 2k, linear, 2 fields, another handle
    src2, tgt2, _ = c1_inputs(6_000, seed=13)
    tgt2 = tgt2[:2_000]
    print(f"noalloc {precision}: 6000 -> 2000, 2 fields", flush=True)
    op2 = d.MovingLeastSquaresOperator(cuda(src2), cuda(tgt2), basis="linear", k=8,
                                       precision=precision)
    x2 = rng.standard_normal((2, len(src2)))
    a2, b2 = _both(lib, op2, x2, len(src2), len(tgt2), 2, 0, 3)
    assert np.array_equal(a2.view(np.uint64), b2.view(np.uint64))
    # and the first handle again, after the smaller call used the scratch
    a3, b3 = _both(lib, op, x, len(src), len(tgt), 5, 0, 0)
    assert np.array_equal(a3.view(np.uint64), a.view(np.uint64))
    assert np.array_equal(b3.view(np.uint64), a.view(np.uint64))


def test_noalloc_allocates_nothing(lib):
    d, _lib, L = lib
    src, tgt, f = c1_inputs(20_000, seed=14)
    op = d.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="linear")
    x = cuda(np.stack([f, 2 * f, f * f]))
    y = torch.zeros((3, len(tgt)), dtype=torch.float64, device="cuda")
    args = (op._h, _vp(x.data_ptr()), _vp(y.data_ptr()), 3, len(src), len(tgt))
    _lib.check(L.dtkx_mls_apply_noalloc(*args))
    torch.cuda.synchronize()
    free0, _ = torch.cuda.mem_get_info()
    for _ in range(5):
        _lib.check(L.dtkx_mls_apply_noalloc(*args))
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 == free0


def test_noalloc_refusals(lib):
    d, _lib, L = lib
    src, tgt, f = c1_inputs(2_000, seed=15)
    op = d.MovingLeastSquaresOperator(cuda(src), cuda(tgt), basis="linear")
    # host arrays would need staging buffers
    xh = np.ascontiguousarray(f)
    yh = np.zeros(len(tgt))
    rc = L.dtkx_mls_apply_noalloc(op._h, _vp(xh.ctypes.data), _vp(yh.ctypes.data), 1,
                                  len(src), len(tgt))
    assert rc == _lib.ERR_INVALID_ARGUMENT
    assert "is_device_pointer" in L.dtkx_last_error().decode()
    # a scratch need above the fixed array: 1.2M sources x 8 fields x 8 B
    n = 1_200_000
    rng = np.random.default_rng(16)
    big = rng.random((n, 3))
    print(f"noalloc refusal: {n} sources, 8 fields", flush=True)
    op = d.MovingLeastSquaresOperator(cuda(big), cuda(big[:10]), basis="linear")
    x = torch.zeros((8, n), dtype=torch.float64, device="cuda")
    y = torch.zeros((8, 10), dtype=torch.float64, device="cuda")
    assert n * 8 * 8 > SCRATCH_BYTES
    rc = L.dtkx_mls_apply_noalloc(op._h, _vp(x.data_ptr()), _vp(y.data_ptr()), 8, n, 10)
    assert rc == _lib.ERR_UNSUPPORTED
    assert "kNoAllocScratchBytes" in L.dtkx_last_error().decode()
    # the allocating apply takes the same call
    _lib.check(L.dtkx_mls_apply(op._h, _vp(x.data_ptr()), _vp(y.data_ptr()), 8, n, 10))

"""GPU parity of the nearest-neighbour operator (dtkx_nn_*, the reference's
NearestNeighborOperator, packages/Meshfree/src/DTK_NearestNeighborOperator_
def.hpp:23-86) against the CPU oracle.

Bars: the nearest-source indices bit-exact against the oracle's k = 1 search
(float and double coordinates, ties by index); applied values are copies,
bit for bit (signed zeros, infinities and NaN payloads included). The
reference's own analytic cases run through the C++ drop-in
(tests/native/test_nn_cpp.cpp, tests/test_cpp_dropin.py).
"""
import ctypes

import numpy as np
import pytest

from conftest import grid_points

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
    return np.ascontiguousarray(a, dtype=np.float64).view(np.uint64)


@pytest.mark.parametrize("geometry", ["float", "double"])
def test_nn_indices_and_values_match_oracle(dtk, oracle_lib, geometry):
    rng = np.random.default_rng(21)
    src = rng.random((50_000, 3))
    tgt = rng.random((20_000, 3)) * 1.1 - 0.05
    g = 0 if geometry == "float" else 1
    gi, _ = oracle_lib.knn(src, tgt, 1, geometry=g)
    op = dtk.NearestNeighborOperator(cuda(src), cuda(tgt), geometry=geometry)
    assert np.array_equal(op.indices().cpu().numpy(), gi[:, 0])
    x = rng.standard_normal((3, len(src)))
    y = torch.empty((3, len(tgt)), dtype=torch.float64, device="cuda")
    op.apply(cuda(x), y)
    assert np.array_equal(bits(y.cpu().numpy()), bits(x[:, gi[:, 0]]))


def test_nn_ties_duplicates_and_special_values(dtk, oracle_lib):
    # lattice targets equidistant from several sources, duplicated sources
    lattice = grid_points((8, 8, 8), (0, 0, 0)) / 7.0
    src = np.concatenate([lattice, lattice[:100]])
    tgt = np.concatenate([lattice + 0.5 / 7.0, lattice[::3]])
    gi, _ = oracle_lib.knn(src, tgt, 1)
    op = dtk.NearestNeighborOperator(cuda(src), cuda(tgt))
    idx = op.indices().cpu().numpy()
    assert np.array_equal(idx, gi[:, 0])
    x = np.linspace(-1, 1, len(src))
    x[:4] = [-0.0, np.inf, -np.inf, np.nan]
    x.view(np.uint64)[4] = 0x7FF800000000BEEF  # a NaN with a payload
    y = torch.empty(len(tgt), dtype=torch.float64, device="cuda")
    op.apply(cuda(x), y)
    assert np.array_equal(bits(y.cpu().numpy()), bits(x[idx]))


def test_nn_layouts_and_host_pointers(dtk, oracle_lib):
    rng = np.random.default_rng(22)
    src = rng.random((3000, 3))
    tgt = rng.random((1000, 3))
    gi, _ = oracle_lib.knn(src, tgt, 1)
    ref = gi[:, 0]
    soa = dtk.NearestNeighborOperator(cuda(src.T.copy()), cuda(tgt.T.copy()),
                                      layout="soa")
    assert np.array_equal(soa.indices().cpu().numpy(), ref)
    host = dtk.NearestNeighborOperator(src, tgt)  # staged host points
    assert np.array_equal(host.indices().cpu().numpy(), ref)
    x = rng.standard_normal(len(src))
    y = np.zeros(len(tgt))
    host.apply(x, y)
    assert np.array_equal(y, x[ref])


def test_nn_padded_host_fields_keep_their_gaps(dtk, oracle_lib):
    # [fields][ld] with ld > n through the C ABI: the entries between the
    # fields of a host output come back unchanged
    from datatransferkit_amd import _lib

    rng = np.random.default_rng(23)
    src = rng.random((500, 3))
    tgt = rng.random((300, 3))
    gi, _ = oracle_lib.knn(src, tgt, 1)
    op = dtk.NearestNeighborOperator(cuda(src), cuda(tgt))
    nf, lds, ldt = 3, 512, 320
    x = rng.standard_normal((nf, lds))
    y = np.full((nf, ldt), 7.25)
    L = _lib.load()
    _lib.check(L.dtkx_nn_apply(op._h, ctypes.c_void_p(x.ctypes.data),
                               ctypes.c_void_p(y.ctypes.data), nf, lds, ldt))
    assert np.array_equal(y[:, :300], x[:, gi[:, 0]])
    assert np.all(y[:, 300:] == 7.25)


def test_nn_preconditions(dtk):
    from datatransferkit_amd import DtkxError, _lib

    with pytest.raises(DtkxError) as e:
        dtk.NearestNeighborOperator(cuda(np.zeros((0, 3))), cuda(np.zeros((4, 3))))
    assert e.value.status == _lib.ERR_EMPTY_SOURCE
    op = dtk.NearestNeighborOperator(cuda(np.random.rand(10, 3)),
                                     cuda(np.random.rand(5, 3)))
    with pytest.raises(DtkxError) as e:
        op.apply(cuda(np.zeros(9)), torch.zeros(5, dtype=torch.float64, device="cuda"))
    assert e.value.status == _lib.ERR_SIZE_MISMATCH
    empty = dtk.NearestNeighborOperator(cuda(np.random.rand(10, 3)),
                                        cuda(np.zeros((0, 3))))
    assert empty.indices().numel() == 0


@pytest.mark.parametrize("shift", [0, 1, 7])
def test_fetch_semantics_one_rank(dtk, shift):
    """The reference's fetch test (tstDetailsCommunicationHelpers.cpp:167-222)
    on one rank is a gather, out[i] = v[indices[i]], of a rank-1 view (v)
    and a rank-2 view (w, DIM = 2 columns). Here the gather is the
    nearest-neighbour apply: target i sits on source indices[i], the cyclic
    shift indices[i] = (shift + i) % P of the test's plan, so its nearest
    source is that one (distance 0, lattice points distinct)."""
    P = 125
    src = grid_points((5, 5, 5), (0, 0, 0)) * 0.25
    indices = (shift + np.arange(P)) % P
    tgt = src[indices]
    op = dtk.NearestNeighborOperator(cuda(src), cuda(tgt))
    assert np.array_equal(op.indices().cpu().numpy(), indices)
    v = np.arange(P, dtype=np.float64)                  # iota
    y = torch.empty(P, dtype=torch.float64, device="cuda")
    op.apply(cuda(v), y)
    assert np.array_equal(y.cpu().numpy(), indices.astype(np.float64))
    # w(i, j) = i + j P, as fields [j][i]; w_ref(i, j) = indices(i) + j P
    w = np.stack([np.arange(P) + j * P for j in range(2)]).astype(np.float64)
    yw = torch.empty((2, P), dtype=torch.float64, device="cuda")
    op.apply(cuda(w), yw)
    ref = np.stack([indices + j * P for j in range(2)]).astype(np.float64)
    assert np.array_equal(yw.cpu().numpy(), ref)

"""Host-side argument checks of the batch wrapper: every case here raises
before the library is called, so it runs without a GPU (CPU tensors)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def dtk():
    import datatransferkit_amd as d

    return d


def pts(n, cols=3):
    return torch.zeros((n, cols), dtype=torch.float64)


def test_from_concatenated_checks_counts(dtk):
    src, tgt = pts(10), pts(6)
    cases = [
        (([6, 4], [3, 3, 0]), 2),     # different problem counts
        (([6, 5], [3, 3]), 2),        # source counts do not sum to the rows
        (([6, 4], [3, 2]), 2),        # target counts do not sum to the rows
        (([6, 4, -0], [3, 4, -1]), 1),  # negative count
    ]
    for (ns, nt), status in cases:
        with pytest.raises(dtk.DtkxError) as e:
            dtk.MovingLeastSquaresBatch.from_concatenated(src, tgt, ns, nt, "linear", 4)
        assert e.value.status == status, (ns, nt)
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch.from_concatenated(src, tgt, [6, 4], [3, 3],
                                                      ["linear"], 4)
    assert e.value.status == 2


def test_batch_point_shapes_and_types(dtk):
    f = dtk.MovingLeastSquaresBatch.from_concatenated
    with pytest.raises(dtk.DtkxError) as e:
        f(pts(10, 2), pts(6), [10], [6], "linear", 4)
    assert e.value.status == 3
    with pytest.raises(dtk.DtkxError) as e:
        f(pts(10).float(), pts(6), [10], [6], "linear", 4)
    assert e.value.status == 1
    with pytest.raises(dtk.DtkxError) as e:
        f(pts(10).t().contiguous().t(), pts(6), [10], [6], "linear", 4)
    assert e.value.status == 1
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch([(pts(10), pts(6, 2), "linear", 4)])
    assert e.value.status == 3
    with pytest.raises(dtk.DtkxError) as e:
        dtk.MovingLeastSquaresBatch([(pts(10), torch.zeros(6, dtype=torch.float64),
                                      "linear", 4)])
    assert e.value.status == 3
    with pytest.raises(KeyError):
        dtk.MovingLeastSquaresBatch([(pts(10), pts(6), "cubic", 4)])


def test_point_address_of_numpy_and_torch():
    from datatransferkit_amd import operators

    f64 = operators._torch_float64()
    t = pts(5)
    a, rows, keep = operators._point_address(t, "p", f64)
    assert (a, rows) == (t.data_ptr(), 5) and keep is t
    n = np.zeros((7, 3))
    a, rows, keep = operators._point_address(n, "p", f64)
    assert rows == 7 and a == keep.ctypes.data
    # a non-contiguous numpy array is copied, and the copy is what is kept
    n2 = np.zeros((3, 7)).T
    a, rows, keep = operators._point_address(n2, "p", f64)
    assert rows == 7 and keep.flags.c_contiguous and a == keep.ctypes.data

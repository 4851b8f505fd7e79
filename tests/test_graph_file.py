"""The DTKXCSR1 operator file (dtkx_mls_write_graph, read_graph).

CPU: the reader against a file written here byte by byte from the format
declared in include/dtkx/mls.h, and its rejections. GPU: a written operator
reads back as its get_graph CSR with global column ids, and keeps the MLS
invariants (rows of k entries, columns in range, coefficients summing to
one: the constant term is reproduced, Impl.hpp:293-336 with basis(0) = e1).
"""
import struct

import numpy as np
import pytest

import datatransferkit_amd as dtk
from datatransferkit_amd import DtkxError


def _write(path, offsets, columns, coeffs, n_cols, col_base, magic=b"DTKXCSR1"):
    cb = coeffs.dtype.itemsize
    with open(path, "wb") as fh:
        fh.write(magic)
        fh.write(struct.pack("<4q", len(offsets) - 1, n_cols, len(columns), col_base))
        fh.write(struct.pack("<2i", cb, 0))
        fh.write(np.asarray(offsets, "<i8").tobytes())
        fh.write(np.asarray(columns, "<i8").tobytes())
        fh.write(coeffs.astype("<f8" if cb == 8 else "<f4").tobytes())


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_reader_round_trip(tmp_path, dtype):
    p = tmp_path / "g.csr"
    off = np.array([0, 2, 2, 5])
    col = np.array([10, 13, 11, 12, 10]) + 1_000_000_000_000
    co = np.array([0.5, 0.5, 0.2, 0.3, 0.5], dtype=dtype)
    _write(p, off, col, co, n_cols=4, col_base=1_000_000_000_010)
    g = dtk.read_graph(p)
    assert (g["n_rows"], g["n_cols"], g["col_base"]) == (3, 4, 1_000_000_000_010)
    assert np.array_equal(g["offsets"], off)
    assert np.array_equal(g["columns"], col)
    assert g["coeffs"].dtype == dtype and np.array_equal(g["coeffs"], co)


def test_reader_rejects_bad_files(tmp_path):
    p = tmp_path / "bad.csr"
    _write(p, [0, 1], [0], np.ones(1), 1, 0, magic=b"NOTACSR!")
    with pytest.raises(DtkxError):
        dtk.read_graph(p)
    _write(p, [0, 2], [0, 1], np.ones(2), 2, 0)
    data = p.read_bytes()
    p.write_bytes(data[:-4])  # truncated coefficients
    with pytest.raises(DtkxError):
        dtk.read_graph(p)
    _write(p, [1, 2], [0, 1], np.ones(2), 2, 0)  # offsets must start at 0
    with pytest.raises(DtkxError):
        dtk.read_graph(p)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["fp64", "fp32_bulk"])
def test_written_operator_reads_back(tmp_path, precision):
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    rng = np.random.default_rng(31)
    src = torch.from_numpy(rng.random((4000, 3))).cuda()
    tgt = torch.from_numpy(0.1 + 0.8 * rng.random((1500, 3))).cuda()
    op = dtk.MovingLeastSquaresOperator(src, tgt, basis="quadratic", rbf="wendland2",
                                        k=16, precision=precision,
                                        solver="cholesky" if precision == "fp32_bulk"
                                        else "classical")
    p = tmp_path / "op.csr"
    op.write_graph(p, col_base=7)
    g = dtk.read_graph(p)
    off, idx, co = (t.cpu().numpy() for t in op.graph())
    assert (g["n_rows"], g["n_cols"], g["col_base"]) == (1500, 4000, 7)
    assert np.array_equal(g["offsets"], off)
    assert np.array_equal(g["columns"], idx.reshape(-1).astype(np.int64) + 7)
    assert np.array_equal(g["coeffs"], co.reshape(-1))
    # invariants: k entries per row, columns in range, partition of unity
    assert np.all(np.diff(g["offsets"]) == 16)
    assert g["columns"].min() >= 7 and g["columns"].max() < 7 + 4000
    # sum_j c_j = (A^+ A)_00: one to eps * cond(A), some 1e5 for quadratic
    # moments at k = 16 (2.8e-11 measured); float coefficients add 32 ulps
    rows = np.add.reduceat(g["coeffs"].astype(np.float64), g["offsets"][:-1])
    tol = 1e-9 if precision == "fp64" else 1e-5
    assert np.abs(rows - 1.0).max() < tol

"""Host-side mirror of DataTransferKit's MovingLeastSquaresOperator.

Same construction / apply surface as the reference
(packages/Meshfree/src/DTK_MovingLeastSquaresOperator_decl.hpp:36-63):

    op = MovingLeastSquaresOperator(source_points, target_points,
                                    rbf="wendland0", basis="linear")
    op.apply(source_values, target_values)

The template policies of the reference (RBF, polynomial basis) become
keyword arguments; ``polynomial_basis`` and ``radial_basis_function``
attributes mirror its public typedefs (:42-45). Points are [n, 3] float64
arrays (torch CUDA tensors are used in place; numpy arrays or CPU tensors
are staged by the library). Preconditions the reference checks with
DTK_REQUIRE raise :class:`DtkxError` here (always on).

Every computation runs in the HIP library (``libdtkx.so``); there is no
CPU path.
"""
from __future__ import annotations

import ctypes
import sys

import numpy as np

from . import _lib
from ._lib import DtkxError, check


def _ptr(a):
    """(pointer, keep-alive) for a torch tensor or numpy array."""
    if a is None:
        return None, None
    if hasattr(a, "data_ptr"):  # torch tensor
        if not a.is_contiguous():
            raise DtkxError(_lib.ERR_INVALID_ARGUMENT, "tensor must be contiguous")
        return ctypes.c_void_p(a.data_ptr()), a
    arr = np.ascontiguousarray(a)
    return ctypes.c_void_p(arr.ctypes.data), arr


def _dtype_name(a):
    return str(a.dtype).replace("torch.", "")


def _check_points(p, name):
    shape = tuple(p.shape)
    if len(shape) != 2:
        raise DtkxError(_lib.ERR_DIMENSION, f"{name} must be [n, dim]")
    if _dtype_name(p) != "float64":
        raise DtkxError(_lib.ERR_INVALID_ARGUMENT, f"{name} must be float64")
    return shape


def _torch_float64():
    """torch.float64, or None when torch has not been imported (then no
    argument is a torch tensor)."""
    torch = sys.modules.get("torch")
    return None if torch is None else torch.float64


def _point_address(p, name, f64):
    """(address, rows, keep-alive) of one [n, 3] float64 point array of a
    batch, with the checks of _check_points and _ptr; f64 =
    _torch_float64(). A contiguous float64 torch tensor takes four
    attribute reads, which is what a batch of thousands of problems spends
    its Python time on."""
    if f64 is not None and getattr(p, "dtype", None) is f64 and p.is_contiguous():
        shape = p.shape
        if len(shape) == 2 and shape[1] == 3:
            return p.data_ptr(), shape[0], p
    shape = _check_points(p, name)
    if shape[1] != 3:
        raise DtkxError(_lib.ERR_DIMENSION, f"{name} must be [n, 3]")
    ptr, keep = _ptr(p)
    return ptr.value or 0, shape[0], keep


# numpy image of _lib.Problem: a batch's problem records are filled column
# by column and handed to the library as one buffer
_PROBLEM_DTYPE = np.dtype({
    "names": [f for f, _ in _lib.Problem._fields_],
    "formats": [np.uint64, np.int64, np.uint64, np.int64, np.int32, np.int32],
    "offsets": [getattr(_lib.Problem, f).offset for f, _ in _lib.Problem._fields_],
    "itemsize": ctypes.sizeof(_lib.Problem),
})


def _fields(vals, n, name):
    """(fields, leading dimension) of a [n] or [fields, n] value array."""
    shape = tuple(vals.shape)
    if len(shape) == 1:
        nf, ld = 1, shape[0]
    elif len(shape) == 2:
        nf, ld = shape
    else:
        raise DtkxError(_lib.ERR_INVALID_ARGUMENT, f"{name} must be 1-D or 2-D")
    if ld != n:
        raise DtkxError(_lib.ERR_SIZE_MISMATCH,
                        f"{name}.extent(0) == {n} (got {ld})")
    return nf, ld


class MovingLeastSquaresOperator:
    """MLS transfer operator on one GPU (one-rank communicator)."""

    def __init__(self, source_points, target_points, *, rbf="wendland0",
                 basis="linear", k=0, precision="fp64", geometry="float",
                 solver="classical", truncation="absolute", rcond=1e-12,
                 max_rotations=0, group_size=0, store_pinv=False,
                 stream=None, neighbors=None):
        ns, ds = _check_points(source_points, "source_points")
        nt, dt = _check_points(target_points, "target_points")
        if ds != dt:
            raise DtkxError(_lib.ERR_DIMENSION,
                            "source_points.extent(1) == target_points.extent(1)")
        o = _lib.default_options()
        o.dim = ds
        o.basis = _lib.BASIS[basis]
        o.rbf = _lib.RBF[rbf]
        o.k = int(k)
        o.precision = _lib.PRECISION[precision]
        o.geometry = _lib.GEOMETRY[geometry]
        o.solver = _lib.SOLVER[solver]
        o.truncation = _lib.TRUNCATION[truncation]
        o.rcond = float(rcond)
        o.max_rotations = int(max_rotations)
        o.group_size = int(group_size)
        o.store_pinv = int(bool(store_pinv))
        o.stream = ctypes.c_void_p(stream) if stream else None
        self.polynomial_basis = basis
        self.radial_basis_function = rbf
        self.precision = precision
        self._n_src, self._n_tgt = ns, nt
        L = _lib.load()
        sp, ka = _ptr(source_points)
        tp, kb = _ptr(target_points)
        h = ctypes.c_void_p()
        if neighbors is None:
            check(L.dtkx_mls_create(sp, ns, tp, nt, ctypes.byref(o),
                                    ctypes.byref(h)))
        else:
            nb = neighbors
            ip, kc = _ptr(nb)
            check(L.dtkx_mls_create_with_neighbors(
                sp, ns, tp, nt, ip, int(nb.shape[1]), ctypes.byref(o),
                ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().dtkx_mls_destroy(h)
            except Exception:
                pass
            self._h = None

    # ---------------------------------------------------------------- apply
    def _fields(self, vals, n, name):
        return _fields(vals, n, name)

    def apply(self, source_values, target_values):
        """target_values = C source_values (fields along the leading axis)."""
        nf, lds = self._fields(source_values, self._n_src, "source_values")
        nf2, ldt = self._fields(target_values, self._n_tgt, "target_values")
        if nf != nf2:
            raise DtkxError(_lib.ERR_SIZE_MISMATCH, "field counts differ")
        fp32 = _dtype_name(source_values) == "float32"
        L = _lib.load()
        xp, ka = _ptr(source_values)
        yp, kb = _ptr(target_values)
        if fp32:
            check(L.dtkx_mls_apply_fp32(self._h, xp, yp, nf, lds, ldt))
        else:
            check(L.dtkx_mls_apply(self._h, xp, yp, nf, lds, ldt))
        return target_values

    def apply_transpose(self, target_bar, source_bar):
        """source_bar = C^T target_bar."""
        nf, ldt = self._fields(target_bar, self._n_tgt, "target_bar")
        nf2, lds = self._fields(source_bar, self._n_src, "source_bar")
        if nf != nf2:
            raise DtkxError(_lib.ERR_SIZE_MISMATCH, "field counts differ")
        xp, ka = _ptr(target_bar)
        yp, kb = _ptr(source_bar)
        check(_lib.load().dtkx_mls_apply_transpose(self._h, xp, yp, nf, ldt, lds))
        return source_bar

    def target_gradient(self, source_values, target_bar, out=None):
        """dL/d(target coordinates) [n_tgt, 3] for L = sum_f <target_bar_f,
        C(t) source_values_f>, neighbour sets frozen (double geometry)."""
        import torch

        nf, lds = self._fields(source_values, self._n_src, "source_values")
        nf2, ldt = self._fields(target_bar, self._n_tgt, "target_bar")
        if nf != nf2:
            raise DtkxError(_lib.ERR_SIZE_MISMATCH, "field counts differ")
        if out is None:
            out = torch.empty((self._n_tgt, 3), dtype=torch.float64,
                              device=getattr(target_bar, "device", "cuda"))
        xp, ka = _ptr(source_values)
        yp, kb = _ptr(target_bar)
        gp, kc = _ptr(out)
        check(_lib.load().dtkx_mls_target_gradient(self._h, xp, yp, nf, lds, ldt, gp))
        return out

    def write_graph(self, path, col_base=0):
        """Write the operator's CSR (caller-order rows, global column ids
        col_base + source index) in the DTKXCSR1 format; see read_graph."""
        check(_lib.load().dtkx_mls_write_graph(self._h, str(path).encode(),
                                               int(col_base)))

    # ------------------------------------------------------------ inspection
    def stats(self) -> dict:
        s = _lib.Stats()
        check(_lib.load().dtkx_mls_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()

    def graph_pointers(self):
        """Device pointers (offsets, indices, coeffs) and nnz."""
        off, idx, co = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
        nnz = ctypes.c_int64()
        check(_lib.load().dtkx_mls_get_graph(
            self._h, ctypes.byref(off), ctypes.byref(idx), ctypes.byref(co),
            ctypes.byref(nnz)))
        return off.value, idx.value, co.value, nnz.value

    def graph(self):
        """Copies of (offsets, indices [n_tgt, k], coeffs [n_tgt, k]) as torch
        CUDA tensors."""
        import torch

        off, idx, co, nnz = self.graph_pointers()
        nt = self._n_tgt
        k = nnz // nt if nt else 0
        offsets = torch.empty(nt + 1, dtype=torch.int32, device="cuda")
        indices = torch.empty((nt, k), dtype=torch.int32, device="cuda")
        cdt = torch.float32 if self.precision == "fp32_bulk" else torch.float64
        coeffs = torch.empty((nt, k), dtype=cdt, device="cuda")
        _memcpy_d2d(offsets, off, 4 * (nt + 1))
        if nnz:
            _memcpy_d2d(indices, idx, 4 * nnz)
            _memcpy_d2d(coeffs, co, coeffs.element_size() * nnz)
        return offsets, indices, coeffs


def _batch_options(rbf, precision, geometry, solver, truncation, rcond,
                   max_rotations, group_size, stream):
    """The dtkx_mls_options of a batch (each problem names its own basis
    and k)."""
    o = _lib.default_options()
    o.basis = _lib.BASIS["linear"]
    o.rbf = _lib.RBF[rbf]
    o.precision = _lib.PRECISION[precision]
    o.geometry = _lib.GEOMETRY[geometry]
    o.solver = _lib.SOLVER[solver]
    o.truncation = _lib.TRUNCATION[truncation]
    o.rcond = float(rcond)
    o.max_rotations = int(max_rotations)
    o.group_size = int(group_size)
    o.stream = ctypes.c_void_p(stream) if stream else None
    return o


class MovingLeastSquaresBatch:
    """Many independent MLS operators built together: one segmented LBVH
    build and one query launch per kernel size for all problems, one setup
    launch per basis, and an apply replayed from a captured HIP graph.

    problems: sequence of (source_points [n_s, 3], target_points [n_t, 3],
    basis, k) with float64 CUDA tensors; options as for
    MovingLeastSquaresOperator (group_size 0 lets the classical solver pick
    its lanes per system by timing). The point tensors are held by the
    batch: rebuild() reads them again after the caller has moved the points
    in place. from_concatenated() builds the same batch from points held in
    two arrays."""

    def __init__(self, problems, *, rbf="wendland0", precision="fp64",
                 geometry="float", solver="classical", truncation="absolute",
                 rcond=1e-12, max_rotations=0, group_size=0, stream=None):
        f64 = _torch_float64()
        bases = _lib.BASIS
        sp, ns, tp, nt, basis_ids, ks, keep = [], [], [], [], [], [], []
        for src, tgt, basis, k in problems:
            a, rows, kp = _point_address(src, "source_points", f64)
            sp.append(a)
            ns.append(rows)
            keep.append(kp)
            a, rows, kp = _point_address(tgt, "target_points", f64)
            tp.append(a)
            nt.append(rows)
            keep.append(kp)
            basis_ids.append(bases[basis])
            ks.append(int(k))
        self._create(sp, ns, tp, nt, basis_ids, ks, keep, _batch_options(
            rbf, precision, geometry, solver, truncation, rcond, max_rotations,
            group_size, stream))
        self.precision = precision

    @classmethod
    def from_concatenated(cls, source_points, target_points, n_source, n_target,
                          basis, k, *, rbf="wendland0", precision="fp64",
                          geometry="float", solver="classical",
                          truncation="absolute", rcond=1e-12, max_rotations=0,
                          group_size=0, stream=None):
        """The batch of the problems whose points are concatenated in
        problem order: source_points [sum n_source, 3] and target_points
        [sum n_target, 3] (float64 CUDA tensors), problem p owning
        n_source[p] sources and n_target[p] targets; basis and k are one
        value for every problem or one per problem. The same batch as the
        constructor's over the per-problem slices, formed without a Python
        step per problem."""
        f64 = _torch_float64()
        n_source = np.asarray(n_source, dtype=np.int64).reshape(-1)
        n_target = np.asarray(n_target, dtype=np.int64).reshape(-1)
        n = len(n_source)
        if len(n_target) != n:
            raise DtkxError(_lib.ERR_SIZE_MISMATCH, "n_source and n_target lengths")
        if np.any(n_source < 0) or np.any(n_target < 0):
            raise DtkxError(_lib.ERR_INVALID_ARGUMENT, "negative point count")
        addr = []
        for pts, counts, name in ((source_points, n_source, "source_points"),
                                  (target_points, n_target, "target_points")):
            a, rows, _ = _point_address(pts, name, f64)
            if rows != int(counts.sum()):
                raise DtkxError(_lib.ERR_SIZE_MISMATCH,
                                f"{name}.extent(0) == {int(counts.sum())} (got {rows})")
            off = np.zeros(n, np.uint64)
            np.cumsum(counts[:-1], out=off[1:])
            # a problem without points may point one past the end (never read)
            addr.append(off * np.uint64(24) + np.uint64(a))
        bases = [_lib.BASIS[basis]] * n if isinstance(basis, str) else \
            [_lib.BASIS[b] for b in basis]
        ks = np.broadcast_to(np.asarray(k, dtype=np.int64), (n,))
        if len(bases) != n:
            raise DtkxError(_lib.ERR_SIZE_MISMATCH, "one basis per problem")
        self = cls.__new__(cls)
        self._create(addr[0], n_source, addr[1], n_target, bases, ks,
                     [source_points, target_points], _batch_options(
                         rbf, precision, geometry, solver, truncation, rcond,
                         max_rotations, group_size, stream))
        self.precision = precision
        return self

    def _create(self, sp, ns, tp, nt, basis_ids, ks, keep, o):
        n = len(ns)
        rec = np.zeros(n, dtype=_PROBLEM_DTYPE)
        rec["src_xyz"], rec["n_src"] = sp, ns
        rec["tgt_xyz"], rec["n_tgt"] = tp, nt
        rec["basis"], rec["k"] = basis_ids, ks
        self._keep = keep
        self._ns = rec["n_src"].tolist()
        self._nt = rec["n_tgt"].tolist()
        # byte offsets of each problem's values in concatenated arrays
        self._x_off = np.zeros(n, np.uint64)
        self._y_off = np.zeros(n, np.uint64)
        np.cumsum(rec["n_src"][:-1], out=self._x_off[1:])
        np.cumsum(rec["n_tgt"][:-1], out=self._y_off[1:])
        self._x_off *= np.uint64(8)
        self._y_off *= np.uint64(8)
        self._n_src_total = int(rec["n_src"].sum())
        self._n_tgt_total = int(rec["n_tgt"].sum())
        h = ctypes.c_void_p()
        arr = (_lib.Problem * n).from_buffer(rec) if n else None
        check(_lib.load().dtkx_mls_batch_create(n, arr, ctypes.byref(o), ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().dtkx_mls_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def apply(self, source_values, target_values):
        """One field per problem: target_values[p] = C_p source_values[p]
        (float64 CUDA tensors of the problems' sizes)."""
        n = len(self._ns)
        if len(source_values) != n or len(target_values) != n:
            raise DtkxError(_lib.ERR_SIZE_MISMATCH, "one array per problem")
        f64 = _torch_float64()
        xs, ys = [], []
        for i, ns, nt, x, y in zip(range(n), self._ns, self._nt,
                                   source_values, target_values):
            if x.shape[-1] != ns or y.shape[-1] != nt:
                raise DtkxError(_lib.ERR_SIZE_MISMATCH, f"problem {i}: value sizes")
            if (x.dtype is not f64 or y.dtype is not f64) and (
                    _dtype_name(x) != "float64" or _dtype_name(y) != "float64"):
                raise DtkxError(_lib.ERR_INVALID_ARGUMENT, "values must be float64")
            if not (x.is_contiguous() and y.is_contiguous()):
                raise DtkxError(_lib.ERR_INVALID_ARGUMENT,
                                f"problem {i}: values must be contiguous")
            if not (x.is_cuda and y.is_cuda):
                raise DtkxError(_lib.ERR_INVALID_ARGUMENT,
                                f"problem {i}: values must be CUDA tensors")
            xs.append(x.data_ptr())
            ys.append(y.data_ptr())
        xs = np.array(xs, dtype=np.uint64)
        ys = np.array(ys, dtype=np.uint64)
        check(_lib.load().dtkx_mls_batch_apply(self._h, xs.ctypes.data, ys.ctypes.data))
        return target_values

    def apply_concatenated(self, source_values, target_values):
        """apply() with the problems' values concatenated in problem order:
        source_values [sum n_src] and target_values [sum n_tgt], contiguous
        float64 CUDA tensors, problem p's values at offset sum_{q<p} n_q.
        The same library call as apply(); its per-problem pointers are
        formed in one array operation, so the host cost of a call does not
        grow with the number of problems."""
        for v, total, name in ((source_values, self._n_src_total, "source_values"),
                               (target_values, self._n_tgt_total, "target_values")):
            if tuple(v.shape) != (total,):
                raise DtkxError(_lib.ERR_SIZE_MISMATCH,
                                f"{name} must be [{total}] (got {tuple(v.shape)})")
            if _dtype_name(v) != "float64":
                raise DtkxError(_lib.ERR_INVALID_ARGUMENT, f"{name} must be float64")
            if not v.is_contiguous():
                raise DtkxError(_lib.ERR_INVALID_ARGUMENT, f"{name} must be contiguous")
            if not v.is_cuda:
                raise DtkxError(_lib.ERR_INVALID_ARGUMENT, f"{name} must be a CUDA tensor")
        xs = self._x_off + np.uint64(source_values.data_ptr())
        # a problem with no targets may point one past the end (never read)
        ys = self._y_off + np.uint64(target_values.data_ptr())
        check(_lib.load().dtkx_mls_batch_apply(self._h, xs.ctypes.data, ys.ctypes.data))
        return target_values

    def rebuild(self):
        """Builds every operator again from the current contents of the
        point tensors given at construction (same sizes, bases and k): one
        replay of the search and setup captured on the first call, with
        results bit-identical to a new batch over the moved points."""
        check(_lib.load().dtkx_mls_batch_rebuild(self._h))
        return self

    def stats(self) -> dict:
        s = _lib.Stats()
        check(_lib.load().dtkx_mls_batch_get_stats(self._h, ctypes.byref(s)))
        return s.as_dict()


def _memcpy_d2d(dst_tensor, src_ptr, nbytes):
    import torch

    # wrap the raw device pointer without taking ownership
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                              ctypes.c_int]
    torch.cuda.synchronize()
    rc = hip.hipMemcpy(ctypes.c_void_p(dst_tensor.data_ptr()),
                       ctypes.c_void_p(src_ptr), nbytes, 3)  # DeviceToDevice
    if rc != 0:
        raise DtkxError(_lib.ERR_HIP, f"hipMemcpy failed ({rc})")


_CSR_MAGIC = b"DTKXCSR1"


def read_graph(path):
    """Read a DTKXCSR1 file (dtkx_mls_write_graph) into numpy arrays:
    dict(n_rows, n_cols, col_base, offsets int64 [n_rows + 1], columns int64
    [nnz], coeffs float64 or float32 [nnz]). Little-endian; the header is
    validated and nothing in the file is executed."""
    with open(path, "rb") as fh:
        if fh.read(8) != _CSR_MAGIC:
            raise DtkxError(_lib.ERR_INVALID_ARGUMENT, f"{path}: not a DTKXCSR1 file")
        n_rows, n_cols, nnz, col_base = (int(v) for v in np.fromfile(fh, "<i8", 4))
        cbytes, _ = (int(v) for v in np.fromfile(fh, "<i4", 2))
        if cbytes not in (4, 8) or min(n_rows, n_cols, nnz) < 0:
            raise DtkxError(_lib.ERR_INVALID_ARGUMENT, f"{path}: bad header")
        offsets = np.fromfile(fh, "<i8", n_rows + 1)
        columns = np.fromfile(fh, "<i8", nnz)
        coeffs = np.fromfile(fh, "<f8" if cbytes == 8 else "<f4", nnz)
    if (len(offsets) != n_rows + 1 or len(columns) != nnz or len(coeffs) != nnz
            or offsets[0] != 0 or offsets[-1] != nnz):
        raise DtkxError(_lib.ERR_INVALID_ARGUMENT, f"{path}: truncated or inconsistent")
    return {"n_rows": n_rows, "n_cols": n_cols, "col_base": col_base,
            "offsets": offsets, "columns": columns, "coeffs": coeffs}


class NearestNeighborOperator:
    """Nearest-neighbour transfer on one GPU: each target takes the value of
    its nearest source (packages/Meshfree/src/DTK_NearestNeighborOperator_
    decl.hpp:24-50, _def.hpp:23-86; ties go to the smallest source index).

        op = NearestNeighborOperator(source_points, target_points)
        op.apply(source_values, target_values)
    """

    def __init__(self, source_points, target_points, *, geometry="float",
                 layout="aos", stream=None):
        if layout == "aos":
            ns, ds = _check_points(source_points, "source_points")
            nt, dt = _check_points(target_points, "target_points")
        else:
            ds, ns = _check_points(source_points, "source_points")
            dt, nt = _check_points(target_points, "target_points")
        if ds != 3 or dt != 3:
            raise DtkxError(_lib.ERR_DIMENSION, "points must be 3-D")
        self._n_src, self._n_tgt = ns, nt
        sp, ka = _ptr(source_points)
        tp, kb = _ptr(target_points)
        h = ctypes.c_void_p()
        check(_lib.load().dtkx_nn_create(
            sp, ns, tp, nt, _lib.GEOMETRY[geometry], _lib.LAYOUT[layout],
            ctypes.c_void_p(stream) if stream else None, ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().dtkx_nn_destroy(h)
            except Exception:
                pass
            self._h = None

    def apply(self, source_values, target_values):
        """target_values[f, t] = source_values[f, nearest(t)]."""
        nf, lds = _fields(source_values, self._n_src, "source_values")
        nf2, ldt = _fields(target_values, self._n_tgt, "target_values")
        if nf != nf2:
            raise DtkxError(_lib.ERR_SIZE_MISMATCH, "field counts differ")
        xp, ka = _ptr(source_values)
        yp, kb = _ptr(target_values)
        check(_lib.load().dtkx_nn_apply(self._h, xp, yp, nf, lds, ldt))
        return target_values

    def indices(self):
        """Copy of the nearest source of every target (int32 CUDA tensor)."""
        import torch

        p, n = ctypes.c_void_p(), ctypes.c_int64()
        check(_lib.load().dtkx_nn_get_indices(self._h, ctypes.byref(p),
                                              ctypes.byref(n)))
        out = torch.empty(n.value, dtype=torch.int32, device="cuda")
        if n.value:
            _memcpy_d2d(out, p.value, 4 * n.value)
        return out


class SplineOperator:
    """Compactly supported RBF spline transfer on one GPU
    (packages/Meshfree/src/DTK_SplineOperator_decl.hpp:40-113,
    _def.hpp:170-295): y = (Q + N) (P + M + P^T)^-1 S x with the linear
    polynomial basis and 4 neighbours per row, solved by restarted GMRES on
    the device to ``tolerance`` (the reference's Belos setting, 1e-10).

        op = SplineOperator(source_points, target_points)
        op.apply(source_values, target_values)
    """

    def __init__(self, source_points, target_points, *, rbf="wendland0",
                 geometry="float", layout="aos", tolerance=1e-10, restart=300,
                 max_iterations=1000, stream=None):
        if layout == "aos":
            ns, ds = _check_points(source_points, "source_points")
            nt, dt = _check_points(target_points, "target_points")
        else:
            ds, ns = _check_points(source_points, "source_points")
            dt, nt = _check_points(target_points, "target_points")
        if ds != 3 or dt != 3:
            raise DtkxError(_lib.ERR_DIMENSION, "points must be 3-D")
        self._n_src, self._n_tgt = ns, nt
        o = _lib.default_spline_options()
        o.rbf = _lib.RBF[rbf]
        o.geometry = _lib.GEOMETRY[geometry]
        o.layout = _lib.LAYOUT[layout]
        o.tolerance = tolerance
        o.restart = restart
        o.max_iterations = max_iterations
        o.stream = stream
        sp, ka = _ptr(source_points)
        tp, kb = _ptr(target_points)
        h = ctypes.c_void_p()
        check(_lib.load().dtkx_spline_create(sp, ns, tp, nt, ctypes.byref(o),
                                             ctypes.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                _lib.load().dtkx_spline_destroy(h)
            except Exception:
                pass
            self._h = None

    def apply(self, source_values, target_values):
        """target_values = A source_values ([n] or [fields, n]); one solve
        per field."""
        nf, lds = _fields(source_values, self._n_src, "source_values")
        nf2, ldt = _fields(target_values, self._n_tgt, "target_values")
        if nf != nf2:
            raise DtkxError(_lib.ERR_SIZE_MISMATCH, "field counts differ")
        xp, ka = _ptr(source_values)
        yp, kb = _ptr(target_values)
        check(_lib.load().dtkx_spline_apply(self._h, xp, yp, nf, lds, ldt))
        return target_values

    def solve(self, source_values, coeffs):
        """coeffs [n_src + 4] = (P + M + P^T)^-1 S source_values."""
        _fields(source_values, self._n_src, "source_values")
        _fields(coeffs, self._n_src + 4, "coeffs")
        xp, ka = _ptr(source_values)
        zp, kb = _ptr(coeffs)
        check(_lib.load().dtkx_spline_solve(self._h, xp, zp))
        return coeffs

    def evaluate(self, coeffs, target_values):
        """target_values = (Q + N) coeffs."""
        _fields(coeffs, self._n_src + 4, "coeffs")
        _fields(target_values, self._n_tgt, "target_values")
        zp, ka = _ptr(coeffs)
        yp, kb = _ptr(target_values)
        check(_lib.load().dtkx_spline_evaluate(self._h, zp, yp))
        return target_values

    def system_apply(self, z, y):
        """y = (P + M + P^T) z, both [n_src + 4]."""
        _fields(z, self._n_src + 4, "z")
        _fields(y, self._n_src + 4, "y")
        zp, ka = _ptr(z)
        yp, kb = _ptr(y)
        check(_lib.load().dtkx_spline_system_apply(self._h, zp, yp))
        return y

    def basis(self, which):
        """(indices [rows, kk] int32, values [rows, kk] float64) CUDA copies
        of M (which="M": the sources' rows) or N ("N": the targets')."""
        import torch

        ip, vp = ctypes.c_void_p(), ctypes.c_void_p()
        rows, kk = ctypes.c_int64(), ctypes.c_int32()
        check(_lib.load().dtkx_spline_get_basis(
            self._h, {"M": 0, "N": 1}[which], ctypes.byref(ip), ctypes.byref(vp),
            ctypes.byref(rows), ctypes.byref(kk)))
        n = rows.value * kk.value
        idx = torch.empty((rows.value, kk.value), dtype=torch.int32, device="cuda")
        val = torch.empty((rows.value, kk.value), dtype=torch.float64, device="cuda")
        if n:
            _memcpy_d2d(idx, ip.value, 4 * n)
            _memcpy_d2d(val, vp.value, 8 * n)
        return idx, val

    def info(self) -> dict:
        """Last solve: iterations, explicit relative residual, converged."""
        s = _lib.SplineInfo()
        check(_lib.load().dtkx_spline_get_info(self._h, ctypes.byref(s)))
        return {"iterations": s.iterations, "residual": s.residual,
                "converged": bool(s.converged), "reorthogonalized": s.reorthogonalized,
                "restarts": s.restarts}


def knn(source_points, target_points, k, geometry="float", layout="aos",
        with_distances=False, stream=None):
    """k nearest sources of each target: indices [n_tgt, min(k, n_src)]
    ascending by (distance, index); optionally the distances."""
    import torch

    ns = source_points.shape[0] if layout == "aos" else source_points.shape[1]
    nt = target_points.shape[0] if layout == "aos" else target_points.shape[1]
    kk = min(k, ns)
    dev = source_points.device if hasattr(source_points, "device") else "cuda"
    idx = torch.empty((nt, kk), dtype=torch.int32, device=dev)
    dist = torch.empty((nt, kk), dtype=torch.float64, device=dev) if with_distances else None
    sp, ka = _ptr(source_points)
    tp, kb = _ptr(target_points)
    ip, kc = _ptr(idx)
    dp, kd = _ptr(dist)
    check(_lib.load().dtkx_knn(sp, ns, tp, nt, k, _lib.GEOMETRY[geometry],
                               _lib.LAYOUT[layout], ip, dp,
                               ctypes.c_void_p(stream) if stream else None))
    return (idx, dist) if with_distances else idx


def svd_pseudo_inverse(matrices, max_rotations=0, stream=None):
    """Batched pseudo-inverse of [n_mat, n, n] float64 matrices (reference
    SVDFunctor semantics). Returns (pinv, n_underdetermined, n_unconverged)."""
    import torch

    n = matrices.shape[-1]
    nm = matrices.numel() // (n * n) if hasattr(matrices, "numel") else matrices.size // (n * n)
    out = torch.empty_like(matrices) if hasattr(matrices, "data_ptr") else np.empty_like(matrices)
    ap, ka = _ptr(matrices)
    op, kb = _ptr(out)
    under = ctypes.c_int64()
    unconv = ctypes.c_int64()
    check(_lib.load().dtkx_svd_pseudo_inverse(
        n, nm, ap, op, ctypes.byref(under), ctypes.byref(unconv),
        int(max_rotations), ctypes.c_void_p(stream) if stream else None))
    return out, under.value, unconv.value


def release_cached_memory():
    """Return the library's cached device blocks (current device) to HIP."""
    check(_lib.load().dtkx_release_cached_memory())

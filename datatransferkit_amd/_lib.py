"""ctypes binding of the dtkx C ABI (include/dtkx/mls.h).

Loads ``datatransferkit_amd/libdtkx.so`` built in-tree by
``datatransferkit_amd/build.py``. There is no fallback: if the library is
missing, importing the operators raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# DTKX_LIBRARY selects another build of the same library (e.g. one built for
# a different --offload-arch); by default the in-tree build is loaded
LIB_PATH = os.environ.get("DTKX_LIBRARY") or os.path.join(HERE, "libdtkx.so")

# status codes
SUCCESS = 0
ERR_INVALID_ARGUMENT = 1
ERR_SIZE_MISMATCH = 2
ERR_DIMENSION = 3
ERR_EMPTY_SOURCE = 4
ERR_HIP = 5
ERR_NOT_CONVERGED = 6
ERR_UNSUPPORTED = 7
ERR_OUT_OF_MEMORY = 8

BASIS = {"constant": 1, "linear": 4, "quadratic": 10}
RBF = {
    "wendland0": 0, "wendland2": 1, "wendland4": 2, "wendland6": 3,
    "wu2": 4, "wu4": 5, "buhmann2": 6, "buhmann3": 7, "buhmann4": 8,
}
PRECISION = {"fp64": 0, "fp32_bulk": 1}
GEOMETRY = {"float": 0, "double": 1}
SOLVER = {"classical": 0, "cyclic": 1, "cholesky": 2}
TRUNCATION = {"absolute": 0, "relative": 1}
LAYOUT = {"aos": 0, "soa": 1}

# every symbol include/dtkx/mls.h declares
EXPORTED = [
    "dtkx_last_error", "dtkx_version", "dtkx_release_cached_memory",
    "dtkx_mls_options_init",
    "dtkx_mls_create", "dtkx_mls_create_with_neighbors", "dtkx_mls_apply",
    "dtkx_mls_apply_fp32", "dtkx_mls_apply_noalloc", "dtkx_mls_apply_transpose",
    "dtkx_mls_target_gradient", "dtkx_mls_get_graph", "dtkx_mls_write_graph",
    "dtkx_mls_get_stats",
    "dtkx_mls_destroy", "dtkx_knn", "dtkx_svd_pseudo_inverse",
    "dtkx_mls_batch_create", "dtkx_mls_batch_apply", "dtkx_mls_batch_rebuild",
    "dtkx_mls_batch_get_stats", "dtkx_mls_batch_destroy",
    "dtkx_nn_create", "dtkx_nn_apply", "dtkx_nn_get_indices", "dtkx_nn_destroy",
    "dtkx_spline_options_init", "dtkx_spline_create", "dtkx_spline_apply",
    "dtkx_spline_solve", "dtkx_spline_evaluate", "dtkx_spline_system_apply",
    "dtkx_spline_get_basis", "dtkx_spline_get_info", "dtkx_spline_destroy",
]


class Options(ctypes.Structure):
    _fields_ = [
        ("dim", ctypes.c_int32), ("basis", ctypes.c_int32),
        ("rbf", ctypes.c_int32), ("k", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("geometry", ctypes.c_int32),
        ("solver", ctypes.c_int32), ("truncation", ctypes.c_int32),
        ("rcond", ctypes.c_double), ("max_rotations", ctypes.c_int32),
        ("layout", ctypes.c_int32), ("group_size", ctypes.c_int32),
        ("store_pinv", ctypes.c_int32), ("stream", ctypes.c_void_p),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("n_source", ctypes.c_int64), ("n_target", ctypes.c_int64),
        ("nnz", ctypes.c_int64), ("basis_size", ctypes.c_int32),
        ("k", ctypes.c_int32), ("n_underdetermined", ctypes.c_int64),
        ("n_unconverged", ctypes.c_int64), ("total_rotations", ctypes.c_int64),
        ("ms_knn", ctypes.c_double), ("ms_setup", ctypes.c_double),
        ("group_size_used", ctypes.c_int32), ("n_stack_overflow", ctypes.c_int32),
        ("ms_knn_query", ctypes.c_double), ("n_eigen_fallback", ctypes.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Problem(ctypes.Structure):
    _fields_ = [
        ("src_xyz", ctypes.c_void_p), ("n_src", ctypes.c_int64),
        ("tgt_xyz", ctypes.c_void_p), ("n_tgt", ctypes.c_int64),
        ("basis", ctypes.c_int32), ("k", ctypes.c_int32),
    ]


class SplineOptions(ctypes.Structure):
    _fields_ = [
        ("rbf", ctypes.c_int32), ("geometry", ctypes.c_int32),
        ("layout", ctypes.c_int32), ("restart", ctypes.c_int32),
        ("max_iterations", ctypes.c_int32), ("reserved", ctypes.c_int32),
        ("tolerance", ctypes.c_double), ("stream", ctypes.c_void_p),
    ]


class SplineInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("reorthogonalized", ctypes.c_int32), ("restarts", ctypes.c_int32),
                ("residual", ctypes.c_double)]


class DtkxError(RuntimeError):
    """Raised for a non-zero dtkx status; .status holds the code."""

    def __init__(self, status, message):
        super().__init__(f"[status {status}] {message}")
        self.status = status


_lib = None

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_i32 = ctypes.c_int32


def load():
    """Load libdtkx.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with "
            "`python -m datatransferkit_amd.build` (no CPU fallback exists)")
    L = ctypes.CDLL(LIB_PATH)
    L.dtkx_last_error.restype = ctypes.c_char_p
    L.dtkx_version.restype = ctypes.c_char_p
    L.dtkx_release_cached_memory.argtypes = []
    L.dtkx_mls_options_init.argtypes = [ctypes.POINTER(Options)]
    L.dtkx_mls_create.argtypes = [_vp, _i64, _vp, _i64, ctypes.POINTER(Options),
                                  ctypes.POINTER(_vp)]
    L.dtkx_mls_create_with_neighbors.argtypes = [
        _vp, _i64, _vp, _i64, _vp, _i32, ctypes.POINTER(Options),
        ctypes.POINTER(_vp)]
    L.dtkx_mls_apply.argtypes = [_vp, _vp, _vp, _i32, _i64, _i64]
    L.dtkx_mls_apply_fp32.argtypes = [_vp, _vp, _vp, _i32, _i64, _i64]
    L.dtkx_mls_apply_noalloc.argtypes = [_vp, _vp, _vp, _i32, _i64, _i64]
    L.dtkx_mls_apply_transpose.argtypes = [_vp, _vp, _vp, _i32, _i64, _i64]
    L.dtkx_mls_target_gradient.argtypes = [_vp, _vp, _vp, _i32, _i64, _i64, _vp]
    L.dtkx_mls_get_graph.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                     ctypes.POINTER(_vp), ctypes.POINTER(_i64)]
    L.dtkx_mls_write_graph.argtypes = [_vp, ctypes.c_char_p, _i64]
    L.dtkx_mls_get_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
    L.dtkx_mls_destroy.argtypes = [_vp]
    L.dtkx_knn.argtypes = [_vp, _i64, _vp, _i64, _i32, _i32, _i32, _vp, _vp, _vp]
    L.dtkx_svd_pseudo_inverse.argtypes = [
        _i32, _i64, _vp, _vp, ctypes.POINTER(_i64), ctypes.POINTER(_i64), _i32, _vp]
    L.dtkx_mls_batch_create.argtypes = [_i32, ctypes.POINTER(Problem),
                                        ctypes.POINTER(Options), ctypes.POINTER(_vp)]
    L.dtkx_mls_batch_apply.argtypes = [_vp, _vp, _vp]
    L.dtkx_mls_batch_rebuild.argtypes = [_vp]
    L.dtkx_mls_batch_get_stats.argtypes = [_vp, ctypes.POINTER(Stats)]
    L.dtkx_mls_batch_destroy.argtypes = [_vp]
    L.dtkx_nn_create.argtypes = [_vp, _i64, _vp, _i64, _i32, _i32, _vp,
                                 ctypes.POINTER(_vp)]
    L.dtkx_nn_apply.argtypes = [_vp, _vp, _vp, _i32, _i64, _i64]
    L.dtkx_nn_get_indices.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_i64)]
    L.dtkx_nn_destroy.argtypes = [_vp]
    L.dtkx_spline_options_init.argtypes = [ctypes.POINTER(SplineOptions)]
    L.dtkx_spline_create.argtypes = [_vp, _i64, _vp, _i64,
                                     ctypes.POINTER(SplineOptions),
                                     ctypes.POINTER(_vp)]
    L.dtkx_spline_apply.argtypes = [_vp, _vp, _vp, _i32, _i64, _i64]
    L.dtkx_spline_solve.argtypes = [_vp, _vp, _vp]
    L.dtkx_spline_evaluate.argtypes = [_vp, _vp, _vp]
    L.dtkx_spline_system_apply.argtypes = [_vp, _vp, _vp]
    L.dtkx_spline_get_basis.argtypes = [_vp, _i32, ctypes.POINTER(_vp),
                                        ctypes.POINTER(_vp), ctypes.POINTER(_i64),
                                        ctypes.POINTER(_i32)]
    L.dtkx_spline_get_info.argtypes = [_vp, ctypes.POINTER(SplineInfo)]
    L.dtkx_spline_destroy.argtypes = [_vp]
    for name in EXPORTED:
        f = getattr(L, name)
        if f.restype is ctypes.c_int or f.restype is None:
            if name not in ("dtkx_last_error", "dtkx_version"):
                f.restype = ctypes.c_int
    _lib = L
    return L


def check(status: int) -> None:
    if status != SUCCESS:
        msg = load().dtkx_last_error().decode(errors="replace")
        raise DtkxError(status, msg)


def default_spline_options() -> SplineOptions:
    o = SplineOptions()
    check(load().dtkx_spline_options_init(ctypes.byref(o)))
    return o


def default_options() -> Options:
    o = Options()
    check(load().dtkx_mls_options_init(ctypes.byref(o)))
    return o

"""CPU oracle for the MLS transfer path -- TEST INFRASTRUCTURE ONLY.

ctypes binding of ``oracle/_build/libdtkx_oracle.so`` (built from
``oracle/dtkx_oracle.cpp`` by ``oracle/Makefile``), a C++ restatement of
DataTransferKit's serial MovingLeastSquaresOperator path. Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg import this
module, always as the checker and never as the measured or shipped path.

Function-level citations of the reference live in ``dtkx_oracle.cpp``.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libdtkx_oracle.so")

# enum values shared with include/dtkx/mls.h
BASIS = {"constant": 1, "linear": 4, "quadratic": 10}
RBF = {
    "wendland0": 0, "wendland2": 1, "wendland4": 2, "wendland6": 3,
    "wu2": 4, "wu4": 5, "buhmann2": 6, "buhmann3": 7, "buhmann4": 8,
}
GEOMETRY_FLOAT, GEOMETRY_DOUBLE = 0, 1
TRUNC_ABSOLUTE, TRUNC_RELATIVE = 0, 1
LAYOUT_AOS, LAYOUT_SOA = 0, 1


class Counters(ctypes.Structure):
    _fields_ = [("n_underdetermined", ctypes.c_int64),
                ("n_unconverged", ctypes.c_int64),
                ("total_rotations", ctypes.c_int64)]


def build() -> None:
    """Compile the oracle library (g++, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)
_i64 = ctypes.c_int64


def _declare(L):
    L.dtkx_oracle_rbf.argtypes = [ctypes.c_int, ctypes.c_double, _dp, _i64, _dp]
    L.dtkx_oracle_basis.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _i64, _dp]
    L.dtkx_oracle_svd_2x2.argtypes = [_dp, _dp, _dp]
    L.dtkx_oracle_svd_2x2.restype = None
    L.dtkx_oracle_svd_pseudo_inverse.argtypes = [
        ctypes.c_int, _i64, _dp, _dp, ctypes.c_int, ctypes.c_double,
        ctypes.c_long, ctypes.POINTER(Counters)]
    L.dtkx_oracle_knn.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int,
                                  ctypes.c_int, ctypes.c_int, _ip, _dp]
    L.dtkx_oracle_knn_brute.argtypes = [_dp, _i64, _dp, _i64, ctypes.c_int,
                                        ctypes.c_int, _ip, _dp]
    L.dtkx_oracle_mls_setup.argtypes = [
        _dp, _i64, _dp, _i64, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
        ctypes.c_long, _dp, _dp, ctypes.POINTER(Counters)]
    L.dtkx_oracle_apply.argtypes = [_i64, ctypes.c_int, _ip, _dp, _dp, _i64,
                                    _dp, _i64, ctypes.c_int]
    L.dtkx_oracle_apply_transpose.argtypes = [
        _i64, ctypes.c_int, _ip, _dp, _dp, _i64, _dp, _i64, _i64, ctypes.c_int]
    L.dtkx_oracle_mls_create.argtypes = [
        _dp, _i64, _dp, _i64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
        ctypes.c_long, _ip, _dp, ctypes.POINTER(Counters)]
    L.dtkx_oracle_basis_weights.argtypes = [
        _dp, _i64, _dp, _i64, ctypes.c_int, _ip, ctypes.c_int, ctypes.c_int,
        ctypes.c_int, _dp]
    L.dtkx_oracle_set_libm_mode.argtypes = [ctypes.c_int]
    L.dtkx_oracle_set_libm_mode.restype = None
    L.dtkx_oracle_get_libm_mode.restype = ctypes.c_int
    L.dtkx_oracle_elementary.argtypes = [ctypes.c_int, _dp, _dp, _dp, _i64]
    L.dtkx_oracle_elementary.restype = None
    L.dtkx_oracle_num_threads.restype = ctypes.c_int
    L.dtkx_oracle_set_num_threads.argtypes = [ctypes.c_int]
    L.dtkx_oracle_set_num_threads.restype = None


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle call failed with status {rc}")


LIBM_PLATFORM, LIBM_CORRECTLY_ROUNDED = 0, 1


def set_libm_mode(mode: int) -> None:
    """0: platform libm (glibc); 1: correctly rounded via __float128."""
    lib().dtkx_oracle_set_libm_mode(int(mode))


def get_libm_mode() -> int:
    return lib().dtkx_oracle_get_libm_mode()


def native_path() -> str:
    """The speed build's path, keyed by this host's CPU model and flags (a
    build for another host's CPU may use instructions this one lacks)."""
    import hashlib

    key = ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith(("model name", "flags")):
                    key += line
                if line.strip() == "":
                    break
    except OSError:
        pass
    tag = hashlib.sha1(key.encode()).hexdigest()[:12]
    return os.path.join(HERE, "_build", f"libdtkx_oracle_native_{tag}.so")


class speed_build:
    """Context manager routing every oracle call to the -march=native,
    FP-contracting build (built on first use on this host). For timing the
    CPU baseline only: it rounds differently, so it is never a checker."""

    def __enter__(self):
        global _lib
        path = native_path()
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", HERE, "native", f"NATIVE_OUT={path}"],
                           check=True)
        self.prev = lib()
        _lib = ctypes.CDLL(path)
        _declare(_lib)
        return self

    def __exit__(self, *exc):
        global _lib
        _lib = self.prev


class libm_mode:
    """Context manager selecting the oracle's elementary-function mode."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.prev = get_libm_mode()
        set_libm_mode(self.mode)

    def __exit__(self, *exc):
        set_libm_mode(self.prev)


def elementary(op: str, x, y=None) -> np.ndarray:
    """atan2(y, x), sin(x), cos(x) or log(x) as the oracle evaluates them."""
    x = _f64(x)
    y = _f64(y) if y is not None else x
    out = np.empty_like(x)
    code = {"atan2": 0, "sin": 1, "cos": 2, "log": 3}[op]
    lib().dtkx_oracle_elementary(code, _d(y), _d(x), _d(out), x.size)
    return out


def num_threads() -> int:
    return lib().dtkx_oracle_num_threads()


def set_num_threads(n: int) -> None:
    lib().dtkx_oracle_set_num_threads(int(n))


def rbf(name: str, radius: float, r) -> np.ndarray:
    r = _f64(r)
    out = np.empty_like(r)
    _check(lib().dtkx_oracle_rbf(RBF[name], radius, _d(r), r.size, _d(out)))
    return out


def basis(m: int, pts, dim: int = 3, float_point: bool = True) -> np.ndarray:
    pts = _f64(pts).reshape(-1, dim)
    out = np.empty((pts.shape[0], m))
    _check(lib().dtkx_oracle_basis(m, dim, int(float_point), _d(pts),
                                   pts.shape[0], _d(out)))
    return out


def svd_2x2(a):
    a = _f64(a).reshape(2, 2)
    L = np.empty((2, 2))
    R = np.empty((2, 2))
    lib().dtkx_oracle_svd_2x2(_d(a), _d(L), _d(R))
    return L, R


def svd_pseudo_inverse(mats, truncation=TRUNC_ABSOLUTE, rcond=0.0, max_rot=0):
    mats = _f64(mats)
    n = mats.shape[-1]
    out = np.empty_like(mats)
    c = Counters()
    nm = mats.size // (n * n)
    _check(lib().dtkx_oracle_svd_pseudo_inverse(
        n, nm, _d(mats), _d(out), truncation, rcond, max_rot, ctypes.byref(c)))
    return out, c


def knn(src, tgt, k, geometry=GEOMETRY_FLOAT, brute=False, layout=LAYOUT_AOS):
    src = _f64(src)
    tgt = _f64(tgt)
    ns = src.size // 3
    nt = tgt.size // 3
    kk = min(k, ns)
    idx = np.empty((nt, kk), dtype=np.int32)
    dist = np.empty((nt, kk))
    if brute:
        assert layout == LAYOUT_AOS
        _check(lib().dtkx_oracle_knn_brute(_d(src), ns, _d(tgt), nt, k,
                                           geometry, _i(idx), _d(dist)))
    else:
        _check(lib().dtkx_oracle_knn(_d(src), ns, _d(tgt), nt, k, geometry,
                                     layout, _i(idx), _d(dist)))
    return idx, dist


def mls_setup(src, tgt, indices, m, rbf_name="wendland0",
              geometry=GEOMETRY_FLOAT, truncation=TRUNC_ABSOLUTE, rcond=0.0,
              max_rot=0, layout=LAYOUT_AOS, want_a0=False):
    src = _f64(src)
    tgt = _f64(tgt)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    ns = src.size // 3
    nt = tgt.size // 3
    kk = indices.shape[1] if indices.ndim == 2 else (indices.size // max(nt, 1))
    coeffs = np.empty((nt, kk))
    a0 = np.empty((nt, m)) if want_a0 else None
    c = Counters()
    _check(lib().dtkx_oracle_mls_setup(
        _d(src), ns, _d(tgt), nt, layout, _i(indices), kk, m, RBF[rbf_name],
        geometry, truncation, rcond, max_rot, _d(coeffs),
        _d(a0) if want_a0 else None, ctypes.byref(c)))
    if want_a0:
        return coeffs, c, a0
    return coeffs, c


def basis_weights(src, tgt, indices, rbf_name="wendland0",
                  geometry=GEOMETRY_FLOAT, layout=LAYOUT_AOS):
    """Values of the spline's basis operator rows: phi [nt, kk] for the
    given neighbour rows (DTK_SplineOperator_def.hpp:91-109)."""
    src = _f64(src)
    tgt = _f64(tgt)
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    nt = tgt.size // 3
    kk = indices.shape[1] if indices.ndim == 2 else 0
    phi = np.empty((nt, kk))
    _check(lib().dtkx_oracle_basis_weights(
        _d(src), src.size // 3, _d(tgt), nt, layout, _i(indices), kk,
        RBF[rbf_name], geometry, _d(phi)))
    return phi


def mls_create(src, tgt, m, k=0, rbf_name="wendland0", geometry=GEOMETRY_FLOAT,
               truncation=TRUNC_ABSOLUTE, rcond=0.0, max_rot=0,
               layout=LAYOUT_AOS):
    """Full constructor: kNN + setup. Returns (indices [nt,kk], coeffs, counters)."""
    src = _f64(src)
    tgt = _f64(tgt)
    ns = src.size // 3
    nt = tgt.size // 3
    kk = min(k if k > 0 else m, ns)
    idx = np.empty((nt, kk), dtype=np.int32)
    coeffs = np.empty((nt, kk))
    c = Counters()
    _check(lib().dtkx_oracle_mls_create(
        _d(src), ns, _d(tgt), nt, layout, k, m, RBF[rbf_name], geometry,
        truncation, rcond, max_rot, _i(idx), _d(coeffs), ctypes.byref(c)))
    return idx, coeffs, c


def apply(indices, coeffs, x):
    """y[f, i] = sum_j coeffs[i, j] * x[f, indices[i, j]] (fields first)."""
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    coeffs = _f64(coeffs)
    x = _f64(x)
    xf = x.reshape(-1, x.shape[-1])
    nt, kk = indices.shape
    y = np.empty((xf.shape[0], nt))
    _check(lib().dtkx_oracle_apply(nt, kk, _i(indices), _d(coeffs), _d(xf),
                                   xf.shape[1], _d(y), nt, xf.shape[0]))
    return y.reshape(x.shape[:-1] + (nt,))


def apply_transpose(indices, coeffs, ybar, ns):
    indices = np.ascontiguousarray(indices, dtype=np.int32)
    coeffs = _f64(coeffs)
    ybar = _f64(ybar)
    yf = ybar.reshape(-1, ybar.shape[-1])
    nt, kk = indices.shape
    xbar = np.empty((yf.shape[0], ns))
    _check(lib().dtkx_oracle_apply_transpose(
        nt, kk, _i(indices), _d(coeffs), _d(yf), nt, _d(xbar), ns, ns,
        yf.shape[0]))
    return xbar.reshape(ybar.shape[:-1] + (ns,))

"""Pins datatransferkit_amd/csrc/crmath.hpp (the GPU kernels' correctly
rounded atan2/sin/cos/log) on the CPU: the same header is compiled for the host
and compared bit-for-bit with the oracle's independent correctly rounded
mode (libquadmath __float128, rounded once) on millions of arguments,
including the special values of C99 atan2."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO

_dp = ctypes.POINTER(ctypes.c_double)


@pytest.fixture(scope="module")
def crm(tmp_path_factory):
    out = tmp_path_factory.mktemp("crm") / "libcrm_host.so"
    src = os.path.join(REPO, "tests", "native", "crmath_host.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-shared",
                    "-fPIC", "-o", str(out), src], check=True)
    return ctypes.CDLL(str(out))


def _call(lib, fn, *arrs):
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]
    out = np.empty_like(arrs[0])
    getattr(lib, fn)(*[a.ctypes.data_as(_dp) for a in arrs], out.ctypes.data_as(_dp),
                     ctypes.c_int64(out.size))
    return out


def _same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64)) or \
        bool(np.all((a == b) & (np.signbit(a) == np.signbit(b)) | (np.isnan(a) & np.isnan(b))))


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_sin_cos_match_quadmath(crm, oracle_lib, fn):
    rng = np.random.default_rng(1)
    x = np.concatenate([
        rng.uniform(-np.pi / 2, np.pi / 2, 1_000_000),
        rng.uniform(-50, 50, 100_000),
        np.pi / 2 + rng.integers(-64, 64, 2000) * 2.0 ** -52,
        rng.uniform(-1e-6, 1e-6, 10_000),
        [0.0, -0.0, np.pi / 2, -np.pi / 2, np.pi / 4, 1e-300, -5e-324],
    ])
    got = _call(crm, "crm_" + fn, x)
    with oracle_lib.libm_mode(1):
        ref = oracle_lib.elementary(fn, x)
    bad = ~((got == ref) & (np.signbit(got) == np.signbit(ref)))
    assert bad.sum() == 0, (x[bad][:5], got[bad][:5], ref[bad][:5])


def test_sincos_equals_sin_and_cos(crm):
    rng = np.random.default_rng(4)
    x = np.concatenate([rng.uniform(-np.pi / 2, np.pi / 2, 200_000), rng.uniform(-40, 40, 20_000),
                        [0.0, -0.0, 1e-300, -1e-9, np.pi / 2, -np.pi / 2]])
    s = np.empty_like(x)
    c = np.empty_like(x)
    crm.crm_sincos(x.ctypes.data_as(_dp), s.ctypes.data_as(_dp), c.ctypes.data_as(_dp),
                   ctypes.c_int64(x.size))
    assert np.array_equal(s.view(np.uint64), _call(crm, "crm_sin", x).view(np.uint64))
    assert np.array_equal(c.view(np.uint64), _call(crm, "crm_cos", x).view(np.uint64))


def test_atan2_matches_quadmath(crm, oracle_lib):
    rng = np.random.default_rng(2)
    n = 1_000_000
    y = rng.standard_normal(n) * 10.0 ** rng.uniform(-12, 12, n)
    x = rng.standard_normal(n) * 10.0 ** rng.uniform(-12, 12, n)
    y2 = rng.standard_normal(20000)
    x2 = y2 * (1 + rng.uniform(-1e-6, 1e-6, 20000))  # near the diagonals
    sp = [0.0, -0.0, 1.0, -1.0, np.inf, -np.inf, 1e-310, -1e-310, 1e300, -1e300]
    ys = np.array([a for a in sp for b in sp])
    xs = np.array([b for a in sp for b in sp])
    Y = np.concatenate([y, y2, -y2, ys, [1e-300, 5e-324, 1.0]])
    X = np.concatenate([x, x2, -x2, xs, [1e300, -1.0, 5e-324]])
    got = _call(crm, "crm_atan2", Y, X)
    with oracle_lib.libm_mode(1):
        ref = oracle_lib.elementary("atan2", X, Y)
    bad = ~(((got == ref) & (np.signbit(got) == np.signbit(ref))) | (np.isnan(got) & np.isnan(ref)))
    assert bad.sum() == 0, (Y[bad][:5], X[bad][:5], got[bad][:5], ref[bad][:5])


def test_log_matches_quadmath(crm, oracle_lib):
    rng = np.random.default_rng(5)
    x = np.concatenate([
        rng.uniform(0.0, 1.0, 1_000_000),             # the RBF's x = d / R
        10.0 ** rng.uniform(-307, 307, 200_000),
        1.0 + rng.uniform(-1e-6, 1e-6, 20_000),       # near the zero of log
        np.sqrt(0.5) * (1 + rng.integers(-64, 64, 2000) * 2.0 ** -52),
        [1.0, 2.0, 0.5, 4.0, 1 - 2 ** -53, 1 + 2 ** -52, 5e-324, 2.2e-308,
         1.7e308, np.inf, 0.0, -0.0, -1.0, np.nan],
    ])
    got = _call(crm, "crm_log", x)
    with oracle_lib.libm_mode(1):
        ref = oracle_lib.elementary("log", x)
    bad = ~(((got == ref) & (np.signbit(got) == np.signbit(ref))) | (np.isnan(got) & np.isnan(ref)))
    assert bad.sum() == 0, (x[bad][:5], got[bad][:5], ref[bad][:5])


def test_glibc_is_not_correctly_rounded(crm, oracle_lib):
    # documents why parity is stated against the correctly rounded mode:
    # the platform libm differs from the correctly rounded value on a
    # fraction of arguments, crmath on none
    rng = np.random.default_rng(3)
    x = rng.uniform(-np.pi / 2, np.pi / 2, 200_000)
    glibc = _call(crm, "libm_sin", x)
    with oracle_lib.libm_mode(1):
        ref = oracle_lib.elementary("sin", x)
    assert (glibc != ref).sum() > 0
    assert (_call(crm, "crm_sin", x) != ref).sum() == 0

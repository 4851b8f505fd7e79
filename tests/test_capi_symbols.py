"""The C ABI library loads and exports every function include/dtkx/mls.h
declares (no compute call: runs without a GPU). Also checks the ctypes
structures match the header's layout and the option defaults."""
import ctypes
import os
import re
import subprocess

import pytest

from conftest import REPO

HEADER = os.path.join(REPO, "include", "dtkx", "mls.h")
LIB = os.path.join(REPO, "datatransferkit_amd", "libdtkx.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dtkx_\w+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_every_declared_symbol_is_exported(lib):
    names = declared_functions()
    assert len(names) >= 18
    for n in names:
        assert hasattr(lib, n), n
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (dtkx_\w+)", out))
    assert set(names) <= exported


def test_python_binding_lists_the_header(lib):
    from datatransferkit_amd import _lib

    assert sorted(_lib.EXPORTED) == declared_functions()


def test_struct_layouts_match_header(tmp_path):
    # compile a C program against the header and print the struct layout
    src = tmp_path / "layout.c"
    src.write_text(r"""
#include <stddef.h>
#include <stdio.h>
#include "dtkx/mls.h"
int main(void) {
  printf("%zu %zu %zu %zu\n", sizeof(dtkx_mls_options), offsetof(dtkx_mls_options, rcond),
         offsetof(dtkx_mls_options, stream), sizeof(dtkx_mls_stats));
  printf("%zu %zu\n", offsetof(dtkx_mls_stats, ms_knn), sizeof(dtkx_mls_problem));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", str(exe), str(src)],
                   check=True)
    a = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    from datatransferkit_amd import _lib

    assert int(a[0]) == ctypes.sizeof(_lib.Options)
    assert int(a[1]) == _lib.Options.rcond.offset
    assert int(a[2]) == _lib.Options.stream.offset
    assert int(a[3]) == ctypes.sizeof(_lib.Stats)
    assert int(a[4]) == _lib.Stats.ms_knn.offset
    assert int(a[5]) == ctypes.sizeof(_lib.Problem)


def test_batch_problem_records_read_back_through_ctypes():
    # the batch wrapper fills dtkx_mls_problem records as a numpy structured
    # array: every field must land where the ctypes struct (checked against
    # the header above) reads it
    import numpy as np
    from datatransferkit_amd import _lib, operators

    n = 3
    rec = np.zeros(n, dtype=operators._PROBLEM_DTYPE)
    rec["src_xyz"] = [0x1000, 0x7FFF_FFFF_F000, 0xDEAD_BEE0]
    rec["n_src"] = [1, 2**40, 7]
    rec["tgt_xyz"] = [0x2000, 0x3000, 0xFFFF_FFFF_FFFF_FFF0]
    rec["n_tgt"] = [5, 6, 2**33]
    rec["basis"] = [1, 4, 10]
    rec["k"] = [-1, 32, 2**31 - 1]
    arr = (_lib.Problem * n).from_buffer(rec)
    assert operators._PROBLEM_DTYPE.itemsize == ctypes.sizeof(_lib.Problem)
    for i in range(n):
        p = arr[i]
        assert (p.src_xyz, p.n_src, p.tgt_xyz, p.n_tgt, p.basis, p.k) == tuple(
            int(rec[f][i]) for f in ("src_xyz", "n_src", "tgt_xyz", "n_tgt", "basis", "k"))


def test_options_defaults_are_reference_semantics(lib):
    from datatransferkit_amd import _lib

    o = _lib.default_options()
    assert (o.dim, o.basis, o.rbf, o.k) == (3, 4, 0, 0)  # Wendland<0>, Linear3, k = m
    assert (o.precision, o.geometry, o.solver, o.truncation) == (0, 0, 0, 0)


def test_error_path_without_gpu(lib):
    # argument validation fails before any device call
    from datatransferkit_amd import _lib

    o = _lib.default_options()
    o.dim = 2
    h = ctypes.c_void_p()
    L = _lib.load()
    rc = L.dtkx_mls_create(None, 0, None, 0, ctypes.byref(o), ctypes.byref(h))
    assert rc == _lib.ERR_DIMENSION
    assert "o.dim == 3" in L.dtkx_last_error().decode()
    o.dim = 3
    rc = L.dtkx_mls_create(None, 0, None, 0, ctypes.byref(o), ctypes.byref(h))
    assert rc == _lib.ERR_EMPTY_SOURCE
    assert ", failed in capi.hip, line" in L.dtkx_last_error().decode()


def test_buhmann4_is_refused(lib):
    # Buhmann<4> as the reference writes it is -67.9 at the edge of its
    # support (DTK_CompactlySupportedRadialBasisFunctions.hpp:139-146)
    from datatransferkit_amd import _lib

    L = _lib.load()
    h = ctypes.c_void_p()
    o = _lib.default_options()
    o.rbf = _lib.RBF["buhmann4"]
    rc = L.dtkx_mls_create(None, 0, None, 0, ctypes.byref(o), ctypes.byref(h))
    assert rc == _lib.ERR_UNSUPPORTED
    assert "DTKX_RBF_BUHMANN4" in L.dtkx_last_error().decode()
    o = _lib.default_spline_options()
    o.rbf = _lib.RBF["buhmann4"]
    assert L.dtkx_spline_create(None, 0, None, 0, ctypes.byref(o),
                                ctypes.byref(h)) == _lib.ERR_UNSUPPORTED
    o = _lib.default_options()
    o.rbf = _lib.RBF["buhmann3"]
    rc = L.dtkx_mls_create(None, 0, None, 0, ctypes.byref(o), ctypes.byref(h))
    assert rc == _lib.ERR_EMPTY_SOURCE


def test_spline_structs_and_defaults(tmp_path, lib):
    """dtkx_spline_options / dtkx_spline_info: the ctypes layouts match the
    header, and the defaults are the reference's (Wendland<0>, float
    points, Belos tolerance 1e-10, DTK_SplineOperator_def.hpp:258)."""
    src = tmp_path / "spl.c"
    src.write_text(r"""
#include <stddef.h>
#include <stdio.h>
#include "dtkx/mls.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu\n", sizeof(dtkx_spline_options),
         offsetof(dtkx_spline_options, tolerance),
         offsetof(dtkx_spline_options, stream), sizeof(dtkx_spline_info),
         offsetof(dtkx_spline_info, residual));
  return 0;
}
""")
    exe = tmp_path / "spl"
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), "-o", str(exe), str(src)],
                   check=True)
    a = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True,
                                        check=True).stdout.split()]
    from datatransferkit_amd import _lib

    assert a == [ctypes.sizeof(_lib.SplineOptions), _lib.SplineOptions.tolerance.offset,
                 _lib.SplineOptions.stream.offset, ctypes.sizeof(_lib.SplineInfo),
                 _lib.SplineInfo.residual.offset]
    o = _lib.default_spline_options()
    assert (o.rbf, o.geometry, o.layout) == (0, 0, 0)
    assert (o.restart, o.max_iterations, o.tolerance) == (300, 1000, 1e-10)


def test_spline_error_paths_without_gpu(lib):
    # option and size checks fail before any device call
    from datatransferkit_amd import _lib

    L = _lib.load()
    h = ctypes.c_void_p()
    o = _lib.default_spline_options()
    o.restart = 0
    assert L.dtkx_spline_create(None, 0, None, 0, ctypes.byref(o),
                                ctypes.byref(h)) == _lib.ERR_INVALID_ARGUMENT
    o = _lib.default_spline_options()
    o.tolerance = float("nan")
    assert L.dtkx_spline_create(None, 0, None, 0, ctypes.byref(o),
                                ctypes.byref(h)) == _lib.ERR_INVALID_ARGUMENT
    o = _lib.default_spline_options()
    assert L.dtkx_spline_create(None, 0, None, 0, ctypes.byref(o),
                                ctypes.byref(h)) == _lib.ERR_EMPTY_SOURCE
    assert "n_src > 0" in L.dtkx_last_error().decode()
    assert L.dtkx_spline_apply(None, None, None, 1, 0, 0) == _lib.ERR_INVALID_ARGUMENT

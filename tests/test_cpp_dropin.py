"""The C++ drop-ins (include/dtkx/MovingLeastSquaresOperator.hpp,
NearestNeighborOperator.hpp, SplineOperator.hpp) compile with a plain C++17
compiler against the C ABI, and on the GPU pass the reference's own operator
tests restated in tests/native/test_mls_cpp.cpp (MLS and Spline) and
test_nn_cpp.cpp. The batch wrapper (MovingLeastSquaresBatch.hpp) is checked
bit for bit against single operators by tests/native/test_batch_cpp.cpp,
which allocates device memory through the HIP runtime."""
import os
import subprocess

import pytest

from conftest import REPO

SRC = os.path.join(REPO, "tests", "native", "test_mls_cpp.cpp")
SRC_NN = os.path.join(REPO, "tests", "native", "test_nn_cpp.cpp")
SRC_BATCH = os.path.join(REPO, "tests", "native", "test_batch_cpp.cpp")
LIBDIR = os.path.join(REPO, "datatransferkit_amd")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
# the batch test's own device allocations go through the HIP runtime
HIP_ARGS = ["-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROCM, "include"),
            "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
            f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]


def _build(tmp_path, src=SRC, name="test_mls_cpp", extra=()):
    exe = tmp_path / name
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-I",
                    os.path.join(REPO, "include"), src, "-L", LIBDIR, "-ldtkx",
                    f"-Wl,-rpath,{LIBDIR}", *extra, "-o", str(exe)], check=True)
    return exe


def test_dropin_header_compiles_and_links(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()
    assert _build(tmp_path, SRC_NN, "test_nn_cpp").exists()
    assert _build(tmp_path, SRC_BATCH, "test_batch_cpp", HIP_ARGS).exists()


@pytest.mark.gpu
def test_dropin_batch_equals_single_operators(tmp_path):
    exe = _build(tmp_path, SRC_BATCH, "test_batch_cpp", HIP_ARGS)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 14


@pytest.mark.gpu
def test_dropin_nearest_neighbor_reference_tests(tmp_path):
    exe = _build(tmp_path, SRC_NN, "test_nn_cpp")
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 5


@pytest.mark.gpu
def test_dropin_reference_operator_tests(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([str(exe), os.path.join(REPO, "tests", "golden")],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
    assert r.stdout.count("PASS") == 20

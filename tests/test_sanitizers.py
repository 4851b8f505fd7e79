"""The sanitizer gate of the reference's tests (a test fails if its output
matches "data race;leak;runtime error",
packages/Meshfree/test/CMakeLists.txt:10), applied to the CPU code here:
the oracle and the host build of crmath.hpp compiled with
AddressSanitizer (leaks included) and UndefinedBehaviorSanitizer, driven
over every exported oracle function (tests/native/sanitize_oracle.cpp).
ThreadSanitizer is not run: libgomp is not instrumented, so OpenMP code
reports false races; the oracle's parallel loops write disjoint rows."""
import os
import re
import shutil
import subprocess

import pytest

from conftest import REPO

FAIL = re.compile(r"data race|leak|runtime error")


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_oracle_under_asan_ubsan(tmp_path):
    exe = tmp_path / "sanitize_oracle"
    cmd = ["g++", "-O1", "-g", "-std=c++17", "-fopenmp", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-fno-omit-frame-pointer", "-I", os.path.join(REPO, "include"),
           os.path.join(REPO, "oracle", "dtkx_oracle.cpp"),
           os.path.join(REPO, "tests", "native", "sanitize_oracle.cpp"),
           "-lquadmath", "-o", str(exe)]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe)], capture_output=True, text=True, env=env, timeout=300)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert not FAIL.search(out), out[-4000:]
    assert "sanitize driver: ok" in out

"""The golden fixtures are exactly what their generator writes: the
reference tests' random inputs regenerated with libstdc++'s
default_random_engine (tests/golden/gen_reference_inputs.cpp)."""
import os
import subprocess

import numpy as np

from conftest import GOLDEN

NAMES = ["svd_full_rank", "svd_rank_deficient"] + [
    f"mls_points_m{m}_{w}" for m in (1, 4, 10) for w in ("src", "tgt")]


def test_fixtures_regenerate_bit_for_bit(tmp_path):
    exe = tmp_path / "gen"
    subprocess.run(["g++", "-O0", "-std=c++17", "-o", str(exe),
                    os.path.join(GOLDEN, "gen_reference_inputs.cpp")], check=True)
    subprocess.run([str(exe), str(tmp_path)], check=True)
    for n in NAMES:
        a = np.fromfile(os.path.join(GOLDEN, n + ".f64"), dtype="<f8")
        b = np.fromfile(os.path.join(tmp_path, n + ".f64"), dtype="<f8")
        assert a.size > 0 and np.array_equal(a.view(np.uint64), b.view(np.uint64)), n


def test_fixture_shapes_and_ranges():
    a = np.fromfile(os.path.join(GOLDEN, "svd_full_rank.f64"))
    assert a.size == 10 * 32 * 32 and np.abs(a).max() <= 1000
    b = np.fromfile(os.path.join(GOLDEN, "svd_rank_deficient.f64")).reshape(10, 40, 40)
    assert np.all(b[:, :, 3] == 0) and np.count_nonzero(b) == 10 * 40 * 39
    s = np.fromfile(os.path.join(GOLDEN, "mls_points_m10_src.f64")).reshape(10, 10, 3)
    t = np.fromfile(os.path.join(GOLDEN, "mls_points_m10_tgt.f64")).reshape(10, 3)
    centers = 2.0 * np.arange(10)
    # sources within radius 0.5 of (2i, 2i, 2i), targets within 0.25
    assert np.all(np.abs(s - centers[:, None, None]) <= 0.5)
    assert np.all(np.abs(t - centers[:, None]) <= 0.25)

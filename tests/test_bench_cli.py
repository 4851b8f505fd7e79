"""The benchmark's command line on the CPU: one GPU only, and its
algorithmic byte models stay what they state."""
import json
import os
import subprocess
import sys

from conftest import REPO


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_spline_solve_bytes_model():
    sys.path.insert(0, REPO)
    import bench

    n_src, restart = 1000, 300
    # one iteration at basis size nv: the system's operator pass
    # (80 bytes per source) and the two basis sweeps of the delayed
    # Gram-Schmidt step over the n_src + 4 entries
    info = {"iterations": 1, "restarts": 0}
    n = n_src + 4
    assert bench.spline_solve_bytes(n_src, info, restart) >= 80 * n_src
    two = bench.spline_solve_bytes(n_src, {"iterations": 2, "restarts": 0}, restart)
    one = bench.spline_solve_bytes(n_src, info, restart)
    # each further iteration sweeps one more basis vector twice
    three = bench.spline_solve_bytes(n_src, {"iterations": 3, "restarts": 0}, restart)
    assert (three - two) - (two - one) == 16 * n

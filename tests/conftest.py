import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)

GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


def load_golden(name, shape=None):
    a = np.fromfile(os.path.join(GOLDEN, name + ".f64"), dtype="<f8")
    return a.reshape(shape) if shape is not None else a


# Test functions of the reference's MLS tests
# (packages/Meshfree/test/tstMeshfreeOperators.cpp:180-199).
def poly_for_basis(m):
    if m == 1:
        return lambda p: 3.0 + 0.0 * p[:, 0]
    if m == 4:
        return lambda p: 4 + 2 * p[:, 0] + 3 * p[:, 1] - 2 * p[:, 2]
    return lambda p: (2 + 3 * p[:, 0] - 5 * p[:, 1] + 2 * p[:, 2] + 3 * p[:, 0] * p[:, 0]
                      + 4 * p[:, 0] * p[:, 1] - 2 * p[:, 0] * p[:, 2] + p[:, 1] * p[:, 1]
                      - 3 * p[:, 1] * p[:, 2] + 4 * p[:, 2] * p[:, 2])


def grid_points(n, offset):
    """makeGridPoints (tstMeshfreeOperators.cpp:105-127): i, j, k nested."""
    return np.array([[offset[0] + i, offset[1] + j, offset[2] + k]
                     for i in range(n[0]) for j in range(n[1]) for k in range(n[2])],
                    dtype=np.float64)


def c1_inputs(n=10_000, seed=20240101):
    """Plain FP64 MLS workload: n sources and n targets i.i.d. uniform in
    [0,1]^3 (SplitMix64-seeded numpy generator), f = sin(2 pi x) cos(2 pi y) + z^2."""
    rng = np.random.default_rng(seed)
    src = rng.random((n, 3))
    tgt = rng.random((n, 3))
    f = np.sin(2 * np.pi * src[:, 0]) * np.cos(2 * np.pi * src[:, 1]) + src[:, 2] ** 2
    return src, tgt, f


@pytest.fixture(scope="session")
def oracle_lib():
    import oracle

    oracle.build()
    return oracle

import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if REPO not in sys.path:
    sys.path.insert(0, REPO)
GOLDEN = os.path.join(REPO, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libtrtt_hip.so)")


@pytest.fixture(scope="session")
def oracle():
    from oracle.oracle import Oracle, build

    build()
    return Oracle()


@pytest.fixture(scope="session")
def golden3d():
    return dict(np.load(os.path.join(GOLDEN, "proj3d.npz")))


@pytest.fixture(scope="session")
def golden2d():
    return dict(np.load(os.path.join(GOLDEN, "proj2d.npz")))


@pytest.fixture(scope="session")
def golden_sparse2d():
    return dict(np.load(os.path.join(GOLDEN, "sparse2d.npz")))


@pytest.fixture(scope="session")
def golden_xform():
    return dict(np.load(os.path.join(GOLDEN, "xform.npz")))


@pytest.fixture(scope="session")
def golden_footprint():
    return dict(np.load(os.path.join(GOLDEN, "footprint.npz")))


def rel_err(a, b):
    """max |a-b| / max |b| (the parity measure used throughout the tests)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b).max() if a.size else 0.0
    s = np.abs(b).max() if b.size else 0.0
    return d / s if s > 0 else d


CASES3D = ["cone_a", "cone_edge", "cone_off", "par_a", "par_edge"]
CASES2D = ["fan_a", "fan_edge", "par_a", "par_edge"]

import os
import sys

import numpy
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built hipng library")


@pytest.fixture(scope="session")
def small_obs():
    """C0-like geometry (100 stations, 1 km disc, dec -45 deg), subsampled for brute force."""
    from rascil_amd import simulation

    bvis = simulation.config0(seed=0)
    uvw = bvis.uvw.reshape(-1, 3)
    rng = numpy.random.default_rng(5)
    rows = numpy.sort(rng.choice(uvw.shape[0], 1500, replace=False))
    return uvw[rows].copy(), bvis


def rel_l2(a, b):
    return float(numpy.linalg.norm(numpy.ravel(a - b)) / numpy.linalg.norm(numpy.ravel(b)))


def rel_max(a, b):
    return float(numpy.max(numpy.abs(a - b)) / numpy.max(numpy.abs(b)))

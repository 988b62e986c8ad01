"""CPU port of the algorithm (oracle/wstack_cpu.py) against the brute-force sums: it is
the step-by-step checker of the GPU kernels and the bench's CPU baseline, so it is held
to epsilon here first."""

import numpy
import pytest

from conftest import rel_l2
from oracle import explicit as ox
from oracle import wstack_cpu
from rascil_amd import wgridder

C = 299792458.0
FREQ2 = numpy.array([1.0e8, 1.31e8])


def _plan(uvw, freq, nx, ny, px, py, eps, wst, nb=1):
    ul = numpy.abs(uvw).max(axis=0) * max(freq) / C
    wmin = numpy.abs(uvw[:, 2]).min() * min(freq) / C
    return wgridder.plan_params(ul[0], ul[1], wmin, ul[2], nx, ny, px, py, eps, wst, "fp64", nb).as_dict()


@pytest.mark.parametrize("wst", [False, True])
@pytest.mark.parametrize("eps", [1e-4, 1e-9])
def test_cpu_port_ms2dirty(small_obs, wst, eps):
    uvw, _ = small_obs
    uvw = uvw[:600]
    rng = numpy.random.default_rng(0)
    ms = rng.normal(size=(600, 2, 2)) + 1j * rng.normal(size=(600, 2, 2))
    wgt = rng.uniform(0.5, 1.5, (600, 2, 2))
    plan = _plan(uvw, FREQ2, 48, 40, 8e-4, 9e-4, eps, wst, 2)
    d = wstack_cpu.cpu_ms2dirty(uvw, FREQ2, ms, wgt, plan)
    for b in range(2):
        ref = ox.explicit_ms2dirty(uvw, FREQ2, ms[..., b], wgt[..., b], 48, 40, 8e-4, 9e-4, wst)
        assert rel_l2(d[b], ref) < eps


@pytest.mark.parametrize("wst", [False, True])
@pytest.mark.parametrize("eps", [1e-4, 1e-9])
def test_cpu_port_dirty2ms(small_obs, wst, eps):
    uvw, _ = small_obs
    uvw = uvw[:600]
    rng = numpy.random.default_rng(1)
    dirty = rng.normal(size=(2, 48, 40))
    wgt = rng.uniform(0.5, 1.5, (600, 2, 2))
    plan = _plan(uvw, FREQ2, 48, 40, 8e-4, 9e-4, eps, wst, 2)
    v = wstack_cpu.cpu_dirty2ms(uvw, FREQ2, dirty, wgt, plan)
    for b in range(2):
        ref = ox.explicit_dirty2ms(uvw, FREQ2, dirty[b], wgt[..., b], 8e-4, 9e-4, wst)
        assert rel_l2(v[..., b], ref) < eps

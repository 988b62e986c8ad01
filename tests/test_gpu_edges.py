"""Degenerate geometries through the C-ABI against the FP64 brute-force sums (oracle): one
visibility, one row over many channels, one channel over many rows, and every visibility at
one w (a plane range of zero width, with w-stacking on). Each reaches the bucket sort, the
run detection and the plane bookkeeping at an extreme."""

import numpy
import pytest

from conftest import rel_l2
from oracle import explicit as ox
from rascil_amd import wgridder

pytestmark = pytest.mark.gpu

MODES = [("fp64", 1e-7, 1e-7), ("mixed", 1e-5, 1e-5)]  # precision, epsilon, tolerance (relative L2)
NX, NY, PX, PY = 64, 48, 1e-3, 1.2e-3


def _check(uvw, freq, prec, eps, tol, wst, seed=0):
    rng = numpy.random.default_rng(seed)
    shape = (uvw.shape[0], len(freq))
    ms = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    wgt = rng.uniform(0.5, 1.5, shape)
    dirty = rng.normal(size=(NX, NY))
    msi, wgi, di = (ms, wgt, dirty) if prec == "fp64" else (ms.astype("c8"), wgt.astype("f4"), dirty.astype("f4"))
    d = wgridder.ms2dirty(uvw, freq, msi, wgi, NX, NY, PX, PY, 0, 0, eps, wst, precision=prec)
    ref = ox.explicit_ms2dirty(uvw, freq, ms, wgt, NX, NY, PX, PY, wst)
    ei = rel_l2(d, ref)
    v = wgridder.dirty2ms(uvw, freq, di, None, PX, PY, 0, 0, eps, wst, precision=prec)
    vref = ox.explicit_dirty2ms(uvw, freq, dirty, None, PX, PY, wst)
    ep = rel_l2(v, vref)
    print(f"{prec} eps={eps} wst={wst} rows={uvw.shape[0]} chans={len(freq)}: invert {ei:.2e} predict {ep:.2e}")
    assert ei <= tol and ep <= tol


@pytest.mark.parametrize("prec,eps,tol", MODES)
@pytest.mark.parametrize("wst", [False, True])
def test_single_visibility(prec, eps, tol, wst):
    _check(numpy.array([[123.4, -56.7, 89.1]]), numpy.array([1.3e8]), prec, eps, tol, wst)


@pytest.mark.parametrize("prec,eps,tol", MODES)
def test_one_row_many_channels(prec, eps, tol):
    _check(numpy.array([[300.0, -200.0, 50.0]]), numpy.linspace(1.0e8, 1.6e8, 64), prec, eps, tol, True)


@pytest.mark.parametrize("prec,eps,tol", MODES)
def test_one_channel_many_rows(prec, eps, tol):
    uvw = numpy.random.default_rng(3).normal(0.0, 150.0, (500, 3))
    _check(uvw, numpy.array([1.1e8]), prec, eps, tol, True)


@pytest.mark.parametrize("prec,eps,tol", MODES)
@pytest.mark.parametrize("w", [0.0, 37.0, -37.0])
def test_all_visibilities_at_one_w(prec, eps, tol, w):
    uvw = numpy.random.default_rng(4).normal(0.0, 150.0, (300, 3))
    uvw[:, 2] = w
    _check(uvw, numpy.array([1.0e8, 1.2e8]) if w == 0.0 else numpy.array([1.0e8]), prec, eps, tol, True)
